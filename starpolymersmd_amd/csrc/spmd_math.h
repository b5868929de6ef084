/*
 * spmd_math.h -- per-element arithmetic of the MPC-MD step, shared by every
 * kernel of the library (host+device inline).
 *
 * Everything that decides a cell index is computed in FP64 from the FP32
 * particle state with the reference's operation order and no FMA
 * contraction (the translation units that include this header set
 * `#pragma clang fp contract(off)`), so cell indices are bit-identical to an
 * FP64 evaluation of star-polymers/lib/Hydrodynamics/MPC.cpp:192-237,421-429.
 */
#ifndef SPMD_MATH_H
#define SPMD_MATH_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SPMD_HD __host__ __device__ __forceinline__

namespace spmd {

/* ---------------------------------------------------------------- Philox
 * Philox4x32-10 (Salmon et al., SC'11): multipliers 0xD2511F53/0xCD9E8D57,
 * key increments 0x9E3779B9/0xBB67AE85, ten rounds. */
SPMD_HD void philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                      uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Stream layout: counter = (index, collision number, replica<<8 | tag,
 * block), key = seed. */
enum : uint32_t { TAG_SHIFT = 1, TAG_AXIS = 2, TAG_GAMMA = 3, TAG_FLUID_INIT = 4, TAG_STAR_INIT = 5, TAG_CLOSURE = 6 };

SPMD_HD void draw4(uint64_t seed, uint32_t replica, uint32_t tag, uint32_t index, uint32_t k,
                   uint32_t block, uint32_t out[4]) {
    philox10(index, k, (replica << 8) | tag, block, (uint32_t)seed, (uint32_t)(seed >> 32), out);
}

/* Uniform on the open interval (0,1) with 2^-32 resolution; exact in FP64. */
SPMD_HD double u01(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386962890625e-10; }

/* Gamma(shape >= 1, scale): Marsaglia & Tsang (ACM TOMS 26(3), 2000) with
 * Box-Muller normals; attempt j uses Philox block j of (cell, k, TAG_GAMMA). */
/* One Marsaglia-Tsang attempt from the Philox words w (d = shape - 1/3,
 * c = 1/sqrt(9d)): true with v^3 when it is accepted. */
SPMD_HD bool gamma_attempt(const uint32_t w[4], double d, double c, double& v3) {
    const double r = sqrt(-2.0 * log(u01(w[0])));
    const double x = r * cos(6.283185307179586 * u01(w[1]));
    double v = 1.0 + c * x;
    if (v <= 0.0) return false;
    v = v * v * v;
    const double u = u01(w[2]);
    const double x2 = x * x;
    v3 = v;
    if (u < 1.0 - 0.0331 * (x2 * x2)) return true;
    return log(u) < 0.5 * x2 + d * (1.0 - v + log(v));
}

SPMD_HD double gamma_mt(uint64_t seed, uint32_t replica, uint32_t cell, uint32_t k, double shape,
                        double scale) {
    const double d = shape - 1.0 / 3.0;
    const double c = 1.0 / sqrt(9.0 * d);
    for (uint32_t j = 0; j < 64; ++j) {
        uint32_t w[4];
        draw4(seed, replica, TAG_GAMMA, cell, k, j, w);
        double v3;
        if (gamma_attempt(w, d, c, v3)) return d * v3 * scale;
    }
    return d * scale;
}

/* ------------------------------------------------------------ box wrap */

/* round(x / L) with round-half-away-from-zero, bit-identical to libm's
 * round() of the correctly rounded quotient.  x * (1/L) is used unless the
 * product lies within 1e-7 of a half-integer, where the rounding decision
 * could differ; there the exact quotient is formed. */
SPMD_HD double round_div(double x, double L, double invL) {
    const double q = x * invL;
    const double t = trunc(q);
    const double f = fabs(q - t);
    if (fabs(f - 0.5) < 1e-7) return round(x / L);
    return f > 0.5 ? t + copysign(1.0, q) : t;
}

struct Box {
    double L[3];       /* even-rounded box (Box.cpp:14-19) */
    double invL[3];
    double half[3];    /* L*0.5 */
    double inner[3];   /* L*0.5*(1 - 1e-9): inside it le_wrap is the identity */
    int iL[3];
    double delrx;      /* Lees-Edwards x offset of the y images */
    double shear;
    int shear_on;
};

/* MPC::LEBC(Particle&) MPC.cpp:421-429 / plain wrap MPC.cpp:473-478, chosen
 * like MPC::wrap(Particle&) MPC.cpp:487-490.  Returns the y image index cy. */
SPMD_HD double le_wrap(const Box& b, double& x, double& y, double& z, double& vx) {
    /* Inside the central image every round_div below is a signed zero (the
     * margin 1e-9 keeps q = c/L away from 1/2 by far more than its
     * rounding), so the wrap only turns -0.0 into +0.0 and subtracts
     * sign(y) 0 shear L from vx: the same bits without the divisions. */
    if (fabs(x) < b.inner[0] && fabs(y) < b.inner[1] && fabs(z) < b.inner[2]) {
        if (b.shear_on) vx -= copysign(0.0, y) * b.shear * b.L[1];
        x += 0.0;
        y += 0.0;
        z += 0.0;
        return 0.0;
    }
    if (b.shear_on) {
        const double cy = round_div(y, b.L[1], b.invL[1]);
        x -= b.L[0] * round_div(x, b.L[0], b.invL[0]);
        x -= cy * b.delrx;
        x -= b.L[0] * round_div(x, b.L[0], b.invL[0]);
        y -= b.L[1] * cy;
        z -= b.L[2] * round_div(z, b.L[2], b.invL[2]);
        vx -= cy * b.shear * b.L[1];
        return cy;
    }
    x -= round_div(x, b.L[0], b.invL[0]) * b.L[0];
    y -= round_div(y, b.L[1], b.invL[1]) * b.L[1];
    z -= round_div(z, b.L[2], b.invL[2]) * b.L[2];
    return 0.0;
}

/* MPC::sort() cell index, MPC.cpp:209 (int truncation, unsigned result). */
SPMD_HD uint32_t cell_index(const Box& b, double x, double y, double z) {
    const int ix = (int)(x + b.half[0]);
    const int iy = (int)(y + b.half[1]);
    const int iz = (int)(z + b.half[2]);
    return (uint32_t)(ix + b.iL[0] * iy + b.iL[0] * b.iL[1] * iz);
}

/* ------------------------------------------------------------ potentials
 * Potentials.h:14-120: force magnitudes F/r (to be multiplied by the
 * separation vector) and potentials, same branches and constants. */
SPMD_HD double fene_force(double r2) { return r2 == 0.0 ? 0.0 : -30.0 / (1.0 - (r2 / 2.25)); }
SPMD_HD double fene_anchor_force(double r2) { return r2 == 0.0 ? 0.0 : -30.0 / (1.0 - (r2 / 20.25)); }
SPMD_HD double fene_potential(double r2) { return r2 == 0.0 ? 0.0 : -15.0 * 2.25 * log(1.0 - r2 / 2.25); }
SPMD_HD double fene_anchor_potential(double r2) {
    return r2 == 0.0 ? 0.0 : -15.0 * 20.25 * log(1.0 - r2 / 20.25);
}
SPMD_HD double rm24_of(double r2) {
    const double rm2 = 1.0 / r2;
    double rm24 = rm2 * rm2 * rm2;
    rm24 *= rm24;
    rm24 *= rm24;
    return rm24;
}
/* 1/x for x in a normal range: hardware reciprocal plus two Newton steps
 * (relative error at the level of one ulp). */
SPMD_HD double rcp_nr(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
#else
    double r = 1.0 / x;
#endif
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return r;
}

/* x = A^-1 b for a symmetric positive-definite 3x3 A, as
 * Matrix3d::llt().solve does (MPC.cpp:373): A = L L^T, L y = b, L^T x = y.
 * A singular cell (all members on a line) gives non-finite x; the caller
 * keeps only finite results. */
SPMD_HD void llt_solve3(const double A[3][3], const double b[3], double x[3]) {
    const double l00 = sqrt(A[0][0]);
    const double l10 = A[1][0] / l00, l20 = A[2][0] / l00;
    const double l11 = sqrt(A[1][1] - l10 * l10);
    const double l21 = (A[2][1] - l20 * l10) / l11;
    const double l22 = sqrt((A[2][2] - l20 * l20) - l21 * l21);
    const double y0 = b[0] / l00;
    const double y1 = (b[1] - l10 * y0) / l11;
    const double y2 = ((b[2] - l20 * y0) - l21 * y1) / l22;
    x[2] = y2 / l22;
    x[1] = (y1 - l21 * x[2]) / l11;
    x[0] = ((y0 - l10 * x[1]) - l20 * x[2]) / l00;
}

/* The MD kernels' forms of the force magnitudes: one reciprocal per pair,
 * no divergent division.  Same functions as pair_force / fene_force up to
 * rounding (the 48-24 core and B-B tail share 192 r^-2 r^-24 (r^-24 - 1/2),
 * the tail scaled by lambda). */
SPMD_HD double pair_force_fast(double r2, bool both_b, double lambda) {
    /* no early exit: beyond both cutoffs the weight below is 0 (r2 > 0) */
    const double rm2 = rcp_nr(r2);
    double rm24 = rm2 * rm2 * rm2;
    rm24 *= rm24;
    rm24 *= rm24;
    const double f = 192.0 * rm2 * rm24 * (rm24 - 0.5);
    const double w = r2 < 1.0594631 ? 1.0 : (both_b && r2 < 2.25 ? lambda : 0.0);
    return f * w;
}
SPMD_HD double fene_force_fast(double r2, double R02) {
    /* -30 / (1 - r2/R0^2) = -30 R0^2 / (R0^2 - r2); r2 == 0 gives 0 like the reference */
    return r2 == 0.0 ? 0.0 : -30.0 * R02 * rcp_nr(R02 - r2);
}

/* non-bonded pair force magnitude; both_b = both beads B-type */
SPMD_HD double pair_force(double r2, bool both_b, double lambda) {
    if (r2 < 1.0594631) {
        const double rm2 = 1.0 / r2;
        double rm24 = rm2 * rm2 * rm2;
        rm24 *= rm24;
        rm24 *= rm24;
        return 192.0 * rm2 * rm24 * (rm24 - 0.5);
    }
    if (both_b && r2 < 2.25) {
        const double rm2 = 1.0 / r2;
        double rm24 = rm2 * rm2 * rm2;
        rm24 *= rm24;
        rm24 *= rm24;
        return 192.0 * lambda * rm2 * rm24 * (rm24 - 0.5);
    }
    return 0.0;
}
SPMD_HD double pair_potential(double r2, bool both_b, double lambda) {
    if (r2 < 1.0594631) {
        double rm24 = 1.0 / r2;
        rm24 *= rm24 * rm24;
        rm24 *= rm24;
        rm24 *= rm24;
        return 4.0 * rm24 * (rm24 - 1.0) + 1.0 - (both_b ? lambda : 0.0);
    }
    if (both_b && r2 < 2.25) {
        double rm24 = 1.0 / r2;
        rm24 *= rm24 * rm24;
        rm24 *= rm24;
        rm24 *= rm24;
        return 4.0 * lambda * (rm24 * (rm24 - 1.0) - 0.0000593);
    }
    return 0.0;
}

}  // namespace spmd

#endif
