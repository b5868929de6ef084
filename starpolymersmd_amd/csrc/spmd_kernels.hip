/*
 * spmd_kernels.hip -- HIP kernels of one MPC-MD collision interval on
 * gfx950 (CDNA4, wave64).  All kernels are memory- or latency-bound; none
 * uses MFMA.
 *
 * Per collision (MPC::step, star-polymers/lib/Hydrodynamics/MPC.cpp:116-142)
 * the library launches, for all replicas at once:
 *   k_prologue     per replica: delrx advance of the pending MPC::step calls,
 *                  the collision's Philox counter and grid shift
 *   k_md           (second stream) one workgroup per replica: the pending
 *                  velocity-Verlet substeps (Thermostat_None::propagate),
 *                  check_bounds, the monomers' cells
 *   k_bin          per solvent particle: recentre, stream, LE wrap, shifted
 *                  cell index, slot in the cell via run-aggregated atomics
 *   k_scan         single-pass exclusive scan of the cell counts
 *   k_perm         per solvent particle: perm[offset[cell] + slot] = index
 *                  (4-byte scattered stores; the particle data stays put)
 *   k_sort_redo    one workgroup per replica: bin, scan and perm again for
 *                  a replica whose MD step recentred (bin and perm ran
 *                  beside k_md assuming it would not); returns at once
 *                  otherwise
 *   k_collide      per tile of cells: gather the tile's particles through
 *                  perm, stream them again, momentum and energy sums in LDS
 *                  (64-bit fixed point, order independent), MBS thermostat,
 *                  SRD rotation; positions and velocities are written to
 *                  the other buffer in cell order (coalesced)
 */
#pragma clang fp contract(off)

#include "spmd_internal.h"

namespace spmd {

namespace {

constexpr double kFix = 4294967296.0;          /* 2^32 */
constexpr double kUnfix = 2.3283064365386962890625e-10;
constexpr double kPi = 3.14159265358979323846;

/* round(v * 2^32) (ties to even) as a 64-bit two's complement value.  One
 * FMA against 1.5 * 2^52 leaves the integer in the low mantissa bits, which
 * is exact for |v| < 2^19; larger |v| (no physical sum comes near) is
 * clamped there. */
__device__ __forceinline__ unsigned long long to_fix(double v) {
    v = fmin(fmax(v, -524287.0), 524287.0);
    const double t = __builtin_fma(v, kFix, 6755399441055744.0);
    return (unsigned long long)(__double_as_longlong(t) - 0x4338000000000000LL);
}
__device__ __forceinline__ double from_fix(unsigned long long v) { return (double)(long long)v * kUnfix; }

/* A value every lane of the wave holds alike, in a scalar register (the
 * compiler cannot tell that a load from a uniform address is uniform when
 * the kernel also writes that memory). */
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

/* A word that device-scope atomics of an earlier kernel built (the row
 * tiles' arrival counts): read with an agent-scope atomic load, like the
 * atomics themselves, never through a copy of its line that a plain access
 * of an earlier kernel left in this XCD's L2. */
__device__ __forceinline__ uint32_t load_count(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Consecutive logical blocks on one XCD: the dispatcher places block b on
 * XCD b % 8, so logical = start(xcd) + b / 8 keeps neighbouring particle
 * chunks and cell tiles in one XCD's L2. */
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    const uint32_t x = b & 7u, i = b >> 3;
    const uint32_t per = nb >> 3, rem = nb & 7u;
    return x * per + (x < rem ? x : rem) + i;
}

__device__ __forceinline__ int find_rep(const uint32_t* base, int n, uint32_t b) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (base[mid] <= b) lo = mid; else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t lane = __lane_id();
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

/* MPC::check_bounds fluid part (MPC.cpp:183-188) when this collision
 * recentres, then MPC::streaming (MPC.cpp:192-197). */
__device__ __forceinline__ void recenter_stream(const Box& b, double h, const RepCtl& c, double& x, double& y,
                                                double& z, double& vx, double& vy, double& vz) {
    if (c.recenter) {
        x -= c.com[0]; y -= c.com[1]; z -= c.com[2];
        vx -= c.com_vel[0]; vy -= c.com_vel[1]; vz -= c.com_vel[2];
        le_wrap(b, x, y, z, vx);
    }
    x += vx * h; y += vy * h; z += vz * h;
    le_wrap(b, x, y, z, vx);
}

/* Cell of the shifted image (MPC::shiftParticles MPC.cpp:301-305 + sort
 * MPC.cpp:209). */
__device__ __forceinline__ uint32_t shifted_cell(const Box& b, const RepCtl& c, double x, double y, double z,
                                                 double vx) {
    double px = x + c.shift[0], py = y + c.shift[1], pz = z + c.shift[2];
    le_wrap(b, px, py, pz, vx);
    return cell_index(b, px, py, pz);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* gamma_mt's variate with FP32 hardware transcendentals: the same Philox
 * words and the same Marsaglia-Tsang attempts, the normal (Box-Muller: one
 * v_log_f32, one v_cos_f32 in revolutions) and both acceptance tests in
 * FP32.  A test whose two sides lie closer than 1e-3 (the FP32 evaluation
 * is good to about 3e-4 in the worst case, |x| = 6.8; typically 1e-6) is
 * decided again by gamma_attempt in FP64, so the accepted attempt is always
 * gamma_mt's and the variate d v^3 agrees with it to the FP32 rounding of
 * v^3 (a few 1e-7 relative).  The FP64 transcendentals were the largest
 * single cost of a collision. */
__device__ __forceinline__ double gamma_fast_w0(uint64_t seed, uint32_t replica, uint32_t cell, uint32_t k,
                                                const uint32_t w0[4], double shape, double scale) {
    const double d = shape - 1.0 / 3.0;
    const double c = 1.0 / sqrt(9.0 * d);
    const float df = (float)d, cf = (float)c;
    constexpr float kLn2 = 0.693147180559945f;
    constexpr float k2m32 = 2.3283064365386963e-10f;
    for (uint32_t j = 0; j < 64; ++j) {
        uint32_t w[4] = {w0[0], w0[1], w0[2], w0[3]};
        if (j) draw4(seed, replica, TAG_GAMMA, cell, k, j, w);
        const float u0 = ((float)w[0] + 0.5f) * k2m32;
        const float u1 = ((float)w[1] + 0.5f) * k2m32;
        const float u2 = ((float)w[2] + 0.5f) * k2m32;
        const float x = __builtin_sqrtf(-2.0f * kLn2 * __builtin_amdgcn_logf(u0)) * __builtin_amdgcn_cosf(u1);
        const float v = 1.0f + cf * x;
        const float v3 = v * v * v, x2 = x * x;
        const float t1 = 1.0f - 0.0331f * (x2 * x2);
        if (fabsf(v) >= 1e-3f && fabsf(u2 - t1) >= 1e-3f) {
            if (v <= 0.0f) continue;
            if (u2 < t1) return d * (double)v3 * scale;
            const float lhs = kLn2 * __builtin_amdgcn_logf(u2);
            const float rhs = 0.5f * x2 + df * (1.0f - v3 + kLn2 * __builtin_amdgcn_logf(v3));
            if (fabsf(lhs - rhs) >= 1e-3f * (1.0f + fabsf(rhs))) {
                if (lhs < rhs) return d * (double)v3 * scale;
                continue;
            }
        }
        double vv;
        if (gamma_attempt(w, d, c, vv)) return d * vv * scale;
    }
    return d * scale;
}

/* gamma_fast_w0 with the first attempt's words drawn here (the caller can
 * draw them itself, e.g. while the member count is still being loaded). */
__device__ __forceinline__ double gamma_fast(uint64_t seed, uint32_t replica, uint32_t cell, uint32_t k,
                                             double shape, double scale) {
    uint32_t w0[4];
    draw4(seed, replica, TAG_GAMMA, cell, k, 0, w0);
    return gamma_fast_w0(seed, replica, cell, k, w0, shape, scale);
}

/* The SRD rotation matrix about the axis (MPC::collide MPC.cpp:241-257)
 * phi = 2 pi u1, cos(theta) = 2 (u2 - 1/2), stored as FP32.  The sine and
 * cosine of phi are the FP32 hardware ones in revolutions (a few 1e-7
 * absolute); sin(theta) = sqrt(1 - cos^2) and the entries stay FP64, since
 * 1 - cos^2 cancels near the poles. */
__device__ __forceinline__ void rotation_f32(double u1, double u2, double c, double s, float* R) {
    const double th = 2. * (u2 - 0.5);
    const double st = sqrt(1 - th * th);
    const double a0 = st * (double)__builtin_amdgcn_cosf((float)u1);
    const double a1 = st * (double)__builtin_amdgcn_sinf((float)u1);
    const double a2 = th;
    R[0] = (float)(a0 * a0 + (1 - a0 * a0) * c);
    R[1] = (float)(a0 * a1 * (1 - c) - a2 * s);
    R[2] = (float)(a0 * a2 * (1 - c) + a1 * s);
    R[3] = (float)(a0 * a1 * (1 - c) + a2 * s);
    R[4] = (float)(a1 * a1 + (1 - a1 * a1) * c);
    R[5] = (float)(a1 * a2 * (1 - c) - a0 * s);
    R[6] = (float)(a0 * a2 * (1 - c) - a1 * s);
    R[7] = (float)(a1 * a2 * (1 - c) + a0 * s);
    R[8] = (float)(a2 * a2 + (1 - a2 * a2) * c);
}

/* Sum over the workgroup in a fixed order (deterministic). */
template <int NT>
__device__ double block_sum(double v, double* scratch) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) scratch[w] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) t += scratch[i];
    return t;
}

}  // namespace

/* ======================================================= MD + prologue */

/* Verlet list radii: the reference keeps every pair within 2.0 (Box.cpp:9-10)
 * and rebuilds once a monomer moved 0.25 = (2.0 - 1.5)/2.  Only B-B pairs
 * interact beyond r^2 = 1.0594631, so the other pairs need a list radius of
 * only 1.0293 + 0.5 for the same rebuild rule; the set of interacting pairs
 * -- and so every force -- is unchanged.  The rebuild tests distances in
 * FP32 against radii padded by 1e-4 (far above FP32 rounding of r^2 at
 * these coordinates), so its list holds every pair the FP64 test would
 * keep; extra pairs beyond the cutoffs contribute nothing. */
constexpr float kListBB2f = 4.0001f;
constexpr float kListAA2f = 2.3401f;   /* (sqrt(1.0594631) + 0.5)^2 = 2.3388 */
constexpr int kMdBatch = 4;            /* independent pairs per lane in flight */
constexpr int kGridMax = 4096;         /* cells of the rebuild grid */

/* Dynamic LDS of k_md for a stride of ns monomers (the host computes the
 * same size with md_lds_bytes), with cap pair-list entries. */
struct MdLds {
    /* positions, and the force accumulators (64-bit fixed point while the
     * forces are summed, FP64 forces after conversion); velocities, Verlet
     * reference positions and masses live in the owning threads' registers */
    double *x, *y, *z, *fx, *fy, *fz, *red;
    /* rebuild scratch over the force accumulators (zeroed again after it):
     * FP32 positions sorted by grid cell, w = bits of (i | B << 16), and
     * each monomer's grid cell */
    float4* sp4;       /* ns */
    uint32_t* gcid;    /* ns */
    uint32_t* gcell;   /* grid cell ends after the fill (kGridMax + 1) */
    uint32_t* cpre;    /* rebuild: pairs found per sorted slot, then their prefix (ns + 1) */
    uint32_t* bl;      /* bonds: i | j << 12 | anchor << 24 (ns) */
    uint32_t* pl;      /* half pair list: i | j << 16 (cap) */
    uint8_t* bt;       /* monomer is B-type (ns) */
    int* flags;
};

__device__ __forceinline__ MdLds md_carve(unsigned char* base, int ns, uint32_t cap) {
    MdLds L;
    double* p = reinterpret_cast<double*>(base);
    L.x = p; L.y = p + ns; L.z = p + 2 * ns;
    L.fx = p + 3 * ns; L.fy = p + 4 * ns; L.fz = p + 5 * ns;
    L.red = p + 6 * ns;                                           /* 80 doubles */
    L.flags = reinterpret_cast<int*>(L.red + 80);                 /* 16 ints */
    L.sp4 = reinterpret_cast<float4*>(L.fx);                      /* 16 ns <= 24 ns bytes */
    L.gcid = reinterpret_cast<uint32_t*>(L.sp4 + ns);             /* + 4 ns */
    L.gcell = reinterpret_cast<uint32_t*>(L.flags + 16);
    L.cpre = L.gcell + kGridMax + 1;
    L.bl = L.cpre + ns + 1;
    L.pl = L.bl + ns;
    L.bt = reinterpret_cast<uint8_t*>(L.pl + cap);
    return L;
}

size_t md_lds_bytes(int ns, uint32_t cap) {
    return (size_t)6 * ns * 8 + 80 * 8 + 16 * 4 + (size_t)(kGridMax + 1) * 4 + (size_t)(ns + 1) * 4 +
           (size_t)ns * 4 + (size_t)cap * 4 + (size_t)ns + 64;
}

template <int NT>
__device__ double block_sum_all(double v, double* scratch) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) scratch[w] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) t += scratch[i];
    return t;
}

/* Forces of all list entries -- the npl half-list pairs (Newton's third
 * law) followed by the nb bonds -- kMdBatch independent entries per lane in
 * flight, added to both monomers' 64-bit fixed-point LDS accumulators (the
 * sum does not depend on the order of the entries). */
__device__ __forceinline__ void md_add_force(const MdLds& S, int i, int j, unsigned long long gx,
                                             unsigned long long gy, unsigned long long gz) {
    unsigned long long* qx = reinterpret_cast<unsigned long long*>(S.fx);
    unsigned long long* qy = reinterpret_cast<unsigned long long*>(S.fy);
    unsigned long long* qz = reinterpret_cast<unsigned long long*>(S.fz);
    atomicAdd(&qx[j], gx); atomicAdd(&qy[j], gy); atomicAdd(&qz[j], gz);
    atomicAdd(&qx[i], 0ull - gx); atomicAdd(&qy[i], 0ull - gy); atomicAdd(&qz[i], 0ull - gz);
}

/* Non-bonded part of Box::calculate_forces_verlet (Box.cpp:243-256) over the
 * Verlet half list pl (entries i | j << 16; an LDS or a global array, known
 * at the call site so the loads are not generic): each pair once, Newton's
 * third law, into the fixed-point accumulators.  The force magnitude is
 * evaluated without branches (pairs beyond their cutoff get weight 0) and
 * kMdBatch list entries are loaded before any is used. */
template <typename PL>
__device__ __forceinline__ void md_pair_forces(const MdLds& S, PL pl, uint32_t npl, double lambda) {
#pragma clang fp contract(fast)
    constexpr int NT = kMdThreads;
    for (uint32_t q0 = threadIdx.x; q0 < npl; q0 += NT * kMdBatch) {
        uint32_t e[kMdBatch];
#pragma unroll
        for (int k = 0; k < kMdBatch; ++k) {
            const uint32_t q = q0 + k * NT;
            e[k] = q < npl ? pl[q] : 0xFFFFFFFFu;
        }
        unsigned long long gx[kMdBatch], gy[kMdBatch], gz[kMdBatch];
        bool ok[kMdBatch];
#pragma unroll
        for (int k = 0; k < kMdBatch; ++k) {
            const bool valid = e[k] != 0xFFFFFFFFu;
            const int i = valid ? (int)(e[k] & 0xFFFF) : 0, j = valid ? (int)(e[k] >> 16) : 0;
            const double dx = S.x[j] - S.x[i], dy = S.y[j] - S.y[i], dz = S.z[j] - S.z[i];
            const double r2 = valid ? (dx * dx + dy * dy) + dz * dz : 4.0;
            const double fa = pair_force_fast(r2, (S.bt[i] & S.bt[j]) != 0, lambda);
            ok[k] = fa != 0.0;
            gx[k] = to_fix(dx * fa); gy[k] = to_fix(dy * fa); gz[k] = to_fix(dz * fa);
        }
#pragma unroll
        for (int k = 0; k < kMdBatch; ++k)
            if (ok[k]) md_add_force(S, (int)(e[k] & 0xFFFF), (int)(e[k] >> 16), gx[k], gy[k], gz[k]);
    }
}

/* Bonded part of Box::calculate_forces_verlet (Box.cpp:257-285): every bond
 * once (entries i | j << 12 | anchor << 24), FENE or anchor FENE. */
__device__ __forceinline__ void md_bond_forces(const MdLds& S, uint32_t nb) {
#pragma clang fp contract(fast)
    for (uint32_t q = threadIdx.x; q < nb; q += kMdThreads) {
        const uint32_t e = S.bl[q];
        const int i = (int)(e & 0xFFF), j = (int)((e >> 12) & 0xFFF);
        const double dx = S.x[j] - S.x[i], dy = S.y[j] - S.y[i], dz = S.z[j] - S.z[i];
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        const double fa = fene_force_fast(r2, ((e >> 24) & 1) ? 20.25 : 2.25);
        if (fa != 0.0) md_add_force(S, i, j, to_fix(dx * fa), to_fix(dy * fa), to_fix(dz * fa));
    }
}

/* Molecule::Epot (Box.cpp:257-277): every Verlet pair once (the reference
 * adds 0.5 U from each side) plus every bond; per-thread fixed point, then
 * a block sum. */
__device__ double md_epot_flat(const MdLds& S, const uint32_t* pl, uint32_t np, uint32_t nb, double lambda) {
    unsigned long long acc = 0ull;
    for (uint32_t q = threadIdx.x; q < np; q += kMdThreads) {
        const uint32_t e = pl[q];
        const int i = e & 0xFFFF, j = e >> 16;
        const double dx = S.x[j] - S.x[i], dy = S.y[j] - S.y[i], dz = S.z[j] - S.z[i];
        acc += to_fix(pair_potential((dx * dx + dy * dy) + dz * dz, S.bt[i] && S.bt[j], lambda));
    }
    for (uint32_t q = threadIdx.x; q < nb; q += kMdThreads) {
        const uint32_t e = S.bl[q];
        const int i = e & 0xFFF, j = (e >> 12) & 0xFFF;
        const double dx = S.x[j] - S.x[i], dy = S.y[j] - S.y[i], dz = S.z[j] - S.z[i];
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        acc += to_fix(((e >> 24) & 1) ? fene_anchor_potential(r2) : fene_potential(r2));
    }
    return block_sum_all<kMdThreads>(from_fix(acc), S.red);
}

/* Exclusive scan of a[0..n) in place, total in a[n] and returned; each
 * thread scans a contiguous chunk, deterministic. */
template <int NT>
__device__ uint32_t block_scan_u32(uint32_t* a, int n, double* scratch_d) {
    uint32_t* wsum = reinterpret_cast<uint32_t*>(scratch_d);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int per = (n + NT - 1) / NT;
    const int i0 = tid * per;
    uint32_t t = 0;
    for (int q = 0; q < per; ++q) t += i0 + q < n ? a[i0 + q] : 0u;
    uint32_t inc = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(inc, o);
        if (lane >= o) inc += u;
    }
    __syncthreads();
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < NT / 64; ++q) {
        const uint32_t u = wsum[q];
        pre += q < w ? u : 0u;
        tot += u;
    }
    uint32_t run = pre + inc - t;
    for (int q = 0; q < per; ++q)
        if (i0 + q < n) { const uint32_t v = a[i0 + q]; a[i0 + q] = run; run += v; }
    if (tid == 0) a[n] = tot;
    __syncthreads();
    return tot;
}
__device__ __forceinline__ uint32_t md_block_scan(uint32_t* a, int n, double* scratch_d) {
    return block_scan_u32<kMdThreads>(a, n, scratch_d);
}

/* exclusive block scan of one value per thread; *total gets the sum */
__device__ __forceinline__ uint32_t md_thread_scan(uint32_t v, double* scratch_d, uint32_t* total) {
    uint32_t* wsum = reinterpret_cast<uint32_t*>(scratch_d);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(inc, o);
        if (lane >= o) inc += u;
    }
    __syncthreads();
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < kMdThreads / 64; ++q) {
        const uint32_t u = wsum[q];
        pre += q < w ? u : 0u;
        tot += u;
    }
    *total = tot;
    return pre + inc - v;
}

/* The list pairs of the monomer at sorted slot p: its candidates are the
 * slots of the 3x3 rows of three neighbouring cells (x-fastest cells, so
 * each row is one contiguous slot range; all 18 row bounds are read before
 * the first candidate).  Pairs j > i inside the list radius go to
 * scr[0..kVlPerMono); returns their number (more means overflow). */
__device__ __forceinline__ uint32_t md_scan_cands(const MdLds& S, int p, const int gd[3], uint32_t* scr) {
    const float4 me = S.sp4[p];
    const uint32_t im = __float_as_uint(me.w);
    const int i = (int)(im & 0xFFFF);
    const bool bi = (im >> 16) != 0;
    const uint32_t cid = S.gcid[i];
    const int cx = (int)(cid % (uint32_t)gd[0]);
    const int cy = (int)((cid / (uint32_t)gd[0]) % (uint32_t)gd[1]);
    const int cz = (int)(cid / ((uint32_t)gd[0] * (uint32_t)gd[1]));
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < gd[0] ? cx + 1 : gd[0] - 1;
    /* half shell: the slots after p up to the end of cell x+1 of its own
     * row, then the whole x-runs of rows (y+1, z) and (y-1..y+1, z+1):
     * every pair of the 27-cell neighbourhood once */
    uint32_t ra0[5], ra1[5];
    {
        const int rowc = gd[0] * (cy + gd[1] * cz);
        ra0[0] = (uint32_t)p + 1u;
        ra1[0] = S.gcell[rowc + x1];
    }
#pragma unroll
    for (int r = 1; r < 5; ++r) {
        const int y = r == 1 ? cy + 1 : cy + (r - 3), z = r == 1 ? cz : cz + 1;
        const bool ok = y >= 0 && y < gd[1] && z < gd[2];
        const int rowc = ok ? gd[0] * (y + gd[1] * z) : 0;
        ra0[r] = ok && rowc + x0 ? S.gcell[rowc + x0 - 1] : 0u;
        ra1[r] = ok ? S.gcell[rowc + x1] : 0u;
    }
    uint32_t n = 0;
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        const uint32_t a1 = ra1[r];
        for (uint32_t q = ra0[r]; q < a1; q += 4) {
            float4 c[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) c[k] = S.sp4[q + k < a1 ? q + k : a1 - 1];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t jm = __float_as_uint(c[k].w);
                const int j = (int)(jm & 0xFFFF);
                const float dx = c[k].x - me.x, dy2 = c[k].y - me.y, dz2 = c[k].z - me.z;
                const float r2 = __builtin_fmaf(dz2, dz2, __builtin_fmaf(dy2, dy2, dx * dx));
                const bool hit = q + k < a1 && r2 <= ((bi && (jm >> 16)) ? kListBB2f : kListAA2f);
                const uint32_t lo = (uint32_t)(i < j ? i : j), hi = (uint32_t)(i < j ? j : i);
                if (hit && n < (uint32_t)kVlPerMono) scr[n] = lo | (hi << 16);
                n += hit ? 1u : 0u;
            }
        }
    }
    return n;
}

/* Box::update_VerletLists (Box.cpp:105-180): the half list of every pair
 * inside the list radius, plus the COM-centred grid bounds check that
 * raises SIGSEGV in the reference.  The monomers are binned on a grid of
 * side >= 2 over their bounding box and sorted by cell; each lane takes
 * the monomers at its sorted slots, scans their half shell (md_scan_cands)
 * and writes their pairs to a global scratch row of kVlPerMono entries per
 * slot; after a block scan of the
 * counts the rows are packed into the list -- in LDS when it fits there,
 * else in the replica's global region.  Leaves the force accumulators
 * zeroed; the owning threads reset their Verlet reference positions.
 * Returns the entry count; *in_lds tells where the list is. */
__device__ uint32_t md_rebuild(const MdLds& S, uint32_t* gpl, uint32_t gcap, uint32_t* gscr, const RepDev& rep,
                               RepCtl* ctl, int N, uint32_t cap, bool* in_lds) {
    constexpr int NT = kMdThreads;
    const int tid = threadIdx.x;
    /* COM of the molecule (Box::center_of_mass_reference_frame) and the
     * bounding box: 9 values per wave through S.red */
    double lc[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) { lc[c] = 0.0; lc[3 + c] = 1e300; lc[6 + c] = -1e300; }
    for (int i = tid; i < N; i += NT) {
        const double v[3] = {S.x[i], S.y[i], S.z[i]};
#pragma unroll
        for (int c = 0; c < 3; ++c) { lc[c] += v[c]; lc[3 + c] = fmin(lc[3 + c], v[c]); lc[6 + c] = fmax(lc[6 + c], v[c]); }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lc[c] += __shfl_xor(lc[c], o);
            lc[3 + c] = fmin(lc[3 + c], __shfl_xor(lc[3 + c], o));
            lc[6 + c] = fmax(lc[6 + c], __shfl_xor(lc[6 + c], o));
        }
    const int w = tid >> 6;
    __syncthreads();
    if ((tid & 63) == 0)
#pragma unroll
        for (int c = 0; c < 9; ++c) S.red[w * 9 + c] = lc[c];
    __syncthreads();
    double cs[3], lo[3], ext[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double sum = 0.0, a = 1e300, b = -1e300;
        for (int q = 0; q < NT / 64; ++q) {
            sum += S.red[q * 9 + c];
            a = fmin(a, S.red[q * 9 + 3 + c]);
            b = fmax(b, S.red[q * 9 + 6 + c]);
        }
        cs[c] = sum / (double)N;
        lo[c] = a;
        ext[c] = b - a;
    }
    if (tid == 0) { ctl->vcom[0] = cs[0]; ctl->vcom[1] = cs[1]; ctl->vcom[2] = cs[2]; ctl->n_rebuild++; }
    bool oob = false;
    for (int i = tid; i < N; i += NT) {
        const double v[3] = {S.x[i], S.y[i], S.z[i]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int cn = (int)((v[c] - cs[c] + rep.box.L[c] * 0.5) / rep.vside[c]);
            oob |= cn < 0 || cn >= rep.vgrid[c];
        }
    }
    if (oob) atomicOr(&ctl->status, ST_VERLET_OOB);
    /* grid: side >= 2 (the largest list radius), grown until the cells fit */
    double side = 2.0;
    int gd[3] = {1, 1, 1};
    for (int it = 0; it < 64; ++it, side *= 1.25) {
        long nc = 1;
        int g[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double q = ext[c] / side;
            g[c] = q < 4096.0 ? (int)q + 1 : 4097;  /* also catches inf/NaN */
            nc *= g[c];
        }
        if (nc <= kGridMax) { gd[0] = g[0]; gd[1] = g[1]; gd[2] = g[2]; break; }
    }  /* no fit (non-finite positions): one cell, every pair tested */
    const int ncell = gd[0] * gd[1] * gd[2];
    const double inv = 1.0 / side;
    for (int c = tid; c <= ncell; c += NT) S.gcell[c] = 0u;
    __syncthreads();
    for (int i = tid; i < N; i += NT) {
        const int cx = max(min((int)((S.x[i] - lo[0]) * inv), gd[0] - 1), 0);
        const int cy = max(min((int)((S.y[i] - lo[1]) * inv), gd[1] - 1), 0);
        const int cz = max(min((int)((S.z[i] - lo[2]) * inv), gd[2] - 1), 0);
        const uint32_t cid = (uint32_t)(cx + gd[0] * (cy + gd[1] * cz));
        S.gcid[i] = cid;
        atomicAdd(&S.gcell[cid], 1u);
    }
    __syncthreads();
    md_block_scan(S.gcell, ncell, S.red);
    for (int i = tid; i < N; i += NT) {
        const uint32_t slot = atomicAdd(&S.gcell[S.gcid[i]], 1u);
        S.sp4[slot] = make_float4((float)S.x[i], (float)S.y[i], (float)S.z[i],
                                  __uint_as_float((uint32_t)i | (S.bt[i] ? 1u << 16 : 0u)));
    }
    __syncthreads();  /* gcell[c] now ends cell c */
    /* pairs per sorted slot into scratch rows, prefix of the counts, pack */
    bool over = false;
    for (int p = tid; p < N; p += NT) {
        uint32_t n = md_scan_cands(S, p, gd, gscr + (size_t)p * kVlPerMono);
        over |= n > (uint32_t)kVlPerMono;
        S.cpre[p] = n < (uint32_t)kVlPerMono ? n : (uint32_t)kVlPerMono;
    }
    if (over) atomicOr(&ctl->status, ST_VERLET_OVERFLOW);
    __threadfence_block();
    __syncthreads();
    const uint32_t total = md_block_scan(S.cpre, N, S.red);
    const uint32_t lcap = cap < gcap ? cap : gcap;  /* an LDS list must also fit its global copy */
    const bool lds = total <= lcap;
    uint32_t* out = lds ? S.pl : gpl;
    const uint32_t ocap = lds ? lcap : gcap;
    for (int p = tid; p < N; p += NT) {
        const uint32_t o = S.cpre[p], n = S.cpre[p + 1] - o;
        const uint32_t* src = gscr + (size_t)p * kVlPerMono;
        for (uint32_t k = 0; k < n; ++k)
            if (o + k < ocap) out[o + k] = src[k];
    }
    if (!lds && total > gcap && tid == 0) atomicOr(&ctl->status, ST_VERLET_OVERFLOW);
    __syncthreads();
    /* the scratch lived in the force accumulators */
    unsigned long long* qx = reinterpret_cast<unsigned long long*>(S.fx);
    unsigned long long* qy = reinterpret_cast<unsigned long long*>(S.fy);
    unsigned long long* qz = reinterpret_cast<unsigned long long*>(S.fz);
    for (int i = tid; i < N; i += NT) { qx[i] = 0ull; qy[i] = 0ull; qz[i] = 0ull; }
    __syncthreads();
    *in_lds = lds;
    return total < ocap ? total : ocap;
}

/* One workgroup per replica (its star): the pending velocity-Verlet
 * substeps (Thermostat_None::propagate, Thermostat_None.cpp:21-38); then,
 * for a collision, check_bounds and the monomers' cells.  Thread t owns
 * monomers t + 512 k (k < R): their velocities, Verlet reference positions
 * and masses stay in its registers, their positions in registers and LDS
 * (the force loops read any monomer's position). */
template <int R>
__global__ __launch_bounds__(kMdThreads) void k_md(DevState d, int n_md, int epot_last, int md_init, int do_coll,
                                                  double dt, int ns, uint32_t cap) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NT = kMdThreads;
    /* claim all 256 VGPRs: two waves per SIMD then fill the register file,
     * so no wave of the concurrently running sort kernels shares this CU */
    asm volatile("v_mov_b32 v255, 0" ::: "v255");
    MdLds S = md_carve(smem, ns, cap);
    const int r = blockIdx.x;
    const int tid = threadIdx.x;
    const RepDev& rep = d.reps[r];
    RepCtl* ctl = d.ctl + r;
    const int N = (int)rep.n_mono;
    const uint32_t mb = rep.mono_base;
    const double dth = dt * 0.50;
    unsigned long long* const qx = reinterpret_cast<unsigned long long*>(S.fx);
    unsigned long long* const qy = reinterpret_cast<unsigned long long*>(S.fy);
    unsigned long long* const qz = reinterpret_cast<unsigned long long*>(S.fz);
    /* the replica's global pair-list region */
    uint32_t* const gpl = d.vl + (size_t)mb * kVlPerMono;
    const uint32_t gcap = (uint32_t)N * kVlPerMono;
    uint32_t* const gscr = d.vl + ((size_t)d.m_tot + mb) * kVlPerMono;  /* rebuild scratch rows */
    const uint32_t nbond = rep.n_bond < (uint32_t)ns ? rep.n_bond : (uint32_t)ns;

    double rx[R], ry[R], rz[R], rvx[R], rvy[R], rvz[R], rpx[R], rpy[R], rpz[R], rim[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * NT;
        rx[k] = ry[k] = rz[k] = rvx[k] = rvy[k] = rvz[k] = rpx[k] = rpy[k] = rpz[k] = rim[k] = 0.0;
        if (i >= N) continue;
        const size_t g = 3 * (size_t)(mb + i);
        rx[k] = d.mpos[g]; ry[k] = d.mpos[g + 1]; rz[k] = d.mpos[g + 2];
        rvx[k] = d.mvel[g]; rvy[k] = d.mvel[g + 1]; rvz[k] = d.mvel[g + 2];
        rpx[k] = d.mvpos[g]; rpy[k] = d.mvpos[g + 1]; rpz[k] = d.mvpos[g + 2];
        rim[k] = dth / d.mmass[mb + i];  /* half-kick factor dt/(2m) */
        S.x[i] = rx[k]; S.y[i] = ry[k]; S.z[i] = rz[k];
        S.fx[i] = d.mforce[g]; S.fy[i] = d.mforce[g + 1]; S.fz[i] = d.mforce[g + 2];
        S.bt[i] = d.mbtype[mb + i];
    }
    for (uint32_t q = tid; q < nbond; q += NT) S.bl[q] = d.bonds[rep.bond_base + q];
    /* Verlet half list of the last rebuild -> LDS when it fits */
    uint32_t npl = ctl->vl_count;
    bool lds_vl = false;
    if (N > 0 && !md_init && npl <= cap) {
        for (uint32_t q = tid; q < npl; q += NT) S.pl[q] = gpl[q];
        lds_vl = true;
    }
    __syncthreads();
    bool rebuilt = false;
    double epot_val = 0.0;
    bool epot_set = false;

    auto rebuild = [&]() {
        bool in_lds = false;
        npl = md_rebuild(S, gpl, gcap, gscr, rep, ctl, N, cap, &in_lds);
        lds_vl = in_lds;
        rebuilt = true;
#pragma unroll
        for (int k = 0; k < R; ++k) { rpx[k] = rx[k]; rpy[k] = ry[k]; rpz[k] = rz[k]; }
    };
    /* Box::calculate_forces_verlet (Box.cpp:243-286) into the fixed-point
     * accumulators (zeroed by the caller); Epot of the same positions.  The
     * caller converts the accumulators. */
    auto forces = [&](bool calc_epot) {
        if (lds_vl) md_pair_forces(S, S.pl, npl, rep.lambda);
        else md_pair_forces(S, (const uint32_t*)gpl, npl, rep.lambda);
        md_bond_forces(S, nbond);
        __syncthreads();
        const uint32_t* pl = lds_vl ? S.pl : gpl;
        epot_val = calc_epot ? md_epot_flat(S, pl, npl, nbond, rep.lambda) : 0.0;
        epot_set = true;
    };
    /* first half kick and drift of a substep with force f, then
     * Box::check_VerletLists (Box.cpp:182-194); the accumulators are
     * cleared for the next forces */
    auto kick_drift = [&](int k, double fx, double fy, double fz) -> bool {
        const int i = tid + k * NT;
        const double h = rim[k];
        rvx[k] += fx * h; rvy[k] += fy * h; rvz[k] += fz * h;
        rx[k] += rvx[k] * dt; ry[k] += rvy[k] * dt; rz[k] += rvz[k] * dt;
        S.x[i] = rx[k]; S.y[i] = ry[k]; S.z[i] = rz[k];
        qx[i] = 0ull; qy[i] = 0ull; qz[i] = 0ull;
        const double dx = rx[k] - rpx[k], dy = ry[k] - rpy[k], dz = rz[k] - rpz[k];
        return sqrt((dx * dx + dy * dy) + dz * dz) > (2.0 - 1.5) * 0.5;
    };

    if (N > 0) {
        if (md_init) {
            rebuild();  /* leaves the accumulators zeroed */
            forces(false);
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int i = tid + k * NT;
                if (i >= N) continue;
                S.fx[i] = from_fix(qx[i]); S.fy[i] = from_fix(qy[i]); S.fz[i] = from_fix(qz[i]);
            }
            __syncthreads();
        }
        bool need = false;
        if (n_md > 0)
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (tid + k * NT < N) {
                    const int i = tid + k * NT;
                    need |= kick_drift(k, S.fx[i], S.fy[i], S.fz[i]);
                }
        for (int s = 0; s < n_md; ++s) {
            if (__syncthreads_or(need)) rebuild();
            forces(epot_last && s == n_md - 1);
            /* forces out of fixed point, second half kick of substep s and,
             * fused, the first half kick and drift of substep s + 1 */
            need = false;
            const bool more = s + 1 < n_md;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int i = tid + k * NT;
                if (i >= N) continue;
                const double fx = from_fix(qx[i]), fy = from_fix(qy[i]), fz = from_fix(qz[i]);
                const double h = rim[k];
                rvx[k] += fx * h; rvy[k] += fy * h; rvz[k] += fz * h;
                if (more) need |= kick_drift(k, fx, fy, fz);
                else { S.fx[i] = fx; S.fy[i] = fy; S.fz[i] = fz; }
            }
        }
        __syncthreads();
        if (rebuilt && lds_vl)
            for (uint32_t q = tid; q < npl && q < gcap; q += NT) gpl[q] = S.pl[q];
        if (rebuilt && tid == 0) ctl->vl_count = npl;
    }
    if (tid == 0 && epot_set) ctl->epot = epot_val;

    bool forces_stored = false;  /* the recentring wrote the forces already */
    if (do_coll) {
        /* MPC::check_bounds (MPC.cpp:160-190) on the monomers; a recentring
         * also marks the speculative solvent sort of this collision for redo */
        bool out = false;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (tid + k * NT >= N) continue;
            out |= fabs(rx[k]) >= (rep.box.L[0] * 0.5 - 0.5);
            out |= fabs(ry[k]) >= (rep.box.L[1] * 0.5 - 0.5);
            out |= fabs(rz[k]) >= (rep.box.L[2] * 0.5 - 0.5);
        }
        const bool rec = __syncthreads_or(out) != 0;
        if (rec) {
            /* Molecule::calculate_center_of_mass(_velocity) (Molecule.cpp:76-92):
             * summed in monomer order, then divided by N, so the COM -- and
             * every solvent image the recentring decides with it -- is
             * bit-identical to the sequential FP64 sum.  The forces go to
             * HBM first; their LDS rows then carry the velocities. */
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int i = tid + k * NT;
                if (i >= N) continue;
                const size_t g = 3 * (size_t)(mb + i);
                d.mforce[g] = S.fx[i]; d.mforce[g + 1] = S.fy[i]; d.mforce[g + 2] = S.fz[i];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int i = tid + k * NT;
                if (i >= N) continue;
                S.x[i] = rx[k]; S.y[i] = ry[k]; S.z[i] = rz[k];
                S.fx[i] = rvx[k]; S.fy[i] = rvy[k]; S.fz[i] = rvz[k];
            }
            __syncthreads();
            if (tid < 6) {
                const double* src = tid == 0 ? S.x : tid == 1 ? S.y : tid == 2 ? S.z
                                  : tid == 3 ? S.fx : tid == 4 ? S.fy : S.fz;
                double acc = 0.0;
#pragma unroll 8
                for (int i = 0; i < N; ++i) acc += src[i];
                S.red[tid] = acc / (double)N;
            }
            __syncthreads();
            double cm[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) cm[c] = S.red[c];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                rx[k] -= cm[0]; ry[k] -= cm[1]; rz[k] -= cm[2];
                rvx[k] -= cm[3]; rvy[k] -= cm[4]; rvz[k] -= cm[5];
            }
            if (tid == 0)
                for (int c = 0; c < 3; ++c) { ctl->com[c] = cm[c]; ctl->com_vel[c] = cm[3 + c]; }
        }
        forces_stored = rec;
        if (tid == 0) {
            ctl->recenter = rec ? 1 : 0;
            ctl->redo = rec ? 1 : 0;
        }
    }

    /* write the monomer state back (their cells for the collision are
     * found by k_sort_redo, once the grid shift is drawn) */
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * NT;
        if (i >= N) continue;
        const size_t g = 3 * (size_t)(mb + i);
        d.mpos[g] = rx[k]; d.mpos[g + 1] = ry[k]; d.mpos[g + 2] = rz[k];
        d.mvel[g] = rvx[k]; d.mvel[g + 1] = rvy[k]; d.mvel[g + 2] = rvz[k];
        if (!forces_stored) { d.mforce[g] = S.fx[i]; d.mforce[g + 1] = S.fy[i]; d.mforce[g + 2] = S.fz[i]; }
        d.mvpos[g] = rpx[k]; d.mvpos[g + 1] = rpy[k]; d.mvpos[g + 2] = rpz[k];
    }
}

/* ========================================================== prologue */



/* Per replica, before anything else of a collision: the delrx advance of
 * every pending MPC::step call (MPC.cpp:117-118), then the collision's
 * Philox counter and grid shift (MPC.cpp:123).  Block 0 resets the scan
 * state. */
__global__ __launch_bounds__(64) void k_prologue(DevState d, int do_coll, int n_delrx, double delrx_dt) {
    const int r = blockIdx.x;
    const RepDev& rep = d.reps[r];
    RepCtl* ctl = d.ctl + r;
    if (threadIdx.x == 0) {
        double dl = ctl->delrx;
        for (int q = 0; q < n_delrx; ++q) {
            dl += rep.box.shear * rep.box.L[1] * delrx_dt;
            dl -= rep.box.L[0] * round(dl / rep.box.L[0]);
        }
        ctl->delrx = dl;
        if (do_coll) {
            const uint32_t k = ctl->k_done;
            double sh[3];
            if (ctl->inj_pending) {
                for (int c = 0; c < 3; ++c) sh[c] = ctl->inj_shift[c];
                ctl->inj_active = 1;
                ctl->inj_pending = 0;
            } else {
                uint32_t w[4];
                draw4(rep.seed, rep.replica, TAG_SHIFT, 0, k, 0, w);
                for (int c = 0; c < 3; ++c) sh[c] = u01(w[c]) - 0.5;
                ctl->inj_active = 0;
            }
            for (int c = 0; c < 3; ++c) ctl->shift[c] = sh[c];
            ctl->k_cur = k;
            ctl->k_done = k + 1;
            if (d.m_tot == 0) {  /* else k_md, running beside this kernel, sets them */
                ctl->recenter = 0;
                ctl->redo = 0;
            }
        }
    }
    if (do_coll && r == 0) {
        for (uint32_t t = threadIdx.x; t < d.n_scan_tiles; t += 64) d.scan_flags[t] = 0ull;
        if (threadIdx.x == 0) { *d.scan_ticket = 0u; *d.grid_bar = 0u; }
    }
}

/* ============================================================== bin */

/* Control words the sort kernels use; in the speculative mode the
 * recentring fields (written concurrently by k_md) are not read. */
__device__ __forceinline__ bool sort_ctl(const DevState& d, int r, int mode, RepCtl& c) {
    const RepCtl* g = d.ctl + r;
    c.delrx = g->delrx;
    c.shift[0] = g->shift[0]; c.shift[1] = g->shift[1]; c.shift[2] = g->shift[2];
    c.recenter = mode == kSortSpec ? 0 : g->recenter;
    if (c.recenter)
        for (int q = 0; q < 3; ++q) { c.com[q] = g->com[q]; c.com_vel[q] = g->com_vel[q]; }
    return true;
}

/* Cells touched by one workgroup's particles are aggregated in an LDS hash
 * table first (open addressing, HS >= 2 x particles, so an insert always
 * finds a slot): a particle's slot in its cell is its order among the
 * workgroup's particles of that cell (LDS atomic) plus the workgroup's base
 * in the cell (one global atomic per distinct cell).  The particles arrive
 * in the previous collision's cell order, so a workgroup touches a few
 * hundred cells. */
template <int PPT>
__global__ __launch_bounds__(256) void k_bin(DevState d, double h, int mode) {
    constexpr int HS = 2 * 256 * PPT;
    constexpr uint32_t kEmpty = 0xFFFFFFFFu;
    __shared__ uint32_t h_key[HS], h_cnt[HS];
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const int r = find_rep(d.pblk_base, d.n_rep, lb);
    const RepDev& rep = d.reps[r];
    RepCtl c;
    if (!sort_ctl(d, r, mode, c)) return;
    Box b = rep.box;
    b.delrx = c.delrx;
    for (int q = threadIdx.x; q < HS; q += 256) { h_key[q] = kEmpty; h_cnt[q] = 0u; }
    __syncthreads();
    const uint64_t p0 = rep.part_base + (uint64_t)(lb - d.pblk_base[r]) * (256 * PPT);
    const uint64_t pend = rep.part_base + rep.n_fluid;
    const int cur = d.cur;
    const float *X = d.x[cur], *Y = d.y[cur], *Z = d.z[cur];
    const float *VX = d.vx[cur], *VY = d.vy[cur], *VZ = d.vz[cur];
    bool bad = false, big = false;
    uint32_t slot[PPT], loc[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const uint64_t i = p0 + (uint64_t)j * 256 + threadIdx.x;
        slot[j] = kEmpty;
        loc[j] = 0u;
        if (i >= pend) continue;
        double x = X[i], y = Y[i], z = Z[i], vx = VX[i], vy = VY[i], vz = VZ[i];
        recenter_stream(b, h, c, x, y, z, vx, vy, vz);
        const uint32_t cell = shifted_cell(b, c, x, y, z, vx);
        if (cell >= rep.ncells) {
            bad = true;
            d.pcell[i] = ~0u;
            continue;
        }
        const uint32_t g = (uint32_t)rep.cell_base + cell;
        d.pcell[i] = g;
        uint32_t sl = (g * 0x9E3779B1u) >> (32 - __builtin_ctz(HS));
        for (int probe = 0; probe < HS; ++probe) {
            const uint32_t k = atomicCAS(&h_key[sl], kEmpty, g);
            if (k == kEmpty || k == g) break;
            sl = (sl + 1) & (HS - 1);
        }
        slot[j] = sl;
        loc[j] = atomicAdd(&h_cnt[sl], 1u);
    }
    __syncthreads();
    /* one global atomic per distinct cell: the count becomes the base */
    for (int q = threadIdx.x; q < HS; q += 256) {
        const uint32_t k = h_key[q];
        if (k != kEmpty) h_cnt[q] = atomicAdd(&d.counts[k], h_cnt[q]);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        if (slot[j] == kEmpty) continue;
        const uint64_t i = p0 + (uint64_t)j * 256 + threadIdx.x;
        const uint32_t rank = h_cnt[slot[j]] + loc[j];
        big |= rank > 0xFFFFu;
        d.rank[i] = (uint16_t)(rank > 0xFFFFu ? 0xFFFFu : rank);
    }
    if (bad) atomicOr(&d.ctl[r].status, ST_CELL_OOB);
    if (big) atomicOr(&d.ctl[r].status, ST_RANK_OVERFLOW);
}

/* ============================================================= scan */

__global__ __launch_bounds__(kScanThreads) void k_scan(DevState d) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wsum[kScanThreads / 64];
    __shared__ uint32_t s_prefix;
    const int tid = threadIdx.x;
    if (tid == 0) s_tile = atomicAdd(d.scan_ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= d.n_scan_tiles) return;  /* ticket not reset: never index past the flags */
    const uint64_t base = (uint64_t)tile * kScanTile + (uint64_t)tid * kScanItems;
    uint32_t v[kScanItems];
    if (base + kScanItems <= d.c_tot) {
        const uint4* p = reinterpret_cast<const uint4*>(d.counts + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const uint4 t = p[q];
            v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) v[q] = base + q < d.c_tot ? d.counts[base + q] : 0u;
    }
    uint32_t run = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) { const uint32_t t = v[q]; v[q] = run; run += t; }
    /* wave inclusive scan of the thread totals */
    const int lane = tid & 63, w = tid >> 6;
    uint32_t inc = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(inc, o);
        if (lane >= o) inc += t;
    }
    if (lane == 63) s_wsum[w] = inc;
    __syncthreads();
    uint32_t wpre = 0, agg = 0;
#pragma unroll
    for (int q = 0; q < kScanThreads / 64; ++q) {
        if (q < w) wpre += s_wsum[q];
        agg += s_wsum[q];
    }
    const uint32_t texcl = wpre + inc - run;
    /* publish this tile's aggregate, then wave 0 sums the aggregates of
     * all earlier tiles 64 at a time (earlier tiles hold earlier tickets,
     * so they are running and will publish; the wait is bounded) */
    if (tid == 0)
        __hip_atomic_store(&d.scan_flags[tile], (1ull << 62) | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (w == 0) {
        uint32_t acc = 0;
        bool timeout = false;
        for (uint32_t b0 = 0; b0 < tile; b0 += 64) {
            const uint32_t j = b0 + lane;
            if (j < tile) {
                uint64_t f;
                uint32_t spins = 0;
                while (((f = __hip_atomic_load(&d.scan_flags[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) >> 62) == 0) {
                    if (++spins > (1u << 22)) { timeout = true; break; }
                    __builtin_amdgcn_s_sleep(1);
                }
                acc += (uint32_t)(f & 0xFFFFFFFFull);
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        if (timeout) atomicOr(&d.ctl[0].status, ST_SCAN_TIMEOUT);
        if (lane == 0) {
            s_prefix = acc;
            if ((uint64_t)(tile + 1) * kScanTile >= d.c_tot) d.offs[d.cur ^ 1][d.c_tot] = acc + agg;
        }
    }
    __syncthreads();
    const uint32_t off = s_prefix + texcl;
    if (base + kScanItems <= d.c_tot) {
        uint4* po = reinterpret_cast<uint4*>(d.offs[d.cur ^ 1] + base);
        uint4* pc = reinterpret_cast<uint4*>(d.counts + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            po[q] = make_uint4(off + v[4 * q], off + v[4 * q + 1], off + v[4 * q + 2], off + v[4 * q + 3]);
            pc[q] = make_uint4(0u, 0u, 0u, 0u);
        }
    } else {
        for (int q = 0; q < kScanItems; ++q)
            if (base + q < d.c_tot) { d.offs[d.cur ^ 1][base + q] = off + v[q]; d.counts[base + q] = 0u; }
    }
}

/* Exclusive scan of one replica's cell counts when its speculative sort is
 * redone (rare: only after a recentring); one workgroup per replica. */
/* The whole sort again (bin, scan, perm) for a replica whose MD step
 * recentred it, by one workgroup: recentring is rare, and one launch that
 * returns at once otherwise keeps the common path short.  The counts were
 * left zero by the speculative scan. */
__global__ __launch_bounds__(1024) void k_sort_redo(DevState d, double h) {
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    const int r = blockIdx.x;
    const RepDev& rep = d.reps[r];
    {
        /* the monomers' cells (MPC::shiftParticles MPC.cpp:306-310 + sort
         * MPC.cpp:221-235: shifted, never wrapped), from the positions
         * k_md left and the grid shift k_prologue drew */
        const double sh0 = d.ctl[r].shift[0], sh1 = d.ctl[r].shift[1], sh2 = d.ctl[r].shift[2];
        bool oob = false;
        for (uint32_t i = threadIdx.x; i < rep.n_mono; i += 1024) {
            const size_t g = 3 * (size_t)(rep.mono_base + i);
            const uint32_t ci = cell_index(rep.box, d.mpos[g] + sh0, d.mpos[g + 1] + sh1, d.mpos[g + 2] + sh2);
            oob |= ci >= rep.ncells;
            d.mcell[rep.mono_base + i] = ci < rep.ncells ? (uint32_t)rep.cell_base + ci : ~0u;
        }
        if (oob) atomicOr(&d.ctl[r].status, ST_CELL_OOB);
    }
    if (!d.ctl[r].redo) return;
    RepCtl c;
    sort_ctl(d, r, kSortNormal, c);
    Box b = rep.box;
    b.delrx = c.delrx;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint64_t p0 = rep.part_base, pend = rep.part_base + rep.n_fluid;
    const int cur = d.cur;
    bool bad = false, big = false;
    for (uint64_t i = p0 + tid; i < pend; i += 1024) {
        double x = d.x[cur][i], y = d.y[cur][i], z = d.z[cur][i];
        double vx = d.vx[cur][i], vy = d.vy[cur][i], vz = d.vz[cur][i];
        recenter_stream(b, h, c, x, y, z, vx, vy, vz);
        const uint32_t cell = shifted_cell(b, c, x, y, z, vx);
        if (cell >= rep.ncells) { bad = true; d.pcell[i] = ~0u; continue; }
        const uint32_t g = (uint32_t)rep.cell_base + cell;
        d.pcell[i] = g;
        const uint32_t rank = atomicAdd(&d.counts[g], 1u);
        big |= rank > 0xFFFFu;
        d.rank[i] = (uint16_t)(rank > 0xFFFFu ? 0xFFFFu : rank);
    }
    if (bad) atomicOr(&d.ctl[r].status, ST_CELL_OOB);
    if (big) atomicOr(&d.ctl[r].status, ST_RANK_OVERFLOW);
    if (tid == 0) s_carry = (uint32_t)rep.part_base;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < rep.ncells; c0 += 1024) {
        const uint32_t cc = c0 + tid;
        const uint32_t g = (uint32_t)rep.cell_base + cc;
        const uint32_t v = cc < rep.ncells ? atomicExch(&d.counts[g], 0u) : 0u;  /* read and re-zero */
        uint32_t inc = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(inc, o);
            if (lane >= o) inc += u;
        }
        if (lane == 63) s_w[w] = inc;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int q = 0; q < 16; ++q) { pre += q < w ? s_w[q] : 0u; tot += s_w[q]; }
        const uint32_t base = s_carry;
        if (cc < rep.ncells) d.offs[d.cur ^ 1][g] = base + pre + inc - v;
        __syncthreads();
        if (tid == 0) s_carry = base + tot;
        __syncthreads();
    }
    __threadfence();
    __syncthreads();
    for (uint64_t i = p0 + tid; i < pend; i += 1024) {
        const uint32_t cell = d.pcell[i];
        if (cell == ~0u) continue;
        const uint32_t dst = d.offs[d.cur ^ 1][cell] + d.rank[i];
        if (dst >= pend || dst < p0) continue;
        d.perm[dst] = (uint32_t)i;
    }
}

/* ============================================================== perm */

/* Sorted slot of every particle: perm[offsets[cell] + rank] = index.  Only
 * 4 bytes per particle move here; k_collide gathers the particle data. */
template <int PPT>
__global__ __launch_bounds__(256) void k_perm(DevState d, int mode) {
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const int r = find_rep(d.pblk_base, d.n_rep, lb);
    const RepDev& rep = d.reps[r];
    const uint64_t p0 = rep.part_base + (uint64_t)(lb - d.pblk_base[r]) * (256 * PPT);
    const uint64_t pend = rep.part_base + rep.n_fluid;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const uint64_t i = p0 + (uint64_t)j * 256 + threadIdx.x;
        if (i >= pend) continue;
        const uint32_t cell = d.pcell[i];
        if (cell == ~0u) continue;
        const uint32_t dst = d.offs[d.cur ^ 1][cell] + d.rank[i];
        if (dst >= pend || dst < rep.part_base) continue;  /* cannot happen when bin and scan agree */
        d.perm[dst] = (uint32_t)i;
    }
}

/* =========================================================== collide */

/* One workgroup per tile of TC consecutive cells of one replica.  The
 * tile's particles are gathered from the current buffer through perm (their
 * sorted slots are contiguous), streamed again exactly as k_bin streamed
 * them, and written to the other buffer in cell order.  Up to CAP of them
 * stay in registers between the passes (larger tiles gather again).  Per
 * cell, in this order (MPC.cpp:130-137): CMV over fluid and monomers,
 * relative kinetic energy, MBS scale sqrt(Gamma/E) (only n >= 2), then
 * v <- CMV + s R (v - CMV).  Sums are 64-bit fixed point in LDS, so their
 * value does not depend on the order the particles arrive in.
 * AMC adds, for replicas with the angular-momentum flag and cells of more
 * than two members (MPC.cpp:258-269, 324-374): the centre of mass CMP, the
 * inertia tensor I and C = sum m r x (u - R u) about it (r = x - CMP, u = v -
 * CMV, before the thermostat), omega = s I^-1 C by Cholesky, and v +=
 * omega x r.  Positions enter as coordinates inside the cell (FP32, exact to
 * 6e-8), so r carries no cancellation. */
template <int TC, int CAP, bool AMC>
__global__ __launch_bounds__(256) void k_collide(DevState d, double h) {
    constexpr int NT = 256;
    constexpr int KM = CAP / NT;       /* particles a thread keeps in registers */
    constexpr int MCAP = 2 * TC + 64;  /* monomers per tile (unit cells hold ~1 at most) */
    constexpr int TA = AMC ? TC : 1;   /* cells of the angular-momentum sums */
    constexpr int KA = AMC ? KM : 1;
    __shared__ uint32_t s_off[TC + 1];
    __shared__ unsigned long long s_px[TC], s_py[TC], s_pz[TC], s_e[TC], s_mm[TC];
    __shared__ uint32_t s_nm[TC];
    __shared__ double s_cmv[TC][3];
    __shared__ double s_g[TC];
    __shared__ float s_R[TC][9];
    __shared__ uint32_t s_ml[MCAP];    /* local cell << 16 | monomer */
    __shared__ uint32_t s_mn;
    __shared__ uint32_t s_bad;
    __shared__ unsigned long long s_am[12][TA];  /* sum m l (3), I (6: xx yy zz xy xz yz), C (3) */
    __shared__ double s_cmp[TA][3], s_w[TA][3];

    const int tid = threadIdx.x;
    const uint32_t lt = xcd_remap(blockIdx.x, gridDim.x);
    const int r = find_rep(d.tile_base, d.n_rep, lt);
    const RepDev& rep = d.reps[r];
    RepCtl* ctl = d.ctl + r;
    RepCtl cc;
    sort_ctl(d, r, kSortNormal, cc);
    Box b = rep.box;
    b.delrx = cc.delrx;
    const uint32_t c_lo = (lt - d.tile_base[r]) * TC;
    const uint32_t nc = min((uint32_t)TC, rep.ncells - c_lo);
    const uint32_t gc0 = (uint32_t)rep.cell_base + c_lo;
    const int cur = d.cur, alt = cur ^ 1;
    const float *X = d.x[cur], *Y = d.y[cur], *Z = d.z[cur];
    const float *VX = d.vx[cur], *VY = d.vy[cur], *VZ = d.vz[cur];
    float *OX = d.x[alt], *OY = d.y[alt], *OZ = d.z[alt];
    float *OVX = d.vx[alt], *OVY = d.vy[alt], *OVZ = d.vz[alt];
    const uint32_t* ids = d.ids[cur];
    uint32_t* oids = d.ids[alt];
    const double shear = rep.box.shear, LyD = rep.box.L[1];

    for (uint32_t i = tid; i <= nc; i += NT) s_off[i] = d.offs[d.cur ^ 1][gc0 + i];
    for (uint32_t i = tid; i < nc; i += NT) {
        s_px[i] = s_py[i] = s_pz[i] = s_e[i] = s_mm[i] = 0ull;
        s_nm[i] = 0u;
    }
    if (tid == 0) { s_mn = 0u; s_bad = 0u; }
    const bool amc = AMC && rep.amc;
    if (amc)
        for (uint32_t i = tid; i < 12u * nc; i += NT) s_am[i / nc][i % nc] = 0ull;
    __syncthreads();
    /* the replica's monomers in this tile's cells (any order: the sums
     * are fixed point) */
    for (uint32_t j = tid; j < rep.n_mono; j += NT) {
        const uint32_t rel = d.mcell[rep.mono_base + j] - gc0;
        if (rel < nc) {
            const uint32_t slot = atomicAdd(&s_mn, 1u);
            if (slot < (uint32_t)MCAP) { s_ml[slot] = (rel << 16) | j; atomicAdd(&s_nm[rel], 1u); }
        }
    }
    __syncthreads();
    /* A scan whose bounded look-back wait expired (reported as
     * ST_SCAN_TIMEOUT) leaves offsets that need not be monotone: the tile
     * then touches nothing, and its slot range is clamped to the replica's
     * particles in any case. */
    if (d.ctl[0].status & ST_SCAN_TIMEOUT) return;
    const uint64_t pr0 = rep.part_base, pr1 = rep.part_base + rep.n_fluid;
    const uint32_t p0 = (uint32_t)min<uint64_t>(max<uint64_t>(s_off[0], pr0), pr1);
    const uint32_t p1 = (uint32_t)min<uint64_t>(max<uint64_t>(s_off[nc], p0), pr1);
    const uint32_t np = p1 - p0;
    const uint32_t nmt = s_mn < (uint32_t)MCAP ? s_mn : (uint32_t)MCAP;
    if (tid == 0 && s_mn > (uint32_t)MCAP) atomicOr(&ctl->status, ST_TILE_MONO);
    const uint32_t k = ctl->k_cur;
    const bool inj = ctl->inj_active != 0;

    /* per cell: the Gamma variate (depends only on n) and the rotation --
     * issued before the particle loads so the waves overlap them */
    for (uint32_t ci = tid; ci < nc; ci += NT) {
        const uint32_t cell = c_lo + ci;
        const uint32_t n = (s_off[ci + 1] - s_off[ci]) + s_nm[ci];
        double g = 0.0;
        if (n >= 2)
            g = inj ? d.inj_gamma[gc0 + ci] : gamma_fast(rep.seed, rep.replica, cell, k, 1.5 * (n - 1), rep.kT);
        s_g[ci] = g;
        double u1, u2;
        if (inj) {
            u1 = d.inj_axis[2 * (size_t)(gc0 + ci)];
            u2 = d.inj_axis[2 * (size_t)(gc0 + ci) + 1];
        } else {
            uint32_t w[4];
            draw4(rep.seed, rep.replica, TAG_AXIS, cell, k, 0, w);
            u1 = u01(w[0]);
            u2 = u01(w[1]);
        }
        rotation_f32(u1, u2, rep.c, rep.s, s_R[ci]);
    }

    /* slot q of the tile: its particle streamed (MPC::streaming + LEBC, as
     * in k_bin), FP32 velocity in the unshifted frame, y image index cy of
     * the shifted position (MPC::shiftParticles' wrap) and local cell; with
     * write_pos the new position (and id) go to the other buffer */
    auto gather = [&](uint32_t q, bool write_pos, uint32_t& lc, int& cy, float& fvx, float& fvy, float& fvz,
                      float& lx, float& ly, float& lz) {
        const uint32_t src = d.perm[p0 + q];
        double x = X[src], y = Y[src], z = Z[src], vx = VX[src], vy = VY[src], vz = VZ[src];
        recenter_stream(b, h, cc, x, y, z, vx, vy, vz);
        if (write_pos) {
            OX[p0 + q] = (float)x; OY[p0 + q] = (float)y; OZ[p0 + q] = (float)z;
            if (ids) oids[p0 + q] = ids[src];
        }
        fvx = (float)vx; fvy = (float)vy; fvz = (float)vz;
        double px = x + cc.shift[0], py = y + cc.shift[1], pz = z + cc.shift[2], vimg = vx;
        cy = (int)le_wrap(b, px, py, pz, vimg);
        lc = cell_index(b, px, py, pz) - c_lo;
        if (lc >= nc) { s_bad = 1u; lc = 0; }
        if (AMC) {  /* coordinates inside the cell, as cell_index truncates */
            const double qx = px + b.half[0], qy = py + b.half[1], qz = pz + b.half[2];
            lx = (float)(qx - (double)(int)qx); ly = (float)(qy - (double)(int)qy); lz = (float)(qz - (double)(int)qz);
        }
    };
    /* a monomer's coordinates inside its cell (shifted, never wrapped) */
    auto mono_local = [&](uint32_t mi, double& lx, double& ly, double& lz) {
        const double qx = d.mpos[3 * mi] + cc.shift[0] + b.half[0];
        const double qy = d.mpos[3 * mi + 1] + cc.shift[1] + b.half[1];
        const double qz = d.mpos[3 * mi + 2] + cc.shift[2] + b.half[2];
        lx = qx - (double)(int)qx; ly = qy - (double)(int)qy; lz = qz - (double)(int)qz;
    };
    uint32_t rlc[KM];
    int rcy[KM];
    float rvx[KM], rvy[KM], rvz[KM];
    float rlx[KA], rly[KA], rlz[KA];
    auto add_l = [&](uint32_t lc, double m, double lx, double ly, double lz) {
        atomicAdd(&s_am[0][lc], to_fix(m * lx));
        atomicAdd(&s_am[1][lc], to_fix(m * ly));
        atomicAdd(&s_am[2][lc], to_fix(m * lz));
    };

    /* MPC::calculateCMV (MPC.cpp:314-322) */
#pragma unroll
    for (int kk = 0; kk < KM; ++kk) {
        const uint32_t q = tid + kk * NT;
        if (q >= np) continue;
        float lx = 0.f, ly = 0.f, lz = 0.f;
        gather(q, true, rlc[kk], rcy[kk], rvx[kk], rvy[kk], rvz[kk], lx, ly, lz);
        const uint32_t lc = rlc[kk];
        atomicAdd(&s_px[lc], to_fix((double)rvx[kk] - (double)rcy[kk] * shear * LyD));
        atomicAdd(&s_py[lc], to_fix((double)rvy[kk]));
        atomicAdd(&s_pz[lc], to_fix((double)rvz[kk]));
        if (AMC) {
            rlx[AMC ? kk : 0] = lx; rly[AMC ? kk : 0] = ly; rlz[AMC ? kk : 0] = lz;
            if (amc) add_l(lc, 1.0, lx, ly, lz);
        }
    }
    for (uint32_t q = tid + CAP; q < np; q += NT) {
        uint32_t lc; int cy; float fx, fy, fz, lx = 0.f, ly = 0.f, lz = 0.f;
        gather(q, true, lc, cy, fx, fy, fz, lx, ly, lz);
        atomicAdd(&s_px[lc], to_fix((double)fx - (double)cy * shear * LyD));
        atomicAdd(&s_py[lc], to_fix((double)fy));
        atomicAdd(&s_pz[lc], to_fix((double)fz));
        if (amc) add_l(lc, 1.0, lx, ly, lz);
    }
    for (uint32_t j = tid; j < nmt; j += NT) {
        const uint32_t e = s_ml[j];
        const uint32_t lc = e >> 16;
        const uint32_t mi = rep.mono_base + (e & 0xFFFFu);
        const double mm = d.mmass[mi];
        atomicAdd(&s_px[lc], to_fix(d.mvel[3 * mi] * mm));
        atomicAdd(&s_py[lc], to_fix(d.mvel[3 * mi + 1] * mm));
        atomicAdd(&s_pz[lc], to_fix(d.mvel[3 * mi + 2] * mm));
        atomicAdd(&s_mm[lc], to_fix(mm));
        if (amc) {
            double lx, ly, lz;
            mono_local(mi, lx, ly, lz);
            add_l(lc, mm, lx, ly, lz);
        }
    }
    __syncthreads();
    if (tid == 0 && s_bad) atomicOr(&ctl->status, ST_SORT_MISMATCH);
    for (uint32_t ci = tid; ci < nc; ci += NT) {
        const double M = (double)(s_off[ci + 1] - s_off[ci]) + from_fix(s_mm[ci]);
        double c0 = from_fix(s_px[ci]), c1 = from_fix(s_py[ci]), c2 = from_fix(s_pz[ci]);
        if (M > 0) { c0 /= M; c1 /= M; c2 /= M; }
        s_cmv[ci][0] = c0; s_cmv[ci][1] = c1; s_cmv[ci][2] = c2;
        if (amc) {  /* MPC::calculateCMP MPC.cpp:324-333, inside the cell */
            const double inv = M > 0 ? 1.0 / M : 0.0;
            s_cmp[ci][0] = from_fix(s_am[0][ci]) * inv;
            s_cmp[ci][1] = from_fix(s_am[1][ci]) * inv;
            s_cmp[ci][2] = from_fix(s_am[2][ci]) * inv;
        }
    }
    __syncthreads();

    /* MPC::thermostat relative kinetic energy (MPC.cpp:277-279) */
    /* the angular-momentum sums of one member: I about CMP and
     * C = m r x (u - R u) (MPC::calculateAngular MPC.cpp:344-372) */
    auto add_amc = [&](uint32_t lc, double m, double lx, double ly, double lz, double ux, double uy, double uz) {
        const double rx = lx - s_cmp[lc][0], ry = ly - s_cmp[lc][1], rz = lz - s_cmp[lc][2];
        const float* R = s_R[lc];
        const double qx = ux - (((double)R[0] * ux + (double)R[1] * uy) + (double)R[2] * uz);
        const double qy = uy - (((double)R[3] * ux + (double)R[4] * uy) + (double)R[5] * uz);
        const double qz = uz - (((double)R[6] * ux + (double)R[7] * uy) + (double)R[8] * uz);
        atomicAdd(&s_am[3][lc], to_fix(m * (ry * ry + rz * rz)));
        atomicAdd(&s_am[4][lc], to_fix(m * (rx * rx + rz * rz)));
        atomicAdd(&s_am[5][lc], to_fix(m * (ry * ry + rx * rx)));
        atomicAdd(&s_am[6][lc], to_fix(-m * rx * ry));
        atomicAdd(&s_am[7][lc], to_fix(-m * rx * rz));
        atomicAdd(&s_am[8][lc], to_fix(-m * ry * rz));
        atomicAdd(&s_am[9][lc], to_fix(m * (ry * qz - rz * qy)));
        atomicAdd(&s_am[10][lc], to_fix(m * (rz * qx - rx * qz)));
        atomicAdd(&s_am[11][lc], to_fix(m * (rx * qy - ry * qx)));
    };
    auto add_e = [&](uint32_t lc, int cy, float fx, float fy, float fz, float lx, float ly, float lz) {
        const double vx = (double)fx - (double)cy * shear * LyD;
        const double dx = vx - s_cmv[lc][0], dy = (double)fy - s_cmv[lc][1], dz = (double)fz - s_cmv[lc][2];
        atomicAdd(&s_e[lc], to_fix((dx * dx + dy * dy) + dz * dz));
        if (amc) add_amc(lc, 1.0, lx, ly, lz, dx, dy, dz);
    };
#pragma unroll
    for (int kk = 0; kk < KM; ++kk)
        if (tid + kk * NT < np)
            add_e(rlc[kk], rcy[kk], rvx[kk], rvy[kk], rvz[kk], AMC ? rlx[AMC ? kk : 0] : 0.f,
                  AMC ? rly[AMC ? kk : 0] : 0.f, AMC ? rlz[AMC ? kk : 0] : 0.f);
    for (uint32_t q = tid + CAP; q < np; q += NT) {
        uint32_t lc; int cy; float fx, fy, fz, lx = 0.f, ly = 0.f, lz = 0.f;
        gather(q, false, lc, cy, fx, fy, fz, lx, ly, lz);
        add_e(lc, cy, fx, fy, fz, lx, ly, lz);
    }
    for (uint32_t j = tid; j < nmt; j += NT) {
        const uint32_t e = s_ml[j];
        const uint32_t lc = e >> 16;
        const uint32_t mi = rep.mono_base + (e & 0xFFFFu);
        const double dx = d.mvel[3 * mi] - s_cmv[lc][0];
        const double dy = d.mvel[3 * mi + 1] - s_cmv[lc][1];
        const double dz = d.mvel[3 * mi + 2] - s_cmv[lc][2];
        atomicAdd(&s_e[lc], to_fix(d.mmass[mi] * ((dx * dx + dy * dy) + dz * dz)));
        if (amc) {
            double lx, ly, lz;
            mono_local(mi, lx, ly, lz);
            add_amc(lc, d.mmass[mi], lx, ly, lz, dx, dy, dz);
        }
    }
    __syncthreads();

    /* MBS scale sqrt(Gamma / E_rel) for n >= 2 (MPC.cpp:287-289); with the
     * angular-momentum flag, omega = s I^-1 C for n > 2 (MPC.cpp:260, 373) */
    for (uint32_t ci = tid; ci < nc; ci += NT) {
        const uint32_t n = (s_off[ci + 1] - s_off[ci]) + s_nm[ci];
        s_g[ci] = n >= 2 ? sqrt(s_g[ci] / (from_fix(s_e[ci]) * 0.5)) : -1.0;
        if (amc) {
            double w[3] = {0.0, 0.0, 0.0};
            if (n > 2) {
                const double A[3][3] = {{from_fix(s_am[3][ci]), from_fix(s_am[6][ci]), from_fix(s_am[7][ci])},
                                        {from_fix(s_am[6][ci]), from_fix(s_am[4][ci]), from_fix(s_am[8][ci])},
                                        {from_fix(s_am[7][ci]), from_fix(s_am[8][ci]), from_fix(s_am[5][ci])}};
                const double c3[3] = {from_fix(s_am[9][ci]), from_fix(s_am[10][ci]), from_fix(s_am[11][ci])};
                llt_solve3(A, c3, w);
                const double sc = s_g[ci] >= 0.0 ? s_g[ci] : 1.0;
                const bool ok = isfinite(w[0]) && isfinite(w[1]) && isfinite(w[2]);
                for (int q = 0; q < 3; ++q) w[q] = ok ? w[q] * sc : 0.0;
            }
            s_w[ci][0] = w[0]; s_w[ci][1] = w[1]; s_w[ci][2] = w[2];
        }
    }
    __syncthreads();

    /* v <- s v + (1-s) CMV, then v <- CMV + R (v - CMV) (MPC.cpp:290-292,
     * 266-267); the image offset is undone as the unshift wrap does */
    auto apply = [&](uint32_t lc, double& vx, double& vy, double& vz) {
        const double c0 = s_cmv[lc][0], c1 = s_cmv[lc][1], c2 = s_cmv[lc][2];
        const double sc = s_g[lc];
        if (sc >= 0.0) {
            vx = sc * vx + (1 - sc) * c0;
            vy = sc * vy + (1 - sc) * c1;
            vz = sc * vz + (1 - sc) * c2;
        }
        const float* R = s_R[lc];
        const double d0 = vx - c0, d1 = vy - c1, d2 = vz - c2;
        vx = c0 + (((double)R[0] * d0 + (double)R[1] * d1) + (double)R[2] * d2);
        vy = c1 + (((double)R[3] * d0 + (double)R[4] * d1) + (double)R[5] * d2);
        vz = c2 + (((double)R[6] * d0 + (double)R[7] * d1) + (double)R[8] * d2);
    };
    /* v += omega x (x - CMP) (MPC.cpp:268) */
    auto add_w = [&](uint32_t lc, double lx, double ly, double lz, double& vx, double& vy, double& vz) {
        const double rx = lx - s_cmp[lc][0], ry = ly - s_cmp[lc][1], rz = lz - s_cmp[lc][2];
        const double w0 = s_w[lc][0], w1 = s_w[lc][1], w2 = s_w[lc][2];
        vx += w1 * rz - w2 * ry;
        vy += w2 * rx - w0 * rz;
        vz += w0 * ry - w1 * rx;
    };
    auto finish = [&](uint32_t q, uint32_t lc, int cy, float fx, float fy, float fz, float lx, float ly, float lz) {
        const double off = (double)cy * shear * LyD;
        double vx = (double)fx - off, vy = fy, vz = fz;
        apply(lc, vx, vy, vz);
        if (amc) add_w(lc, lx, ly, lz, vx, vy, vz);
        vx += off;
        OVX[p0 + q] = (float)vx; OVY[p0 + q] = (float)vy; OVZ[p0 + q] = (float)vz;
    };
#pragma unroll
    for (int kk = 0; kk < KM; ++kk) {
        const uint32_t q = tid + kk * NT;
        if (q < np)
            finish(q, rlc[kk], rcy[kk], rvx[kk], rvy[kk], rvz[kk], AMC ? rlx[AMC ? kk : 0] : 0.f,
                   AMC ? rly[AMC ? kk : 0] : 0.f, AMC ? rlz[AMC ? kk : 0] : 0.f);
    }
    for (uint32_t q = tid + CAP; q < np; q += NT) {
        uint32_t lc; int cy; float fx, fy, fz, lx = 0.f, ly = 0.f, lz = 0.f;
        gather(q, false, lc, cy, fx, fy, fz, lx, ly, lz);
        finish(q, lc, cy, fx, fy, fz, lx, ly, lz);
    }
    for (uint32_t j = tid; j < nmt; j += NT) {
        const uint32_t e = s_ml[j];
        const uint32_t lc = e >> 16;
        const uint32_t mi = rep.mono_base + (e & 0xFFFFu);
        double vx = d.mvel[3 * mi], vy = d.mvel[3 * mi + 1], vz = d.mvel[3 * mi + 2];
        apply(lc, vx, vy, vz);
        if (amc) {
            double lx, ly, lz;
            mono_local(mi, lx, ly, lz);
            add_w(lc, lx, ly, lz, vx, vy, vz);
        }
        d.mvel[3 * mi] = vx; d.mvel[3 * mi + 1] = vy; d.mvel[3 * mi + 2] = vz;
    }
}

/* ===================================================== row-tile path */

/* A row tile is nr consecutive y-rows (whole x-rows) of one z-plane of a
 * replica, so its cells are one contiguous range of the cell order.  Per
 * collision:
 *   k_advect        per block of particles of the current buffer: streaming
 *                   and the new shifted cell (bit-identical to k_bin's), and
 *                   the particle's index appended to the bucket of its
 *                   destination row tile (an LDS hash of the few tiles a
 *                   block reaches, then one global atomic per tile to
 *                   reserve its slots)
 *   k_rows_redo     after the MD: the monomers' cells and the first slot of
 *                   each row tile in the new buffer (scan of the bucket
 *                   counts); a replica the MD recentred is binned again
 *   k_collide_rows  one workgroup per row tile: its bucket, the particles
 *                   gathered and streamed again, ranked per cell in LDS,
 *                   the collision with the cell sums over 2-4 lanes per
 *                   cell, and the new buffer written in cell order with its
 *                   cell offsets.
 * No cell-level scan or permutation: 4 bytes per particle go through the
 * buckets.  The buffer being in cell order (any collision leaves it so)
 * keeps a tile's gathers within a few rows of it. */

struct RowTile {
    int z, ty, y0, ny;
    uint32_t c_lo, nc;     /* first cell (replica-local) and cells */
};

__device__ __forceinline__ RowTile row_tile(const RepDev& rep, uint32_t t) {
    RowTile T;
    T.z = (int)(t / (uint32_t)rep.tpp);
    T.ty = (int)t - T.z * rep.tpp;
    T.y0 = T.ty * rep.nr;
    T.ny = min(rep.nr, rep.box.iL[1] - T.y0);
    T.c_lo = (uint32_t)rep.box.iL[0] * (uint32_t)(T.y0 + rep.box.iL[1] * T.z);
    T.nc = (uint32_t)(rep.box.iL[0] * T.ny);
    return T;
}

/* Row tile (replica-local) of a replica-local cell index, and the cell's
 * index inside the tile (row_tile's c_lo + lc is the cell). */
__device__ __forceinline__ void tile_of_cell(const RepDev& rep, uint32_t cell, uint32_t& tile, uint32_t& lc) {
    const uint32_t Lx = (uint32_t)rep.box.iL[0], nr = (uint32_t)rep.nr;
    const uint32_t row = cell / Lx;
    const uint32_t iz = row / (uint32_t)rep.box.iL[1], iy = row - iz * (uint32_t)rep.box.iL[1];
    const uint32_t ty = iy / nr;
    tile = iz * (uint32_t)rep.tpp + ty;
    lc = cell - Lx * (ty * nr + (uint32_t)rep.box.iL[1] * iz);
}

constexpr int kBinHash = 64;      /* distinct destination tiles a block aggregates in LDS */
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;

/* Bin one block of particles [p0, p1) (NT threads, PPT per thread) into the
 * row-tile buckets: bucket + tile * cap holds the indices of the particles
 * whose new cell lies in the tile, cnt[tile] their number.  A particle's
 * slot is its rank among the block's particles of that tile (LDS hash,
 * open addressing) plus the block's base in the tile (one global atomic
 * per distinct tile and block); blocks that reach more than kBinHash tiles
 * (a buffer not in cell order) take a global atomic per particle for the
 * rest.  Slots from cap on are counted but not stored (the collide reports
 * the overflow).  Beside each index, blc holds the particle's cell in its
 * tile (one byte: a tile has at most 256 cells), from which k_cell_gamma
 * counts the cells' members. */
template <int NT, int PPT>
__device__ void bin_block(const DevState& d, int r, uint64_t p0, uint64_t p1, const RepCtl& c, double h,
                          uint32_t* cnt, uint32_t* bucket, uint8_t* blc, uint32_t* s_key, uint32_t* s_cnt,
                          bool& over) {
    const RepDev& rep = d.reps[r];
    Box b = rep.box;
    b.delrx = c.delrx;
    const int tid = threadIdx.x;
    const uint32_t cap = (uint32_t)d.rt_cap;
    __syncthreads();  /* the tables of a previous block are read out */
    for (int q = tid; q < kBinHash; q += NT) { s_key[q] = kHashEmpty; s_cnt[q] = 0u; }
    __syncthreads();
    const int cur = d.cur;
    const float *X = d.x[cur], *Y = d.y[cur], *Z = d.z[cur];
    const float *VX = d.vx[cur], *VY = d.vy[cur], *VZ = d.vz[cur];
    float fx[PPT], fy[PPT], fz[PPT], fvx[PPT], fvy[PPT], fvz[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const uint64_t i = p0 + (uint64_t)(j * NT + tid);
        if (i >= p1) continue;
        fx[j] = X[i]; fy[j] = Y[i]; fz[j] = Z[i];
        fvx[j] = VX[i]; fvy[j] = VY[i]; fvz[j] = VZ[i];
    }
    bool bad = false;
    /* global row tile; hash entry << 16 | rank, or kDirect | the slot a
     * global atomic gave */
    constexpr uint32_t kDirect = 0x80000000u;
    uint32_t tl[PPT], slot[PPT];
    uint8_t lcl[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        tl[j] = kHashEmpty;
        const uint64_t i = p0 + (uint64_t)(j * NT + tid);
        if (i >= p1) continue;
        double x = fx[j], y = fy[j], z = fz[j], vx = fvx[j], vy = fvy[j], vz = fvz[j];
        recenter_stream(b, h, c, x, y, z, vx, vy, vz);
        const uint32_t cell = shifted_cell(b, c, x, y, z, vx);
        if (cell >= rep.ncells) { bad = true; continue; }
        uint32_t lt, lc;
        tile_of_cell(rep, cell, lt, lc);
        const uint32_t g = rep.rt_base + lt;
        tl[j] = g;
        lcl[j] = (uint8_t)lc;
        uint32_t hs = (g * 0x9E3779B1u) >> 26;
        int probe = 0;
        /* a key, once set, never changes: a lane that reads its tile at
         * the home slot (set by an earlier lane) skips the CAS, whose
         * same-address lanes serialise */
        const bool home = __hip_atomic_load(&s_key[hs], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == g;
        for (; !home && probe < kBinHash; ++probe) {
            const uint32_t k = atomicCAS(&s_key[hs], kHashEmpty, g);
            if (k == kHashEmpty || k == g) break;
            hs = (hs + 1) & (kBinHash - 1);
        }
        if (probe == kBinHash) {
            slot[j] = kDirect | min(atomicAdd(&cnt[g], 1u), cap);
        } else {
            slot[j] = (hs << 16) | atomicAdd(&s_cnt[hs], 1u);
        }
    }
    __syncthreads();
    for (int q = tid; q < kBinHash; q += NT) {
        const uint32_t k = s_key[q];
        if (k != kHashEmpty) s_cnt[q] = atomicAdd(&cnt[k], s_cnt[q]);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        if (tl[j] == kHashEmpty) continue;
        const uint64_t i = p0 + (uint64_t)(j * NT + tid);
        const uint32_t sl = (slot[j] & kDirect) ? slot[j] & ~kDirect : s_cnt[slot[j] >> 16] + (slot[j] & 0xFFFFu);
        if (sl < cap) {
            bucket[(size_t)tl[j] * cap + sl] = (uint32_t)i;
            blc[(size_t)tl[j] * cap + sl] = lcl[j];
        } else {
            over = true;
        }
    }
    if (bad) atomicOr(&d.ctl[r].status, ST_CELL_OOB);
}

template <int PPT>
__global__ __launch_bounds__(256) void k_advect(DevState d, double h, int mode) {
    __shared__ uint32_t s_key[kBinHash], s_cnt[kBinHash];
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const int r = find_rep(d.pblk_base, d.n_rep, lb);
    const RepDev& rep = d.reps[r];
    RepCtl c;
    sort_ctl(d, r, mode, c);
    /* the recentring redo's counts of an earlier collision are read out:
     * clear them for this one's redo (which runs after this kernel) */
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < d.n_rtiles; i += gridDim.x * 256u) d.bcount_r[i] = 0u;
    const uint64_t p0 = rep.part_base + (uint64_t)(lb - d.pblk_base[r]) * (256 * PPT);
    const uint64_t p1 = min(p0 + 256 * PPT, rep.part_base + rep.n_fluid);
    bool over = false;
    bin_block<256, PPT>(d, r, p0, p1, c, h, d.bcount[d.cur], d.bucket, d.blc, s_key, s_cnt, over);
    (void)over;  /* an overflowing tile reports it from its count */
}

/* The monomer cells of replica r in an LDS hash of hs entries (cell ->
 * monomers in it; open addressing, hs >= 2 x monomers, so an insert always
 * finds a slot); with write_mcell, each monomer's cell to mcell
 * (MPC::shiftParticles MPC.cpp:306-310 + sort MPC.cpp:221-235: shifted,
 * never wrapped).  The grid shift sh comes from the caller.  Every
 * monomer's position is loaded before any is inserted (the blockDim of 256
 * takes kMaxMono / 256 per thread).  Ends with the table complete for the
 * whole workgroup. */
constexpr uint32_t kMonoHash = 2u * kMaxMono;
constexpr int kMonoPerThread = kMaxMono / 256;

__device__ __forceinline__ uint32_t mono_hash_size(uint32_t n_mono) {
    uint32_t hs = 64u;
    while (hs < 2u * n_mono) hs <<= 1;
    return hs;
}

__device__ __forceinline__ uint32_t mono_slot(uint32_t gc, uint32_t hs) {
    return (gc * 0x9E3779B1u) >> (32 - __builtin_ctz(hs));
}

__device__ void mono_cells(const DevState& d, int r, const double sh[3], bool write_mcell, uint32_t hs,
                           uint32_t* s_mkey, uint32_t* s_mcnt) {
    const RepDev& rep = d.reps[r];
    uint32_t gcs[kMonoPerThread];
    bool oob = false;
#pragma unroll
    for (int u = 0; u < kMonoPerThread; ++u) {
        const uint32_t i = threadIdx.x + 256u * u;
        gcs[u] = ~0u;
        if (i >= rep.n_mono) continue;
        const size_t g = 3 * (size_t)(rep.mono_base + i);
        const uint32_t ci = cell_index(rep.box, d.mpos[g] + sh[0], d.mpos[g + 1] + sh[1], d.mpos[g + 2] + sh[2]);
        oob |= ci >= rep.ncells;
        gcs[u] = ci < rep.ncells ? (uint32_t)rep.cell_base + ci : ~0u;
        if (write_mcell) d.mcell[rep.mono_base + i] = gcs[u];
    }
    __syncthreads();  /* the table of an earlier replica is read out */
    for (uint32_t q = threadIdx.x; q < hs; q += 256u) { s_mkey[q] = kHashEmpty; s_mcnt[q] = 0u; }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kMonoPerThread; ++u) {
        const uint32_t gc = gcs[u];
        if (gc == ~0u) continue;
        uint32_t hc = mono_slot(gc, hs);
        while (true) {
            const uint32_t k = atomicCAS(&s_mkey[hc], kHashEmpty, gc);
            if (k == kHashEmpty || k == gc) break;
            hc = (hc + 1) & (hs - 1);
        }
        atomicAdd(&s_mcnt[hc], 1u);
    }
    if (oob && write_mcell) atomicOr(&d.ctl[r].status, ST_CELL_OOB);
    __syncthreads();
}

/* Monomers of global cell gc in the table (0 if none). */
__device__ __forceinline__ uint32_t mono_count(uint32_t gc, uint32_t hs, const uint32_t* s_mkey,
                                               const uint32_t* s_mcnt) {
    uint32_t hc = mono_slot(gc, hs);
    for (uint32_t probe = 0; probe < hs; ++probe) {
        const uint32_t k = s_mkey[hc];
        if (k == gc) return s_mcnt[hc];
        if (k == kHashEmpty) return 0u;
        hc = (hc + 1) & (hs - 1);
    }
    return 0u;
}

/* A cell's Gamma variate of this collision for n members (the MBS
 * thermostat, MPC.cpp:287-289, with k_collide's expressions), and n. */
__device__ __forceinline__ void draw_cell_gamma(const DevState& d, const RepDev& rep, uint32_t gc, uint32_t n,
                                                uint32_t k, bool inj) {
    d.cg_n[gc] = n;
    d.cg_g[gc] = n >= 2 && !inj ? gamma_fast(rep.seed, rep.replica, gc - (uint32_t)rep.cell_base, k, 1.5 * (n - 1),
                                             rep.kT)
                                : 0.0;
}

/* Row tile lt of replica r: its cells' members counted from the n cell
 * bytes of its bucket (an LDS histogram, s_h of 256), plus the monomers of
 * the table when one is given, and every cell's variate drawn for its
 * count. */
template <int NT>
__device__ void tile_gammas(const DevState& d, int r, uint32_t lt, uint32_t n, const uint8_t* lcs, uint32_t* s_h,
                            uint32_t hs, const uint32_t* s_mkey, const uint32_t* s_mcnt) {
    const RepDev& rep = d.reps[r];
    const RowTile T = row_tile(rep, lt - rep.rt_base);
    const uint32_t k = d.ctl[r].k_cur;
    const bool inj = d.ctl[r].inj_active != 0;
    __syncthreads();  /* the histogram of an earlier tile is read out */
    for (uint32_t c = threadIdx.x; c < 256u; c += NT) s_h[c] = 0u;
    __syncthreads();
    n = min(n, (uint32_t)d.rt_cap);
    const uint32_t* w = reinterpret_cast<const uint32_t*>(lcs);  /* a tile's bytes start 4-aligned */
    for (uint32_t q = threadIdx.x; q < (n + 3u) / 4u; q += NT) {
        const uint32_t v = w[q];
#pragma unroll
        for (uint32_t b = 0; b < 4u; ++b)
            if (4u * q + b < n) atomicAdd(&s_h[(v >> (8u * b)) & 0xFFu], 1u);
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < T.nc; c += NT) {
        const uint32_t gc = (uint32_t)rep.cell_base + T.c_lo + c;
        draw_cell_gamma(d, rep, gc, s_h[c] + (s_mkey ? mono_count(gc, hs, s_mkey, s_mcnt) : 0u), k, inj);
    }
}

/* After k_advect, beside the MD:
 *   - blocks < n_rep: the first slot of every row tile of replica r in the
 *     new buffer (exclusive scan of the bucket counts, 32 per lane a
 *     pass); a recentred replica's collide sums its redo counts itself
 *   - the n_rtiles blocks after them, one wave per row tile: every cell's
 *     Gamma variate of this collision drawn ahead of the collide for its
 *     solvent count (assuming no recentring, as k_advect does); the redo
 *     kernel draws the cells holding monomers again, and all cells of a
 *     recentred replica; the collide checks that each variate was drawn
 *     for the cell's member count. */
__global__ __launch_bounds__(64) void k_cell_gamma(DevState d) {
    __shared__ uint32_t s_h[256];
    if (blockIdx.x < (uint32_t)d.n_rep) {
        /* the scans first: dispatched before the tile blocks, they run
         * beside them (32 counts a lane: two passes for 4,096 tiles) */
        constexpr int KS = 32;
        const int r = (int)blockIdx.x;
        const uint32_t t0 = d.rtile_base[r], nt = d.rtile_base[r + 1] - t0;
        const uint32_t* tc = d.bcount[d.cur] + t0;
        const int lane = threadIdx.x;
        uint32_t carry = (uint32_t)d.reps[r].part_base;
        for (uint32_t c0 = 0; c0 < nt; c0 += 64u * KS) {
            uint32_t v[KS];
            const uint32_t i0 = c0 + (uint32_t)lane * KS;
#pragma unroll
            for (int q = 0; q < KS; ++q) v[q] = i0 + q < nt ? load_count(tc + i0 + q) : 0u;
            uint32_t run = 0u;
#pragma unroll
            for (int q = 0; q < KS; ++q) { const uint32_t t = v[q]; v[q] = run; run += t; }
            uint32_t inc = run;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t u = __shfl_up(inc, o);
                if (lane >= o) inc += u;
            }
            const uint32_t base = carry + inc - run;
#pragma unroll
            for (int q = 0; q < KS; ++q)
                if (i0 + q < nt) d.tbase[t0 + i0 + q] = base + v[q];
            carry += __shfl(inc, 63);
        }
        return;
    }
    const uint32_t b = blockIdx.x - (uint32_t)d.n_rep;
    const int r = find_rep(d.rtile_base, d.n_rep, b);
    tile_gammas<64>(d, r, b, load_count(d.bcount[d.cur] + b), d.blc + (size_t)b * (uint32_t)d.rt_cap, s_h, 0u, nullptr, nullptr);
    if (b == d.reps[r].rt_base && threadIdx.x == 0) d.ctl[r].cg_k = d.ctl[r].k_cur + 1u;
}

/* After k_advect, k_cell_gamma and the MD kernel:
 *   - the monomers' cells of this collision, and the variate of each cell
 *     holding monomers drawn again for its solvent count plus its monomers
 *     (G / n_rep workgroups per replica split its cells)
 *   - a replica whose MD recentred it (MPC::check_bounds MPC.cpp:160-190)
 *     binned again, with the recentring, into the redo buckets (the grid
 *     strides over its particle blocks), then -- after a grid-wide barrier
 *     (the 256 workgroups are all resident; the wait is bounded) -- all its
 *     cells' variates drawn for the new counts. */
template <int PPT>
__global__ __launch_bounds__(256) void k_rows_redo(DevState d, double h) {
    __shared__ uint32_t s_key[kBinHash], s_cnt[kBinHash];
    __shared__ uint32_t s_mkey[kMonoHash], s_mcnt[kMonoHash];
    __shared__ uint32_t s_h[256];
    const int G = (int)gridDim.x;
    const int wpr = d.n_rep >= G ? 1 : G / d.n_rep;
    const int sub = (int)blockIdx.x % wpr;
    bool any_redo = false;
    for (int r = 0; r < d.n_rep; ++r) any_redo |= d.ctl[r].redo != 0;
    for (int r = (int)blockIdx.x / wpr; r < d.n_rep; r += G / wpr) {
        const RepDev& rep = d.reps[r];
        if (rep.n_mono == 0) continue;
        const RepCtl* c = d.ctl + r;
        const double sh[3] = {c->shift[0], c->shift[1], c->shift[2]};
        const uint32_t k = c->k_cur;
        const bool inj = c->inj_active != 0, redo = c->redo != 0;
        const uint32_t hs = mono_hash_size(rep.n_mono);
        mono_cells(d, r, sh, sub == 0, hs, s_mkey, s_mcnt);
        if (sub == 0)  /* the row tiles holding monomers, tagged for the collide */
            for (uint32_t q = threadIdx.x; q < hs; q += 256) {
                const uint32_t gc = s_mkey[q];
                if (gc == kHashEmpty) continue;
                uint32_t t, lc;
                tile_of_cell(rep, gc - (uint32_t)rep.cell_base, t, lc);
                d.ttag[rep.rt_base + t] = k + 1u;
            }
        if (redo) continue;  /* all its cells are drawn below */
        for (uint32_t q = threadIdx.x; q < hs; q += 256) {
            const uint32_t gc = s_mkey[q];
            if (gc == kHashEmpty || (int)(((gc * 0x9E3779B1u) >> 8) % (uint32_t)wpr) != sub) continue;
            /* the first attempt's words drawn while the solvent count loads */
            const uint32_t nf = d.cg_n[gc];
            uint32_t w0[4];
            draw4(rep.seed, rep.replica, TAG_GAMMA, gc - (uint32_t)rep.cell_base, k, 0, w0);
            const uint32_t n = nf + s_mcnt[q];
            d.cg_n[gc] = n;
            d.cg_g[gc] = n >= 2 && !inj ? gamma_fast_w0(rep.seed, rep.replica, gc - (uint32_t)rep.cell_base, k, w0,
                                                        1.5 * (n - 1), rep.kT)
                                        : 0.0;
        }
    }
    if (!any_redo) return;
    for (int r = 0; r < d.n_rep; ++r) {
        if (!d.ctl[r].redo) continue;
        const RepDev& rep = d.reps[r];
        RepCtl c;
        sort_ctl(d, r, kSortNormal, c);
        const uint32_t nb = d.pblk_base[r + 1] - d.pblk_base[r];
        for (uint32_t q = blockIdx.x; q < nb; q += gridDim.x) {
            const uint64_t p0 = rep.part_base + (uint64_t)q * (256 * PPT);
            const uint64_t p1 = min(p0 + 256 * PPT, rep.part_base + rep.n_fluid);
            bool over = false;
            /* at most 2 particles per thread at a time: the registers of this
             * kernel are those of its common path */
            constexpr int RP = PPT > 2 ? 2 : PPT;
            for (uint64_t s0 = p0; s0 < p1; s0 += 256 * RP)
                bin_block<256, RP>(d, r, s0, min(s0 + 256 * RP, p1), c, h, d.bcount_r, d.bucket_r, d.blc_r, s_key,
                                   s_cnt, over);
        }
    }
    /* grid barrier: every workgroup's rebinning stored */
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(d.grid_bar, 1u);
        uint32_t spins = 0;
        while (__hip_atomic_load(d.grid_bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < (uint32_t)G) {
            if (++spins > (1u << 22)) { atomicOr(&d.ctl[0].status, ST_SCAN_TIMEOUT); break; }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    for (int r = 0; r < d.n_rep; ++r) {
        if (!d.ctl[r].redo) continue;
        const RepDev& rep = d.reps[r];
        const uint32_t hs = mono_hash_size(rep.n_mono);
        if (rep.n_mono) {
            const double sh[3] = {d.ctl[r].shift[0], d.ctl[r].shift[1], d.ctl[r].shift[2]};
            mono_cells(d, r, sh, false, hs, s_mkey, s_mcnt);
        }
        const uint32_t t0 = d.rtile_base[r], nt = d.rtile_base[r + 1] - t0;
        for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
            const uint32_t lt = t0 + t;
            const uint32_t n = load_count(d.bcount_r + lt);
            tile_gammas<256>(d, r, lt, n, d.blc_r + (size_t)lt * (uint32_t)d.rt_cap, s_h, hs,
                             rep.n_mono ? s_mkey : nullptr, s_mcnt);
        }
    }
}

/* Row-window collide: one workgroup of NT threads per row tile of the new
 * cells (at most TCM cells, CAP particles).  The arithmetic of every
 * particle and cell is k_collide's, expression for expression, with the
 * cell sums as 64-bit fixed point (so their order does not matter): the
 * two paths give the same bits.  Phases:
 *   (a) the tile's count and first slot; its monomers (as k_collide)
 *   (b) its bucket: particle indices, the particles gathered, streamed (as
 *       k_advect), their cell and rank in the cell (LDS atomic)
 *   (c) exclusive scan of the cell counts (the rotations were drawn in (a))
 *   (d) per particle: position (and id) to the new buffer at its slot,
 *       velocity and image index staged in LDS at the slot, its momentum
 *       added to the cell's sums (LDS atomics); per cell the Gamma variate
 *       drawn ahead; the monomers' momenta
 *   (e) per cell: CMV (with angular momentum CMP); per particle and
 *       monomer: relative kinetic energy (and the inertia tensor and C)
 *       added to the cell's sums; per cell: the MBS scale (and omega)
 *   (f) per slot: v <- CMV + s R (v - CMV) (+ omega x r), written coalesced;
 *       cell offsets of the new buffer; monomer velocities; the bucket
 *       counts of the next collision cleared. */
template <int NT, int CAP, int TCM, bool AMC>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(AMC ? 1 : 6)))
void k_collide_rows(DevState d, double h) {
    static_assert(TCM <= NT && NT % TCM == 0 && CAP % NT == 0, "tile shape");
    constexpr int MCAP = 2 * TCM + 64;
    constexpr int PER = CAP / NT;
    constexpr int NW = NT / 64;
    constexpr int TA = AMC ? TCM : 1;
    constexpr int CA = AMC ? CAP : 1;
    __shared__ float s_vx[CAP], s_vy[CAP], s_vz[CAP];
    __shared__ float s_lx[CA], s_ly[CA], s_lz[CA];
    __shared__ int8_t s_cy[CAP];
    __shared__ uint8_t s_lcs[CAP];
    __shared__ uint32_t s_cnt[TCM], s_off[TCM + 1], s_nm[TCM];
    __shared__ unsigned long long s_p[3][TCM];   /* monomers: m v (3) */
    __shared__ unsigned long long s_mm[TCM];     /* monomer mass */
    __shared__ unsigned long long s_me[TCM];     /* monomer relative kinetic energy */
    __shared__ double s_cmv[TCM][3], s_g[TCM];
    __shared__ float s_R[TCM][9];
    __shared__ uint32_t s_ml[MCAP];
    __shared__ uint32_t s_wsum[NW];
    __shared__ uint32_t s_mn, s_bad, s_pre;
    __shared__ unsigned long long s_am[12][TA];
    __shared__ double s_cmp[TA][3], s_w[TA][3];

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t lt = xcd_remap(blockIdx.x, gridDim.x);
    const int r = find_rep(d.rtile_base, d.n_rep, lt);
    const RepDev& rep = d.reps[r];
    RepCtl* ctl = d.ctl + r;
    RepCtl cc;
    sort_ctl(d, r, kSortNormal, cc);
    Box b = rep.box;
    b.delrx = cc.delrx;
    const uint32_t t = lt - rep.rt_base;
    const RowTile T = row_tile(rep, t);
    const uint32_t nc = T.nc;
    const uint32_t gc0 = (uint32_t)rep.cell_base + T.c_lo;
    const int cur = d.cur, alt = cur ^ 1;
    const bool rd = ctl->redo != 0;
    const uint32_t* cnt = rd ? d.bcount_r : d.bcount[cur];
    const uint32_t* bk = (rd ? d.bucket_r : d.bucket) + (size_t)lt * (uint32_t)d.rt_cap;
    const uint32_t pr0 = (uint32_t)rep.part_base, pr1 = pr0 + rep.n_fluid;
    const float *X = d.x[cur], *Y = d.y[cur], *Z = d.z[cur];
    const float *VX = d.vx[cur], *VY = d.vy[cur], *VZ = d.vz[cur];
    float *OX = d.x[alt], *OY = d.y[alt], *OZ = d.z[alt];
    float *OVX = d.vx[alt], *OVY = d.vy[alt], *OVZ = d.vz[alt];
    const uint32_t* ids = d.ids[cur];
    uint32_t* oids = d.ids[alt];
    const double shear = rep.box.shear, LyD = rep.box.L[1];
    const bool amc = AMC && rep.amc;

    /* (a) */
    const uint32_t n = load_count(cnt + lt);
    const uint32_t tb0 = d.tbase[lt];                /* the tile's first slot in the new buffer */
    const uint32_t tg = uniform_u32(d.ttag[lt]);     /* tagged by the redo kernel: holds monomers */
    /* the bucket's indices, issued beside the count (the tile's CAP
     * entries are all in bounds; those past the count are dropped below,
     * before any is used), and this lane's cell's drawn variate and count:
     * no later phase waits on a load these could have started */
    uint32_t my_i[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) my_i[u] = bk[tid + u * NT];
    uint32_t pre_n = 0u;
    double pre_g = 0.0;
    if ((uint32_t)tid < nc) {
        pre_n = d.cg_n[gc0 + tid];
        pre_g = d.cg_g[gc0 + tid];
    }
    for (uint32_t i = tid; i < nc; i += NT) {
        s_cnt[i] = 0u;
        s_nm[i] = 0u;
        s_p[0][i] = s_p[1][i] = s_p[2][i] = 0ull;
        s_mm[i] = 0ull;
        s_me[i] = 0ull;
    }
    if (amc)
        for (uint32_t i = tid; i < 12u * nc; i += NT) s_am[i / nc][i % nc] = 0ull;
    if (tid == 0) { s_mn = 0u; s_bad = 0u; }
    if (rd) {  /* recentred replica (rare): the redo counts of all earlier tiles */
        uint32_t pre = 0u;
        for (uint32_t q = tid; q < t; q += NT) pre += load_count(cnt + rep.rt_base + q);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o);
        if (lane == 0) s_wsum[wv] = pre;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t p = 0u;
        if (rd)
            for (int q = 0; q < NW; ++q) p += s_wsum[q];
        s_pre = rd ? pr0 + p : tb0;
    }
    /* the replica's monomers in this tile's cells, if the redo kernel
     * tagged the tile as holding any this collision */
    const bool tagged = tg == ctl->k_cur + 1u;
    for (uint32_t j = tid; tagged && j < rep.n_mono; j += NT) {
        const uint32_t rel = d.mcell[rep.mono_base + j] - gc0;
        if (rel < nc) {
            const uint32_t slot = atomicAdd(&s_mn, 1u);
            if (slot < (uint32_t)MCAP) { s_ml[slot] = (rel << 16) | j; atomicAdd(&s_nm[rel], 1u); }
        }
    }
    const uint32_t k = ctl->k_cur;
    const bool inj = ctl->inj_active != 0;
    /* per cell: the rotation about a random axis (MPC::collide MPC.cpp:241-257) */
    for (uint32_t ci = tid; ci < nc; ci += NT) {
        const uint32_t cell = T.c_lo + ci;
        double u1, u2;
        if (inj) {
            u1 = d.inj_axis[2 * (size_t)(gc0 + ci)];
            u2 = d.inj_axis[2 * (size_t)(gc0 + ci) + 1];
        } else {
            uint32_t w[4];
            draw4(rep.seed, rep.replica, TAG_AXIS, cell, k, 0, w);
            u1 = u01(w[0]);
            u2 = u01(w[1]);
        }
        rotation_f32(u1, u2, rep.c, rep.s, s_R[ci]);
    }
    /* uniform exits, status set: the tile cannot hold its particles */
    if (n > (uint32_t)CAP) {
        if (tid == 0) atomicOr(&ctl->status, ST_TILE_OVERFLOW);
        return;
    }

    /* (b) */
    float fx[PER], fy[PER], fz[PER], fvx[PER], fvy[PER], fvz[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const uint32_t e = (uint32_t)(tid + u * NT);
        if (e >= n || my_i[u] < pr0 || my_i[u] >= pr1) my_i[u] = ~0u;  /* the tile's, this replica's */
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        if (my_i[u] == ~0u) continue;
        const uint32_t i = my_i[u];
        fx[u] = X[i]; fy[u] = Y[i]; fz[u] = Z[i];
        fvx[u] = VX[i]; fvy[u] = VY[i]; fvz[u] = VZ[i];
    }
    __syncthreads();  /* cell counters cleared, monomers listed */
    uint32_t my_c[PER];  /* local cell << 16 | rank; image index and local coordinates */
    int my_cy[PER];
    float my_l[AMC ? PER : 1][3];
    bool bad = false;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        my_c[u] = ~0u;
        if (my_i[u] == ~0u) continue;
        double x = fx[u], y = fy[u], z = fz[u], vx = fvx[u], vy = fvy[u], vz = fvz[u];
        recenter_stream(b, h, cc, x, y, z, vx, vy, vz);
        fx[u] = (float)x; fy[u] = (float)y; fz[u] = (float)z;
        fvx[u] = (float)vx; fvy[u] = (float)vy; fvz[u] = (float)vz;
        double px = x + cc.shift[0], py = y + cc.shift[1], pz = z + cc.shift[2], vimg = vx;
        my_cy[u] = (int)le_wrap(b, px, py, pz, vimg);
        const uint32_t lc = cell_index(b, px, py, pz) - T.c_lo;
        if (lc >= nc) { bad = true; continue; }
        my_c[u] = (lc << 16) | atomicAdd(&s_cnt[lc], 1u);
        if (AMC) {  /* coordinates inside the cell, as cell_index truncates */
            const double qx = px + b.half[0], qy = py + b.half[1], qz = pz + b.half[2];
            my_l[AMC ? u : 0][0] = (float)(qx - (double)(int)qx);
            my_l[AMC ? u : 0][1] = (float)(qy - (double)(int)qy);
            my_l[AMC ? u : 0][2] = (float)(qz - (double)(int)qz);
        }
    }
    if (bad) s_bad = 1u;
    __syncthreads();

    /* (c) */
    {
        const uint32_t v = (uint32_t)tid < nc ? s_cnt[tid] : 0u;
        uint32_t inc = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(inc, o);
            if (lane >= o) inc += u;
        }
        if (lane == 63) s_wsum[wv] = inc;
        __syncthreads();
        uint32_t wpre = 0u;
        for (int q = 0; q < wv; ++q) wpre += s_wsum[q];
        if ((uint32_t)tid < nc) s_off[tid] = wpre + inc - v;
        if ((uint32_t)tid == nc - 1u) s_off[nc] = wpre + inc;
    }
    __syncthreads();  /* every cell's offset before any particle takes its slot */
    const uint32_t base = s_pre;
    const uint32_t nmt = s_mn < (uint32_t)MCAP ? s_mn : (uint32_t)MCAP;
    if (tid == 0) {
        if (s_bad) atomicOr(&ctl->status, ST_SORT_MISMATCH);
        if (s_mn > (uint32_t)MCAP) atomicOr(&ctl->status, ST_TILE_MONO);
    }
    /* the counts would place the tile beyond its replica: status set, exit */
    if (base < pr0 || (uint64_t)base + n > (uint64_t)pr1) {
        if (tid == 0) atomicOr(&ctl->status, ST_SORT_MISMATCH);
        return;
    }
    /* (d) per particle: position (and id) to the new buffer at its slot,
     * velocity and image index staged in LDS at the slot; its momentum
     * (MPC::calculateCMV MPC.cpp:314-322, the image offset taken off) and,
     * with angular momentum, its coordinates inside the cell added to the
     * cell's fixed-point sums (LDS atomics: particles arrive in bucket
     * order, so a wave's adds spread over many cells) */
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        if (my_c[u] == ~0u) continue;
        const uint32_t lc = my_c[u] >> 16;
        const uint32_t slot = s_off[lc] + (my_c[u] & 0xFFFFu);
        if (slot >= n) {  /* cannot happen: the counts and ranks agree */
            atomicOr(&ctl->status, ST_SORT_MISMATCH);
            my_c[u] = ~0u;
            continue;
        }
        OX[base + slot] = fx[u]; OY[base + slot] = fy[u]; OZ[base + slot] = fz[u];
        if (ids) oids[base + slot] = ids[my_i[u]];
        s_vx[slot] = fvx[u]; s_vy[slot] = fvy[u]; s_vz[slot] = fvz[u];
        s_cy[slot] = (int8_t)my_cy[u];
        s_lcs[slot] = (uint8_t)lc;
        if (AMC) {
            s_lx[AMC ? slot : 0] = my_l[AMC ? u : 0][0];
            s_ly[AMC ? slot : 0] = my_l[AMC ? u : 0][1];
            s_lz[AMC ? slot : 0] = my_l[AMC ? u : 0][2];
        }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        if (my_c[u] == ~0u) continue;
        const uint32_t lc = my_c[u] >> 16;
        atomicAdd(&s_p[0][lc], to_fix((double)fvx[u] - (double)my_cy[u] * shear * LyD));
        atomicAdd(&s_p[1][lc], to_fix((double)fvy[u]));
        atomicAdd(&s_p[2][lc], to_fix((double)fvz[u]));
        if (amc) {
            atomicAdd(&s_am[0][lc], to_fix(1.0 * (double)my_l[AMC ? u : 0][0]));
            atomicAdd(&s_am[1][lc], to_fix(1.0 * (double)my_l[AMC ? u : 0][1]));
            atomicAdd(&s_am[2][lc], to_fix(1.0 * (double)my_l[AMC ? u : 0][2]));
        }
    }
    /* per cell, one lane each: the Gamma variate, which depends only on the
     * member count, drawn ahead for it (k_cell_gamma, the redo kernel); a
     * count it was not drawn for is an internal error, reported */
    const bool drawn = ctl->cg_k == k + 1u;
    bool miss = false;
    if ((uint32_t)tid < nc) {  /* TCM <= NT: one cell a lane */
        const uint32_t c = tid;
        const uint32_t nn = s_cnt[c] + s_nm[c];
        double g = 0.0;
        if (nn >= 2) {
            if (inj) {
                g = d.inj_gamma[gc0 + c];
            } else {
                miss |= !drawn || pre_n != nn;
                g = pre_g;
            }
        }
        s_g[c] = g;
    }
    if (miss) atomicOr(&ctl->status, ST_GAMMA_MISS);
    /* a monomer's coordinates inside its cell (shifted, never wrapped) */
    auto mono_local = [&](uint32_t mi, double& lx, double& ly, double& lz) {
        const double qx = d.mpos[3 * mi] + cc.shift[0] + b.half[0];
        const double qy = d.mpos[3 * mi + 1] + cc.shift[1] + b.half[1];
        const double qz = d.mpos[3 * mi + 2] + cc.shift[2] + b.half[2];
        lx = qx - (double)(int)qx; ly = qy - (double)(int)qy; lz = qz - (double)(int)qz;
    };
    /* MPC::calculateCMV (MPC.cpp:314-322): the monomers' m v and m */
    for (uint32_t j = tid; j < nmt; j += NT) {
        const uint32_t e = s_ml[j];
        const uint32_t lc = e >> 16;
        const uint32_t mi = rep.mono_base + (e & 0xFFFFu);
        const double mm = d.mmass[mi];
        atomicAdd(&s_p[0][lc], to_fix(d.mvel[3 * mi] * mm));
        atomicAdd(&s_p[1][lc], to_fix(d.mvel[3 * mi + 1] * mm));
        atomicAdd(&s_p[2][lc], to_fix(d.mvel[3 * mi + 2] * mm));
        atomicAdd(&s_mm[lc], to_fix(mm));
        if (amc) {
            double lx, ly, lz;
            mono_local(mi, lx, ly, lz);
            atomicAdd(&s_am[0][lc], to_fix(mm * lx));
            atomicAdd(&s_am[1][lc], to_fix(mm * ly));
            atomicAdd(&s_am[2][lc], to_fix(mm * lz));
        }
    }
    __syncthreads();

    /* (e) per cell, one lane each: CMV (and with angular momentum CMP,
     * MPC::calculateCMP MPC.cpp:324-333, inside the cell) from the sums */
    for (uint32_t ci = tid; ci < nc; ci += NT) {
        const double M = (double)s_cnt[ci] + from_fix(s_mm[ci]);
        double c0 = from_fix(s_p[0][ci]), c1 = from_fix(s_p[1][ci]), c2 = from_fix(s_p[2][ci]);
        if (M > 0) { c0 /= M; c1 /= M; c2 /= M; }
        s_cmv[ci][0] = c0; s_cmv[ci][1] = c1; s_cmv[ci][2] = c2;
        if (amc) {
            const double inv = M > 0 ? 1.0 / M : 0.0;
            s_cmp[ci][0] = from_fix(s_am[0][ci]) * inv;
            s_cmp[ci][1] = from_fix(s_am[1][ci]) * inv;
            s_cmp[ci][2] = from_fix(s_am[2][ci]) * inv;
        }
    }
    __syncthreads();
    /* the angular-momentum terms of one member (MPC::calculateAngular
     * MPC.cpp:344-372): I about CMP and C = m r x (u - R u) */
    auto amc_terms = [&](uint32_t lc, double m, double lx, double ly, double lz, double ux, double uy, double uz,
                         unsigned long long* acc) {
        const double rx = lx - s_cmp[lc][0], ry = ly - s_cmp[lc][1], rz = lz - s_cmp[lc][2];
        const float* R = s_R[lc];
        const double qx = ux - (((double)R[0] * ux + (double)R[1] * uy) + (double)R[2] * uz);
        const double qy = uy - (((double)R[3] * ux + (double)R[4] * uy) + (double)R[5] * uz);
        const double qz = uz - (((double)R[6] * ux + (double)R[7] * uy) + (double)R[8] * uz);
        acc[0] = to_fix(m * (ry * ry + rz * rz));
        acc[1] = to_fix(m * (rx * rx + rz * rz));
        acc[2] = to_fix(m * (ry * ry + rx * rx));
        acc[3] = to_fix(-m * rx * ry);
        acc[4] = to_fix(-m * rx * rz);
        acc[5] = to_fix(-m * ry * rz);
        acc[6] = to_fix(m * (ry * qz - rz * qy));
        acc[7] = to_fix(m * (rz * qx - rx * qz));
        acc[8] = to_fix(m * (rx * qy - ry * qx));
    };
    /* per particle (its own, bucket order, read back from its slot): the
     * relative kinetic energy of the MBS thermostat (MPC.cpp:277-279) and
     * the angular-momentum terms */
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        if (my_c[u] == ~0u) continue;
        const uint32_t lc = my_c[u] >> 16;
        const uint32_t q = s_off[lc] + (my_c[u] & 0xFFFFu);
        const double vx = (double)s_vx[q] - (double)s_cy[q] * shear * LyD;
        const double dx = vx - s_cmv[lc][0], dy = (double)s_vy[q] - s_cmv[lc][1], dz = (double)s_vz[q] - s_cmv[lc][2];
        atomicAdd(&s_me[lc], to_fix((dx * dx + dy * dy) + dz * dz));
        if (amc) {
            unsigned long long t9[9];
            amc_terms(lc, 1.0, s_lx[AMC ? q : 0], s_ly[AMC ? q : 0], s_lz[AMC ? q : 0], dx, dy, dz, t9);
#pragma unroll
            for (int j = 0; j < 9; ++j) atomicAdd(&s_am[3 + j][lc], t9[j]);
        }
    }
    /* the monomers' energies and terms, against the same CMV */
    for (uint32_t j = tid; j < nmt; j += NT) {
        const uint32_t me = s_ml[j];
        const uint32_t lc = me >> 16;
        const uint32_t mi = rep.mono_base + (me & 0xFFFFu);
        const double dx = d.mvel[3 * mi] - s_cmv[lc][0];
        const double dy = d.mvel[3 * mi + 1] - s_cmv[lc][1];
        const double dz = d.mvel[3 * mi + 2] - s_cmv[lc][2];
        atomicAdd(&s_me[lc], to_fix(d.mmass[mi] * ((dx * dx + dy * dy) + dz * dz)));
        if (amc) {
            double lx, ly, lz;
            mono_local(mi, lx, ly, lz);
            unsigned long long t9[9];
            amc_terms(lc, d.mmass[mi], lx, ly, lz, dx, dy, dz, t9);
            for (int q = 0; q < 9; ++q) atomicAdd(&s_am[3 + q][lc], t9[q]);
        }
    }
    __syncthreads();
    /* MBS scale sqrt(Gamma / E_rel) for n >= 2 (MPC.cpp:287-289); with the
     * angular-momentum flag, omega = s I^-1 C for n > 2 (MPC.cpp:260, 373) */
    for (uint32_t ci = tid; ci < nc; ci += NT) {
        const uint32_t nn = s_cnt[ci] + s_nm[ci];
        const double sc = nn >= 2 ? sqrt(s_g[ci] / (from_fix(s_me[ci]) * 0.5)) : -1.0;
        s_g[ci] = sc;
        if (amc) {
            double w[3] = {0.0, 0.0, 0.0};
            if (nn > 2) {
                unsigned long long A6[9];
#pragma unroll
                for (int j = 0; j < 9; ++j) A6[j] = s_am[3 + j][ci];
                const double A[3][3] = {{from_fix(A6[0]), from_fix(A6[3]), from_fix(A6[4])},
                                        {from_fix(A6[3]), from_fix(A6[1]), from_fix(A6[5])},
                                        {from_fix(A6[4]), from_fix(A6[5]), from_fix(A6[2])}};
                const double c3[3] = {from_fix(A6[6]), from_fix(A6[7]), from_fix(A6[8])};
                llt_solve3(A, c3, w);
                const double scl = sc >= 0.0 ? sc : 1.0;
                const bool ok = isfinite(w[0]) && isfinite(w[1]) && isfinite(w[2]);
                for (int q = 0; q < 3; ++q) w[q] = ok ? w[q] * scl : 0.0;
            }
            s_w[ci][0] = w[0]; s_w[ci][1] = w[1]; s_w[ci][2] = w[2];
        }
    }
    __syncthreads();
    /* (f) v <- s v + (1-s) CMV, then v <- CMV + R (v - CMV) (MPC.cpp:290-292,
     * 266-267); the image offset is undone as the unshift wrap does */
    auto apply = [&](uint32_t lc, double& vx, double& vy, double& vz) {
        const double m0 = s_cmv[lc][0], m1 = s_cmv[lc][1], m2 = s_cmv[lc][2];
        const double sc = s_g[lc];
        if (sc >= 0.0) {
            vx = sc * vx + (1 - sc) * m0;
            vy = sc * vy + (1 - sc) * m1;
            vz = sc * vz + (1 - sc) * m2;
        }
        const float* R = s_R[lc];
        const double d0 = vx - m0, d1 = vy - m1, d2 = vz - m2;
        vx = m0 + (((double)R[0] * d0 + (double)R[1] * d1) + (double)R[2] * d2);
        vy = m1 + (((double)R[3] * d0 + (double)R[4] * d1) + (double)R[5] * d2);
        vz = m2 + (((double)R[6] * d0 + (double)R[7] * d1) + (double)R[8] * d2);
    };
    /* v += omega x (x - CMP) (MPC.cpp:268) */
    auto add_w = [&](uint32_t lc, double lx, double ly, double lz, double& vx, double& vy, double& vz) {
        const double rx = lx - s_cmp[lc][0], ry = ly - s_cmp[lc][1], rz = lz - s_cmp[lc][2];
        const double w0 = s_w[lc][0], w1 = s_w[lc][1], w2 = s_w[lc][2];
        vx += w1 * rz - w2 * ry;
        vy += w2 * rx - w0 * rz;
        vz += w0 * ry - w1 * rx;
    };
    for (uint32_t q = tid; q < n; q += NT) {
        const uint32_t lc = s_lcs[q];
        const double offv = (double)s_cy[q] * shear * LyD;
        double vx = (double)s_vx[q] - offv, vy = s_vy[q], vz = s_vz[q];
        apply(lc, vx, vy, vz);
        if (amc) add_w(lc, s_lx[AMC ? q : 0], s_ly[AMC ? q : 0], s_lz[AMC ? q : 0], vx, vy, vz);
        vx += offv;
        OVX[base + q] = (float)vx; OVY[base + q] = (float)vy; OVZ[base + q] = (float)vz;
    }
    uint32_t* offn = d.offs[alt];
    for (uint32_t c = tid; c < nc; c += NT) offn[gc0 + c] = base + s_off[c];
    if (tid == 0) {
        if (T.c_lo + nc == rep.ncells) offn[rep.cell_base + rep.ncells] = base + n;
        d.bcount[alt][lt] = 0u;
    }
    for (uint32_t j = tid; j < nmt; j += NT) {
        const uint32_t me = s_ml[j];
        const uint32_t lc = me >> 16;
        const uint32_t mi = rep.mono_base + (me & 0xFFFFu);
        double vx = d.mvel[3 * mi], vy = d.mvel[3 * mi + 1], vz = d.mvel[3 * mi + 2];
        apply(lc, vx, vy, vz);
        if (amc) {
            double lx, ly, lz;
            mono_local(mi, lx, ly, lz);
            add_w(lc, lx, ly, lz, vx, vy, vz);
        }
        d.mvel[3 * mi] = vx; d.mvel[3 * mi + 1] = vy; d.mvel[3 * mi + 2] = vz;
    }
}

/* ======================================================== fluid init */

/* MPC::initialize (MPC.cpp:34-55): uniform positions and U(-1/2,1/2)
 * velocities from Philox, LE wrap, then the mean velocity removed. */
__device__ __forceinline__ void init_particle(const RepDev& rep, const Box& b, uint32_t li, double& x, double& y,
                                              double& z, double& vx, double& vy, double& vz) {
    uint32_t w0[4], w1[4];
    draw4(rep.seed, rep.replica, TAG_FLUID_INIT, li, 0, 0, w0);
    draw4(rep.seed, rep.replica, TAG_FLUID_INIT, li, 0, 1, w1);
    x = b.L[0] * (u01(w0[0]) - 0.5); vx = u01(w1[0]) - 0.5;
    y = b.L[1] * (u01(w0[1]) - 0.5); vy = u01(w1[1]) - 0.5;
    z = b.L[2] * (u01(w0[2]) - 0.5); vz = u01(w1[2]) - 0.5;
    le_wrap(b, x, y, z, vx);
}

__global__ __launch_bounds__(256) void k_fluid_init(DevState d, int pass) {
    __shared__ double sred[4];
    const uint32_t lb = blockIdx.x;
    const int r = find_rep(d.pblk_base, d.n_rep, lb);
    const RepDev& rep = d.reps[r];
    Box b = rep.box;
    b.delrx = d.ctl[r].delrx;
    const uint32_t ppb = 256u * (uint32_t)d.ppt;
    const uint32_t l0 = (lb - d.pblk_base[r]) * ppb;
    const int cur = d.cur;
    double sv[3] = {0.0, 0.0, 0.0};
    const double* mean = d.red + 4ull * (d.n_pblk + r);
    for (uint32_t j = 0; j < (uint32_t)d.ppt; ++j) {
        const uint32_t li = l0 + j * 256u + threadIdx.x;
        if (li >= rep.n_fluid) continue;
        double x, y, z, vx, vy, vz;
        init_particle(rep, b, li, x, y, z, vx, vy, vz);
        const uint64_t i = rep.part_base + li;
        if (pass == 0) {
            sv[0] += vx; sv[1] += vy; sv[2] += vz;
        } else {
            d.x[cur][i] = (float)x; d.y[cur][i] = (float)y; d.z[cur][i] = (float)z;
            d.vx[cur][i] = (float)(vx - mean[0]);
            d.vy[cur][i] = (float)(vy - mean[1]);
            d.vz[cur][i] = (float)(vz - mean[2]);
            if (d.ids[cur]) d.ids[cur][i] = li;
        }
    }
    if (pass == 0) {
        for (int c = 0; c < 3; ++c) {
            const double t = block_sum<256>(sv[c], sred);
            if (threadIdx.x == 0) d.red[4ull * lb + c] = t;
        }
    }
}

/* per replica: sum of its block partials in block order, mean velocity */
__global__ __launch_bounds__(64) void k_fluid_mean(DevState d) {
    const int r = blockIdx.x;
    if (threadIdx.x != 0) return;
    const RepDev& rep = d.reps[r];
    double s[3] = {0.0, 0.0, 0.0};
    for (uint32_t bk = d.pblk_base[r]; bk < d.pblk_base[r + 1]; ++bk)
        for (int c = 0; c < 3; ++c) s[c] += d.red[4ull * bk + c];
    for (int c = 0; c < 3; ++c) d.red[4ull * (d.n_pblk + r) + c] = s[c] / (double)rep.n_fluid;
}

/* ======================================================= temperature */

/* MPC::calculateEkinInCell summed over cells (MPC.cpp:384-403) with the
 * filled-cell count (MPC.cpp:412-418), per collide tile; partials per tile
 * in red[2*tile] (ekin, filled). */
template <int TC>
__global__ __launch_bounds__(256) void k_temperature(DevState d, int replica) {
    __shared__ double sred[4];
    __shared__ unsigned long long s_q[4][TC];  /* monomer sum m v (3) and m per cell, fixed point */
    const uint32_t lt = d.tile_base[replica] + blockIdx.x;
    const RepDev& rep = d.reps[replica];
    const uint32_t c_lo = blockIdx.x * TC;
    const uint32_t gc0 = (uint32_t)rep.cell_base + c_lo;
    const uint32_t nc = min((uint32_t)TC, rep.ncells - c_lo);
    const int cur = d.cur;
    const float *VX = d.vx[cur], *VY = d.vy[cur], *VZ = d.vz[cur];
    for (uint32_t i = threadIdx.x; i < 4 * TC; i += 256) (&s_q[0][0])[i] = 0ull;
    __syncthreads();
    /* monomers in the cells of the last collision */
    for (uint32_t j = threadIdx.x; j < rep.n_mono; j += 256) {
        const uint32_t rel = d.mcell[rep.mono_base + j] - gc0;
        if (rel < nc) {
            const uint32_t mi = rep.mono_base + j;
            const double mm = d.mmass[mi];
            atomicAdd(&s_q[0][rel], to_fix(d.mvel[3 * mi] * mm));
            atomicAdd(&s_q[1][rel], to_fix(d.mvel[3 * mi + 1] * mm));
            atomicAdd(&s_q[2][rel], to_fix(d.mvel[3 * mi + 2] * mm));
            atomicAdd(&s_q[3][rel], to_fix(mm));
        }
    }
    __syncthreads();
    double ek = 0.0, filled = 0.0;
    for (uint32_t ci = threadIdx.x; ci < nc; ci += 256) {
        const uint32_t a = d.offs[d.cur][gc0 + ci], e = d.offs[d.cur][gc0 + ci + 1];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, M = 0.0;
        for (uint32_t p = a; p < e; ++p) { s0 += VX[p]; s1 += VY[p]; s2 += VZ[p]; M += 1.0; }
        s0 += from_fix(s_q[0][ci]); s1 += from_fix(s_q[1][ci]); s2 += from_fix(s_q[2][ci]);
        M += from_fix(s_q[3][ci]);
        if (M > 0) { s0 /= M; s1 /= M; s2 /= M; }
        double e2 = 0.0;
        for (uint32_t p = a; p < e; ++p) {
            const double dx = VX[p] - s0, dy = VY[p] - s1, dz = VZ[p] - s2;
            e2 += (dx * dx + dy * dy) + dz * dz;
        }
        ek += e2 * 0.5;
        filled += e > a ? 1.0 : 0.0;
    }
    const double te = block_sum<256>(ek, sred);
    const double tf = block_sum<256>(filled, sred);
    if (threadIdx.x == 0) { d.red[2ull * lt] = te; d.red[2ull * lt + 1] = tf; }
}

/* =================================================== learned closure */

/* Features of the closure network, one workgroup per replica (its star),
 * written in global monomer order, 24 per monomer:
 *   0-2  r = x - x_com                 3-5  v - v_com
 *   6-8  u = v - (shear y, 0, 0)       (velocity relative to the imposed flow)
 *   9-11 bond vector to the neighbour nearer the core, 12-14 to the one
 *        farther out (zero where absent; the core's are zero)
 *   15   monomers within r < 1.5       16 B type   17 core (degree > 2)
 *   18   shear rate
 *   19-21 mean velocity of the monomers within r < 1.5 minus v (zero if none)
 *   22   1/m           23 bonds of the monomer
 * With fmin/fmax, a monomer with any feature outside its trained range
 * widened by 25% counts in RepCtl::closure_oor.
 *
 * Neighbours (15, 19-21) come from cells of side kNbSide (1.5 and a
 * margin that keeps every pair the FP32 distance test accepts in adjacent
 * cells) around the star's centre, folded onto a periodic 16^3 grid of
 * buckets: a counting sort of the monomers by bucket into LDS copies of
 * their positions and velocities, then, per monomer, 9 slot ranges (the
 * three x-neighbour buckets of a row are adjacent unless the row wraps).
 * Cells 16 apart share a bucket; their monomers are >= 22.5 apart and fail
 * the distance test, and the 27 cells around a monomer never share one, so
 * nothing is counted twice.  The velocity sums are 64-bit fixed point, so
 * the order inside a bucket (set by LDS atomics) does not change the
 * result.  Thread t owns monomers t + kClThreads k: their cells and sorted
 * slots stay in its registers between the phases. */
constexpr int kNbHash = 4096;
constexpr float kNbSide = 1.5f * 1.001f;
constexpr int kClThreads = 1024;
constexpr int kClPer = kMaxMono / kClThreads;
constexpr size_t kClLds = (size_t)kMaxMono * (16 + 16 + 16) + (size_t)(kNbHash + 1) * 4;

__device__ __forceinline__ uint32_t nb_bucket(int cx, int cy, int cz) {
    return (uint32_t)(cx & 15) | ((uint32_t)(cy & 15) << 4) | ((uint32_t)(cz & 15) << 8);
}

/* The six centre-of-mass sums of the workgroup at once: each is the fixed
 * order block_sum takes (wave sums, then the waves' partials in order), so
 * the values are block_sum's bit for bit.  Thread c < 6 adds up sum c;
 * scratch holds 6 (NT / 64 + 1) doubles. */
template <int NT>
__device__ void block_sum6(double (&v)[6], double* scratch) {
    constexpr int W = NT / 64;
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] = wave_sum(v[c]);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < 6; ++c) scratch[c * W + w] = v[c];
    __syncthreads();
    if (threadIdx.x < 6) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < W; ++i) t += scratch[threadIdx.x * W + i];
        scratch[6 * W + threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] = scratch[6 * W + c];
}

__global__ __launch_bounds__(kClThreads) void k_closure_features(DevState d, float* feat, const float* fmin,
                                                                  const float* fmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NT = kClThreads;
    const int r = blockIdx.x, tid = threadIdx.x;
    const RepDev& rep = d.reps[r];
    const int N = (int)rep.n_mono;
    const uint32_t mb = rep.mono_base;
    float4* spos = reinterpret_cast<float4*>(smem);              /* N, bucket order: position, key */
    float4* svel = spos + kMaxMono;                              /* N, bucket order: velocity */
    uint32_t* prv = reinterpret_cast<uint32_t*>(svel + kMaxMono);  /* N */
    uint32_t* nxt = prv + kMaxMono;                              /* N */
    uint32_t* deg = nxt + kMaxMono;                              /* N */
    uint32_t* smap = deg + kMaxMono;                             /* N: each monomer's sorted slot */
    uint32_t* hb = smap + kMaxMono;                              /* kNbHash + 1: bucket ends */
    __shared__ double sred[6 * (NT / 64) + 6];
    __shared__ float sbound[2 * kClosureFeatures];  /* trained range widened by 25%: low, high */
    /* each owned monomer's state, loaded once */
    double x[kClPer], y[kClPer], z[kClPer], vx[kClPer], vy[kClPer], vz[kClPer];
    float im[kClPer], bt[kClPer];
    int cx[kClPer], cy[kClPer], cz[kClPer];
    uint32_t slot[kClPer];
    double cs[6] = {0, 0, 0, 0, 0, 0};
    for (int c = tid; c <= kNbHash; c += NT) hb[c] = 0u;
    if (fmin && tid < kClosureFeatures) {
        const float lo = fmin[tid], hi = fmax[tid], pad = 0.25f * (hi - lo);
        sbound[tid] = lo - pad;
        sbound[kClosureFeatures + tid] = hi + pad;
    }
#pragma unroll
    for (int k = 0; k < kClPer; ++k) {
        const int i = tid + k * NT;
        x[k] = y[k] = z[k] = vx[k] = vy[k] = vz[k] = 0.0;
        im[k] = bt[k] = 0.0f;
        if (i < N) {
            const size_t g = 3 * (size_t)(mb + i);
            x[k] = d.mpos[g]; y[k] = d.mpos[g + 1]; z[k] = d.mpos[g + 2];
            vx[k] = d.mvel[g]; vy[k] = d.mvel[g + 1]; vz[k] = d.mvel[g + 2];
            im[k] = (float)(1.0 / d.mmass[mb + i]);
            bt[k] = d.mbtype[mb + i] ? 1.0f : 0.0f;
            cs[0] += x[k]; cs[1] += y[k]; cs[2] += z[k];
            cs[3] += vx[k]; cs[4] += vy[k]; cs[5] += vz[k];
            prv[i] = (uint32_t)i; nxt[i] = (uint32_t)i; deg[i] = 0u;
        }
    }
    double com[6] = {cs[0], cs[1], cs[2], cs[3], cs[4], cs[5]};
    block_sum6<NT>(com, sred);
#pragma unroll
    for (int c = 0; c < 6; ++c) com[c] /= (double)(N > 0 ? N : 1);
    const float ox = (float)com[0], oy = (float)com[1], oz = (float)com[2], inv = 1.0f / kNbSide;
    for (uint32_t q = tid; q < rep.n_bond; q += NT) {
        const uint32_t e = d.bonds[rep.bond_base + q];
        const uint32_t i = e & 0xFFF, j = (e >> 12) & 0xFFF;  /* i < j: i nearer the core */
        atomicAdd(&deg[i], 1u); atomicAdd(&deg[j], 1u);
        prv[j] = i;
        nxt[i] = j;
    }
#pragma unroll
    for (int k = 0; k < kClPer; ++k)
        if (tid + k * NT < N) {
            cx[k] = (int)floorf(((float)x[k] - ox) * inv);
            cy[k] = (int)floorf(((float)y[k] - oy) * inv);
            cz[k] = (int)floorf(((float)z[k] - oz) * inv);
            atomicAdd(&hb[nb_bucket(cx[k], cy[k], cz[k])], 1u);
        }
    __syncthreads();
    block_scan_u32<NT>(hb, kNbHash, sred);
#pragma unroll
    for (int k = 0; k < kClPer; ++k) {
        const int i = tid + k * NT;
        if (i < N) {
            slot[k] = atomicAdd(&hb[nb_bucket(cx[k], cy[k], cz[k])], 1u);
            spos[slot[k]] = make_float4((float)x[k], (float)y[k], (float)z[k], 0.0f);
            svel[slot[k]] = make_float4((float)vx[k], (float)vy[k], (float)vz[k], 0.0f);
            smap[i] = slot[k];
        }
    }
    __syncthreads();  /* hb[b] now ends bucket b */
    const double shear = rep.box.shear;
    uint32_t oor = 0;
    static_assert(kClPer == 2, "two monomers per thread");
    /* the loop stays rolled (its body is large); the owned values are
     * selected, never indexed by k, so they stay in registers */
#define SEL(a) (k ? a[1] : a[0])
#pragma unroll 1
    for (int k = 0; k < kClPer; ++k) {
        const int i = tid + k * NT;
        if (i >= N) continue;
        const uint32_t mys = SEL(slot);
        const double X = SEL(x), Y = SEL(y), Z = SEL(z), VX = SEL(vx), VY = SEL(vy), VZ = SEL(vz);
        const float mx = (float)X, my = (float)Y, mz = (float)Z;
        const int qx = SEL(cx), qy = SEL(cy), qz = SEL(cz);
        int cnt = 0;
        unsigned long long nvx = 0ull, nvy = 0ull, nvz = 0ull;
        auto take = [&](uint32_t t, const float4& o) {
            const float dx = o.x - mx, dy = o.y - my, dz = o.z - mz;
            if (t != mys && __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) < 2.25f) {
                const float4 ov = svel[t];
                ++cnt;
                nvx += to_fix(ov.x); nvy += to_fix(ov.y); nvz += to_fix(ov.z);
            }
        };
        const uint32_t x0 = (uint32_t)(qx - 1) & 15u;
#pragma unroll 1
        for (int c = 0; c < 18; ++c) {
            /* row (dy, dz) = (c/2 % 3 - 1, c/6 - 1); part c%2: the run from
             * x0, then (if the row wraps) the run from bucket 0 */
            const uint32_t row = ((uint32_t)(qy + (c >> 1) % 3 - 1) & 15u) << 4 |
                                 ((uint32_t)(qz + c / 6 - 1) & 15u) << 8;
            uint32_t b0, b1;
            if (!(c & 1)) { b0 = row + x0; b1 = row + (x0 + 2 < 16u ? x0 + 2 : 15u); }
            else if (x0 + 2 >= 16u) { b0 = row; b1 = row + x0 + 2 - 16u; }
            else continue;
            const uint32_t s1 = hb[b1];
            uint32_t t = b0 ? hb[b0 - 1] : 0u;
            /* two candidates' loads in flight */
            for (; t + 1 < s1; t += 2) {
                const float4 o0 = spos[t], o1 = spos[t + 1];
                take(t, o0);
                take(t + 1, o1);
            }
            if (t < s1) take(t, spos[t]);
        }
        const bool core = deg[i] > 2u;
        float f[kClosureFeatures];
        f[0] = (float)(X - com[0]); f[1] = (float)(Y - com[1]); f[2] = (float)(Z - com[2]);
        f[3] = (float)(VX - com[3]); f[4] = (float)(VY - com[4]); f[5] = (float)(VZ - com[5]);
        const double ux = VX - shear * Y;
        f[6] = (float)ux; f[7] = (float)VY; f[8] = (float)VZ;
        /* the bonded neighbours' positions from their sorted LDS copies
         * (the same FP32 values as their FP64 positions rounded) */
        const float4 pp = spos[smap[core ? i : (int)prv[i]]], pn = spos[smap[core ? i : (int)nxt[i]]];
        f[9] = pp.x - mx; f[10] = pp.y - my; f[11] = pp.z - mz;
        f[12] = pn.x - mx; f[13] = pn.y - my; f[14] = pn.z - mz;
        f[15] = (float)cnt;
        f[16] = SEL(bt);
        f[17] = core ? 1.0f : 0.0f;
        f[18] = (float)shear;
        const float mvx = (float)VX, mvy = (float)VY, mvz = (float)VZ;
        const double ic = cnt ? 1.0 / (double)cnt : 0.0;
        f[19] = cnt ? (float)(from_fix(nvx) * ic) - mvx : 0.0f;
        f[20] = cnt ? (float)(from_fix(nvy) * ic) - mvy : 0.0f;
        f[21] = cnt ? (float)(from_fix(nvz) * ic) - mvz : 0.0f;
        f[22] = SEL(im);
        f[23] = (float)deg[i];
        /* six 16-byte stores per row (rows are 96 bytes, 16-byte aligned),
         * each group checked against the widened range as it goes (bounds
         * read from LDS through an opaque offset: hoisted, the 48 bounds
         * would hold registers through the neighbour scan) */
        float4* o = reinterpret_cast<float4*>(feat + (size_t)(mb + i) * kClosureFeatures);
        uint32_t bo = 0;
        asm volatile("" : "+v"(bo));
        bool out = false;
#pragma unroll
        for (int c = 0; c < kClosureFeatures / 4; ++c) {
            o[c] = make_float4(f[4 * c], f[4 * c + 1], f[4 * c + 2], f[4 * c + 3]);
            if (fmin)
#pragma unroll
                for (int j = 4 * c; j < 4 * c + 4; ++j)
                    out |= f[j] < sbound[bo + j] || f[j] > sbound[bo + kClosureFeatures + j];
        }
        oor += out ? 1u : 0u;
    }
#undef SEL
    if (fmin && oor) atomicAdd(&d.ctl[r].closure_oor, oor);
}

/* replica of global monomer m (per lane: a wave may span replicas) */
__device__ __forceinline__ int find_rep_mono(const DevState& d, uint32_t m) {
    int lo = 0, hi = d.n_rep;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (d.reps[mid].mono_base <= m) lo = mid; else hi = mid;
    }
    return lo;
}

/* The closure's solvent kick in place of the MPC collision: per monomer
 * v += mu + sigma xi, mu = out[0..2] sdy, sigma = exp(out[3..5] / 2) sdy,
 * xi standard normal from Philox (index = global monomer, counter k =
 * interval, tag TAG_CLOSURE) through Box-Muller. */
__global__ __launch_bounds__(256) void k_closure_kick(DevState d, const float* out, float sy0, float sy1, float sy2,
                                                       uint32_t k) {
    const uint32_t m = blockIdx.x * 256 + threadIdx.x;
    if (m >= d.m_tot) return;
    const int r = find_rep_mono(d, m);
    const RepDev& rep = d.reps[r];
    uint32_t w[4];
    draw4(rep.seed, rep.replica, TAG_CLOSURE, m - rep.mono_base, k, 0, w);
    /* Box-Muller normals in FP32 with the hardware transcendentals: noise,
     * whose few-ulp errors are far below anything the sampled distribution
     * can show; the velocity update itself stays FP64 */
    const float r0 = sqrtf(-2.0f * __logf((float)u01(w[0]))), r1 = sqrtf(-2.0f * __logf((float)u01(w[2])));
    const float a1 = 6.283185307f * (float)u01(w[1]), a3 = 6.283185307f * (float)u01(w[3]);
    const float xi[3] = {r0 * __cosf(a1), r0 * __sinf(a1), r1 * __cosf(a3)};
    const float* o = out + (size_t)m * 6;
    const float sy[3] = {sy0, sy1, sy2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float s = fminf(fmaxf(o[3 + c], -12.0f), 12.0f);
        d.mvel[3 * (size_t)m + c] += ((double)o[c] + (double)(__expf(0.5f * s) * xi[c])) * sy[c];
    }
}

/* ======================================================= observables */

/* Box::calculate_gyration_tensor (Box.cpp:330-352) and
 * Box::calculate_rotation_frequency (Box.cpp:354-379) of one replica's
 * star in its centre-of-mass frame (Box::center_of_mass_reference_frame,
 * Box.cpp:81-102): out[0..8] = G (row-major), out[9..11] = omega = I^-1 L,
 * out[12..17] = COM position and velocity (Box::COM_Pos, COM_Vel), out[18]
 * = kinetic energy in the COM frame. */
__global__ __launch_bounds__(256) void k_star_moments(DevState d, int replica, double* out) {
    __shared__ double sred[4];
    const RepDev& rep = d.reps[replica];
    const int N = (int)rep.n_mono;
    const uint32_t mb = rep.mono_base;
    double c[6] = {0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < N; i += 256) {
        const size_t g = 3 * (size_t)(mb + i);
        for (int q = 0; q < 3; ++q) { c[q] += d.mpos[g + q]; c[3 + q] += d.mvel[g + q]; }
    }
    double com[6];
    for (int q = 0; q < 6; ++q) com[q] = block_sum<256>(c[q], sred) / (double)(N > 0 ? N : 1);
    double m[15];  /* G (6: xx xy xz yy yz zz), I (6), L (3) */
    for (int q = 0; q < 15; ++q) m[q] = 0.0;
    for (int i = threadIdx.x; i < N; i += 256) {
        const size_t g = 3 * (size_t)(mb + i);
        const double x = d.mpos[g] - com[0], y = d.mpos[g + 1] - com[1], z = d.mpos[g + 2] - com[2];
        const double vx = d.mvel[g] - com[3], vy = d.mvel[g + 1] - com[4], vz = d.mvel[g + 2] - com[5];
        m[0] += x * x; m[1] += x * y; m[2] += x * z; m[3] += y * y; m[4] += y * z; m[5] += z * z;
        const double r2 = (x * x + y * y) + z * z;
        m[6] += r2 - x * x; m[7] -= x * y; m[8] -= x * z; m[9] += r2 - y * y; m[10] -= y * z; m[11] += r2 - z * z;
        m[12] += y * vz - z * vy; m[13] += z * vx - x * vz; m[14] += x * vy - y * vx;
    }
    for (int q = 0; q < 15; ++q) m[q] = block_sum<256>(m[q], sred);
    if (threadIdx.x == 0) {
        const double inv = 1.0 / (double)(N > 0 ? N : 1);
        const double G[9] = {m[0], m[1], m[2], m[1], m[3], m[4], m[2], m[4], m[5]};
        for (int q = 0; q < 9; ++q) out[q] = G[q] * inv;
        /* omega = I^-1 L by Cramer's rule (I symmetric positive definite for
         * a non-collinear star) */
        const double a = m[6], b = m[7], cc = m[8], e = m[9], f = m[10], k = m[11];
        const double det = a * (e * k - f * f) - b * (b * k - f * cc) + cc * (b * f - e * cc);
        const double L0 = m[12], L1 = m[13], L2 = m[14];
        const double w0 = (L0 * (e * k - f * f) - b * (L1 * k - f * L2) + cc * (L1 * f - e * L2)) / det;
        const double w1 = (a * (L1 * k - f * L2) - L0 * (b * k - f * cc) + cc * (b * L2 - L1 * cc)) / det;
        const double w2 = (a * (e * L2 - L1 * f) - b * (b * L2 - L1 * cc) + L0 * (b * f - e * cc)) / det;
        out[9] = w0; out[10] = w1; out[11] = w2;
        for (int q = 0; q < 6; ++q) out[12 + q] = com[q];
    }
    /* Molecule::calculate_Ekin Molecule.cpp:60-74: sum m |v - <v>|^2 / 2 */
    double ek = 0.0;
    for (int i = threadIdx.x; i < N; i += 256) {
        const size_t g = 3 * (size_t)(mb + i);
        const double vx = d.mvel[g] - com[3], vy = d.mvel[g + 1] - com[4], vz = d.mvel[g + 2] - com[5];
        ek += d.mmass[mb + i] * ((vx * vx + vy * vy) + vz * vz);
    }
    ek = block_sum<256>(ek, sred);
    if (threadIdx.x == 0) out[18] = ek / 2.0;
}

/* Box::calculate_patches_new Box.cpp:434-507 for one replica's star (arm a,
 * bead k at a*(A+B)+k, beads k > A solvophobic): two arms are bonded when
 * some pair of their B beads has a negative pair energy (Box::calculate_epot
 * Box.cpp:386-399 -- arms share no FENE bond); patches are the connected
 * groups of more than one arm.  out[0] = number of patches, out[1] = mean
 * arms per patch (0 without patches). */
__global__ __launch_bounds__(256) void k_star_patches(DevState d, int replica, int arms, int A, int B,
                                                      double* out) {
    __shared__ uint32_t s_bond[kPatchArms][kPatchArms / 32];
    __shared__ int s_lab[kPatchArms];
    __shared__ int s_stack[kPatchArms];
    const RepDev& rep = d.reps[replica];
    const double* P = d.mpos + 3ull * rep.mono_base;
    const double lambda = rep.lambda;
    for (int i = threadIdx.x; i < kPatchArms * (kPatchArms / 32); i += 256) (&s_bond[0][0])[i] = 0u;
    __syncthreads();
    const int per = A + B, npair = arms * (arms - 1) / 2;
    for (int q = threadIdx.x; q < npair; q += 256) {
        int i = 0, rem = q;
        while (rem >= arms - 1 - i) { rem -= arms - 1 - i; ++i; }
        const int k = i + 1 + rem;
        bool hit = false;
        for (int j = A + 1; j <= per && !hit; ++j) {
            const double* pj = P + 3ull * (uint32_t)(i * per + j);
            for (int l = A + 1; l <= per; ++l) {
                const double* pl = P + 3ull * (uint32_t)(k * per + l);
                const double dx = pl[0] - pj[0], dy = pl[1] - pj[1], dz = pl[2] - pj[2];
                if (pair_potential((dx * dx + dy * dy) + dz * dz, true, lambda) < 0.0) { hit = true; break; }
            }
        }
        if (hit) {
            atomicOr(&s_bond[i][k >> 5], 1u << (k & 31));
            atomicOr(&s_bond[k][i >> 5], 1u << (i & 31));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int a = 0; a < arms; ++a) s_lab[a] = -1;
        int n_patch = 0, in_patch = 0;
        for (int a = 0; a < arms; ++a) {
            if (s_lab[a] >= 0) continue;
            int top = 0, size = 0;
            s_lab[a] = a;
            s_stack[top++] = a;
            while (top > 0) {
                const int u = s_stack[--top];
                ++size;
                for (int v = 0; v < arms; ++v)
                    if (s_lab[v] < 0 && ((s_bond[u][v >> 5] >> (v & 31)) & 1u)) { s_lab[v] = a; s_stack[top++] = v; }
            }
            if (size > 1) { ++n_patch; in_patch += size; }
        }
        out[0] = (double)n_patch;
        out[1] = n_patch > 0 ? (double)in_patch / (double)n_patch : 0.0;
    }
}

/* Solvent observers of one replica in one pass over its particles (current
 * buffer), summed in 64-bit fixed point so that results do not depend on
 * the order of the atomics:
 *   kObsProfile  VelocityX::operator() Velocity_x.h:76-96: r - COM_Pos and
 *                v - COM_Vel through MPC::wrap(pos, vel) (MPC.cpp:492-495),
 *                binned by floor(x/w) and floor(y/w): sum vx, sum vy, count
 *   kObsFourier  Fourier_Autocorr::operator() Fourier_Autocorr.h:55-63:
 *                sum_i v_i cos(k_j x_ij) and sum_i v_i sin(k_j x_ij) with
 *                k_j = 2 pi / L_j, j = x, y, z
 *   kObsSpeed    VelocityDistribution::operator() Vel_distribution.h:35-40:
 *                counts of floor(|v|/w_v), the last bin taking the rest */
__global__ __launch_bounds__(256) void k_fluid_observe(DevState d, ObsArgs a) {
    __shared__ unsigned long long s_prof[2][kObsBins][3];
    __shared__ unsigned long long s_four[18];
    __shared__ uint32_t s_speed[kObsSpeedBins + 1];
    const RepDev& rep = d.reps[a.replica];
    Box b = rep.box;
    b.delrx = d.ctl[a.replica].delrx;
    const int nb0 = a.nb[0], nb1 = a.nb[1];
    for (int i = threadIdx.x; i < 3 * nb0; i += 256) (&s_prof[0][0][0])[i] = 0ull;
    for (int i = threadIdx.x; i < 3 * nb1; i += 256) (&s_prof[1][0][0])[i] = 0ull;
    for (int i = threadIdx.x; i < 18; i += 256) s_four[i] = 0ull;
    for (int i = threadIdx.x; i <= a.nsb; i += 256) s_speed[i] = 0u;
    __syncthreads();
    double com[6] = {0, 0, 0, 0, 0, 0};
    if (a.com)
        for (int q = 0; q < 6; ++q) com[q] = a.com[q];
    const int cur = d.cur;
    const size_t base = rep.part_base;
    const float *X = d.x[cur] + base, *Y = d.y[cur] + base, *Z = d.z[cur] + base;
    const float *VX = d.vx[cur] + base, *VY = d.vy[cur] + base, *VZ = d.vz[cur] + base;
    const double kk[3] = {2.0 * kPi / b.L[0], 2.0 * kPi / b.L[1], 2.0 * kPi / b.L[2]};
    double f[18];
    for (int q = 0; q < 18; ++q) f[q] = 0.0;
    const uint32_t n = rep.n_fluid;
    for (int t = 0; t < kObsPerThread; ++t) {
        const uint32_t i = (blockIdx.x * kObsPerThread + t) * 256u + threadIdx.x;
        if (i >= n) break;
        const double p[3] = {X[i], Y[i], Z[i]};
        const double v[3] = {VX[i], VY[i], VZ[i]};
        if (a.flags & kObsProfile) {
            double x = p[0] - com[0], y = p[1] - com[1], z = p[2] - com[2], ux = v[0] - com[3];
            const double uy = v[1] - com[4];
            le_wrap(b, x, y, z, ux);
            const int bx = min(max((int)floor(x / a.w) - a.b0[0], 0), nb0 - 1);
            const int by = min(max((int)floor(y / a.w) - a.b0[1], 0), nb1 - 1);
            const unsigned long long qx = to_fix(ux), qy = to_fix(uy);
            atomicAdd(&s_prof[0][bx][0], qx);
            atomicAdd(&s_prof[0][bx][1], qy);
            atomicAdd(&s_prof[0][bx][2], 1ull);
            atomicAdd(&s_prof[1][by][0], qx);
            atomicAdd(&s_prof[1][by][1], qy);
            atomicAdd(&s_prof[1][by][2], 1ull);
        }
        if (a.flags & kObsFourier) {
            for (int j = 0; j < 3; ++j) {
                double sn, cs;
                sincos(kk[j] * p[j], &sn, &cs);
                for (int c = 0; c < 3; ++c) { f[6 * j + 2 * c] += v[c] * cs; f[6 * j + 2 * c + 1] += v[c] * sn; }
            }
        }
        if (a.flags & kObsSpeed) {
            const double sp = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
            const double q = floor(sp / a.vw);
            atomicAdd(&s_speed[q < (double)a.nsb ? (int)q : a.nsb], 1u);
        }
    }
    if (a.flags & kObsFourier)
        for (int q = 0; q < 18; ++q) atomicAdd(&s_four[q], to_fix(f[q]));
    __syncthreads();
    if (a.flags & kObsProfile) {
        for (int i = threadIdx.x; i < 3 * nb0; i += 256) {
            const unsigned long long v = (&s_prof[0][0][0])[i];
            if (v) atomicAdd(a.prof + i, v);
        }
        for (int i = threadIdx.x; i < 3 * nb1; i += 256) {
            const unsigned long long v = (&s_prof[1][0][0])[i];
            if (v) atomicAdd(a.prof + 3 * kObsBins + i, v);
        }
    }
    if (a.flags & kObsFourier)
        for (int i = threadIdx.x; i < 18; i += 256) atomicAdd(a.four + i, s_four[i]);
    if (a.flags & kObsSpeed)
        for (int i = threadIdx.x; i <= a.nsb; i += 256)
            if (s_speed[i]) atomicAdd(a.speed + i, (unsigned long long)s_speed[i]);
}

/* ========================================================= launchers */

hipError_t launch_prologue(const DevState& d, hipStream_t st, int do_collision, int n_delrx, double delrx_dt) {
    hipLaunchKernelGGL(k_prologue, dim3(d.n_rep), dim3(64), 0, st, d, do_collision, n_delrx, delrx_dt);
    return hipGetLastError();
}

template <int R>
static hipError_t launch_md_r(const DevState& d, hipStream_t st, int n_md, int epot_last, int md_init,
                              int do_collision, double md_dt, int ns, uint32_t cap, size_t lds) {
    static size_t attr_set = 0;
    if (lds > attr_set) {
        const hipError_t e = hipFuncSetAttribute((const void*)k_md<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return e;
        attr_set = lds;
    }
    hipLaunchKernelGGL(k_md<R>, dim3(d.n_rep), dim3(kMdThreads), lds, st, d, n_md, epot_last, md_init,
                       do_collision, md_dt, ns, cap);
    return hipGetLastError();
}

hipError_t launch_md(const DevState& d, hipStream_t st, int n_md, int epot_last, int md_init, int do_collision,
                     double md_dt) {
    if (d.m_tot == 0) return hipSuccess;
    const int ns = d.md_ns;
    const uint32_t cap = d.md_vl_cap;
    const size_t lds = md_lds_bytes(ns, cap);
    if (ns <= kMdThreads)
        return launch_md_r<1>(d, st, n_md, epot_last, md_init, do_collision, md_dt, ns, cap, lds);
    if (ns <= 2 * kMdThreads)
        return launch_md_r<2>(d, st, n_md, epot_last, md_init, do_collision, md_dt, ns, cap, lds);
    return launch_md_r<4>(d, st, n_md, epot_last, md_init, do_collision, md_dt, ns, cap, lds);
}

hipError_t launch_bin(const DevState& d, hipStream_t st, double h, int mode) {
    if (d.n_pblk == 0) return hipSuccess;
    const dim3 g(d.n_pblk), b(256);
    switch (d.ppt) {
    case 1: hipLaunchKernelGGL(k_bin<1>, g, b, 0, st, d, h, mode); break;
    case 2: hipLaunchKernelGGL(k_bin<2>, g, b, 0, st, d, h, mode); break;
    case 4: hipLaunchKernelGGL(k_bin<4>, g, b, 0, st, d, h, mode); break;
    case 8: hipLaunchKernelGGL(k_bin<8>, g, b, 0, st, d, h, mode); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_scan(const DevState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_scan, dim3(d.n_scan_tiles), dim3(kScanThreads), 0, st, d);
    return hipGetLastError();
}

hipError_t launch_sort_redo(const DevState& d, hipStream_t st, double h) {
    hipLaunchKernelGGL(k_sort_redo, dim3(d.n_rep), dim3(1024), 0, st, d, h);
    return hipGetLastError();
}

hipError_t launch_perm(const DevState& d, hipStream_t st, int mode) {
    if (d.n_pblk == 0) return hipSuccess;
    const dim3 g(d.n_pblk), b(256);
    switch (d.ppt) {
    case 1: hipLaunchKernelGGL(k_perm<1>, g, b, 0, st, d, mode); break;
    case 2: hipLaunchKernelGGL(k_perm<2>, g, b, 0, st, d, mode); break;
    case 4: hipLaunchKernelGGL(k_perm<4>, g, b, 0, st, d, mode); break;
    case 8: hipLaunchKernelGGL(k_perm<8>, g, b, 0, st, d, mode); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_collide(const DevState& d, hipStream_t st, double h) {
    const dim3 g(d.n_tiles), b(256);
    if (d.any_amc) {
        switch (d.tile_cells) {
        case 64: hipLaunchKernelGGL((k_collide<64, 768, true>), g, b, 0, st, d, h); break;
        case 128: hipLaunchKernelGGL((k_collide<128, 1536, true>), g, b, 0, st, d, h); break;
        case 256: hipLaunchKernelGGL((k_collide<256, 3072, true>), g, b, 0, st, d, h); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (d.tile_cells) {
    case 64: hipLaunchKernelGGL((k_collide<64, 768, false>), g, b, 0, st, d, h); break;
    case 128: hipLaunchKernelGGL((k_collide<128, 1536, false>), g, b, 0, st, d, h); break;
    case 256: hipLaunchKernelGGL((k_collide<256, 3072, false>), g, b, 0, st, d, h); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_advect(const DevState& d, hipStream_t st, double h, int mode) {
    if (!d.rt_cap || !d.n_rtiles) return hipErrorInvalidValue;
    switch (d.ppt) {
    case 1: hipLaunchKernelGGL(k_advect<1>, dim3(d.n_pblk), dim3(256), 0, st, d, h, mode); break;
    case 2: hipLaunchKernelGGL(k_advect<2>, dim3(d.n_pblk), dim3(256), 0, st, d, h, mode); break;
    case 4: hipLaunchKernelGGL(k_advect<4>, dim3(d.n_pblk), dim3(256), 0, st, d, h, mode); break;
    case 8: hipLaunchKernelGGL(k_advect<8>, dim3(d.n_pblk), dim3(256), 0, st, d, h, mode); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_cell_gamma(const DevState& d, hipStream_t st) {
    if (!d.n_rtiles) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cell_gamma, dim3(d.n_rtiles + (uint32_t)d.n_rep), dim3(64), 0, st, d);
    return hipGetLastError();
}

hipError_t launch_rows_redo(const DevState& d, hipStream_t st, double h) {
    switch (d.ppt) {
    case 1: hipLaunchKernelGGL(k_rows_redo<1>, dim3(256), dim3(256), 0, st, d, h); break;
    case 2: hipLaunchKernelGGL(k_rows_redo<2>, dim3(256), dim3(256), 0, st, d, h); break;
    case 4: hipLaunchKernelGGL(k_rows_redo<4>, dim3(256), dim3(256), 0, st, d, h); break;
    case 8: hipLaunchKernelGGL(k_rows_redo<8>, dim3(256), dim3(256), 0, st, d, h); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int NT, int CAP, int TCM>
static void launch_collide_rows_t(const DevState& d, hipStream_t st, double h) {
    if (d.any_amc)
        hipLaunchKernelGGL((k_collide_rows<NT, CAP, TCM, true>), dim3(d.n_rtiles), dim3(NT), 0, st, d, h);
    else
        hipLaunchKernelGGL((k_collide_rows<NT, CAP, TCM, false>), dim3(d.n_rtiles), dim3(NT), 0, st, d, h);
}

hipError_t launch_collide_rows(const DevState& d, hipStream_t st, double h) {
    if (!d.n_rtiles) return hipErrorInvalidValue;
    if (d.rt_cap == 1024 && d.rt_tcm == 64) launch_collide_rows_t<256, 1024, 64>(d, st, h);
    else if (d.rt_cap == 1024 && d.rt_tcm == 128) launch_collide_rows_t<256, 1024, 128>(d, st, h);
    else if (d.rt_cap == 1024 && d.rt_tcm == 256) launch_collide_rows_t<256, 1024, 256>(d, st, h);
    else if (d.rt_cap == 2048 && d.rt_tcm == 64) launch_collide_rows_t<512, 2048, 64>(d, st, h);
    else if (d.rt_cap == 2048 && d.rt_tcm == 128) launch_collide_rows_t<512, 2048, 128>(d, st, h);
    else if (d.rt_cap == 2048 && d.rt_tcm == 256) launch_collide_rows_t<512, 2048, 256>(d, st, h);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_fluid_init(const DevState& d, hipStream_t st) {
    if (d.n_pblk == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fluid_init, dim3(d.n_pblk), dim3(256), 0, st, d, 0);
    hipLaunchKernelGGL(k_fluid_mean, dim3(d.n_rep), dim3(64), 0, st, d);
    hipLaunchKernelGGL(k_fluid_init, dim3(d.n_pblk), dim3(256), 0, st, d, 1);
    return hipGetLastError();
}

/* MPC::calculateFluidVelocity (MPC.cpp:334-342) of cells c0 .. c0+n-1 of
 * one replica: the mean velocity of the solvent particles the last
 * collision sorted into the cell (its slice of the current buffer), 0 for
 * an empty cell; summed in slice order in FP64. */
__global__ __launch_bounds__(256) void k_cell_velocity(DevState d, int replica, uint32_t c0, uint32_t n, double* out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = (uint32_t)d.reps[replica].cell_base + c0 + i;
    const uint32_t a = d.offs[d.cur][g], e = d.offs[d.cur][g + 1];
    const int cur = d.cur;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t p = a; p < e; ++p) { s0 += d.vx[cur][p]; s1 += d.vy[cur][p]; s2 += d.vz[cur][p]; }
    if (e > a) { s0 /= (double)(e - a); s1 /= (double)(e - a); s2 /= (double)(e - a); }
    out[3ull * i] = s0; out[3ull * i + 1] = s1; out[3ull * i + 2] = s2;
}

hipError_t launch_cell_velocity(const DevState& d, hipStream_t st, int replica, uint32_t c0, uint32_t n, double* out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cell_velocity, dim3((n + 255) / 256), dim3(256), 0, st, d, replica, c0, n, out);
    return hipGetLastError();
}

hipError_t launch_temperature(const DevState& d, hipStream_t st, int replica, uint32_t ntiles) {
    switch (d.tile_cells) {
    case 64: hipLaunchKernelGGL(k_temperature<64>, dim3(ntiles), dim3(256), 0, st, d, replica); break;
    case 128: hipLaunchKernelGGL(k_temperature<128>, dim3(ntiles), dim3(256), 0, st, d, replica); break;
    case 256: hipLaunchKernelGGL(k_temperature<256>, dim3(ntiles), dim3(256), 0, st, d, replica); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_closure_features(const DevState& d, hipStream_t st, float* feat, const float* fmin,
                                   const float* fmax) {
    if (d.m_tot == 0) return hipSuccess;
    const size_t lds = kClLds;
    static bool attr = false;
    if (!attr) {
        const hipError_t e = hipFuncSetAttribute((const void*)k_closure_features,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(k_closure_features, dim3(d.n_rep), dim3(kClThreads), lds, st, d, feat, fmin, fmax);
    return hipGetLastError();
}

hipError_t launch_closure_kick(const DevState& d, hipStream_t st, const float* out, float sy0, float sy1, float sy2,
                               uint32_t k) {
    if (d.m_tot == 0) return hipSuccess;
    hipLaunchKernelGGL(k_closure_kick, dim3((d.m_tot + 255) / 256), dim3(256), 0, st, d, out, sy0, sy1, sy2, k);
    return hipGetLastError();
}

hipError_t launch_star_moments(const DevState& d, hipStream_t st, int replica, double* out) {
    hipLaunchKernelGGL(k_star_moments, dim3(1), dim3(256), 0, st, d, replica, out);
    return hipGetLastError();
}

hipError_t launch_star_patches(const DevState& d, hipStream_t st, int replica, int arms, int A, int B, double* out) {
    if (arms < 1 || arms > kPatchArms) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_star_patches, dim3(1), dim3(256), 0, st, d, replica, arms, A, B, out);
    return hipGetLastError();
}

/* Vel_Autocorr_Fluid::operator() and calculate (Vel_Autocorr_Fluid.h:44-78)
 * for replica r: each solvent particle's velocity goes into slot `slot` of
 * its window of E samples (hist, by particle id: E x 3 floats); with calc
 * (the window is full), the products of the oldest sample (slot `oldest`)
 * with every sample of the window, oldest first, are summed over the
 * particles into sums[lag] -- 64-bit fixed point (exact, so the order of
 * particles and workgroups does not matter), one global atomic per lag and
 * workgroup. */
__global__ __launch_bounds__(256) void k_fluid_vacf(DevState d, int r, float* hist, uint32_t E, uint32_t slot,
                                                    uint32_t oldest, int calc, unsigned long long* sums) {
    __shared__ unsigned long long s_sum[kVacfMaxEntries];
    const RepDev& rep = d.reps[r];
    const int tid = threadIdx.x;
    const uint32_t li = blockIdx.x * 256u + (uint32_t)tid;
    const bool live = li < rep.n_fluid;
    if (calc)
        for (uint32_t j = tid; j < E; j += 256) s_sum[j] = 0ull;
    float* h = nullptr;
    if (live) {
        const int cur = d.cur;
        const uint64_t i = rep.part_base + li;
        const uint32_t id = d.ids[cur][i];
        h = hist + (size_t)id * E * 3;
        h[3 * slot] = d.vx[cur][i]; h[3 * slot + 1] = d.vy[cur][i]; h[3 * slot + 2] = d.vz[cur][i];
    }
    if (!calc) return;
    __syncthreads();
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if (live) { ox = h[3 * oldest]; oy = h[3 * oldest + 1]; oz = h[3 * oldest + 2]; }
    for (uint32_t j = 0; j < E; ++j) {
        unsigned long long p = 0ull;
        if (live) {
            uint32_t sl = oldest + j;
            if (sl >= E) sl -= E;
            p = to_fix((ox * (double)h[3 * sl] + oy * (double)h[3 * sl + 1]) + oz * (double)h[3 * sl + 2]);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o);
        if ((tid & 63) == 0) atomicAdd(&s_sum[j], p);
    }
    __syncthreads();
    for (uint32_t j = tid; j < E; j += 256)
        if (s_sum[j]) atomicAdd(&sums[j], s_sum[j]);
}

hipError_t launch_fluid_vacf(const DevState& d, hipStream_t st, int r, uint32_t n_fluid, float* hist, uint32_t E,
                             uint32_t slot, uint32_t oldest, int calc, unsigned long long* sums) {
    if (E == 0 || E > (uint32_t)kVacfMaxEntries || slot >= E || oldest >= E) return hipErrorInvalidValue;
    const uint32_t nblk = (n_fluid + 255) / 256;
    if (nblk == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fluid_vacf, dim3(nblk), dim3(256), 0, st, d, r, hist, E, slot, oldest, calc, sums);
    return hipGetLastError();
}

/* n words of p set to zero (a kernel rather than a memset node, so that the
 * call stays correct when a caller captures it into a graph) */
__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) p[i] = 0ull;
}

hipError_t launch_fill_u64(unsigned long long* p, uint32_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t nblk = (n + 255) / 256 < 64 ? (n + 255) / 256 : 64;
    hipLaunchKernelGGL(k_fill_u64, dim3(nblk), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

/* The velocity profile's fixed-point bins acc[2][kObsBins][3] (sum vx, sum
 * vy at 2^-32, count) as doubles in the caller's layout
 * out[(a * max_bins + b) * 3 + c], bins b < nb[a]; other entries untouched */
__global__ __launch_bounds__(256) void k_profile_export(const unsigned long long* acc, int nb0, int nb1,
                                                        uint32_t max_bins, double* out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t a = i >= (uint32_t)nb0 ? 1u : 0u;
    const uint32_t b = a ? i - (uint32_t)nb0 : i;
    if (b >= (uint32_t)(a ? nb1 : nb0)) return;
    const unsigned long long* q = acc + 3ull * (a * kObsBins + b);
    double* o = out + 3ull * ((size_t)a * max_bins + b);
    o[0] = (double)(long long)q[0] * kUnfix;
    o[1] = (double)(long long)q[1] * kUnfix;
    o[2] = (double)q[2];
}

hipError_t launch_profile_export(const unsigned long long* acc, int nb0, int nb1, uint32_t max_bins, double* out,
                                 hipStream_t st) {
    if (nb0 < 1 || nb0 > kObsBins || nb1 < 1 || nb1 > kObsBins || (uint32_t)nb0 > max_bins ||
        (uint32_t)nb1 > max_bins)
        return hipErrorInvalidValue;
    const uint32_t n = (uint32_t)(nb0 + nb1);
    hipLaunchKernelGGL(k_profile_export, dim3((n + 255) / 256), dim3(256), 0, st, acc, nb0, nb1, max_bins, out);
    return hipGetLastError();
}

hipError_t launch_fluid_observe(const DevState& d, hipStream_t st, const ObsArgs& a, uint32_t n_fluid) {
    if (a.nb[0] < 1 || a.nb[0] > kObsBins || a.nb[1] < 1 || a.nb[1] > kObsBins || a.nsb < 0 ||
        a.nsb > kObsSpeedBins)
        return hipErrorInvalidValue;
    const uint32_t per = 256u * kObsPerThread;
    const uint32_t nblk = (n_fluid + per - 1) / per;
    if (nblk == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fluid_observe, dim3(nblk), dim3(256), 0, st, d, a);
    return hipGetLastError();
}

}  // namespace spmd
