/*
 * oracle.cpp -- FP64 CPU oracle for the MPC-MD hot path.  TEST
 * INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it as the checker; the HIP library never does.
 *
 * A line-by-line restatement, in FP64 and in the reference's operation
 * order, of:
 *   MPC ctor / initialize / step / streaming / shift / sort / CMV /
 *     thermostat / collide / check_bounds / LEBC / wrap / temperature
 *                                 star-polymers/lib/Hydrodynamics/MPC.cpp
 *   Box Verlet lists, forces, COM frame, gyration, rotation frequency
 *                                 star-polymers/lib/Box.cpp
 *   velocity Verlet               star-polymers/lib/Thermostats/Thermostat_None.cpp
 *   star builder, COM, Ekin       star-polymers/lib/Molecules/Molecule.cpp
 *   potentials                    star-polymers/lib/Molecules/Potentials.h
 *   RNG (reference mode)          star-polymers/lib/Rand.cpp:6-43
 * Each function below cites the lines it follows.  Vector algebra that the
 * reference does with Eigen 3 (absent here) is written out per component in
 * the association order Eigen evaluates 3-vectors: ((a0+a1)+a2).
 *
 * Built with g++ -O2 (no -ffast-math, no FMA) so every FP64 operation is a
 * correctly rounded IEEE operation, which the bit-exact cell-index tests
 * rely on.
 */
#include <omp.h>
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <forward_list>
#include <map>
#include <array>
#include <complex>
#include <random>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

/* ---------------------------------------------------------------- Philox
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
 * as easy as 1, 2, 3"), restated from the paper: multipliers 0xD2511F53,
 * 0xCD9E8D57, Weyl key increments 0x9E3779B9, 0xBB67AE85, ten rounds. */
inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

inline void philox10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

/* Stream layout shared with the HIP library: counter = (index, collision
 * number, replica<<8 | tag, block). */
enum : uint32_t { TAG_SHIFT = 1, TAG_AXIS = 2, TAG_GAMMA = 3, TAG_FLUID_INIT = 4, TAG_STAR_INIT = 5 };

inline double u01(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386962890625e-10; }

inline void philox_draw(uint64_t seed, uint32_t replica, uint32_t tag, uint32_t index,
                        uint32_t k, uint32_t block, uint32_t out[4]) {
    uint32_t ctr[4] = {index, k, (replica << 8) | tag, block};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    philox10(ctr, key, out);
}

/* Gamma(shape >= 1, scale) by Marsaglia & Tsang, "A simple method for
 * generating gamma variables", ACM TOMS 26(3), 2000; normals by Box-Muller.
 * Attempt j consumes Philox block j of the (cell, k, TAG_GAMMA) stream. */
double philox_gamma(uint64_t seed, uint32_t replica, uint32_t cell, uint32_t k,
                    double shape, double scale) {
    const double d = shape - 1.0 / 3.0;
    const double c = 1.0 / std::sqrt(9.0 * d);
    for (uint32_t j = 0; j < 64; ++j) {
        uint32_t w[4];
        philox_draw(seed, replica, TAG_GAMMA, cell, k, j, w);
        const double r = std::sqrt(-2.0 * std::log(u01(w[0])));
        const double x = r * std::cos(6.283185307179586 * u01(w[1]));
        double v = 1.0 + c * x;
        if (v <= 0.0) continue;
        v = v * v * v;
        const double u = u01(w[2]);
        const double x2 = x * x;
        if (u < 1.0 - 0.0331 * (x2 * x2)) return d * v * scale;
        if (std::log(u) < 0.5 * x2 + d * (1.0 - v + std::log(v))) return d * v * scale;
    }
    return d * scale;
}

/* ------------------------------------------------------------ Potentials
 * Potentials.h:14-120, same branches and operation order. */
double fene_potential(double r2) { return r2 == 0.0 ? 0.0 : -15.0 * 2.25 * std::log(1.0 - r2 / 2.25); }
double fene_force(double r2) { return r2 == 0.0 ? 0.0 : -30.0 / (1.0 - (r2 / 2.25)); }
double fene_anchor_potential(double r2) { return r2 == 0.0 ? 0.0 : -15.0 * 20.25 * std::log(1.0 - r2 / 20.25); }
double fene_anchor_force(double r2) { return r2 == 0.0 ? 0.0 : -30.0 / (1.0 - (r2 / 20.25)); }
double aa_potential(double r2) {
    double p = 0.0;
    if (r2 < 1.0594631) {
        double rm24 = 1.0 / r2;
        rm24 *= rm24 * rm24;
        rm24 *= rm24;
        rm24 *= rm24;
        p = 4.0 * rm24 * (rm24 - 1.0) + 1.0;
    }
    return p;
}
double aa_force(double r2) {
    double f = 0.0;
    if (r2 < 1.0594631) {
        double rm2 = 1.0 / r2;
        double rm24 = rm2 * rm2 * rm2;
        rm24 *= rm24;
        rm24 *= rm24;
        f = 192.0 * rm2 * rm24 * (rm24 - 0.5);
    }
    return f;
}
double bb_potential(double r2, double lambda) {
    double p = 0.0;
    if (r2 < 1.0594631) {
        double rm24 = 1.0 / r2;
        rm24 *= rm24 * rm24;
        rm24 *= rm24;
        rm24 *= rm24;
        p = 4.0 * rm24 * (rm24 - 1.0) + 1.0 - lambda;
    } else if (r2 < 2.25) {
        double rm24 = 1.0 / r2;
        rm24 *= rm24 * rm24;
        rm24 *= rm24;
        rm24 *= rm24;
        p = 4.0 * lambda * (rm24 * (rm24 - 1.0) - 0.0000593);
    }
    return p;
}
double bb_force(double r2, double lambda) {
    double f = 0.0;
    if (r2 < 1.0594631) {
        double rm2 = 1.0 / r2;
        double rm24 = rm2 * rm2 * rm2;
        rm24 *= rm24;
        rm24 *= rm24;
        f = 192.0 * rm2 * rm24 * (rm24 - 0.5);
    } else if (r2 < 2.25) {
        double rm2 = 1.0 / r2;
        double rm24 = rm2 * rm2 * rm2;
        rm24 *= rm24;
        rm24 *= rm24;
        f = 192.0 * lambda * rm2 * rm24 * (rm24 - 0.5);
    }
    return f;
}

struct V3 {
    double v[3];
    double& operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};
inline double dot3(const V3& a, const V3& b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

struct Mono {
    V3 pos{}, vel{}, force{}, vpos{};
    double mass = 1.0;
    bool type = false;    /* AmphiType: 1 = B (solvophobic) */
    bool anchor = false;
    std::forward_list<int> neighbors;  /* directed bonds (Particle.h:57) */
    std::forward_list<int> verlet;     /* Verlet list (Particle.h:58) */
    uint32_t cell = 0;
};

}  // namespace

struct orc_sys {
    orc_params p{};
    int L[3]{};
    uint32_t number_of_cells = 0;  /* MPC.cpp:16: the maximum index, L^3-1 */
    uint64_t n_fluid = 0;
    double c = 0, s = 0, delrx = 0;
    bool shear_on = false;
    int64_t k_coll = 0;            /* collisions done (Philox counter) */

    std::vector<double> pos, vel;  /* AoS fluid, 3 per particle */
    std::vector<uint32_t> fcell;

    std::vector<Mono> mono;
    uint32_t A = 0, B = 0, arms = 0;
    double epot = 0.0;
    V3 com_pos{}, com_vel{};
    std::vector<V3> com_frame_pos, com_frame_vel;  /* Molecules_com_reference_frame */

    /* Verlet grid (Box.cpp:20-26) */
    double verlet_radius = 2.0, verlet_radius2 = 4.0, cutoff = 1.5;
    int cell_size[3]{};
    double cell_side[3]{};

    /* cell lists of the last sort(): fluid indices >= 0, monomer m as -1-m */
    std::vector<std::vector<int64_t>> cell_list;
    std::vector<uint32_t> cell_fluid_count;

    /* reference RNG (Rand.cpp:6-11) */
    std::mt19937_64 gen{1};
    std::uniform_real_distribution<double> dis_uniform;
    std::gamma_distribution<double> dis_gamma;

    /* injected randoms */
    V3 inj_shift{};
    std::vector<double> inj_axis, inj_gamma;
    V3 last_shift{};
    int last_recentered = 0;

    double uniform() { return dis_uniform(gen); }  /* Rand::real_uniform */
    double gamma(double a, double b) {               /* Rand::real_gamma, Rand.cpp:40-43 */
        dis_gamma.param(std::gamma_distribution<double>::param_type(a, b));
        return dis_gamma(gen);
    }
};

namespace {

/* MPC::LEBC(Particle&) MPC.cpp:421-429, and the plain wrap MPC.cpp:473-478,
 * selected by shear_on as in MPC::wrap(Particle&) MPC.cpp:487-490. */
void wrap(const orc_sys& s, double* x, double* v) {
    const double Lx = s.L[0], Ly = s.L[1], Lz = s.L[2];
    if (s.shear_on) {
        double cy = std::round(x[1] / Ly);
        x[0] -= Lx * std::round(x[0] / Lx);
        x[0] -= cy * s.delrx;
        x[0] -= Lx * std::round(x[0] / Lx);
        x[1] -= Ly * cy;
        x[2] -= Lz * std::round(x[2] / Lz);
        v[0] -= cy * s.p.shear * Ly;
    } else {
        for (int i = 0; i < 3; ++i) x[i] -= std::round(x[i] / s.L[i]) * s.L[i];
    }
}

V3 com_position(const orc_sys& s) {  /* Molecule::calculate_center_of_mass, Molecule.cpp:76-83 */
    V3 c{};
    for (auto& m : s.mono)
        for (int i = 0; i < 3; ++i) c[i] += m.pos[i];
    for (int i = 0; i < 3; ++i) c[i] /= (double)s.mono.size();
    return c;
}
V3 com_velocity(const orc_sys& s) {  /* Molecule.cpp:85-92 */
    V3 c{};
    for (auto& m : s.mono)
        for (int i = 0; i < 3; ++i) c[i] += m.vel[i];
    for (int i = 0; i < 3; ++i) c[i] /= (double)s.mono.size();
    return c;
}

double mol_ekin(const orc_sys& s) {  /* Molecule::calculate_Ekin, Molecule.cpp:60-74 */
    V3 cv{};
    for (auto& m : s.mono)
        for (int i = 0; i < 3; ++i) cv[i] += m.vel[i];
    for (int i = 0; i < 3; ++i) cv[i] /= (double)s.mono.size();
    double e = 0.0;
    for (auto& m : s.mono) {
        V3 r{{m.vel[0] - cv[0], m.vel[1] - cv[1], m.vel[2] - cv[2]}};
        e += m.mass * dot3(r, r);
    }
    return e / 2.0;
}

/* Box::center_of_mass_reference_frame, Box.cpp:81-102 (one molecule). */
void com_reference_frame(orc_sys& s) {
    if (s.mono.empty()) return;
    V3 c{}, cv{};
    for (auto& m : s.mono)
        for (int i = 0; i < 3; ++i) { c[i] += m.pos[i]; cv[i] += m.vel[i]; }
    for (int i = 0; i < 3; ++i) { c[i] /= (double)s.mono.size(); cv[i] /= (double)s.mono.size(); }
    s.com_frame_pos.resize(s.mono.size());
    s.com_frame_vel.resize(s.mono.size());
    for (size_t k = 0; k < s.mono.size(); ++k)
        for (int i = 0; i < 3; ++i) {
            s.com_frame_pos[k][i] = s.mono[k].pos[i] - c[i];
            s.com_frame_vel[k][i] = s.mono[k].vel[i] - cv[i];
        }
    s.com_pos = c;
    s.com_vel = cv;
}

/* Box::update_VerletLists, Box.cpp:105-180.  Returns -1 where the reference
 * raises SIGSEGV (monomer outside the COM-centred grid). */
int update_verlet(orc_sys& s) {
    const int nx = s.cell_size[0], ny = s.cell_size[1], nz = s.cell_size[2];
    std::vector<std::forward_list<int>> grid((size_t)nx * ny * nz);
    for (auto& m : s.mono) m.verlet.clear();
    com_reference_frame(s);
    int cn[3];
    for (size_t k = 0; k < s.mono.size(); ++k) {
        auto& m = s.mono[k];
        for (int i = 0; i < 3; ++i)
            cn[i] = (int)((m.pos[i] - s.com_pos[i] + s.L[i] * 0.5) / s.cell_side[i]);
        m.vpos = m.pos;
        if (cn[0] < 0 || cn[0] >= nx || cn[1] < 0 || cn[1] >= ny || cn[2] < 0 || cn[2] >= nz) return -1;
        grid[((size_t)cn[0] * ny + cn[1]) * nz + cn[2]].push_front((int)k);
    }
    for (size_t k = 0; k < s.mono.size(); ++k) {
        auto& m = s.mono[k];
        for (int i = 0; i < 3; ++i)
            cn[i] = (int)((m.pos[i] - s.com_pos[i] + s.L[i] * 0.5) / s.cell_side[i]);
        for (int j = cn[0] - 1; j < cn[0] + 2; ++j) {
            if (j < 0 || j >= nx) continue;
            for (int kk = cn[1] - 1; kk < cn[1] + 2; ++kk) {
                if (kk < 0 || kk >= ny) continue;
                for (int l = cn[2] - 1; l < cn[2] + 2; ++l) {
                    if (l < 0 || l >= nz) continue;
                    for (int o : grid[((size_t)j * ny + kk) * nz + l]) {
                        if (o == (int)k) continue;
                        V3 d{{s.mono[o].pos[0] - m.pos[0], s.mono[o].pos[1] - m.pos[1],
                              s.mono[o].pos[2] - m.pos[2]}};
                        if (dot3(d, d) <= s.verlet_radius2) m.verlet.push_front(o);
                    }
                }
            }
        }
    }
    return 0;
}

/* Box::check_VerletLists, Box.cpp:182-194. */
int check_verlet(orc_sys& s) {
    for (auto& m : s.mono) {
        V3 d{{m.pos[0] - m.vpos[0], m.pos[1] - m.vpos[1], m.pos[2] - m.vpos[2]}};
        if (std::sqrt(dot3(d, d)) > (s.verlet_radius - s.cutoff) * 0.5) return update_verlet(s);
    }
    return 0;
}

/* Box::calculate_forces_verlet, Box.cpp:243-286. */
void forces_verlet(orc_sys& s, bool calc_epot) {
    s.epot = 0.0;
    for (auto& m : s.mono) m.force = V3{};
    for (size_t k = 0; k < s.mono.size(); ++k) {
        auto& m = s.mono[k];
        for (int o : m.verlet) {
            const auto& other = s.mono[o];
            V3 d{{other.pos[0] - m.pos[0], other.pos[1] - m.pos[1], other.pos[2] - m.pos[2]}};
            double r2 = dot3(d, d);
            double fa;
            if (m.type && other.type) {
                if (calc_epot) s.epot += 0.5 * bb_potential(r2, s.p.lambda);
                fa = bb_force(r2, s.p.lambda);
            } else {
                if (calc_epot) s.epot += 0.5 * aa_potential(r2);
                fa = aa_force(r2);
            }
            for (int i = 0; i < 3; ++i) m.force[i] -= d[i] * fa;
        }
        for (int o : m.neighbors) {
            auto& nb = s.mono[o];
            V3 d{{nb.pos[0] - m.pos[0], nb.pos[1] - m.pos[1], nb.pos[2] - m.pos[2]}};
            double r2 = dot3(d, d);
            double fa;
            if (m.anchor) {
                if (calc_epot) s.epot += fene_anchor_potential(r2);
                fa = fene_anchor_force(r2);
            } else {
                if (calc_epot) s.epot += fene_potential(r2);
                fa = fene_force(r2);
            }
            for (int i = 0; i < 3; ++i) {
                double f = d[i] * fa;
                m.force[i] -= f;
                nb.force[i] += f;
            }
        }
    }
}

/* MPC::check_bounds, MPC.cpp:160-190. */
void check_bounds(orc_sys& s) {
    s.last_recentered = 0;
    bool out = false;
    for (auto& m : s.mono) {
        for (int i = 0; i < 3; ++i)
            if (std::fabs(m.pos[i]) >= (s.L[i] * 0.5 - 0.5)) { out = true; break; }
        if (out) break;
    }
    if (!out) return;
    s.last_recentered = 1;
    V3 c = com_position(s), cv = com_velocity(s);
    for (auto& m : s.mono)
        for (int i = 0; i < 3; ++i) { m.pos[i] -= c[i]; m.vel[i] -= cv[i]; }
    for (uint64_t p = 0; p < s.n_fluid; ++p) {
        double* x = &s.pos[3 * p];
        double* v = &s.vel[3 * p];
        for (int i = 0; i < 3; ++i) { x[i] -= c[i]; v[i] -= cv[i]; }
        wrap(s, x, v);
    }
}

/* MPC::sort, MPC.cpp:201-237. */
int sort_cells(orc_sys& s) {
    for (auto& l : s.cell_list) l.clear();
    std::fill(s.cell_fluid_count.begin(), s.cell_fluid_count.end(), 0u);
    const int Lx = s.L[0], Ly = s.L[1], Lz = s.L[2];
#pragma omp parallel for schedule(static)
    for (uint64_t p = 0; p < s.n_fluid; ++p) {
        const double* x = &s.pos[3 * p];
        s.fcell[p] = (uint32_t)((int)(x[0] + Lx * 0.5) + Lx * (int)(x[1] + Ly * 0.5) +
                                Lx * Ly * (int)(x[2] + Lz * 0.5));
    }
    for (uint64_t p = 0; p < s.n_fluid; ++p) {  /* in particle order, as the reference pushes */
        const uint32_t ci = s.fcell[p];
        if (ci > s.number_of_cells) return -1;
        s.cell_list[ci].push_back((int64_t)p);
        s.cell_fluid_count[ci]++;
    }
    for (size_t k = 0; k < s.mono.size(); ++k) {
        const V3& x = s.mono[k].pos;
        uint32_t ci = (uint32_t)((int)(x[0] + Lx * 0.5) + Lx * (int)(x[1] + Ly * 0.5) +
                                 Lx * Ly * (int)(x[2] + Lz * 0.5));
        s.mono[k].cell = ci;
        if (ci > s.number_of_cells) return -1;
        s.cell_list[ci].push_back(-1 - (int64_t)k);
    }
    return 0;
}

inline double* member_vel(orc_sys& s, int64_t id) {
    return id >= 0 ? &s.vel[3 * id] : s.mono[-1 - id].vel.v;
}
inline double member_mass(const orc_sys& s, int64_t id) { return id >= 0 ? 1.0 : s.mono[-1 - id].mass; }
inline const double* member_pos(const orc_sys& s, int64_t id) {
    return id >= 0 ? &s.pos[3 * id] : s.mono[-1 - id].pos.v;
}

/* The 3x3 symmetric positive-definite solve of Matrix3d::llt().solve
 * (MPC.cpp:373): Cholesky A = L L^T, then L y = b and L^T x = y. */
void llt_solve3(const double A[3][3], const double b[3], double x[3]) {
    double L[3][3] = {};
    for (int j = 0; j < 3; ++j) {
        double d = A[j][j];
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        L[j][j] = std::sqrt(d);
        for (int i = j + 1; i < 3; ++i) {
            double v = A[i][j];
            for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
            L[i][j] = v / L[j][j];
        }
    }
    double y[3];
    for (int i = 0; i < 3; ++i) {
        double v = b[i];
        for (int k = 0; k < i; ++k) v -= L[i][k] * y[k];
        y[i] = v / L[i][i];
    }
    for (int i = 2; i >= 0; --i) {
        double v = y[i];
        for (int k = i + 1; k < 3; ++k) v -= L[k][i] * x[k];
        x[i] = v / L[i][i];
    }
}

/* MPC::calculateCMV, MPC.cpp:314-322. */
V3 calc_cmv(orc_sys& s, uint32_t cell) {
    V3 cmv{};
    double tm = 0.0;
    for (int64_t id : s.cell_list[cell]) {
        const double* v = member_vel(s, id);
        double m = member_mass(s, id);
        for (int i = 0; i < 3; ++i) cmv[i] += v[i] * m;
        tm += m;
    }
    if (tm > 0) for (int i = 0; i < 3; ++i) cmv[i] /= tm;
    return cmv;
}

/* MPC::thermostat, MPC.cpp:273-299. */
void thermostat(orc_sys& s, uint32_t cell, const V3& cmv) {
    double ek = 0.0;
    unsigned count = 0;
    for (int64_t id : s.cell_list[cell]) {
        const double* v = member_vel(s, id);
        V3 d{{v[0] - cmv[0], v[1] - cmv[1], v[2] - cmv[2]}};
        ek += member_mass(s, id) * dot3(d, d);
        count++;
    }
    if (count < 2) return;
    ek *= 0.5;
    double g;
    switch (s.p.rng_mode) {
    case ORC_RNG_REFERENCE: g = s.gamma(1.5 * (count - 1), s.p.kT); break;
    case ORC_RNG_PHILOX:
        g = philox_gamma(s.p.seed, s.p.replica, cell, (uint32_t)s.k_coll, 1.5 * (count - 1), s.p.kT);
        break;
    default: g = s.inj_gamma[cell]; break;
    }
    double scaling = std::sqrt(g / ek);
    for (int64_t id : s.cell_list[cell]) {
        double* v = member_vel(s, id);
        for (int i = 0; i < 3; ++i) v[i] = scaling * v[i] + (1 - scaling) * cmv[i];
    }
}

/* MPC::collide, MPC.cpp:239-270; with the angular-momentum flag and more
 * than two members, MPC::calculateCMP (MPC.cpp:324-333) and
 * MPC::calculateAngular (MPC.cpp:344-374) add omega x (r - CMP). */
void collide(orc_sys& s, uint32_t cell, const V3& cmv) {
    double u1, u2;
    switch (s.p.rng_mode) {
    case ORC_RNG_REFERENCE: u1 = s.uniform(); u2 = s.uniform(); break;
    case ORC_RNG_PHILOX: {
        uint32_t w[4];
        philox_draw(s.p.seed, s.p.replica, TAG_AXIS, cell, (uint32_t)s.k_coll, 0, w);
        u1 = u01(w[0]); u2 = u01(w[1]);
        break;
    }
    default: u1 = s.inj_axis[2 * cell]; u2 = s.inj_axis[2 * cell + 1]; break;
    }
    double phi = 2. * kPi * u1;
    double theta = 2. * (u2 - 0.5);
    double a0 = std::sqrt(1 - theta * theta) * std::cos(phi);
    double a1 = std::sqrt(1 - theta * theta) * std::sin(phi);
    double a2 = theta;
    const double c = s.c, sn = s.s;
    double R[3][3];
    R[0][0] = a0 * a0 + (1 - a0 * a0) * c;
    R[0][1] = a0 * a1 * (1 - c) - a2 * sn;
    R[0][2] = a0 * a2 * (1 - c) + a1 * sn;
    R[1][0] = a0 * a1 * (1 - c) + a2 * sn;
    R[1][1] = a1 * a1 + (1 - a1 * a1) * c;
    R[1][2] = a1 * a2 * (1 - c) - a0 * sn;
    R[2][0] = a0 * a2 * (1 - c) - a1 * sn;
    R[2][1] = a1 * a2 * (1 - c) + a0 * sn;
    R[2][2] = a2 * a2 + (1 - a2 * a2) * c;
    const bool amc = s.p.amc && s.cell_list[cell].size() > 2;
    double cmp[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    if (amc) {
        double tm = 0.0;
        for (int64_t id : s.cell_list[cell]) {
            const double* x = member_pos(s, id);
            const double m = member_mass(s, id);
            for (int i = 0; i < 3; ++i) cmp[i] += x[i] * m;
            tm += m;
        }
        for (int i = 0; i < 3; ++i) cmp[i] /= tm;
        double C[3] = {0, 0, 0}, I[3][3] = {};
        for (int64_t id : s.cell_list[cell]) {
            const double* x = member_pos(s, id);
            const double* v = member_vel(s, id);
            const double m = member_mass(s, id);
            const double r[3] = {x[0] - cmp[0], x[1] - cmp[1], x[2] - cmp[2]};
            const double u[3] = {v[0] - cmv[0], v[1] - cmv[1], v[2] - cmv[2]};
            double q[3];  /* u - R u */
            for (int i = 0; i < 3; ++i) q[i] = u[i] - ((R[i][0] * u[0] + R[i][1] * u[1]) + R[i][2] * u[2]);
            C[0] += m * (r[1] * q[2] - r[2] * q[1]);
            C[1] += m * (r[2] * q[0] - r[0] * q[2]);
            C[2] += m * (r[0] * q[1] - r[1] * q[0]);
            I[0][0] += m * (r[1] * r[1] + r[2] * r[2]);
            I[1][1] += m * (r[0] * r[0] + r[2] * r[2]);
            I[2][2] += m * (r[1] * r[1] + r[0] * r[0]);
            I[0][1] -= m * r[0] * r[1];
            I[0][2] -= m * r[0] * r[2];
            I[1][2] -= m * r[1] * r[2];
        }
        I[1][0] = I[0][1];
        I[2][0] = I[0][2];
        I[2][1] = I[1][2];
        llt_solve3(I, C, w);
    }
    for (int64_t id : s.cell_list[cell]) {
        double* v = member_vel(s, id);
        double d0 = v[0] - cmv[0], d1 = v[1] - cmv[1], d2 = v[2] - cmv[2];
        double n0 = cmv[0] + ((R[0][0] * d0 + R[0][1] * d1) + R[0][2] * d2);
        double n1 = cmv[1] + ((R[1][0] * d0 + R[1][1] * d1) + R[1][2] * d2);
        double n2 = cmv[2] + ((R[2][0] * d0 + R[2][1] * d1) + R[2][2] * d2);
        if (amc) {
            const double* x = member_pos(s, id);
            const double r[3] = {x[0] - cmp[0], x[1] - cmp[1], x[2] - cmp[2]};
            n0 += w[1] * r[2] - w[2] * r[1];
            n1 += w[2] * r[0] - w[0] * r[2];
            n2 += w[0] * r[1] - w[1] * r[0];
        }
        v[0] = n0; v[1] = n1; v[2] = n2;
    }
}

/* MPC::shiftParticles, MPC.cpp:301-311. */
void shift_particles(orc_sys& s, const V3& sh) {
#pragma omp parallel for schedule(static)
    for (uint64_t p = 0; p < s.n_fluid; ++p) {
        double* x = &s.pos[3 * p];
        for (int i = 0; i < 3; ++i) x[i] += sh[i];
        wrap(s, x, &s.vel[3 * p]);
    }
    for (auto& m : s.mono)
        for (int i = 0; i < 3; ++i) m.pos[i] += sh[i];
}

}  // namespace

extern "C" {

orc_sys* orc_create(const orc_params* p) {
    auto* s = new orc_sys;
    s->p = *p;
    for (int i = 0; i < 3; ++i) {  /* Box.cpp:14-19 */
        int l = p->L[i];
        if (l % 2) l++;
        s->L[i] = l;
        s->cell_size[i] = int(l / s->verlet_radius);  /* Box.cpp:20-25 */
        s->cell_side[i] = l / (double)s->cell_size[i];
    }
    /* MPC ctor, MPC.cpp:10-32 */
    s->number_of_cells = (uint32_t)(s->L[0] * s->L[1] * s->L[2] - 1);
    s->n_fluid = (uint64_t)p->n_c * (s->number_of_cells + 1ull);
    s->c = std::cos(p->alpha_deg * kPi / 180.0);
    s->s = std::sin(p->alpha_deg * kPi / 180.0);
    s->delrx = 0.;
    s->shear_on = p->shear != 0.;
    s->pos.assign(3 * s->n_fluid, 0.0);
    s->vel.assign(3 * s->n_fluid, 0.0);
    s->fcell.assign(s->n_fluid, 0u);
    s->cell_list.assign(s->number_of_cells + 1ull, {});
    s->cell_fluid_count.assign(s->number_of_cells + 1ull, 0u);
    return s;
}

void orc_destroy(orc_sys* s) { delete s; }

/* MPC::initialize(), MPC.cpp:34-55. */
void orc_fluid_initialize(orc_sys* s) {
    V3 cmv{};
    for (uint64_t p = 0; p < s->n_fluid; ++p) {
        double* x = &s->pos[3 * p];
        double* v = &s->vel[3 * p];
        if (s->p.rng_mode == ORC_RNG_REFERENCE) {
            for (int i = 0; i < 3; ++i) {
                x[i] = s->L[i] * (s->uniform() - 0.5);
                v[i] = s->uniform() - 0.5;
            }
        } else {
            uint32_t w0[4], w1[4];
            philox_draw(s->p.seed, s->p.replica, TAG_FLUID_INIT, (uint32_t)p, 0, 0, w0);
            philox_draw(s->p.seed, s->p.replica, TAG_FLUID_INIT, (uint32_t)p, 0, 1, w1);
            for (int i = 0; i < 3; ++i) {
                x[i] = s->L[i] * (u01(w0[i]) - 0.5);
                v[i] = u01(w1[i]) - 0.5;
            }
        }
        wrap(*s, x, v);
        for (int i = 0; i < 3; ++i) cmv[i] += v[i];
    }
    for (int i = 0; i < 3; ++i) cmv[i] /= (double)s->n_fluid;
    for (uint64_t p = 0; p < s->n_fluid; ++p)
        for (int i = 0; i < 3; ++i) s->vel[3 * p + i] -= cmv[i];
}

uint64_t orc_num_fluid(const orc_sys* s) { return s->n_fluid; }

/* OpenMP threads of the streaming, shift, cell-index and per-cell loops
 * (counter-based randomness only; the reference's generator keeps them
 * sequential); results do not depend on it */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

void orc_set_fluid(orc_sys* s, const double* pos, const double* vel, uint64_t n) {
    s->n_fluid = n;
    s->pos.assign(pos, pos + 3 * n);
    s->vel.assign(vel, vel + 3 * n);
    s->fcell.assign(n, 0u);
}

void orc_get_fluid(const orc_sys* s, double* pos, double* vel) {
    if (pos) std::memcpy(pos, s->pos.data(), sizeof(double) * 3 * s->n_fluid);
    if (vel) std::memcpy(vel, s->vel.data(), sizeof(double) * 3 * s->n_fluid);
}

/* Box::add_star (Box.cpp:39-46) -> Molecule::initialize_open_star
 * (Molecule.cpp:126-176) -> Box::set_center_of_mass_to_zero (Box.cpp:74-79)
 * -> center_of_mass_reference_frame. */
int orc_add_star(orc_sys* s, uint32_t A, uint32_t B, uint32_t arms, double mass, double bond,
                 double anchor_bond) {
    const uint32_t per_arm = A + B;
    if (per_arm < 2 || arms == 0) return -1;
    const uint32_t N = per_arm * arms + 1;
    s->mono.assign(N, Mono{});
    for (auto& m : s->mono) m.mass = mass;
    s->A = A; s->B = B; s->arms = arms;
    V3 avg{};
    unsigned n = (unsigned)std::ceil(std::sqrt(arms));
    unsigned arm_count = 0;
    s->mono[0].anchor = true;
    for (unsigned i = 0; i < n; i++) {
        double phi = 2. * kPi * i / n;
        for (unsigned j = 1; j < n + 1; j++) {
            if (arm_count >= arms) break;
            double theta = kPi * j / (n + 1);
            V3 dir{{std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta), std::cos(theta)}};
            for (unsigned k = 1; k <= per_arm; k++) {
                unsigned idx = arm_count * per_arm + k;
                if (k > A) s->mono[idx].type = true;
                if (k == 1) {
                    for (int q = 0; q < 3; ++q) s->mono[idx].pos[q] = anchor_bond * dir[q];
                    s->mono[0].neighbors.push_front((int)idx);
                    s->mono[idx].neighbors.push_front((int)idx + 1);
                } else {
                    for (int q = 0; q < 3; ++q) s->mono[idx].pos[q] = (anchor_bond + (k - 1) * bond) * dir[q];
                    if (k != per_arm) s->mono[idx].neighbors.push_front((int)idx + 1);
                }
                if (s->p.rng_mode == ORC_RNG_REFERENCE) {
                    for (int l = 0; l < 3; l++) s->mono[idx].vel[l] = s->uniform() - 0.5;
                } else {
                    uint32_t w[4];
                    philox_draw(s->p.seed, s->p.replica, TAG_STAR_INIT, idx, 0, 0, w);
                    for (int l = 0; l < 3; l++) s->mono[idx].vel[l] = u01(w[l]) - 0.5;
                }
                for (int q = 0; q < 3; ++q) avg[q] += s->mono[idx].vel[q];
            }
            arm_count++;
        }
    }
    for (int q = 0; q < 3; ++q) avg[q] /= (double)N;
    for (auto& m : s->mono)
        for (int q = 0; q < 3; ++q) m.vel[q] -= avg[q];
    double ek = mol_ekin(*s);
    double scale = std::sqrt(3. * (double)N * s->p.kT / (ek * 2.0));
    for (auto& m : s->mono)
        for (int q = 0; q < 3; ++q) m.vel[q] *= scale;
    V3 com = com_position(*s);
    for (auto& m : s->mono)
        for (int q = 0; q < 3; ++q) m.pos[q] -= com[q];
    com_reference_frame(*s);
    return 0;
}

uint32_t orc_num_monomers(const orc_sys* s) { return (uint32_t)s->mono.size(); }

void orc_set_monomers(orc_sys* s, const double* pos, const double* vel) {
    for (size_t k = 0; k < s->mono.size(); ++k)
        for (int i = 0; i < 3; ++i) {
            if (pos) s->mono[k].pos[i] = pos[3 * k + i];
            if (vel) s->mono[k].vel[i] = vel[3 * k + i];
        }
}

void orc_get_monomers(const orc_sys* s, double* pos, double* vel, double* force) {
    for (size_t k = 0; k < s->mono.size(); ++k)
        for (int i = 0; i < 3; ++i) {
            if (pos) pos[3 * k + i] = s->mono[k].pos[i];
            if (vel) vel[3 * k + i] = s->mono[k].vel[i];
            if (force) force[3 * k + i] = s->mono[k].force[i];
        }
}

void orc_get_monomer_props(const orc_sys* s, double* mass, uint8_t* type, uint8_t* anchor) {
    for (size_t k = 0; k < s->mono.size(); ++k) {
        if (mass) mass[k] = s->mono[k].mass;
        if (type) type[k] = s->mono[k].type;
        if (anchor) anchor[k] = s->mono[k].anchor;
    }
}

uint32_t orc_get_bonds(const orc_sys* s, uint32_t* pairs, uint32_t cap) {
    uint32_t nb = 0;
    for (size_t k = 0; k < s->mono.size(); ++k)
        for (int o : s->mono[k].neighbors) {
            if (pairs && nb < cap) { pairs[2 * nb] = (uint32_t)k; pairs[2 * nb + 1] = (uint32_t)o; }
            nb++;
        }
    return nb;
}

/* Thermostat_None ctor, Thermostat_None.cpp:5-12. */
int orc_md_init(orc_sys* s) {
    if (update_verlet(*s)) return -1;
    forces_verlet(*s, false);
    return 0;
}

/* Thermostat_None::propagate, Thermostat_None.cpp:21-38. */
int orc_md_propagate(orc_sys* s, int calc_epot) {
    const double dt = s->p.md_dt, dth = dt * 0.50;
    for (auto& m : s->mono)
        for (int i = 0; i < 3; ++i) {
            m.vel[i] += (m.force[i] / m.mass) * dth;
            m.pos[i] += m.vel[i] * dt;
        }
    if (check_verlet(*s)) return -1;
    forces_verlet(*s, calc_epot != 0);
    for (auto& m : s->mono)
        for (int i = 0; i < 3; ++i) m.vel[i] += (m.force[i] / m.mass) * dth;
    return 0;
}

double orc_epot(const orc_sys* s) { return s->epot; }

/* MPC::step, MPC.cpp:116-142. */
int orc_mpc_step(orc_sys* s, int64_t t, double dt) {
    s->delrx += s->p.shear * s->L[1] * dt;
    s->delrx -= s->L[0] * std::round(s->delrx / s->L[0]);
    if (t % s->p.step_update) return 0;
    check_bounds(*s);
    const bool par = s->p.rng_mode != ORC_RNG_REFERENCE;  /* mt19937 draws stay in order */
#pragma omp parallel for schedule(static) if (par)
    for (uint64_t p = 0; p < s->n_fluid; ++p) {  /* MPC::streaming, MPC.cpp:192-197 */
        double* x = &s->pos[3 * p];
        double* v = &s->vel[3 * p];
        for (int i = 0; i < 3; ++i) x[i] += v[i] * dt;
        wrap(*s, x, v);
    }
    V3 sh;
    switch (s->p.rng_mode) {
    case ORC_RNG_REFERENCE: {
        /* MPC.cpp:123 builds Vector3d(u()-.5, u()-.5, u()-.5); C++ leaves
         * the argument evaluation order unspecified and g++ (x86-64)
         * evaluates these right to left, so the first draw is z. */
        double first = s->uniform() - 0.5;
        double second = s->uniform() - 0.5;
        double third = s->uniform() - 0.5;
        sh = V3{{third, second, first}};
        break;
    }
    case ORC_RNG_PHILOX: {
        uint32_t w[4];
        philox_draw(s->p.seed, s->p.replica, TAG_SHIFT, 0, (uint32_t)s->k_coll, 0, w);
        sh = V3{{u01(w[0]) - 0.5, u01(w[1]) - 0.5, u01(w[2]) - 0.5}};
        break;
    }
    default: sh = s->inj_shift; break;
    }
    s->last_shift = sh;
    shift_particles(*s, sh);
    if (sort_cells(*s)) return -1;
    /* cells are disjoint and, with counter-based randomness, independent:
     * any thread order gives the same result */
#pragma omp parallel for schedule(dynamic, 1024) if (par)
    for (uint32_t cell = 0; cell <= s->number_of_cells; ++cell) {
        V3 cmv = calc_cmv(*s, cell);
        thermostat(*s, cell, cmv);
        collide(*s, cell, cmv);
    }
    V3 neg{{-sh[0], -sh[1], -sh[2]}};
    shift_particles(*s, neg);
    s->k_coll++;
    return 0;
}

void orc_inject(orc_sys* s, const double* shift, const double* axis_u, const double* gamma) {
    for (int i = 0; i < 3; ++i) s->inj_shift[i] = shift[i];
    size_t nc = s->number_of_cells + 1ull;
    s->inj_axis.assign(axis_u, axis_u + 2 * nc);
    s->inj_gamma.assign(gamma, gamma + nc);
}

double orc_get_delrx(const orc_sys* s) { return s->delrx; }
void orc_set_delrx(orc_sys* s, double d) { s->delrx = d; }
void orc_last_shift(const orc_sys* s, double* shift) {
    for (int i = 0; i < 3; ++i) shift[i] = s->last_shift[i];
}
int orc_last_recentered(const orc_sys* s) { return s->last_recentered; }

void orc_get_cell_index(const orc_sys* s, uint32_t* fluid, uint32_t* mono) {
    if (fluid) std::memcpy(fluid, s->fcell.data(), sizeof(uint32_t) * s->n_fluid);
    if (mono)
        for (size_t k = 0; k < s->mono.size(); ++k) mono[k] = s->mono[k].cell;
}

/* MPC::calculateEkinInCell / EkinTotal / CurrentTemperature / filledCells,
 * MPC.cpp:384-418: fluid-only relative KE about a CMV that includes
 * monomers, over the cell lists of the last sort(). */
double orc_temperature(orc_sys* s) {
    double ekin = 0.0;
    uint32_t filled = 0;
    for (uint32_t cell = 0; cell <= s->number_of_cells; ++cell) {
        V3 cmv = calc_cmv(*s, cell);
        double e = 0.0;
        uint32_t nf = s->cell_fluid_count[cell];
        for (uint32_t i = 0; i < nf; ++i) {
            const double* v = &s->vel[3 * s->cell_list[cell][i]];
            V3 d{{v[0] - cmv[0], v[1] - cmv[1], v[2] - cmv[2]}};
            e += 1.0 * dot3(d, d);
        }
        ekin += e * 0.5;
        if (nf > 0) filled++;
    }
    return 2. * ekin / (3. * ((double)s->n_fluid - filled));
}

void orc_com_frame(orc_sys* s) { com_reference_frame(*s); }

/* Box::calculate_gyration_tensor, Box.cpp:330-353 (one molecule);
 * G9 column-major like Eigen's Matrix3d::data(). */
double orc_gyration_tensor(orc_sys* s, double* G9) {
    double G[3][3] = {};
    for (auto& p : s->com_frame_pos)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) G[a][b] += p[a] * p[b];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) G[a][b] /= (double)s->mono.size();
    if (G9)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) G9[b * 3 + a] = G[a][b];
    return std::sqrt(G[0][0] + G[1][1] + G[2][2]);
}

/* Box::calculate_rotation_frequency, Box.cpp:355-379: omega = I^-1 L with
 * I = sum r^2 1 - r r^T, L = sum r x v (unit masses, as in the reference);
 * the symmetric 3x3 solve is done by cofactors. */
void orc_rotation_frequency(orc_sys* s, double* omega3) {
    double I[3][3] = {};
    double Lm[3] = {};
    for (size_t k = 0; k < s->com_frame_pos.size(); ++k) {
        const V3& r = s->com_frame_pos[k];
        const V3& v = s->com_frame_vel[k];
        double rsq = dot3(r, r);
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++) {
                I[i][j] -= r[i] * r[j];
                if (i == j) I[i][j] += rsq;
            }
        Lm[0] += r[1] * v[2] - r[2] * v[1];
        Lm[1] += r[2] * v[0] - r[0] * v[2];
        Lm[2] += r[0] * v[1] - r[1] * v[0];
    }
    I[1][0] = I[0][1]; I[2][1] = I[1][2]; I[2][0] = I[0][2];
    double c00 = I[1][1] * I[2][2] - I[1][2] * I[2][1];
    double c01 = I[1][2] * I[2][0] - I[1][0] * I[2][2];
    double c02 = I[1][0] * I[2][1] - I[1][1] * I[2][0];
    double det = I[0][0] * c00 + I[0][1] * c01 + I[0][2] * c02;
    double inv[3][3];
    inv[0][0] = c00 / det;
    inv[0][1] = (I[0][2] * I[2][1] - I[0][1] * I[2][2]) / det;
    inv[0][2] = (I[0][1] * I[1][2] - I[0][2] * I[1][1]) / det;
    inv[1][0] = c01 / det;
    inv[1][1] = (I[0][0] * I[2][2] - I[0][2] * I[2][0]) / det;
    inv[1][2] = (I[0][2] * I[1][0] - I[0][0] * I[1][2]) / det;
    inv[2][0] = c02 / det;
    inv[2][1] = (I[0][1] * I[2][0] - I[0][0] * I[2][1]) / det;
    inv[2][2] = (I[0][0] * I[1][1] - I[0][1] * I[1][0]) / det;
    for (int i = 0; i < 3; i++) omega3[i] = inv[i][0] * Lm[0] + inv[i][1] * Lm[1] + inv[i][2] * Lm[2];
}

double orc_ekin_monomers(orc_sys* s) { return mol_ekin(*s); }

/* VelocityX::operator() (Velocity_x.h:76-96) over every fluid particle, with
 * Box::COM_Pos/COM_Vel refreshed first as the driver does before applying
 * it (star-polymers.cpp:467-486; zero without a molecule).  The two
 * std::map's are returned in key order: n_out[a] entries of map a (0: keyed
 * by x, 1: by y), keys[a*cap+i], sums[(a*cap+i)*3+c] = sum vx, sum vy,
 * count. */
void orc_velocity_profile(orc_sys* s, double width, uint32_t cap, uint32_t* n_out, double* keys, double* sums) {
    com_reference_frame(*s);
    std::map<double, std::array<double, 3>> maps[2];
    for (uint64_t i = 0; i < s->n_fluid; ++i) {
        double x[3], v[3];
        for (int c = 0; c < 3; ++c) {
            x[c] = s->pos[3 * i + c] - s->com_pos[c];
            v[c] = s->vel[3 * i + c] - s->com_vel[c];
        }
        wrap(*s, x, v);
        for (int a = 0; a < 2; ++a) {
            auto& e = maps[a][std::floor(x[a] / width) * width];
            e[0] += v[0];
            e[1] += v[1];
            e[2] += 1.0;
        }
    }
    for (int a = 0; a < 2; ++a) {
        uint32_t k = 0;
        for (auto& kv : maps[a]) {
            if (k < cap) {
                keys[a * cap + k] = kv.first;
                for (int c = 0; c < 3; ++c) sums[3 * ((size_t)a * cap + k) + c] = kv.second[c];
            }
            ++k;
        }
        n_out[a] = k;
    }
}

/* Fourier_Autocorr::operator() (Fourier_Autocorr.h:55-72) over the fluid:
 * V_j = (1/N) sum_i v_i exp(i 2 pi x_ij / L), then V_j -= e_j (e_j . V_j);
 * modes[(3j+c)*2 + {0,1}] = Re, Im. */
void orc_fourier_modes(orc_sys* s, double* modes) {
    std::complex<double> V[3][3] = {};
    const std::complex<double> I(0., 1.);
    for (int j = 0; j < 3; ++j) {
        const double nk = 2 * M_PI / s->L[j];
        for (uint64_t i = 0; i < s->n_fluid; ++i) {
            const std::complex<double> ph = std::exp(I * (nk * s->pos[3 * i + j]));
            for (int c = 0; c < 3; ++c) V[j][c] += s->vel[3 * i + c] * ph;
        }
        for (int c = 0; c < 3; ++c) V[j][c] /= (double)s->n_fluid;
        V[j][j] = 0.0;
        for (int c = 0; c < 3; ++c) {
            modes[(3 * j + c) * 2] = V[j][c].real();
            modes[(3 * j + c) * 2 + 1] = V[j][c].imag();
        }
    }
}

/* VelocityDistribution::operator() (Vel_distribution.h:35-40): the map
 * floor(|v|/width)*width -> count, returned as counts of bins 0..nbins-1
 * and the rest in counts[nbins]. */
void orc_speed_histogram(orc_sys* s, double width, uint32_t nbins, uint64_t* counts) {
    std::map<double, double> dist;
    for (uint64_t i = 0; i < s->n_fluid; ++i) {
        const V3 v{{s->vel[3 * i], s->vel[3 * i + 1], s->vel[3 * i + 2]}};
        dist[std::floor(std::sqrt(dot3(v, v)) / width) * width] += 1.0;
    }
    for (uint32_t b = 0; b <= nbins; ++b) counts[b] = 0;
    for (auto& kv : dist) {
        const double b = std::round(kv.first / width);
        counts[b < nbins ? (uint32_t)b : nbins] += (uint64_t)kv.second;
    }
}

/* Box::calculate_patches_new, Box.cpp:434-507 (one molecule): arm i and arm
 * k are bonded if some B bead of i and B bead of k have negative pair
 * energy (Box::calculate_epot, Box.cpp:386-399); clusters are merged in the
 * reference's order (cluster i absorbs every later cluster that holds an
 * arm bonded to one of its arms, while walking its growing member list). */
void orc_patches(orc_sys* s, double* n_patch, double* patch_size) {
    const unsigned f = s->arms, per = s->A + s->B;
    std::vector<std::vector<bool>> bond(f, std::vector<bool>(f, false));
    auto epot = [&](unsigned a, unsigned b) {
        const Mono& p1 = s->mono[a];
        const Mono& p2 = s->mono[b];
        V3 d{{p2.pos[0] - p1.pos[0], p2.pos[1] - p1.pos[1], p2.pos[2] - p1.pos[2]}};
        const double r2 = dot3(d, d);
        double e = (p1.type && p2.type) ? bb_potential(r2, s->p.lambda) : aa_potential(r2);
        for (int n : p1.neighbors)
            if (n == (int)b) e += (p1.anchor || p2.anchor) ? fene_anchor_potential(r2) : fene_potential(r2);
        return e;
    };
    for (unsigned i = 0; i < f; ++i)
        for (unsigned k = 0; k < f; ++k) {
            if (i == k || bond[i][k]) continue;
            for (unsigned j = s->A + 1; j <= per && !bond[i][k]; ++j)
                for (unsigned l = s->A + 1; l <= per; ++l)
                    if (epot(i * per + j, k * per + l) < 0.0) { bond[i][k] = true; break; }
        }
    std::vector<std::vector<unsigned>> cl(f);
    for (unsigned i = 0; i < f; ++i) cl[i].push_back(i);
    for (unsigned i = 0; i < f; ++i) {
        for (size_t pos = 0; pos < cl[i].size(); ++pos) {
            const unsigned first = cl[i][pos];
            for (unsigned j = i + 1; j < f; ++j) {
                bool hit = false;
                for (unsigned arm : cl[j]) hit = hit || bond[first][arm];
                if (hit) {
                    cl[i].insert(cl[i].end(), cl[j].begin(), cl[j].end());
                    cl[j].clear();
                }
            }
        }
    }
    double np = 0.0, sz = 0.0;
    for (auto& c : cl)
        if (c.size() > 1) { np += 1.0; sz += (double)c.size(); }
    if (np > 1) sz /= np;
    *n_patch = np;
    *patch_size = sz;
}

double orc_fene_force(double r2) { return fene_force(r2); }
double orc_fene_anchor_force(double r2) { return fene_anchor_force(r2); }
double orc_aa_force(double r2) { return aa_force(r2); }
double orc_bb_force(double r2, double lambda) { return bb_force(r2, lambda); }
double orc_fene_potential(double r2) { return fene_potential(r2); }
double orc_fene_anchor_potential(double r2) { return fene_anchor_potential(r2); }
double orc_aa_potential(double r2) { return aa_potential(r2); }
double orc_bb_potential(double r2, double lambda) { return bb_potential(r2, lambda); }

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox10(ctr, key, out);
}

double orc_philox_gamma(uint64_t seed, uint32_t replica, uint32_t cell, uint32_t k, double shape,
                        double scale) {
    return philox_gamma(seed, replica, cell, k, shape, scale);
}

void orc_reference_rng_sequence(const int* ops, const double* a, const double* b, int n,
                                double* out) {
    orc_sys tmp;
    for (int i = 0; i < n; ++i) out[i] = ops[i] == 0 ? tmp.uniform() : tmp.gamma(a[i], b[i]);
}

}  // extern "C"
