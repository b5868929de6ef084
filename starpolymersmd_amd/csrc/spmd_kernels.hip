/*
 * spmd_kernels.hip -- HIP kernels of one MPC-MD collision interval on
 * gfx950 (CDNA4, wave64).  All kernels are memory- or latency-bound; none
 * uses MFMA.
 *
 * Per collision (MPC::step, star-polymers/lib/Hydrodynamics/MPC.cpp:116-142)
 * the library launches, on one stream, for all replicas at once:
 *   k_md_prologue  one workgroup per replica: the pending velocity-Verlet
 *                  substeps (Thermostat_None::propagate), the delrx advance,
 *                  check_bounds, the grid shift, monomer cell keys (sorted)
 *   k_bin          per solvent particle: recentre, stream, LE wrap, shifted
 *                  cell index, slot in the cell via a wave-aggregated atomic
 *   k_scan         decoupled look-back exclusive scan of the cell counts
 *   k_scatter      per solvent particle: recompute the streamed state and
 *                  its cell, write it to its slot of the cell-sorted arrays
 *   k_collide      per tile of cells: momentum and energy sums in LDS
 *                  (64-bit fixed point, order independent), MBS thermostat,
 *                  SRD rotation, write-back
 */
#pragma clang fp contract(off)

#include "spmd_internal.h"

namespace spmd {

namespace {

constexpr double kFix = 4294967296.0;          /* 2^32 */
constexpr double kUnfix = 2.3283064365386962890625e-10;
constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ unsigned long long to_fix(double v) {
    return (unsigned long long)__double2ll_rn(v * kFix);
}
__device__ __forceinline__ double from_fix(unsigned long long v) { return (double)(long long)v * kUnfix; }

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

/* Slot of each lane's particle inside its cell: one atomicAdd per distinct
 * cell in the wave (leader peeling over a ballot). */
__device__ __forceinline__ uint32_t wave_rank(uint32_t* counts, uint32_t key, bool ok) {
    uint64_t active = __ballot(ok);
    uint32_t rank = 0;
    const uint64_t lt = lanemask_lt();
    while (active) {
        const int leader = __ffsll((unsigned long long)active) - 1;
        const uint32_t lk = __shfl(key, leader);
        const uint64_t same = __ballot(ok && key == lk) & active;
        uint32_t base = 0;
        if ((int)__lane_id() == leader) base = atomicAdd(&counts[lk], (uint32_t)__popcll(same));
        base = __shfl(base, leader);
        if (ok && key == lk) rank = base + (uint32_t)__popcll(same & lt);
        active &= ~same;
    }
    return rank;
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

__global__ __launch_bounds__(kMdThreads) void k_md_prologue(DevState d, int n_md, int epot_last, int md_init,
                                                           int do_coll, int n_delrx, double delrx_dt, double dt) {
    constexpr int PER = kMaxMono / kMdThreads;
    __shared__ double sx[kMaxMono], sy[kMaxMono], sz[kMaxMono];
    __shared__ uint8_t sbt[kMaxMono];
    __shared__ double sred[kMdThreads / 64];
    __shared__ double s_shift[3];

    const int r = blockIdx.x;
    const int tid = threadIdx.x;
    const RepDev& rep = d.reps[r];
    RepCtl* ctl = d.ctl + r;
    const int N = (int)rep.n_mono;
    const uint32_t mb = rep.mono_base;
    const double dth = dt * 0.50;

    double x[PER][3], v[PER][3], f[PER][3], m[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = tid + k * kMdThreads;
        if (i < N) {
            for (int c = 0; c < 3; ++c) {
                x[k][c] = d.mpos[3 * (mb + i) + c];
                v[k][c] = d.mvel[3 * (mb + i) + c];
                f[k][c] = d.mforce[3 * (mb + i) + c];
            }
            m[k] = d.mmass[mb + i];
            sx[i] = x[k][0]; sy[i] = x[k][1]; sz[i] = x[k][2];
            sbt[i] = d.mbtype[mb + i];
        }
    }
    __syncthreads();

    /* Box::update_VerletLists (Box.cpp:105-180): every pair with r^2 <= 4
     * -- the COM-centred 27-cell grid of side >= 2 selects exactly these --
     * plus the grid bounds check that raises SIGSEGV in the reference. */
    auto rebuild = [&]() {
        double cs[3];
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < PER; ++k)
                if (tid + k * kMdThreads < N) acc += x[k][c];
            cs[c] = block_sum<kMdThreads>(acc, sred) / (double)N;
        }
        if (tid == 0) { ctl->vcom[0] = cs[0]; ctl->vcom[1] = cs[1]; ctl->vcom[2] = cs[2]; ctl->n_rebuild++; }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = tid + k * kMdThreads;
            if (i >= N) continue;
            bool oob = false;
            for (int c = 0; c < 3; ++c) {
                const int cn = (int)((x[k][c] - cs[c] + rep.box.L[c] * 0.5) / rep.vside[c]);
                oob |= cn < 0 || cn >= rep.vgrid[c];
            }
            if (oob) atomicOr(&ctl->status, ST_VERLET_OOB);
            int cnt = 0;
            uint16_t* lst = d.vl + (size_t)(mb + i) * kVcap;
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                const double dx = sx[j] - x[k][0], dy = sy[j] - x[k][1], dz = sz[j] - x[k][2];
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                if (r2 <= 4.0) {
                    if (cnt < kVcap) lst[cnt] = (uint16_t)j;
                    cnt++;
                }
            }
            if (cnt > kVcap) atomicOr(&ctl->status, ST_VERLET_OVERFLOW);
            d.vl_n[mb + i] = (uint16_t)(cnt < kVcap ? cnt : kVcap);
            for (int c = 0; c < 3; ++c) d.mvpos[3 * (mb + i) + c] = x[k][c];
        }
    };

    /* Box::calculate_forces_verlet (Box.cpp:243-286). */
    auto forces = [&](bool calc_epot) {
        double ep = 0.0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = tid + k * kMdThreads;
            if (i >= N) continue;
            double fx = 0.0, fy = 0.0, fz = 0.0;
            const bool bi = sbt[i] != 0;
            const int nvl = d.vl_n[mb + i];
            const uint16_t* lst = d.vl + (size_t)(mb + i) * kVcap;
            for (int q = 0; q < nvl; ++q) {
                const int j = lst[q];
                const double dx = sx[j] - x[k][0], dy = sy[j] - x[k][1], dz = sz[j] - x[k][2];
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                const bool bb = bi && sbt[j] != 0;
                const double fa = pair_force(r2, bb, rep.lambda);
                fx -= dx * fa; fy -= dy * fa; fz -= dz * fa;
                if (calc_epot) ep += 0.5 * pair_potential(r2, bb, rep.lambda);
            }
            const uint32_t a0 = d.adj_off[rep.adj_base + i], a1 = d.adj_off[rep.adj_base + i + 1];
            for (uint32_t q = a0; q < a1; ++q) {
                const uint32_t e = d.adj[q];
                const int j = (int)(e & ~kBondAnchor);
                const bool anchor = (e & kBondAnchor) != 0;
                const double dx = sx[j] - x[k][0], dy = sy[j] - x[k][1], dz = sz[j] - x[k][2];
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                const double fa = anchor ? fene_anchor_force(r2) : fene_force(r2);
                fx -= dx * fa; fy -= dy * fa; fz -= dz * fa;
                if (calc_epot) ep += 0.5 * (anchor ? fene_anchor_potential(r2) : fene_potential(r2));
            }
            f[k][0] = fx; f[k][1] = fy; f[k][2] = fz;
        }
        const double tot = block_sum<kMdThreads>(ep, sred);
        if (tid == 0) ctl->epot = calc_epot ? tot : 0.0;
    };

    if (N > 0) {
        if (md_init) {
            rebuild();
            forces(false);
        }
        for (int s = 0; s < n_md; ++s) {
            /* Thermostat_None::propagate (Thermostat_None.cpp:21-38) */
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = tid + k * kMdThreads;
                if (i >= N) continue;
                for (int c = 0; c < 3; ++c) {
                    v[k][c] += (f[k][c] / m[k]) * dth;
                    x[k][c] += v[k][c] * dt;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = tid + k * kMdThreads;
                if (i < N) { sx[i] = x[k][0]; sy[i] = x[k][1]; sz[i] = x[k][2]; }
            }
            /* Box::check_VerletLists (Box.cpp:182-194) */
            bool need = false;
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = tid + k * kMdThreads;
                if (i >= N) continue;
                const double dx = x[k][0] - d.mvpos[3 * (mb + i)];
                const double dy = x[k][1] - d.mvpos[3 * (mb + i) + 1];
                const double dz = x[k][2] - d.mvpos[3 * (mb + i) + 2];
                need |= sqrt((dx * dx + dy * dy) + dz * dz) > (2.0 - 1.5) * 0.5;
            }
            if (__syncthreads_or(need)) {
                rebuild();
                __syncthreads();
            }
            forces(epot_last && s == n_md - 1);
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = tid + k * kMdThreads;
                if (i >= N) continue;
                for (int c = 0; c < 3; ++c) v[k][c] += (f[k][c] / m[k]) * dth;
            }
            __syncthreads();
        }
    }

    if (tid == 0) {
        /* MPC::step delrx advance, MPC.cpp:117-118, once per call */
        double dl = ctl->delrx;
        for (int q = 0; q < n_delrx; ++q) {
            dl += rep.box.shear * rep.box.L[1] * delrx_dt;
            dl -= rep.box.L[0] * round(dl / rep.box.L[0]);
        }
        ctl->delrx = dl;
    }

    if (do_coll) {
        /* MPC::check_bounds, MPC.cpp:160-190 */
        bool out = false;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = tid + k * kMdThreads;
            if (i >= N) continue;
            for (int c = 0; c < 3; ++c) out |= fabs(x[k][c]) >= (rep.box.L[c] * 0.5 - 0.5);
        }
        const bool rec = __syncthreads_or(out) != 0;
        if (rec) {
            double cp[3], cv[3];
            for (int c = 0; c < 3; ++c) {
                double ap = 0.0, av = 0.0;
#pragma unroll
                for (int k = 0; k < PER; ++k)
                    if (tid + k * kMdThreads < N) { ap += x[k][c]; av += v[k][c]; }
                cp[c] = block_sum<kMdThreads>(ap, sred) / (double)N;
                cv[c] = block_sum<kMdThreads>(av, sred) / (double)N;
            }
#pragma unroll
            for (int k = 0; k < PER; ++k)
                if (tid + k * kMdThreads < N)
                    for (int c = 0; c < 3; ++c) { x[k][c] -= cp[c]; v[k][c] -= cv[c]; }
            if (tid == 0)
                for (int c = 0; c < 3; ++c) { ctl->com[c] = cp[c]; ctl->com_vel[c] = cv[c]; }
        }
        if (tid == 0) {
            ctl->recenter = rec ? 1 : 0;
            const uint32_t k = ctl->k_done;
            if (ctl->inj_pending) {
                for (int c = 0; c < 3; ++c) s_shift[c] = ctl->inj_shift[c];
                ctl->inj_active = 1;
                ctl->inj_pending = 0;
            } else {
                uint32_t w[4];
                draw4(rep.seed, rep.replica, TAG_SHIFT, 0, k, 0, w);
                for (int c = 0; c < 3; ++c) s_shift[c] = u01(w[c]) - 0.5;
                ctl->inj_active = 0;
            }
            for (int c = 0; c < 3; ++c) ctl->shift[c] = s_shift[c];
            ctl->k_cur = k;
            ctl->k_done = k + 1;
        }
        __syncthreads();
        /* monomer cell keys (MPC::shiftParticles MPC.cpp:306-310 + sort
         * MPC.cpp:221-235: shifted, never wrapped), bitonic-sorted */
        uint64_t* keys = reinterpret_cast<uint64_t*>(sx);
        int np2 = 1;
        while (np2 < N) np2 <<= 1;
        for (int i = tid; i < np2; i += kMdThreads) keys[i] = ~0ull;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = tid + k * kMdThreads;
            if (i >= N) continue;
            const uint32_t ci = cell_index(rep.box, x[k][0] + s_shift[0], x[k][1] + s_shift[1], x[k][2] + s_shift[2]);
            if (ci >= rep.ncells) atomicOr(&ctl->status, ST_CELL_OOB);
            const uint64_t g = ci < rep.ncells ? rep.cell_base + ci : 0xFFFFFFFFull;
            keys[i] = (g << 32) | (uint32_t)i;
        }
        __syncthreads();
        for (int kk = 2; kk <= np2; kk <<= 1) {
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int i = tid; i < np2; i += kMdThreads) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const uint64_t a = keys[i], b2 = keys[ixj];
                        const bool up = (i & kk) == 0;
                        if ((a > b2) == up) { keys[i] = b2; keys[ixj] = a; }
                    }
                }
                __syncthreads();
            }
        }
        for (int i = tid; i < N; i += kMdThreads) d.msorted[mb + i] = keys[i];
        if (r == 0) {
            for (uint32_t t = tid; t < d.n_scan_tiles; t += kMdThreads) d.scan_flags[t] = 0ull;
            if (tid == 0) *d.scan_ticket = 0u;
        }
    }

#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = tid + k * kMdThreads;
        if (i >= N) continue;
        for (int c = 0; c < 3; ++c) {
            d.mpos[3 * (mb + i) + c] = x[k][c];
            d.mvel[3 * (mb + i) + c] = v[k][c];
            d.mforce[3 * (mb + i) + c] = f[k][c];
        }
    }
}

/* ============================================================== bin */

template <int PPT>
__global__ __launch_bounds__(256) void k_bin(DevState d, double h) {
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const int r = find_rep(d.pblk_base, d.n_rep, lb);
    const RepDev& rep = d.reps[r];
    const RepCtl c = d.ctl[r];
    Box b = rep.box;
    b.delrx = c.delrx;
    const uint64_t p0 = rep.part_base + (uint64_t)(lb - d.pblk_base[r]) * (256 * PPT);
    const uint64_t pend = rep.part_base + rep.n_fluid;
    const int cur = d.cur;
    const float *X = d.x[cur], *Y = d.y[cur], *Z = d.z[cur];
    const float *VX = d.vx[cur], *VY = d.vy[cur], *VZ = d.vz[cur];
    bool bad = false, big = false;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const uint64_t i = p0 + (uint64_t)j * 256 + threadIdx.x;
        const bool valid = i < pend;
        uint32_t cell = 0;
        bool ok = false;
        if (valid) {
            double x = X[i], y = Y[i], z = Z[i], vx = VX[i], vy = VY[i], vz = VZ[i];
            recenter_stream(b, h, c, x, y, z, vx, vy, vz);
            cell = shifted_cell(b, c, x, y, z, vx);
            ok = cell < rep.ncells;
            bad |= !ok;
        }
        const uint32_t rank = wave_rank(d.counts, (uint32_t)rep.cell_base + cell, ok);
        big |= rank > 0xFFFFu;
        if (valid) d.rank[i] = (uint16_t)(rank > 0xFFFFu ? 0xFFFFu : rank);
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
    if (tid == 0) {
        uint32_t prefix = 0;
        if (tile == 0) {
            __hip_atomic_store(&d.scan_flags[0], (2ull << 62) | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&d.scan_flags[tile], (1ull << 62) | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            int64_t p = (int64_t)tile - 1;
            uint32_t spins = 0;
            while (p >= 0) {
                const uint64_t f = __hip_atomic_load(&d.scan_flags[p], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t st = f >> 62;
                if (st == 0) {
                    if (++spins > (1u << 24)) {
                        atomicOr(&d.ctl[0].status, ST_SCAN_TIMEOUT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix += (uint32_t)(f & 0xFFFFFFFFull);
                if (st == 2) break;
                --p;
            }
            __hip_atomic_store(&d.scan_flags[tile], (2ull << 62) | (uint64_t)(prefix + agg), __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_prefix = prefix;
        if (base <= d.c_tot && (uint64_t)(tile + 1) * kScanTile >= d.c_tot) d.offsets[d.c_tot] = prefix + agg;
    }
    __syncthreads();
    const uint32_t off = s_prefix + texcl;
    if (base + kScanItems <= d.c_tot) {
        uint4* po = reinterpret_cast<uint4*>(d.offsets + base);
        uint4* pc = reinterpret_cast<uint4*>(d.counts + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            po[q] = make_uint4(off + v[4 * q], off + v[4 * q + 1], off + v[4 * q + 2], off + v[4 * q + 3]);
            pc[q] = make_uint4(0u, 0u, 0u, 0u);
        }
    } else {
        for (int q = 0; q < kScanItems; ++q)
            if (base + q < d.c_tot) { d.offsets[base + q] = off + v[q]; d.counts[base + q] = 0u; }
    }
}

/* =========================================================== scatter */

template <int PPT>
__global__ __launch_bounds__(256) void k_scatter(DevState d, double h) {
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const int r = find_rep(d.pblk_base, d.n_rep, lb);
    const RepDev& rep = d.reps[r];
    const RepCtl c = d.ctl[r];
    Box b = rep.box;
    b.delrx = c.delrx;
    const uint64_t p0 = rep.part_base + (uint64_t)(lb - d.pblk_base[r]) * (256 * PPT);
    const uint64_t pend = rep.part_base + rep.n_fluid;
    const int cur = d.cur, alt = cur ^ 1;
    const float *X = d.x[cur], *Y = d.y[cur], *Z = d.z[cur];
    const float *VX = d.vx[cur], *VY = d.vy[cur], *VZ = d.vz[cur];
    float *OX = d.x[alt], *OY = d.y[alt], *OZ = d.z[alt];
    float *OVX = d.vx[alt], *OVY = d.vy[alt], *OVZ = d.vz[alt];
    const uint32_t* ids = d.ids[cur];
    uint32_t* oids = d.ids[alt];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const uint64_t i = p0 + (uint64_t)j * 256 + threadIdx.x;
        if (i >= pend) continue;
        double x = X[i], y = Y[i], z = Z[i], vx = VX[i], vy = VY[i], vz = VZ[i];
        recenter_stream(b, h, c, x, y, z, vx, vy, vz);
        const uint32_t cell = shifted_cell(b, c, x, y, z, vx);
        if (cell >= rep.ncells) continue;
        const uint32_t dst = d.offsets[(uint32_t)rep.cell_base + cell] + d.rank[i];
        if (dst >= pend) continue;  /* cannot happen when bin and scatter agree */
        OX[dst] = (float)x; OY[dst] = (float)y; OZ[dst] = (float)z;
        OVX[dst] = (float)vx; OVY[dst] = (float)vy; OVZ[dst] = (float)vz;
        if (ids) oids[dst] = ids[i];
    }
}

/* =========================================================== collide */

/* One workgroup per tile of TC consecutive cells of one replica.  Particles
 * of the tile are contiguous in the sorted arrays; up to CAP of them are
 * staged in LDS, larger tiles re-read global memory on each pass. */
template <int TC, int CAP>
__global__ __launch_bounds__(256) void k_collide(DevState d) {
    constexpr int NT = 256;
    __shared__ uint32_t s_off[TC + 1];
    __shared__ unsigned long long s_px[TC], s_py[TC], s_pz[TC], s_e[TC], s_mm[TC];
    __shared__ uint32_t s_nm[TC];
    __shared__ double s_cmv[TC][3];
    __shared__ double s_R[TC][9];
    __shared__ double s_scale[TC];
    __shared__ float s_vx[CAP], s_vy[CAP], s_vz[CAP];
    __shared__ uint16_t s_lc[CAP];
    __shared__ uint32_t s_mrange[2];

    const int tid = threadIdx.x;
    const uint32_t lt = xcd_remap(blockIdx.x, gridDim.x);
    const int r = find_rep(d.tile_base, d.n_rep, lt);
    const RepDev& rep = d.reps[r];
    const RepCtl* ctl = d.ctl + r;
    const uint32_t c_lo = (lt - d.tile_base[r]) * TC;
    const uint32_t nc = min((uint32_t)TC, rep.ncells - c_lo);
    const uint32_t gc0 = (uint32_t)rep.cell_base + c_lo;
    const int cur = d.cur ^ 1; /* the scatter wrote here */
    float *VX = d.vx[cur], *VY = d.vy[cur], *VZ = d.vz[cur];
    const float* Y = d.y[cur];
    const int Ly = rep.box.iL[1], Lx = rep.box.iL[0];
    const bool shear_on = rep.box.shear_on != 0;
    const double shear_Ly_unit = rep.box.shear;  /* cy*shear*Ly evaluated as (cy*shear)*Ly */
    const double LyD = rep.box.L[1];

    for (uint32_t i = tid; i <= nc; i += NT) s_off[i] = d.offsets[gc0 + i];
    for (uint32_t i = tid; i < nc; i += NT) {
        s_px[i] = s_py[i] = s_pz[i] = s_e[i] = s_mm[i] = 0ull;
        s_nm[i] = 0u;
    }
    if (tid == 0) {
        /* monomers of this tile: binary search in the sorted keys */
        const uint64_t* ms = d.msorted + rep.mono_base;
        const uint32_t nm = rep.n_mono;
        uint32_t lo = 0, hi = nm;
        const uint64_t klo = (uint64_t)gc0 << 32;
        while (lo < hi) { const uint32_t mid = (lo + hi) >> 1; if (ms[mid] < klo) lo = mid + 1; else hi = mid; }
        uint32_t lo2 = lo, hi2 = nm;
        const uint64_t khi = (uint64_t)(gc0 + nc) << 32;
        while (lo2 < hi2) { const uint32_t mid = (lo2 + hi2) >> 1; if (ms[mid] < khi) lo2 = mid + 1; else hi2 = mid; }
        s_mrange[0] = lo; s_mrange[1] = lo2;
    }
    __syncthreads();
    const uint32_t p0 = s_off[0], np = s_off[nc] - p0;
    const bool fast = np <= (uint32_t)CAP;
    const uint32_t m0 = s_mrange[0], m1 = s_mrange[1];

    auto cell_of = [&](uint32_t q) -> uint32_t {
        if (fast) return s_lc[q] & 0x0FFFu;
        uint32_t lo = 0, hi = nc;  /* last cell with s_off[c] <= p0+q */
        const uint32_t p = p0 + q;
        while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (s_off[mid] <= p) lo = mid; else hi = mid; }
        return lo;
    };
    /* y image index of the shifted image: only cells on the y faces can
     * hold particles that the +shift wrap moved across (valid for Ly >= 4) */
    auto image_cy = [&](uint32_t q, uint32_t lc) -> int {
        if (!shear_on) return 0;
        if (fast) return (int)(s_lc[q] >> 12) - 1;
        const uint32_t gcell = c_lo + lc;
        const int iy = (int)((gcell / (uint32_t)Lx) % (uint32_t)Ly);
        if (iy == 0 && Y[p0 + q] > 0.0f) return 1;
        if (iy == Ly - 1 && Y[p0 + q] < 0.0f) return -1;
        return 0;
    };
    auto load_v = [&](uint32_t q, double& vx, double& vy, double& vz) {
        if (fast) { vx = s_vx[q]; vy = s_vy[q]; vz = s_vz[q]; }
        else { vx = VX[p0 + q]; vy = VY[p0 + q]; vz = VZ[p0 + q]; }
    };

    if (fast) {
        for (uint32_t ci = tid; ci < nc; ci += NT) {
            const uint32_t a = s_off[ci] - p0, e = s_off[ci + 1] - p0;
            int tag = 1;
            if (shear_on) {
                const uint32_t gcell = c_lo + ci;
                const int iy = (int)((gcell / (uint32_t)Lx) % (uint32_t)Ly);
                tag = iy == 0 ? 3 : (iy == Ly - 1 ? 4 : 1);
            }
            for (uint32_t q = a; q < e; ++q) s_lc[q] = (uint16_t)(ci | (tag << 12));
        }
        __syncthreads();
        for (uint32_t q = tid; q < np; q += NT) {
            s_vx[q] = VX[p0 + q]; s_vy[q] = VY[p0 + q]; s_vz[q] = VZ[p0 + q];
            const uint32_t t = s_lc[q] >> 12;
            int cy = 0;
            if (t == 3 && Y[p0 + q] > 0.0f) cy = 1;
            if (t == 4 && Y[p0 + q] < 0.0f) cy = -1;
            s_lc[q] = (uint16_t)((s_lc[q] & 0x0FFFu) | ((cy + 1) << 12));
        }
        __syncthreads();
    }

    /* MPC::calculateCMV (MPC.cpp:314-322): sum m*v and m per cell */
    for (uint32_t q = tid; q < np; q += NT) {
        const uint32_t lc = cell_of(q);
        const int cy = image_cy(q, lc);
        double vx, vy, vz;
        load_v(q, vx, vy, vz);
        vx -= (double)cy * shear_Ly_unit * LyD;
        atomicAdd(&s_px[lc], to_fix(vx));
        atomicAdd(&s_py[lc], to_fix(vy));
        atomicAdd(&s_pz[lc], to_fix(vz));
    }
    for (uint32_t j = m0 + tid; j < m1; j += NT) {
        const uint64_t key = d.msorted[rep.mono_base + j];
        const uint32_t lc = (uint32_t)(key >> 32) - gc0;
        const uint32_t mi = rep.mono_base + (uint32_t)key;
        const double mm = d.mmass[mi];
        atomicAdd(&s_px[lc], to_fix(d.mvel[3 * mi] * mm));
        atomicAdd(&s_py[lc], to_fix(d.mvel[3 * mi + 1] * mm));
        atomicAdd(&s_pz[lc], to_fix(d.mvel[3 * mi + 2] * mm));
        atomicAdd(&s_mm[lc], to_fix(mm));
        atomicAdd(&s_nm[lc], 1u);
    }
    __syncthreads();
    for (uint32_t ci = tid; ci < nc; ci += NT) {
        const double M = (double)(s_off[ci + 1] - s_off[ci]) + from_fix(s_mm[ci]);
        double c0 = from_fix(s_px[ci]), c1 = from_fix(s_py[ci]), c2 = from_fix(s_pz[ci]);
        if (M > 0) { c0 /= M; c1 /= M; c2 /= M; }
        s_cmv[ci][0] = c0; s_cmv[ci][1] = c1; s_cmv[ci][2] = c2;
    }
    __syncthreads();

    /* MPC::thermostat relative kinetic energy (MPC.cpp:277-279) */
    for (uint32_t q = tid; q < np; q += NT) {
        const uint32_t lc = cell_of(q);
        const int cy = image_cy(q, lc);
        double vx, vy, vz;
        load_v(q, vx, vy, vz);
        vx -= (double)cy * shear_Ly_unit * LyD;
        const double dx = vx - s_cmv[lc][0], dy = vy - s_cmv[lc][1], dz = vz - s_cmv[lc][2];
        atomicAdd(&s_e[lc], to_fix((dx * dx + dy * dy) + dz * dz));
    }
    for (uint32_t j = m0 + tid; j < m1; j += NT) {
        const uint64_t key = d.msorted[rep.mono_base + j];
        const uint32_t lc = (uint32_t)(key >> 32) - gc0;
        const uint32_t mi = rep.mono_base + (uint32_t)key;
        const double dx = d.mvel[3 * mi] - s_cmv[lc][0];
        const double dy = d.mvel[3 * mi + 1] - s_cmv[lc][1];
        const double dz = d.mvel[3 * mi + 2] - s_cmv[lc][2];
        atomicAdd(&s_e[lc], to_fix(d.mmass[mi] * ((dx * dx + dy * dy) + dz * dz)));
    }
    __syncthreads();

    /* per cell: MBS scale (MPC.cpp:287-289) and rotation (MPC.cpp:244-257) */
    const uint32_t k = ctl->k_cur;
    const bool inj = ctl->inj_active != 0;
    for (uint32_t ci = tid; ci < nc; ci += NT) {
        const uint32_t cell = c_lo + ci;
        const uint32_t n = (s_off[ci + 1] - s_off[ci]) + s_nm[ci];
        double scale = 1.0;
        if (n >= 2) {
            const double ek = from_fix(s_e[ci]) * 0.5;
            const double g = inj ? d.inj_gamma[gc0 + ci]
                                 : gamma_mt(rep.seed, rep.replica, cell, k, 1.5 * (n - 1), rep.kT);
            scale = sqrt(g / ek);
        }
        s_scale[ci] = n >= 2 ? scale : -1.0;
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
        const double phi = 2. * kPi * u1;
        const double th = 2. * (u2 - 0.5);
        const double a0 = sqrt(1 - th * th) * cos(phi);
        const double a1 = sqrt(1 - th * th) * sin(phi);
        const double a2 = th;
        const double cc = rep.c, ss = rep.s;
        double* R = s_R[ci];
        R[0] = a0 * a0 + (1 - a0 * a0) * cc;
        R[1] = a0 * a1 * (1 - cc) - a2 * ss;
        R[2] = a0 * a2 * (1 - cc) + a1 * ss;
        R[3] = a0 * a1 * (1 - cc) + a2 * ss;
        R[4] = a1 * a1 + (1 - a1 * a1) * cc;
        R[5] = a1 * a2 * (1 - cc) - a0 * ss;
        R[6] = a0 * a2 * (1 - cc) - a1 * ss;
        R[7] = a1 * a2 * (1 - cc) + a0 * ss;
        R[8] = a2 * a2 + (1 - a2 * a2) * cc;
    }
    __syncthreads();

    /* v <- s v + (1-s) CMV, then v <- CMV + R (v - CMV) (MPC.cpp:290-292,
     * 266-267); the image offset is undone as the unshift wrap does */
    auto apply = [&](uint32_t lc, double& vx, double& vy, double& vz) {
        const double c0 = s_cmv[lc][0], c1 = s_cmv[lc][1], c2 = s_cmv[lc][2];
        const double sc = s_scale[lc];
        if (sc >= 0.0) {
            vx = sc * vx + (1 - sc) * c0;
            vy = sc * vy + (1 - sc) * c1;
            vz = sc * vz + (1 - sc) * c2;
        }
        const double* R = s_R[lc];
        const double d0 = vx - c0, d1 = vy - c1, d2 = vz - c2;
        vx = c0 + ((R[0] * d0 + R[1] * d1) + R[2] * d2);
        vy = c1 + ((R[3] * d0 + R[4] * d1) + R[5] * d2);
        vz = c2 + ((R[6] * d0 + R[7] * d1) + R[8] * d2);
    };
    for (uint32_t q = tid; q < np; q += NT) {
        const uint32_t lc = cell_of(q);
        const int cy = image_cy(q, lc);
        double vx, vy, vz;
        load_v(q, vx, vy, vz);
        const double off = (double)cy * shear_Ly_unit * LyD;
        vx -= off;
        apply(lc, vx, vy, vz);
        vx += off;
        VX[p0 + q] = (float)vx; VY[p0 + q] = (float)vy; VZ[p0 + q] = (float)vz;
    }
    for (uint32_t j = m0 + tid; j < m1; j += NT) {
        const uint64_t key = d.msorted[rep.mono_base + j];
        const uint32_t lc = (uint32_t)(key >> 32) - gc0;
        const uint32_t mi = rep.mono_base + (uint32_t)key;
        double vx = d.mvel[3 * mi], vy = d.mvel[3 * mi + 1], vz = d.mvel[3 * mi + 2];
        apply(lc, vx, vy, vz);
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
    const uint32_t lt = d.tile_base[replica] + blockIdx.x;
    const RepDev& rep = d.reps[replica];
    const uint32_t c_lo = blockIdx.x * TC;
    const uint32_t gc0 = (uint32_t)rep.cell_base + c_lo;
    const uint32_t nc = min((uint32_t)TC, rep.ncells - c_lo);
    const int cur = d.cur;
    const float *VX = d.vx[cur], *VY = d.vy[cur], *VZ = d.vz[cur];
    const uint64_t* ms = d.msorted + rep.mono_base;
    double ek = 0.0, filled = 0.0;
    for (uint32_t ci = threadIdx.x; ci < nc; ci += 256) {
        const uint32_t a = d.offsets[gc0 + ci], e = d.offsets[gc0 + ci + 1];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, M = 0.0;
        for (uint32_t p = a; p < e; ++p) { s0 += VX[p]; s1 += VY[p]; s2 += VZ[p]; M += 1.0; }
        /* monomers of this cell from the last collision's sorted keys */
        const uint64_t key = (uint64_t)(gc0 + ci) << 32;
        uint32_t lo = 0, hi = rep.n_mono;
        while (lo < hi) { const uint32_t mid = (lo + hi) >> 1; if (ms[mid] < key) lo = mid + 1; else hi = mid; }
        for (uint32_t j = lo; j < rep.n_mono && (ms[j] >> 32) == (uint64_t)(gc0 + ci); ++j) {
            const uint32_t mi = rep.mono_base + (uint32_t)ms[j];
            const double mm = d.mmass[mi];
            s0 += d.mvel[3 * mi] * mm; s1 += d.mvel[3 * mi + 1] * mm; s2 += d.mvel[3 * mi + 2] * mm;
            M += mm;
        }
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

/* ========================================================= launchers */

hipError_t launch_md_prologue(const DevState& d, hipStream_t st, int n_md, int epot_last, int md_init,
                              int do_collision, int n_delrx, double delrx_dt, double md_dt) {
    hipLaunchKernelGGL(k_md_prologue, dim3(d.n_rep), dim3(kMdThreads), 0, st, d, n_md, epot_last, md_init,
                       do_collision, n_delrx, delrx_dt, md_dt);
    return hipGetLastError();
}

hipError_t launch_bin(const DevState& d, hipStream_t st, double h) {
    const dim3 g(d.n_pblk), b(256);
    switch (d.ppt) {
    case 1: hipLaunchKernelGGL(k_bin<1>, g, b, 0, st, d, h); break;
    case 2: hipLaunchKernelGGL(k_bin<2>, g, b, 0, st, d, h); break;
    case 4: hipLaunchKernelGGL(k_bin<4>, g, b, 0, st, d, h); break;
    case 8: hipLaunchKernelGGL(k_bin<8>, g, b, 0, st, d, h); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_scan(const DevState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_scan, dim3(d.n_scan_tiles), dim3(kScanThreads), 0, st, d);
    return hipGetLastError();
}

hipError_t launch_scatter(const DevState& d, hipStream_t st, double h) {
    const dim3 g(d.n_pblk), b(256);
    switch (d.ppt) {
    case 1: hipLaunchKernelGGL(k_scatter<1>, g, b, 0, st, d, h); break;
    case 2: hipLaunchKernelGGL(k_scatter<2>, g, b, 0, st, d, h); break;
    case 4: hipLaunchKernelGGL(k_scatter<4>, g, b, 0, st, d, h); break;
    case 8: hipLaunchKernelGGL(k_scatter<8>, g, b, 0, st, d, h); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_collide(const DevState& d, hipStream_t st) {
    const dim3 g(d.n_tiles), b(256);
    switch (d.tile_cells) {
    case 64: hipLaunchKernelGGL((k_collide<64, 1024>), g, b, 0, st, d); break;
    case 128: hipLaunchKernelGGL((k_collide<128, 2048>), g, b, 0, st, d); break;
    case 256: hipLaunchKernelGGL((k_collide<256, 3584>), g, b, 0, st, d); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fluid_init(const DevState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_fluid_init, dim3(d.n_pblk), dim3(256), 0, st, d, 0);
    hipLaunchKernelGGL(k_fluid_mean, dim3(d.n_rep), dim3(64), 0, st, d);
    hipLaunchKernelGGL(k_fluid_init, dim3(d.n_pblk), dim3(256), 0, st, d, 1);
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

}  // namespace spmd
