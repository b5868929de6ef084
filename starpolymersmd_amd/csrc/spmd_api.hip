/*
 * spmd_api.hip -- C ABI implementation (include/spmd.h): handle and device
 * memory management, the lazy batching of Thermostat_None::propagate calls,
 * MPC::step semantics (delrx cadence, collision every step_update calls),
 * HIP-graph replay of whole collision intervals, parity hooks, observables.
 */
#pragma clang fp contract(off)

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spmd.h"
#include "spmd_internal.h"

using namespace spmd;

namespace {

struct StarHost {
    uint32_t n = 0;
    uint32_t arms = 0, A = 0, B = 0;
    std::vector<double> pos, vel, force, mass;
    std::vector<uint8_t> btype, anchor;
    std::vector<std::pair<uint32_t, uint32_t>> bonds;  /* directed (owner, other) */
};

struct ProfEvents {
    std::vector<hipEvent_t> ev;  /* start/stop pairs */
};

constexpr int kGraphIntervals = 8;  /* row-tile intervals per graph launch */
constexpr int kProfKinds = 9;
const char* kProfNames[kProfKinds] = {"md", "bin", "scan", "perm", "collide", "prologue", "advect", "redo", "gamma"};

/* Row tiles: a tile's mean occupancy stays below these (cells x density),
 * so that its particle capacity (1024 / 2048) lies more than 12 standard
 * deviations of a Poisson count above the mean. */
constexpr double kRowTileMeanSmall = 640.0;
constexpr double kRowTileMean = 1300.0;

/* Device scratch of spmd_velocity_profile_noalloc: the profile's fixed-point
 * bins [2][kObsBins][3], then kStarMomentsWords doubles for the star's
 * moments (its COM).  One copy per GPU, held until the process ends
 * (49,408 bytes); every call on a GPU shares it. */
constexpr size_t kStarMomentsWords = 32;
constexpr size_t kScratchWords = 6ull * kObsBins + kStarMomentsWords;
static_assert(kStarMomentsWords >= (size_t)kStarMoments, "moments do not fit the scratch");
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

/* Its address on each device, looked up once per device with that device
 * current (the host-side name `scratch` is not a device address). */
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

struct spmd_sys {
    std::vector<spmd_params> params;
    std::vector<RepDev> reps;
    std::vector<StarHost> stars;
    DevState d{};
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;    /* MD kernel beside the speculative solvent sort */
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int device = 0;
    std::string err;
    bool tracking = false;
    bool mono_dirty = true;        /* device monomer buffers need (re)building */
    bool md_inited = false;        /* spmd_md_init ran: a monomer rebuild redoes it */
    unsigned long long* scratch = nullptr;  /* this device's copy of the static scratch */
    uint32_t step_update = 1;
    double md_dt = 0.01;
    /* lazily batched work */
    int pending_md = 0;
    int pending_epot = 0;
    int pending_delrx = 0;
    double pending_dt = 0.0;
    bool collided = false;
    /* graphs of one collision interval, per starting buffer; on the row-tile
     * path also of kGraphIntervals intervals in a row (one launch for all) */
    hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  /* [path][starting buffer] */
    hipGraphExec_t graph_run[2] = {nullptr, nullptr};                       /* [starting buffer] */
    /* collision path: 0 automatic (row tiles when the buffer is in cell
     * order and the replicas allow it), 1 always the sort path */
    int path_pref = 0;
    bool tc_dirty = true;          /* row-tile arrival counters must be cleared first */
    double graph_h = -1.0;
    /* profiling */
    bool prof = false;
    ProfEvents pev[kProfKinds];
    std::vector<size_t> alloc_bytes;
    /* learned closure workspace */
    float* cl_feat = nullptr;      /* m_tot x 24 */
    float* cl_out = nullptr;       /* m_tot x 6 */
    float* cl_range = nullptr;     /* fmin[24], fmax[24] */
    uint32_t cl_cap = 0;
    uint32_t cl_k = 0;             /* closure interval counter (Philox) */
    double* mom = nullptr;         /* star moments (kStarMoments) and patches (2) */
    unsigned long long* obs = nullptr;  /* solvent observer sums */
    double* cellv = nullptr;       /* per-cell mean solvent velocities (spmd_cell_fluid_velocity) */
    size_t cellv_n = 0;
    /* fluid velocity autocorrelation per replica: windows by particle id,
     * lag sums (64-bit fixed point), samples taken and sums made */
    struct Vacf {
        float* hist = nullptr;
        unsigned long long* sums = nullptr;
        uint32_t entries = 0;
        uint64_t calls = 0, counter = 0;
    };
    std::vector<Vacf> vacf;
};

namespace {

int fail(spmd_sys* s, int code, const std::string& msg) {
    if (s) s->err = msg;
    return code;
}

#define HIPCHK(s, call)                                                                        \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail((s), SPMD_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(spmd_sys* s, T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(s, SPMD_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return SPMD_OK;
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void clear_graphs(spmd_sys* s) {
    for (auto& gp : s->graph)
        for (auto& g : gp)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    for (auto& g : s->graph_run)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

int upload_tables(spmd_sys* s) {
    DevState& d = s->d;
    const uint32_t ppb = 256u * (uint32_t)d.ppt;
    std::vector<uint32_t> pb(d.n_rep + 1, 0), tb(d.n_rep + 1, 0);
    for (int r = 0; r < d.n_rep; ++r) {
        pb[r + 1] = pb[r] + (s->reps[r].n_fluid + ppb - 1) / ppb;
        tb[r + 1] = tb[r] + (s->reps[r].ncells + d.tile_cells - 1) / d.tile_cells;
    }
    d.n_pblk = pb[d.n_rep];
    d.n_tiles = tb[d.n_rep];
    HIPCHK(s, hipMemcpy(d.pblk_base, pb.data(), pb.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(d.tile_base, tb.data(), tb.size() * 4, hipMemcpyHostToDevice));
    size_t need = std::max<size_t>(4ull * (d.n_pblk + d.n_rep), 2ull * d.n_tiles);
    dfree(d.red);
    if (int rc = dalloc(s, &d.red, need)) return rc;
    /* row tiles: about tile_cells cells (whole x-rows) each, fewer rows
     * where the mean occupancy would pass kRowTileMean; the path is off
     * for the handle if a single row holds more */
    std::vector<uint32_t> rb(d.n_rep + 1, 0);
    double max_mean = 0.0;
    int max_tc = 0;
    bool ok = true;
    for (int r = 0; r < d.n_rep; ++r) {
        RepDev& rp = s->reps[r];
        const int Lx = rp.box.iL[0], Ly = rp.box.iL[1], Lz = rp.box.iL[2];
        const double rho = (double)rp.n_fluid / (double)rp.ncells;
        int nr = std::max(1, d.tile_cells / Lx);
        nr = std::min(std::min(nr, Ly), std::min(kRowsMax, std::max(1, kRowTileCells / Lx)));
        while (nr > 1 && nr * Lx * rho > kRowTileMean) --nr;
        if (Lx > kRowTileCells || Lx * rho > kRowTileMean) ok = false;
        rp.nr = nr;
        rp.tpp = (Ly + nr - 1) / nr;
        rp.rt_base = rb[r];
        rb[r + 1] = rb[r] + (uint32_t)(rp.tpp * Lz);
        max_mean = std::max(max_mean, nr * Lx * rho);
        max_tc = std::max(max_tc, nr * Lx);
    }
    d.rt_cap = !ok ? 0 : (max_mean <= kRowTileMeanSmall ? 1024 : 2048);
    d.rt_tcm = max_tc <= 64 ? 64 : max_tc <= 128 ? 128 : 256;
    d.n_rtiles = rb[d.n_rep];
    for (int b = 0; b < 2; ++b) {
        dfree(d.bcount[b]);
        if (int rc = dalloc(s, &d.bcount[b], d.n_rtiles)) return rc;
        HIPCHK(s, hipMemset(d.bcount[b], 0, 4ull * d.n_rtiles));
    }
    dfree(d.bcount_r);
    dfree(d.tbase);
    dfree(d.bucket);
    dfree(d.bucket_r);
    dfree(d.blc);
    dfree(d.blc_r);
    if (int rc = dalloc(s, &d.bcount_r, d.n_rtiles)) return rc;
    if (int rc = dalloc(s, &d.tbase, d.n_rtiles)) return rc;
    dfree(d.ttag);
    if (int rc = dalloc(s, &d.ttag, d.n_rtiles)) return rc;
    HIPCHK(s, hipMemset(d.ttag, 0, 4ull * d.n_rtiles));
    const size_t nb = d.rt_cap ? (size_t)d.n_rtiles * (size_t)d.rt_cap : 4;
    if (int rc = dalloc(s, &d.bucket, nb)) return rc;
    if (int rc = dalloc(s, &d.bucket_r, nb)) return rc;
    if (int rc = dalloc(s, &d.blc, nb)) return rc;
    if (int rc = dalloc(s, &d.blc_r, nb)) return rc;
    HIPCHK(s, hipMemset(d.bcount_r, 0, 4ull * d.n_rtiles));
    HIPCHK(s, hipMemset(d.tbase, 0, 4ull * d.n_rtiles));
    HIPCHK(s, hipMemcpy(d.rtile_base, rb.data(), rb.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(d.reps, s->reps.data(), s->reps.size() * sizeof(RepDev), hipMemcpyHostToDevice));
    s->tc_dirty = false;
    clear_graphs(s);
    return SPMD_OK;
}

/* The row-tile path needs the current buffer in cell order (any collision
 * leaves it so) and the replicas' rows within a tile's capacity. */
bool use_rows(const spmd_sys* s) { return s->d.rt_cap != 0 && s->collided && s->path_pref == 0; }

/* The arrival counters start from zero: after a collision of the sort
 * path they hold stale counts (outside any graph capture). */
int prep_rows(spmd_sys* s) {
    if (!s->tc_dirty) return SPMD_OK;
    const DevState& d = s->d;
    for (int b = 0; b < 2; ++b) HIPCHK(s, hipMemsetAsync(d.bcount[b], 0, 4ull * d.n_rtiles, s->stream));
    s->tc_dirty = false;
    return SPMD_OK;
}

/* (Re)build the device monomer buffers of all replicas from the host
 * stars: monomer arrays, bond adjacency, Verlet lists. */
int build_monomers(spmd_sys* s) {
    DevState& d = s->d;
    uint32_t mt = 0;
    std::vector<uint32_t> bonds;
    std::vector<double> pos, vel, force, mass;
    std::vector<uint8_t> bt;
    for (int r = 0; r < d.n_rep; ++r) {
        StarHost& st = s->stars[r];
        RepDev& rp = s->reps[r];
        rp.mono_base = mt;
        rp.n_mono = st.n;
        /* each bond once, lower index first; the anchor FENE applies when
         * the bond's owner (its first monomer) is the core (Box.cpp:270-277) */
        rp.bond_base = (uint32_t)bonds.size();
        for (auto& b : st.bonds) {
            const uint32_t i = std::min(b.first, b.second), j = std::max(b.first, b.second);
            bonds.push_back(i | (j << 12) | (st.anchor[b.first] ? (1u << 24) : 0u));
        }
        rp.n_bond = (uint32_t)bonds.size() - rp.bond_base;
        pos.insert(pos.end(), st.pos.begin(), st.pos.end());
        vel.insert(vel.end(), st.vel.begin(), st.vel.end());
        force.insert(force.end(), st.force.begin(), st.force.end());
        mass.insert(mass.end(), st.mass.begin(), st.mass.end());
        bt.insert(bt.end(), st.btype.begin(), st.btype.end());
        mt += st.n;
    }
    d.m_tot = mt;
    /* MD kernel LDS: stride = largest star; Verlet capacity = what is left
     * of the 160 KiB workgroup limit */
    int ns = 1;
    for (auto& st : s->stars) ns = std::max<int>(ns, (int)st.n);
    ns = (ns + 1) & ~1;
    d.md_ns = ns;
    const size_t budget = 160u * 1024u - 4096u;  /* leave room for the kernel's static LDS */
    const size_t fixed = md_lds_bytes(ns, 0);
    d.md_vl_cap = fixed < budget ? (uint32_t)((budget - fixed) / 4) : 0u;
    dfree(d.mpos); dfree(d.mvel); dfree(d.mforce); dfree(d.mvpos); dfree(d.mmass); dfree(d.mbtype);
    dfree(d.vl); dfree(d.bonds); dfree(d.mcell);
    if (int rc = dalloc(s, &d.mpos, 3ull * mt)) return rc;
    if (int rc = dalloc(s, &d.mvel, 3ull * mt)) return rc;
    if (int rc = dalloc(s, &d.mforce, 3ull * mt)) return rc;
    if (int rc = dalloc(s, &d.mvpos, 3ull * mt)) return rc;
    if (int rc = dalloc(s, &d.mmass, mt)) return rc;
    if (int rc = dalloc(s, &d.mbtype, mt)) return rc;
    if (int rc = dalloc(s, &d.vl, (size_t)mt * kVlPerMono * 2)) return rc;  /* lists + rebuild scratch */
    if (int rc = dalloc(s, &d.bonds, std::max<size_t>(bonds.size(), 1))) return rc;
    if (int rc = dalloc(s, &d.mcell, mt)) return rc;
    if (mt) {
        HIPCHK(s, hipMemcpy(d.mpos, pos.data(), pos.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(s, hipMemcpy(d.mvel, vel.data(), vel.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(s, hipMemcpy(d.mforce, force.data(), force.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(s, hipMemcpy(d.mvpos, pos.data(), pos.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(s, hipMemcpy(d.mmass, mass.data(), mass.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(s, hipMemcpy(d.mbtype, bt.data(), bt.size(), hipMemcpyHostToDevice));
    }
    for (int r = 0; r < d.n_rep; ++r) {  /* no Verlet list until the next rebuild */
        const uint32_t zero = 0u;
        HIPCHK(s, hipMemcpy(&d.ctl[r].vl_count, &zero, 4, hipMemcpyHostToDevice));
    }
    if (!bonds.empty()) HIPCHK(s, hipMemcpy(d.bonds, bonds.data(), bonds.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(d.reps, s->reps.data(), s->reps.size() * sizeof(RepDev), hipMemcpyHostToDevice));
    s->mono_dirty = false;
    clear_graphs(s);
    /* The rebuild emptied every replica's Verlet list and reset its
     * reference positions; after spmd_md_init (the Thermostat_None ctor,
     * Thermostat_None.cpp:5-12) the lists and forces are made again here,
     * so MD never runs on an empty list. */
    if (s->md_inited && mt) {
        HIPCHK(s, launch_md(d, s->stream, 0, 0, 1, 0, s->md_dt));
        HIPCHK(s, hipStreamSynchronize(s->stream));
    }
    return SPMD_OK;
}

/* Pull device monomer state back into the host stars (before a rebuild). */
int download_monomers(spmd_sys* s) {
    if (s->mono_dirty || s->d.m_tot == 0) return SPMD_OK;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    for (int r = 0; r < s->d.n_rep; ++r) {
        StarHost& st = s->stars[r];
        const uint32_t mb = s->reps[r].mono_base;
        if (!st.n) continue;
        HIPCHK(s, hipMemcpy(st.pos.data(), s->d.mpos + 3ull * mb, 24ull * st.n, hipMemcpyDeviceToHost));
        HIPCHK(s, hipMemcpy(st.vel.data(), s->d.mvel + 3ull * mb, 24ull * st.n, hipMemcpyDeviceToHost));
        HIPCHK(s, hipMemcpy(st.force.data(), s->d.mforce + 3ull * mb, 24ull * st.n, hipMemcpyDeviceToHost));
    }
    return SPMD_OK;
}

int ensure_monomers(spmd_sys* s) { return s->mono_dirty ? build_monomers(s) : SPMD_OK; }

void prof_begin(spmd_sys* s, int k, hipStream_t st) {
    if (!s->prof) return;
    hipEvent_t e0 = nullptr;
    if (hipEventCreate(&e0) == hipSuccess) (void)hipEventRecord(e0, st);
    s->pev[k].ev.push_back(e0);
}
void prof_end(spmd_sys* s, int k, hipStream_t st) {
    if (!s->prof) return;
    hipEvent_t e1 = nullptr;
    if (hipEventCreate(&e1) == hipSuccess) (void)hipEventRecord(e1, st);
    s->pev[k].ev.push_back(e1);
}

#define LAUNCH(s, k, st, call)                                                               \
    do {                                                                                      \
        prof_begin((s), (k), (st));                                                           \
        hipError_t e_ = (call);                                                               \
        prof_end((s), (k), (st));                                                             \
        if (e_ != hipSuccess) return fail((s), SPMD_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

/* One collision for all replicas (MPC.cpp:121-141), preceded by the
 * pending MD substeps and delrx advances.  The MD kernel runs on the side
 * stream beside the solvent sort, which assumes no recentring; if the MD's
 * check_bounds recentres a replica, the redo kernels sort that replica
 * again (they return at once otherwise).  Used for direct calls and for
 * graph capture alike. */
hipError_t enqueue_collision_ops(spmd_sys* s, int n_md, int epot_last, int n_delrx, double delrx_dt, double h) {
    const DevState& d = s->d;
    hipError_t e;
    if (d.m_tot > 0) {
        /* the MD substeps need nothing of this collision's prologue: they
         * start first, beside the prologue and the speculative sort */
        if ((e = hipEventRecord(s->ev_fork, s->stream)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(s->side, s->ev_fork, 0)) != hipSuccess) return e;
        prof_begin(s, 0, s->side);
        if ((e = launch_md(d, s->side, n_md, epot_last, 0, 1, s->md_dt)) != hipSuccess) return e;
        prof_end(s, 0, s->side);
        if ((e = hipEventRecord(s->ev_join, s->side)) != hipSuccess) return e;
    }
    prof_begin(s, 5, s->stream);
    if ((e = launch_prologue(d, s->stream, 1, n_delrx, delrx_dt)) != hipSuccess) return e;
    prof_end(s, 5, s->stream);
    if (use_rows(s)) {
        /* row-tile path: advect and the cells' Gamma variates (beside the
         * MD, assuming no recentring), the monomers' cells and a recentred
         * replica's advect again after it, then the row-window collide */
        prof_begin(s, 6, s->stream);
        if ((e = launch_advect(d, s->stream, h, d.m_tot ? kSortSpec : kSortNormal)) != hipSuccess) return e;
        prof_end(s, 6, s->stream);
        prof_begin(s, 8, s->stream);
        if ((e = launch_cell_gamma(d, s->stream)) != hipSuccess) return e;
        prof_end(s, 8, s->stream);
        if (d.m_tot > 0)
            if ((e = hipStreamWaitEvent(s->stream, s->ev_join, 0)) != hipSuccess) return e;
        prof_begin(s, 7, s->stream);
        if ((e = launch_rows_redo(d, s->stream, h)) != hipSuccess) return e;
        prof_end(s, 7, s->stream);
        prof_begin(s, 4, s->stream);
        if ((e = launch_collide_rows(d, s->stream, h)) != hipSuccess) return e;
        prof_end(s, 4, s->stream);
        return hipSuccess;
    }
    s->tc_dirty = true;
    if (d.m_tot == 0) {
        prof_begin(s, 1, s->stream);
        if ((e = launch_bin(d, s->stream, h, kSortNormal)) != hipSuccess) return e;
        prof_end(s, 1, s->stream);
        prof_begin(s, 2, s->stream);
        if ((e = launch_scan(d, s->stream)) != hipSuccess) return e;
        prof_end(s, 2, s->stream);
        prof_begin(s, 3, s->stream);
        if ((e = launch_perm(d, s->stream, kSortNormal)) != hipSuccess) return e;
        prof_end(s, 3, s->stream);
    } else {
        prof_begin(s, 1, s->stream);
        if ((e = launch_bin(d, s->stream, h, kSortSpec)) != hipSuccess) return e;
        prof_end(s, 1, s->stream);
        prof_begin(s, 2, s->stream);
        if ((e = launch_scan(d, s->stream)) != hipSuccess) return e;
        prof_end(s, 2, s->stream);
        prof_begin(s, 3, s->stream);
        if ((e = launch_perm(d, s->stream, kSortSpec)) != hipSuccess) return e;
        prof_end(s, 3, s->stream);
        if ((e = hipStreamWaitEvent(s->stream, s->ev_join, 0)) != hipSuccess) return e;
        if ((e = launch_sort_redo(d, s->stream, h)) != hipSuccess) return e;
    }
    prof_begin(s, 4, s->stream);
    if ((e = launch_collide(d, s->stream, h)) != hipSuccess) return e;
    prof_end(s, 4, s->stream);
    return hipSuccess;
}

int enqueue_collision(spmd_sys* s, int n_md, int epot_last, int n_delrx, double delrx_dt, double h) {
    if (use_rows(s))
        if (int rc = prep_rows(s)) return rc;
    const hipError_t e = enqueue_collision_ops(s, n_md, epot_last, n_delrx, delrx_dt, h);
    if (e != hipSuccess) return fail(s, SPMD_ERR_HIP, std::string("collision launch: ") + hipGetErrorString(e));
    s->d.cur ^= 1;
    s->collided = true;
    return SPMD_OK;
}

int flush_md(spmd_sys* s) {
    if (s->pending_md == 0 && s->pending_delrx == 0) return SPMD_OK;
    if (int rc = ensure_monomers(s)) return rc;
    LAUNCH(s, 5, s->stream, launch_prologue(s->d, s->stream, 0, s->pending_delrx, s->pending_dt));
    LAUNCH(s, 0, s->stream, launch_md(s->d, s->stream, s->pending_md, s->pending_epot, 0, 0, s->md_dt));
    s->pending_md = 0;
    s->pending_delrx = 0;
    return SPMD_OK;
}

int check_status(spmd_sys* s) {
    std::vector<RepCtl> c(s->d.n_rep);
    HIPCHK(s, hipMemcpy(c.data(), s->d.ctl, c.size() * sizeof(RepCtl), hipMemcpyDeviceToHost));
    uint32_t st = 0;
    int rbad = -1;
    for (int r = 0; r < s->d.n_rep; ++r)
        if (c[r].status) { st |= c[r].status; if (rbad < 0) rbad = r; }
    if (!st) return SPMD_OK;
    char buf[160];
    std::snprintf(buf, sizeof buf, "device status 0x%x (first replica %d)", st, rbad);
    if (st & ST_SCAN_TIMEOUT) return fail(s, SPMD_ERR_SCAN, std::string("cell-offset scan wait expired; ") + buf);
    if (st & ST_CELL_OOB) return fail(s, SPMD_ERR_CELL_OOB, std::string("cell index out of range; ") + buf);
    if (st & ST_VERLET_OOB) return fail(s, SPMD_ERR_VERLET_OOB, std::string("monomer outside Verlet grid; ") + buf);
    if (st & ST_VERLET_OVERFLOW) return fail(s, SPMD_ERR_VERLET_OVERFLOW, std::string("Verlet list overflow; ") + buf);
    if (st & ST_SORT_MISMATCH)
        return fail(s, SPMD_ERR_SCAN, std::string("collide found a particle outside its sorted cell; ") + buf);
    if (st & ST_GAMMA_MISS)
        return fail(s, SPMD_ERR_SCAN, std::string("a cell's Gamma variate was not drawn for its member count; ") + buf);
    if (st & ST_TILE_OVERFLOW)
        return fail(s, SPMD_ERR_RANK_OVERFLOW, std::string("more particles arrived in a row tile than it holds; ") + buf);
    if (st & ST_TILE_MONO)
        return fail(s, SPMD_ERR_RANK_OVERFLOW, std::string("more monomers in one collide tile than it holds; ") + buf);
    return fail(s, SPMD_ERR_RANK_OVERFLOW, std::string("cell occupancy overflow; ") + buf);
}

bool valid_rep(const spmd_sys* s, uint32_t r) { return s && r < (uint32_t)s->d.n_rep; }

/* Molecule::initialize_open_star (Molecule.cpp:126-176) with Philox
 * velocities, then Box::set_center_of_mass_to_zero (Box.cpp:74-79). */
void build_star(StarHost& st, const RepDev& rep, const spmd_params& p, uint32_t A, uint32_t B, uint32_t arms,
                double mass, double bond, double anchor_bond) {
    const double kPi = 3.14159265358979323846;
    const uint32_t per_arm = A + B, N = per_arm * arms + 1;
    st.n = N;
    st.arms = arms;
    st.A = A;
    st.B = B;
    st.pos.assign(3ull * N, 0.0);
    st.vel.assign(3ull * N, 0.0);
    st.force.assign(3ull * N, 0.0);
    st.mass.assign(N, mass);
    st.btype.assign(N, 0);
    st.anchor.assign(N, 0);
    st.bonds.clear();
    st.anchor[0] = 1;
    double avg[3] = {0, 0, 0};
    const unsigned n = (unsigned)std::ceil(std::sqrt(arms));
    unsigned arm_count = 0;
    std::vector<std::vector<uint32_t>> nb(N);  /* Neighbors via push_front: newest first */
    for (unsigned i = 0; i < n; i++) {
        const double phi = 2. * kPi * i / n;
        for (unsigned j = 1; j < n + 1; j++) {
            if (arm_count >= arms) break;
            const double theta = kPi * j / (n + 1);
            const double dir[3] = {std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta), std::cos(theta)};
            for (unsigned k = 1; k <= per_arm; k++) {
                const unsigned idx = arm_count * per_arm + k;
                if (k > A) st.btype[idx] = 1;
                if (k == 1) {
                    for (int q = 0; q < 3; ++q) st.pos[3 * idx + q] = anchor_bond * dir[q];
                    nb[0].insert(nb[0].begin(), idx);
                    nb[idx].insert(nb[idx].begin(), idx + 1);
                } else {
                    for (int q = 0; q < 3; ++q) st.pos[3 * idx + q] = (anchor_bond + (k - 1) * bond) * dir[q];
                    if (k != per_arm) nb[idx].insert(nb[idx].begin(), idx + 1);
                }
                uint32_t w[4];
                draw4(p.seed, rep.replica, TAG_STAR_INIT, idx, 0, 0, w);
                for (int l = 0; l < 3; l++) st.vel[3 * idx + l] = u01(w[l]) - 0.5;
                for (int q = 0; q < 3; ++q) avg[q] += st.vel[3 * idx + q];
            }
            arm_count++;
        }
    }
    for (uint32_t i = 0; i < N; ++i)
        for (uint32_t o : nb[i]) st.bonds.emplace_back(i, o);
    for (int q = 0; q < 3; ++q) avg[q] /= (double)N;
    for (uint32_t i = 0; i < N; ++i)
        for (int q = 0; q < 3; ++q) st.vel[3 * i + q] -= avg[q];
    /* Molecule::calculate_Ekin (Molecule.cpp:60-74) */
    double cv[3] = {0, 0, 0};
    for (uint32_t i = 0; i < N; ++i)
        for (int q = 0; q < 3; ++q) cv[q] += st.vel[3 * i + q];
    for (int q = 0; q < 3; ++q) cv[q] /= (double)N;
    double ek = 0.0;
    for (uint32_t i = 0; i < N; ++i) {
        const double r0 = st.vel[3 * i] - cv[0], r1 = st.vel[3 * i + 1] - cv[1], r2 = st.vel[3 * i + 2] - cv[2];
        ek += mass * ((r0 * r0 + r1 * r1) + r2 * r2);
    }
    ek /= 2.0;
    const double scale = std::sqrt(3. * (double)N * p.kT / (ek * 2.0));
    for (auto& v : st.vel) v *= scale;
    double com[3] = {0, 0, 0};
    for (uint32_t i = 0; i < N; ++i)
        for (int q = 0; q < 3; ++q) com[q] += st.pos[3 * i + q];
    for (int q = 0; q < 3; ++q) com[q] /= (double)N;
    for (uint32_t i = 0; i < N; ++i)
        for (int q = 0; q < 3; ++q) st.pos[3 * i + q] -= com[q];
}

/* Random-coil arms (no reference counterpart; the reference only builds
 * straight arms, which for long arms reach past the COM-centred Verlet grid
 * of Box.cpp:129-142).  Topology, bead types, masses and velocities are
 * those of build_star; bead k >= 2 of an arm is placed at distance `bond`
 * from bead k-1 in a Philox-drawn direction, accepted when it keeps r^2 >=
 * 1 to every other bead and does not fold back towards the core. */
void coil_positions(StarHost& st, const RepDev& rep, const spmd_params& p, uint32_t A, uint32_t B, uint32_t arms,
                    double bond, double anchor_bond) {
    const double kPi = 3.14159265358979323846;
    const uint32_t per_arm = A + B, N = st.n;
    std::vector<double> P(3ull * N, 0.0);
    std::vector<uint8_t> placed(N, 0);
    placed[0] = 1;
    /* first beads on a Fibonacci sphere of radius anchor_bond (even spacing:
     * the reference's (phi, theta) grid crowds them near the poles for many
     * arms), all placed before any arm grows */
    const double golden = kPi * (3.0 - std::sqrt(5.0));
    for (unsigned a = 0; a < arms; a++) {
        const double z = 1.0 - (2.0 * a + 1.0) / arms;
        const double rxy = std::sqrt(1.0 - z * z), phi = golden * a;
        const double dir[3] = {rxy * std::cos(phi), rxy * std::sin(phi), z};
        const unsigned idx = a * per_arm + 1;
        for (int q = 0; q < 3; ++q) P[3 * idx + q] = anchor_bond * dir[q];
        placed[idx] = 1;
    }
    auto min_d2 = [&](const double* c, uint32_t skip) {
        double m = 1e300;
        for (uint32_t q = 0; q < N; ++q) {
            if (!placed[q] || q == skip) continue;
            const double dx = c[0] - P[3 * q], dy = c[1] - P[3 * q + 1], dz = c[2] - P[3 * q + 2];
            const double r2 = (dx * dx + dy * dy) + dz * dz;
            if (r2 < m) m = r2;
        }
        return m;
    };
    /* Bead idx next to bead idx-1, r^2 >= 1 to every other placed bead:
     * 4096 Philox draws that must not step inward by more than 0.5 (1024 at
     * the nominal bond, then 10% longer bonds per 1024, at most 1.3x), then
     * 4096 without that rule; the best draw is kept.  `stream` selects an
     * independent set of draws. */
    auto place = [&](unsigned idx, unsigned stream) -> bool {
        placed[idx] = 0;
        const double* prev = &P[3 * (idx - 1)];
        const double rprev = std::sqrt((prev[0] * prev[0] + prev[1] * prev[1]) + prev[2] * prev[2]);
        double best[3] = {0, 0, 0}, best_m = -1.0;
        for (uint32_t att = 0; att < 8192; ++att) {
            uint32_t w[4];
            draw4(p.seed, rep.replica, TAG_STAR_INIT, idx, 1 + stream, att, w);
            const double z = 2.0 * u01(w[0]) - 1.0, ph = 2. * kPi * u01(w[1]);
            const double sz = std::sqrt(1.0 - z * z);
            const double bl = bond * (1.0 + 0.1 * ((att % 4096) / 1024));
            const double c[3] = {prev[0] + bl * sz * std::cos(ph), prev[1] + bl * sz * std::sin(ph),
                                 prev[2] + bl * z};
            const double rc = std::sqrt((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]);
            const double m = (rc >= rprev - 0.5 || att >= 4096) ? min_d2(c, idx - 1) : -0.5;
            if (m > best_m) { best_m = m; best[0] = c[0]; best[1] = c[1]; best[2] = c[2]; }
            if (m >= 1.0) break;
        }
        for (int q = 0; q < 3; ++q) P[3 * idx + q] = best[q];
        placed[idx] = 1;
        return best_m >= 1.0;
    };
    /* grow all arms one bead at a time; a bead that finds no room moves its
     * predecessor (new draws) and tries again */
    for (unsigned k = 2; k <= per_arm; k++) {
        for (unsigned a = 0; a < arms; a++) {
            const unsigned idx = a * per_arm + k;
            bool ok = place(idx, 0);
            for (unsigned s2 = 1; !ok && s2 <= 16 && k > 2; ++s2) {
                placed[idx] = 0;
                place(idx - 1, s2);
                ok = place(idx, s2);
            }
        }
    }
    double com[3] = {0, 0, 0};
    for (uint32_t i = 0; i < N; ++i)
        for (int q = 0; q < 3; ++q) com[q] += P[3 * i + q];
    for (int q = 0; q < 3; ++q) com[q] /= (double)N;
    for (uint32_t i = 0; i < N; ++i)
        for (int q = 0; q < 3; ++q) st.pos[3 * i + q] = P[3 * i + q] - com[q];
}

int check_device(spmd_sys* s) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(s, SPMD_ERR_NO_DEVICE, "no HIP device");
    HIPCHK(s, hipGetDevice(&s->device));
    hipDeviceProp_t prop;
    HIPCHK(s, hipGetDeviceProperties(&prop, s->device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(s, SPMD_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", library built for gfx950");
    return SPMD_OK;
}

/* The current device's address of the static scratch (looked up on the
 * first handle made on that device). */
int scratch_setup(spmd_sys* s) {
    if (s->device < 0 || s->device >= kMaxDevices) return fail(s, SPMD_ERR_NO_DEVICE, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[s->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(s, SPMD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[s->device] = static_cast<unsigned long long*>(p);
    }
    s->scratch = g_scratch_ptr[s->device];
    return SPMD_OK;
}

}  // namespace

extern "C" {

const char* spmd_version(void) { return "starpolymersmd-amd 0.1 (gfx950)"; }

const char* spmd_error_string(int code) {
    switch (code) {
    case SPMD_OK: return "ok";
    case SPMD_ERR_ARG: return "invalid argument";
    case SPMD_ERR_HIP: return "HIP runtime error";
    case SPMD_ERR_OOM: return "device out of memory";
    case SPMD_ERR_CELL_OOB: return "cell index out of range";
    case SPMD_ERR_VERLET_OOB: return "monomer outside the Verlet grid";
    case SPMD_ERR_VERLET_OVERFLOW: return "Verlet list overflow";
    case SPMD_ERR_UNSUPPORTED: return "unsupported option";
    case SPMD_ERR_STATE: return "call out of order";
    case SPMD_ERR_RANK_OVERFLOW: return "cell occupancy overflow";
    case SPMD_ERR_NO_DEVICE: return "no usable gfx950 device";
    case SPMD_ERR_SCAN: return "cell-offset scan did not complete";
    default: return "unknown error";
    }
}

const char* spmd_last_error(const spmd_sys* s) { return s ? s->err.c_str() : "null handle"; }

int spmd_create_batch(const spmd_params* p, uint32_t n_rep, spmd_sys** out) {
    if (!p || !out || n_rep == 0) return SPMD_ERR_ARG;
    *out = nullptr;
    auto* s = new spmd_sys;
    if (int rc = check_device(s)) { delete s; return rc; }
    if (int rc = scratch_setup(s)) { delete s; return rc; }
    s->params.assign(p, p + n_rep);
    s->step_update = p[0].step_update ? p[0].step_update : 1;
    s->md_dt = p[0].md_dt;
    DevState& d = s->d;
    d.n_rep = (int)n_rep;
    d.ppt = 4;
    d.tile_cells = 128;
    uint64_t nt = 0, ct = 0;
    for (uint32_t r = 0; r < n_rep; ++r) {
        const spmd_params& q = p[r];
        if (q.step_update != p[0].step_update || q.md_dt != p[0].md_dt || q.n_c < 0 || q.kT <= 0.0) {
            delete s;
            return SPMD_ERR_ARG;
        }
        RepDev rp{};
        rp.amc = q.angular_momentum != 0;
        d.any_amc |= rp.amc;
        for (int i = 0; i < 3; ++i) {
            int l = q.L[i];
            if (l % 2) l++;                      /* Box.cpp:14-16 */
            if (l < 4 || l > 1024) { delete s; return SPMD_ERR_ARG; }
            rp.box.L[i] = l;
            rp.box.invL[i] = 1.0 / l;
            rp.box.half[i] = l * 0.5;
            rp.box.inner[i] = l * 0.5 * (1.0 - 1e-9);
            rp.box.iL[i] = l;
            rp.vgrid[i] = int(l / 2.0);          /* Box.cpp:20-25 */
            rp.vside[i] = l / (double)rp.vgrid[i];
        }
        rp.box.shear = q.shear;
        rp.box.shear_on = q.shear != 0.;
        rp.kT = q.kT;
        rp.c = std::cos(q.alpha_deg * 3.14159265358979323846 / 180.0);
        rp.s = std::sin(q.alpha_deg * 3.14159265358979323846 / 180.0);
        rp.lambda = q.lambda;
        rp.seed = q.seed;
        rp.replica = r;
        rp.ncells = (uint32_t)(rp.box.iL[0] * rp.box.iL[1] * rp.box.iL[2]);
        if (rp.ncells > 4096u * 4096u) { delete s; return SPMD_ERR_ARG; }
        rp.cell_base = ct;
        rp.part_base = nt;
        const uint64_t nf = (uint64_t)q.n_c * rp.ncells;
        rp.n_fluid = (uint32_t)nf;
        ct += rp.ncells;
        nt += nf;
        s->reps.push_back(rp);
    }
    if (nt >= (1ull << 32) || ct >= (1ull << 31)) { delete s; return SPMD_ERR_ARG; }
    s->stars.resize(n_rep);
    d.n_tot = nt;
    d.c_tot = ct;
    d.n_scan_tiles = (uint32_t)((ct + kScanTile - 1) / kScanTile);
    HIPCHK(s, hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIPCHK(s, hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    HIPCHK(s, hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming));
    HIPCHK(s, hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming));
    int rc = SPMD_OK;
    for (int b = 0; b < 2 && !rc; ++b) {
        rc = rc ? rc : dalloc(s, &d.x[b], nt);
        rc = rc ? rc : dalloc(s, &d.y[b], nt);
        rc = rc ? rc : dalloc(s, &d.z[b], nt);
        rc = rc ? rc : dalloc(s, &d.vx[b], nt);
        rc = rc ? rc : dalloc(s, &d.vy[b], nt);
        rc = rc ? rc : dalloc(s, &d.vz[b], nt);
    }
    rc = rc ? rc : dalloc(s, &d.rank, nt);
    rc = rc ? rc : dalloc(s, &d.pcell, nt);
    rc = rc ? rc : dalloc(s, &d.perm, nt);
    rc = rc ? rc : dalloc(s, &d.counts, ct + 16);
    rc = rc ? rc : dalloc(s, &d.offs[0], ct + 16);
    rc = rc ? rc : dalloc(s, &d.offs[1], ct + 16);
    rc = rc ? rc : dalloc(s, &d.cg_n, ct);
    rc = rc ? rc : dalloc(s, &d.cg_g, ct);
    rc = rc ? rc : dalloc(s, &d.rtile_base, n_rep + 1);
    rc = rc ? rc : dalloc(s, &d.scan_flags, d.n_scan_tiles);
    rc = rc ? rc : dalloc(s, &d.scan_ticket, 1);
    rc = rc ? rc : dalloc(s, &d.grid_bar, 1);
    rc = rc ? rc : dalloc(s, &d.reps, n_rep);
    rc = rc ? rc : dalloc(s, &d.ctl, n_rep);
    rc = rc ? rc : dalloc(s, &d.pblk_base, n_rep + 1);
    rc = rc ? rc : dalloc(s, &d.tile_base, n_rep + 1);
    if (rc) { spmd_destroy(s); return rc; }
    HIPCHK(s, hipMemset(d.counts, 0, (ct + 16) * 4));
    HIPCHK(s, hipMemset(d.offs[0], 0, (ct + 16) * 4));
    HIPCHK(s, hipMemset(d.offs[1], 0, (ct + 16) * 4));
    HIPCHK(s, hipMemset(d.cg_n, 0, ct * 4));
    HIPCHK(s, hipMemset(d.cg_g, 0, ct * 8));
    HIPCHK(s, hipMemset(d.scan_flags, 0, d.n_scan_tiles * 8ull));
    HIPCHK(s, hipMemset(d.scan_ticket, 0, 4));
    HIPCHK(s, hipMemset(d.grid_bar, 0, 4));
    for (int b = 0; b < 2; ++b) {
        HIPCHK(s, hipMemset(d.x[b], 0, nt * 4)); HIPCHK(s, hipMemset(d.y[b], 0, nt * 4));
        HIPCHK(s, hipMemset(d.z[b], 0, nt * 4)); HIPCHK(s, hipMemset(d.vx[b], 0, nt * 4));
        HIPCHK(s, hipMemset(d.vy[b], 0, nt * 4)); HIPCHK(s, hipMemset(d.vz[b], 0, nt * 4));
    }
    std::vector<RepCtl> c(n_rep);
    std::memset(c.data(), 0, c.size() * sizeof(RepCtl));
    HIPCHK(s, hipMemcpy(d.ctl, c.data(), c.size() * sizeof(RepCtl), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(d.reps, s->reps.data(), s->reps.size() * sizeof(RepDev), hipMemcpyHostToDevice));
    if ((rc = upload_tables(s))) { spmd_destroy(s); return rc; }
    if ((rc = build_monomers(s))) { spmd_destroy(s); return rc; }
    *out = s;
    return SPMD_OK;
}

int spmd_create(const spmd_params* p, spmd_sys** out) { return spmd_create_batch(p, 1, out); }

int spmd_destroy(spmd_sys* s) {
    if (!s) return SPMD_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    clear_graphs(s);
    for (auto& pe : s->pev)
        for (auto e : pe.ev) if (e) (void)hipEventDestroy(e);
    DevState& d = s->d;
    for (int b = 0; b < 2; ++b) {
        dfree(d.x[b]); dfree(d.y[b]); dfree(d.z[b]); dfree(d.vx[b]); dfree(d.vy[b]); dfree(d.vz[b]);
        dfree(d.ids[b]);
    }
    dfree(d.rank); dfree(d.pcell); dfree(d.perm); dfree(d.counts); dfree(d.scan_flags); dfree(d.scan_ticket);
    dfree(d.grid_bar);
    for (int b = 0; b < 2; ++b) { dfree(d.offs[b]); dfree(d.bcount[b]); }
    dfree(d.bcount_r); dfree(d.tbase); dfree(d.ttag); dfree(d.bucket); dfree(d.bucket_r); dfree(d.rtile_base);
    dfree(d.blc); dfree(d.blc_r); dfree(d.cg_n); dfree(d.cg_g);
    dfree(d.mpos); dfree(d.mvel); dfree(d.mforce); dfree(d.mvpos); dfree(d.mmass); dfree(d.mbtype);
    dfree(d.vl); dfree(d.bonds); dfree(d.mcell);
    dfree(d.inj_axis); dfree(d.inj_gamma); dfree(d.reps); dfree(d.ctl); dfree(d.pblk_base);
    dfree(d.tile_base); dfree(d.red);
    dfree(s->cl_feat); dfree(s->cl_out); dfree(s->cl_range); dfree(s->mom); dfree(s->obs); dfree(s->cellv);
    for (auto& v : s->vacf) { dfree(v.hist); dfree(v.sums); }
    if (s->side) (void)hipStreamSynchronize(s->side);
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->ev_join) (void)hipEventDestroy(s->ev_join);
    if (s->side) (void)hipStreamDestroy(s->side);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return SPMD_OK;
}

int spmd_num_replicas(const spmd_sys* s, uint32_t* n) {
    if (!s || !n) return SPMD_ERR_ARG;
    *n = (uint32_t)s->d.n_rep;
    return SPMD_OK;
}

int spmd_set_md_dt(spmd_sys* s, double dt) {
    if (!s || !(dt > 0)) return SPMD_ERR_ARG;
    if (s->pending_md) if (int rc = flush_md(s)) return rc;
    s->md_dt = dt;
    clear_graphs(s);
    return SPMD_OK;
}

int spmd_set_tuning(spmd_sys* s, int ppt, int tile_cells) {
    if (!s) return SPMD_ERR_ARG;
    if (ppt != 1 && ppt != 2 && ppt != 4 && ppt != 8) return fail(s, SPMD_ERR_ARG, "ppt must be 1,2,4,8");
    if (tile_cells != 64 && tile_cells != 128 && tile_cells != 256)
        return fail(s, SPMD_ERR_ARG, "tile_cells must be 64,128,256");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->d.ppt = ppt;
    s->d.tile_cells = tile_cells;
    return upload_tables(s);
}

int spmd_set_collision_path(spmd_sys* s, int path) {
    if (!s || (path != SPMD_PATH_AUTO && path != SPMD_PATH_SORT)) return SPMD_ERR_ARG;
    s->path_pref = path == SPMD_PATH_SORT ? 1 : 0;
    return SPMD_OK;
}

int spmd_get_collision_path(const spmd_sys* s, int* next_path) {
    if (!s || !next_path) return SPMD_ERR_ARG;
    *next_path = use_rows(s) ? SPMD_PATH_ROWS : SPMD_PATH_SORT;
    return SPMD_OK;
}

int spmd_get_tuning(const spmd_sys* s, int* ppt, int* tile_cells) {
    if (!s) return SPMD_ERR_ARG;
    if (ppt) *ppt = s->d.ppt;
    if (tile_cells) *tile_cells = s->d.tile_cells;
    return SPMD_OK;
}

int spmd_fluid_initialize(spmd_sys* s) {
    if (!s) return SPMD_ERR_ARG;
    if (int rc = flush_md(s)) return rc;
    LAUNCH(s, 1, s->stream, launch_fluid_init(s->d, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->collided = false;
    return SPMD_OK;
}

int spmd_num_fluid(const spmd_sys* s, uint32_t r, uint64_t* n) {
    if (!valid_rep(s, r) || !n) return SPMD_ERR_ARG;
    *n = s->reps[r].n_fluid;
    return SPMD_OK;
}

int spmd_set_fluid(spmd_sys* s, uint32_t r, const float* pos, const float* vel, uint64_t n) {
    if (!valid_rep(s, r) || !pos || !vel) return SPMD_ERR_ARG;
    if (n != s->reps[r].n_fluid) return fail(s, SPMD_ERR_ARG, "fluid size must equal N_c*L^3");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    const DevState& d = s->d;
    const uint64_t b = s->reps[r].part_base;
    std::vector<float> t(n);
    float* dst[6] = {d.x[d.cur], d.y[d.cur], d.z[d.cur], d.vx[d.cur], d.vy[d.cur], d.vz[d.cur]};
    for (int c = 0; c < 6; ++c) {
        const float* src = c < 3 ? pos : vel;
        for (uint64_t i = 0; i < n; ++i) t[i] = src[3 * i + (c % 3)];
        HIPCHK(s, hipMemcpy(dst[c] + b, t.data(), n * 4, hipMemcpyHostToDevice));
    }
    if (d.ids[d.cur]) {
        std::vector<uint32_t> id(n);
        for (uint64_t i = 0; i < n; ++i) id[i] = (uint32_t)i;
        HIPCHK(s, hipMemcpy(d.ids[d.cur] + b, id.data(), n * 4, hipMemcpyHostToDevice));
    }
    s->collided = false;
    return SPMD_OK;
}

int spmd_get_fluid(spmd_sys* s, uint32_t r, float* pos, float* vel, uint32_t* ids) {
    if (!valid_rep(s, r)) return SPMD_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    const DevState& d = s->d;
    const uint64_t n = s->reps[r].n_fluid, b = s->reps[r].part_base;
    std::vector<float> t(n);
    const float* src[6] = {d.x[d.cur], d.y[d.cur], d.z[d.cur], d.vx[d.cur], d.vy[d.cur], d.vz[d.cur]};
    for (int c = 0; c < 6; ++c) {
        float* dst = c < 3 ? pos : vel;
        if (!dst) continue;
        HIPCHK(s, hipMemcpy(t.data(), src[c] + b, n * 4, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < n; ++i) dst[3 * i + (c % 3)] = t[i];
    }
    if (ids) {
        if (!d.ids[d.cur]) return fail(s, SPMD_ERR_STATE, "particle tracking is off");
        HIPCHK(s, hipMemcpy(ids, d.ids[d.cur] + b, n * 4, hipMemcpyDeviceToHost));
    }
    return SPMD_OK;
}

int spmd_set_tracking(spmd_sys* s, int on) {
    if (!s) return SPMD_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    DevState& d = s->d;
    if (on && !d.ids[0]) {
        for (int b = 0; b < 2; ++b)
            if (int rc = dalloc(s, &d.ids[b], d.n_tot)) return rc;
        std::vector<uint32_t> id(d.n_tot);
        for (int r = 0; r < d.n_rep; ++r)
            for (uint32_t i = 0; i < s->reps[r].n_fluid; ++i) id[s->reps[r].part_base + i] = i;
        HIPCHK(s, hipMemcpy(d.ids[d.cur], id.data(), d.n_tot * 4, hipMemcpyHostToDevice));
    } else if (!on) {
        dfree(d.ids[0]);
        dfree(d.ids[1]);
    }
    s->tracking = on != 0;
    clear_graphs(s);
    return SPMD_OK;
}

int spmd_add_star(spmd_sys* s, uint32_t r, uint32_t A, uint32_t B, uint32_t arms, double mass, double bond,
                  double anchor_bond) {
    if (!valid_rep(s, r) || arms == 0 || A + B < 2 || !(mass > 0)) return SPMD_ERR_ARG;
    if ((A + B) * arms + 1 > (uint32_t)kMaxMono) return fail(s, SPMD_ERR_ARG, "star larger than 2048 monomers");
    if (int rc = flush_md(s)) return rc;
    if (int rc = download_monomers(s)) return rc;
    build_star(s->stars[r], s->reps[r], s->params[r], A, B, arms, mass, bond, anchor_bond);
    s->mono_dirty = true;
    return ensure_monomers(s);
}

int spmd_add_star_coil(spmd_sys* s, uint32_t r, uint32_t A, uint32_t B, uint32_t arms, double mass, double bond,
                       double anchor_bond) {
    if (!valid_rep(s, r) || arms == 0 || A + B < 2 || !(mass > 0) || !(bond > 0)) return SPMD_ERR_ARG;
    if ((A + B) * arms + 1 > (uint32_t)kMaxMono) return fail(s, SPMD_ERR_ARG, "star larger than 2048 monomers");
    if (int rc = flush_md(s)) return rc;
    if (int rc = download_monomers(s)) return rc;
    build_star(s->stars[r], s->reps[r], s->params[r], A, B, arms, mass, bond, anchor_bond);
    coil_positions(s->stars[r], s->reps[r], s->params[r], A, B, arms, bond, anchor_bond);
    s->mono_dirty = true;
    return ensure_monomers(s);
}

int spmd_num_monomers(const spmd_sys* s, uint32_t r, uint32_t* n) {
    if (!valid_rep(s, r) || !n) return SPMD_ERR_ARG;
    *n = s->stars[r].n;
    return SPMD_OK;
}

int spmd_set_monomers(spmd_sys* s, uint32_t r, const double* pos, const double* vel) {
    if (!valid_rep(s, r)) return SPMD_ERR_ARG;
    if (int rc = flush_md(s)) return rc;
    if (int rc = ensure_monomers(s)) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    const uint32_t n = s->stars[r].n, mb = s->reps[r].mono_base;
    if (!n) return SPMD_OK;
    if (pos) {
        HIPCHK(s, hipMemcpy(s->d.mpos + 3ull * mb, pos, 24ull * n, hipMemcpyHostToDevice));
        std::memcpy(s->stars[r].pos.data(), pos, 24ull * n);
    }
    if (vel) {
        HIPCHK(s, hipMemcpy(s->d.mvel + 3ull * mb, vel, 24ull * n, hipMemcpyHostToDevice));
        std::memcpy(s->stars[r].vel.data(), vel, 24ull * n);
    }
    return SPMD_OK;
}

int spmd_get_monomer_masses(const spmd_sys* s, uint32_t r, double* mass) {
    if (!s || r >= (uint32_t)s->d.n_rep || !mass) return SPMD_ERR_ARG;
    const StarHost& st = s->stars[r];
    std::copy(st.mass.begin(), st.mass.begin() + st.n, mass);
    return SPMD_OK;
}

int spmd_get_monomers(spmd_sys* s, uint32_t r, double* pos, double* vel, double* force) {
    if (!valid_rep(s, r)) return SPMD_ERR_ARG;
    if (int rc = flush_md(s)) return rc;
    if (int rc = ensure_monomers(s)) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    const uint32_t n = s->stars[r].n, mb = s->reps[r].mono_base;
    if (!n) return SPMD_OK;
    if (pos) HIPCHK(s, hipMemcpy(pos, s->d.mpos + 3ull * mb, 24ull * n, hipMemcpyDeviceToHost));
    if (vel) HIPCHK(s, hipMemcpy(vel, s->d.mvel + 3ull * mb, 24ull * n, hipMemcpyDeviceToHost));
    if (force) HIPCHK(s, hipMemcpy(force, s->d.mforce + 3ull * mb, 24ull * n, hipMemcpyDeviceToHost));
    return SPMD_OK;
}

int spmd_md_init(spmd_sys* s) {
    if (!s) return SPMD_ERR_ARG;
    if (int rc = flush_md(s)) return rc;
    if (int rc = ensure_monomers(s)) return rc;
    LAUNCH(s, 0, s->stream, launch_md(s->d, s->stream, 0, 0, 1, 0, s->md_dt));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->md_inited = true;
    return check_status(s);
}

int spmd_md_propagate(spmd_sys* s, int calc_epot) {
    if (!s) return SPMD_ERR_ARG;
    s->pending_md++;
    s->pending_epot = calc_epot != 0;
    return SPMD_OK;
}

int spmd_mpc_step(spmd_sys* s, int64_t t, double dt) {
    if (!s) return SPMD_ERR_ARG;
    if (s->pending_delrx && dt != s->pending_dt) {
        const int keep = s->pending_md;
        s->pending_md = 0;
        if (int rc = flush_md(s)) return rc;
        s->pending_md = keep;
    }
    s->pending_dt = dt;
    s->pending_delrx++;
    if (t % (int64_t)s->step_update) return SPMD_OK;
    if (int rc = ensure_monomers(s)) return rc;
    const int rc = enqueue_collision(s, s->pending_md, s->pending_epot, s->pending_delrx, dt, dt);
    s->pending_md = 0;
    s->pending_delrx = 0;
    return rc;
}

int spmd_run(spmd_sys* s, int64_t t0, int64_t n, double mpc_dt) {
    if (!s || n < 0) return SPMD_ERR_ARG;
    const int64_t S = s->step_update;
    int64_t t = t0;
    const int64_t end = t0 + n;
    /* single calls until the start of a collision interval: the call
     * after a collision call, with nothing pending */
    auto at_start = [&]() {
        return s->pending_md == 0 && s->pending_delrx == 0 && t >= 1 && (t - 1) % S == 0;
    };
    while (t < end && !at_start()) {
        if (int rc = spmd_md_propagate(s, 0)) return rc;
        if (int rc = spmd_mpc_step(s, t, mpc_dt)) return rc;
        ++t;
    }
    if (int rc = ensure_monomers(s)) return rc;
    /* whole intervals t = kS+1 .. (k+1)S: replay the captured graph */
    while (end - t >= S && at_start()) {
        /* the first collision after the fluid was set orders the buffer
         * (sort path): run it directly, the steady path from the graph */
        if (s->prof || (!s->collided && s->d.rt_cap && s->path_pref == 0)) {
            if (int rc = enqueue_collision(s, (int)S, 0, (int)S, mpc_dt, mpc_dt)) return rc;
        } else {
            const int c = s->d.cur;
            const int path = use_rows(s) ? 1 : 0;
            if (path)
                if (int rc = prep_rows(s)) return rc;
            /* the row-tile path repeats itself: kGraphIntervals intervals in
             * one graph while as many remain */
            const int K = path && (end - t) / S >= kGraphIntervals ? kGraphIntervals : 1;
            if (s->graph_h != mpc_dt) clear_graphs(s);
            hipGraphExec_t& gx = K > 1 ? s->graph_run[c] : s->graph[path][c];
            if (!gx) {
                hipGraph_t g;
                HIPCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
                hipError_t e = hipSuccess;
                for (int q = 0; q < K && e == hipSuccess; ++q) {
                    e = enqueue_collision_ops(s, (int)S, 0, (int)S, mpc_dt, mpc_dt);
                    s->d.cur ^= 1;
                }
                s->d.cur = c;
                hipError_t e2 = hipStreamEndCapture(s->stream, &g);
                if (e != hipSuccess || e2 != hipSuccess)
                    return fail(s, SPMD_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e ? e : e2));
                HIPCHK(s, hipGraphInstantiate(&gx, g, nullptr, nullptr, 0));
                (void)hipGraphDestroy(g);
                s->graph_h = mpc_dt;
            }
            HIPCHK(s, hipGraphLaunch(gx, s->stream));
            if (!path) s->tc_dirty = true;
            s->d.cur ^= K & 1;
            s->collided = true;
            t += (K - 1) * S;
        }
        t += S;
    }
    for (; t < end; ++t) {
        if (int rc = spmd_md_propagate(s, 0)) return rc;
        if (int rc = spmd_mpc_step(s, t, mpc_dt)) return rc;
    }
    return SPMD_OK;
}

int spmd_sync(spmd_sys* s) {
    if (!s) return SPMD_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return check_status(s);
}

int spmd_set_delrx(spmd_sys* s, uint32_t r, double delrx) {
    if (!valid_rep(s, r)) return SPMD_ERR_ARG;
    if (int rc = flush_md(s)) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(&s->d.ctl[r].delrx, &delrx, 8, hipMemcpyHostToDevice));
    return SPMD_OK;
}

int spmd_get_delrx(spmd_sys* s, uint32_t r, double* delrx) {
    if (!valid_rep(s, r) || !delrx) return SPMD_ERR_ARG;
    if (int rc = flush_md(s)) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(delrx, &s->d.ctl[r].delrx, 8, hipMemcpyDeviceToHost));
    return SPMD_OK;
}

int spmd_inject_collision_randoms(spmd_sys* s, uint32_t r, const double* shift, const double* axis_u,
                                  const double* gamma) {
    if (!valid_rep(s, r) || !shift || !axis_u || !gamma) return SPMD_ERR_ARG;
    DevState& d = s->d;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (!d.inj_axis) {
        if (int rc = dalloc(s, &d.inj_axis, 2 * d.c_tot)) return rc;
        if (int rc = dalloc(s, &d.inj_gamma, d.c_tot)) return rc;
        clear_graphs(s);
    }
    const RepDev& rp = s->reps[r];
    HIPCHK(s, hipMemcpy(d.inj_axis + 2 * rp.cell_base, axis_u, 16ull * rp.ncells, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(d.inj_gamma + rp.cell_base, gamma, 8ull * rp.ncells, hipMemcpyHostToDevice));
    RepCtl c;
    HIPCHK(s, hipMemcpy(&c, d.ctl + r, sizeof c, hipMemcpyDeviceToHost));
    for (int i = 0; i < 3; ++i) c.inj_shift[i] = shift[i];
    c.inj_pending = 1;
    HIPCHK(s, hipMemcpy(d.ctl + r, &c, sizeof c, hipMemcpyHostToDevice));
    return SPMD_OK;
}

int spmd_last_collision(spmd_sys* s, uint32_t r, double* shift, int* recentered, double* com_pos,
                        double* com_vel) {
    if (!valid_rep(s, r)) return SPMD_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    RepCtl c;
    HIPCHK(s, hipMemcpy(&c, s->d.ctl + r, sizeof c, hipMemcpyDeviceToHost));
    for (int i = 0; i < 3; ++i) {
        if (shift) shift[i] = c.shift[i];
        if (com_pos) com_pos[i] = c.com[i];
        if (com_vel) com_vel[i] = c.com_vel[i];
    }
    if (recentered) *recentered = c.recenter;
    return SPMD_OK;
}

int spmd_get_cell_index(spmd_sys* s, uint32_t r, uint32_t* fluid_cells, uint32_t* mono_cells) {
    if (!valid_rep(s, r)) return SPMD_ERR_ARG;
    if (!s->collided) return fail(s, SPMD_ERR_STATE, "no collision since the fluid was set");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    const DevState& d = s->d;
    const RepDev& rp = s->reps[r];
    if (fluid_cells) {
        if (!d.ids[d.cur]) return fail(s, SPMD_ERR_STATE, "particle tracking is off");
        std::vector<uint32_t> off(rp.ncells + 1), ids(rp.n_fluid);
        HIPCHK(s, hipMemcpy(off.data(), d.offs[d.cur] + rp.cell_base, 4ull * (rp.ncells + 1), hipMemcpyDeviceToHost));
        HIPCHK(s, hipMemcpy(ids.data(), d.ids[d.cur] + rp.part_base, 4ull * rp.n_fluid, hipMemcpyDeviceToHost));
        for (uint32_t c = 0; c < rp.ncells; ++c)
            for (uint32_t p = off[c]; p < off[c + 1]; ++p) fluid_cells[ids[p - rp.part_base]] = c;
    }
    if (mono_cells && rp.n_mono) {
        std::vector<uint32_t> mc(rp.n_mono);
        HIPCHK(s, hipMemcpy(mc.data(), d.mcell + rp.mono_base, 4ull * rp.n_mono, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < rp.n_mono; ++i) mono_cells[i] = (uint32_t)(mc[i] - rp.cell_base);
    }
    return SPMD_OK;
}

int spmd_temperature(spmd_sys* s, uint32_t r, double* T) {
    if (!valid_rep(s, r) || !T) return SPMD_ERR_ARG;
    if (!s->collided) return fail(s, SPMD_ERR_STATE, "no collision since the fluid was set");
    if (int rc = flush_md(s)) return rc;
    const RepDev& rp = s->reps[r];
    const uint32_t nt = (rp.ncells + s->d.tile_cells - 1) / s->d.tile_cells;
    std::vector<uint32_t> tb(2);
    HIPCHK(s, hipMemcpy(tb.data(), s->d.tile_base + r, 8, hipMemcpyDeviceToHost));
    HIPCHK(s, launch_temperature(s->d, s->stream, (int)r, nt));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    std::vector<double> red(2ull * nt);
    HIPCHK(s, hipMemcpy(red.data(), s->d.red + 2ull * tb[0], red.size() * 8, hipMemcpyDeviceToHost));
    double ek = 0.0, filled = 0.0;
    for (uint32_t i = 0; i < nt; ++i) { ek += red[2 * i]; filled += red[2 * i + 1]; }
    *T = 2. * ek / (3. * ((double)rp.n_fluid - filled));
    return SPMD_OK;
}

int spmd_cell_fluid_velocity(spmd_sys* s, uint32_t r, int32_t z_start, int32_t z_stop, double* out) {
    if (!valid_rep(s, r) || !out) return SPMD_ERR_ARG;
    const RepDev& rp = s->reps[r];
    const int Lz = rp.box.iL[2];
    if (z_start < 0 || z_stop < z_start || z_stop >= Lz) return fail(s, SPMD_ERR_ARG, "z range outside the box");
    const uint32_t plane = (uint32_t)rp.box.iL[0] * (uint32_t)rp.box.iL[1];
    const uint32_t n = plane * (uint32_t)(z_stop - z_start + 1);
    if (!s->collided) {  /* no sort since the fluid was set: empty cell lists (MPC.cpp:334-342) */
        std::fill(out, out + 3ull * n, 0.0);
        return SPMD_OK;
    }
    if (3ull * n > s->cellv_n) {
        dfree(s->cellv);
        if (int rc = dalloc(s, &s->cellv, 3ull * n)) return rc;
        s->cellv_n = 3ull * n;
    }
    HIPCHK(s, launch_cell_velocity(s->d, s->stream, (int)r, plane * (uint32_t)z_start, n, s->cellv));
    HIPCHK(s, hipMemcpyAsync(out, s->cellv, 24ull * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return SPMD_OK;
}

int spmd_epot(spmd_sys* s, uint32_t r, double* epot) {
    if (!valid_rep(s, r) || !epot) return SPMD_ERR_ARG;
    if (int rc = flush_md(s)) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(epot, &s->d.ctl[r].epot, 8, hipMemcpyDeviceToHost));
    return SPMD_OK;
}

/* ------------------------------------------------------ learned closure */

static int closure_ws(spmd_sys* s) {
    if (int rc = ensure_monomers(s)) return rc;
    const uint32_t m = s->d.m_tot;
    if (m > s->cl_cap) {
        dfree(s->cl_feat); dfree(s->cl_out);
        if (int rc = dalloc(s, &s->cl_feat, (size_t)m * kClosureFeatures)) return rc;
        if (int rc = dalloc(s, &s->cl_out, (size_t)m * 6)) return rc;
        s->cl_cap = m;
    }
    if (!s->cl_range)
        if (int rc = dalloc(s, &s->cl_range, 2 * kClosureFeatures)) return rc;
    return SPMD_OK;
}

int spmd_closure_features(spmd_sys* s, float* feat_dev) {
    if (!s || !feat_dev) return SPMD_ERR_ARG;
    if (reinterpret_cast<uintptr_t>(feat_dev) & 15u) return fail(s, SPMD_ERR_ARG, "feature buffer must be 16-byte aligned");
    if (int rc = flush_md(s)) return rc;
    if (int rc = ensure_monomers(s)) return rc;
    LAUNCH(s, 0, s->stream, launch_closure_features(s->d, s->stream, feat_dev, nullptr, nullptr));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return check_status(s);
}

int spmd_closure_run(spmd_sys* s, spmd_mlp* m, int64_t n_intervals, const float* fmin, const float* fmax,
                     uint32_t* oor_host) {
    if (!s || !m || n_intervals < 0 || (!fmin) != (!fmax)) return SPMD_ERR_ARG;
    if (int rc = flush_md(s)) return rc;
    if (int rc = closure_ws(s)) return rc;
    if (s->d.m_tot == 0) return fail(s, SPMD_ERR_STATE, "no monomers");
    float mu[24], sd[24], sy[3];
    if (int rc = spmd_mlp_get_norm(m, mu, sd, sy)) return fail(s, rc, "reading the network's target scale");
    const float* dmin = nullptr;
    const float* dmax = nullptr;
    if (fmin) {
        HIPCHK(s, hipMemcpy(s->cl_range, fmin, kClosureFeatures * 4ull, hipMemcpyHostToDevice));
        HIPCHK(s, hipMemcpy(s->cl_range + kClosureFeatures, fmax, kClosureFeatures * 4ull, hipMemcpyHostToDevice));
        dmin = s->cl_range;
        dmax = s->cl_range + kClosureFeatures;
        for (int r = 0; r < s->d.n_rep; ++r) HIPCHK(s, hipMemset(&s->d.ctl[r].closure_oor, 0, 4));
    }
    const int S = (int)s->step_update;
    for (int64_t it = 0; it < n_intervals; ++it) {
        LAUNCH(s, 0, s->stream, launch_md(s->d, s->stream, S, 0, 0, 1, s->md_dt));
        LAUNCH(s, 0, s->stream, launch_closure_features(s->d, s->stream, s->cl_feat, dmin, dmax));
        if (int rc = spmd_mlp_forward_on(m, s->cl_feat, s->cl_out, s->d.m_tot, s->stream))
            return fail(s, rc, std::string("closure network: ") + spmd_mlp_last_error(m));
        LAUNCH(s, 0, s->stream, launch_closure_kick(s->d, s->stream, s->cl_out, sy[0], sy[1], sy[2], s->cl_k));
        s->cl_k++;
    }
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (oor_host) {
        std::vector<RepCtl> c(s->d.n_rep);
        HIPCHK(s, hipMemcpy(c.data(), s->d.ctl, c.size() * sizeof(RepCtl), hipMemcpyDeviceToHost));
        for (int r = 0; r < s->d.n_rep; ++r) oor_host[r] = fmin ? c[r].closure_oor : 0u;
    }
    return check_status(s);
}

/* star moments (kStarMoments doubles, see launch_star_moments) of replica r
 * into s->mom and the host array out */
static int star_moments(spmd_sys* s, uint32_t r, double* out) {
    if (int rc = flush_md(s)) return rc;
    if (int rc = ensure_monomers(s)) return rc;
    if (!s->mom)
        if (int rc = dalloc(s, &s->mom, kStarMoments + 2)) return rc;
    HIPCHK(s, launch_star_moments(s->d, s->stream, (int)r, s->mom));
    if (out) {
        HIPCHK(s, hipMemcpyAsync(out, s->mom, kStarMoments * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
    }
    return SPMD_OK;
}

int spmd_gyration_tensor(spmd_sys* s, uint32_t r, double* g9) {
    if (!valid_rep(s, r) || !g9) return SPMD_ERR_ARG;
    double out[kStarMoments];
    if (int rc = star_moments(s, r, out)) return rc;
    for (int q = 0; q < 9; ++q) g9[q] = out[q];
    return SPMD_OK;
}

int spmd_rotation_frequency(spmd_sys* s, uint32_t r, double* w3) {
    if (!valid_rep(s, r) || !w3) return SPMD_ERR_ARG;
    double out[kStarMoments];
    if (int rc = star_moments(s, r, out)) return rc;
    for (int q = 0; q < 3; ++q) w3[q] = out[9 + q];
    return SPMD_OK;
}

int spmd_patches(spmd_sys* s, uint32_t r, double* n_patch, double* patch_size) {
    if (!valid_rep(s, r)) return SPMD_ERR_ARG;
    const StarHost& st = s->stars[r];
    if (st.arms == 0) return fail(s, SPMD_ERR_STATE, "replica has no star");
    if (st.arms > (uint32_t)kPatchArms) return fail(s, SPMD_ERR_ARG, "patch analysis supports at most 128 arms");
    if (int rc = flush_md(s)) return rc;
    if (int rc = ensure_monomers(s)) return rc;
    if (!s->mom)
        if (int rc = dalloc(s, &s->mom, kStarMoments + 2)) return rc;
    HIPCHK(s, launch_star_patches(s->d, s->stream, (int)r, (int)st.arms, (int)st.A, (int)st.B,
                                  s->mom + kStarMoments));
    double out[2];
    HIPCHK(s, hipMemcpyAsync(out, s->mom + kStarMoments, sizeof out, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (n_patch) *n_patch = out[0];
    if (patch_size) *patch_size = out[1];
    return SPMD_OK;
}

int spmd_statistics(spmd_sys* s, uint32_t r, double* row) {
    if (!valid_rep(s, r) || !row) return SPMD_ERR_ARG;
    const StarHost& st = s->stars[r];
    if (st.n == 0) return fail(s, SPMD_ERR_STATE, "replica has no star");
    double m[kStarMoments];
    if (int rc = star_moments(s, r, m)) return rc;
    double np = 0.0, sp = 0.0;
    if (int rc = spmd_patches(s, r, &np, &sp)) return rc;
    double epot = 0.0;
    HIPCHK(s, hipMemcpy(&epot, &s->d.ctl[r].epot, 8, hipMemcpyDeviceToHost));
    const double N = (double)st.n;
    row[0] = epot / N;                                   /* Box::print_Epot */
    row[1] = m[18];                                      /* Box::print_Ekin_and_Temperature */
    row[2] = m[18] * 2. / (3. * (N - 1.));
    row[3] = np;
    row[4] = sp;
    row[5] = std::sqrt((m[0] + m[4]) + m[8]);            /* R_gyr of calculate_gyration_tensor */
    for (int q = 0; q < 9; ++q) row[6 + q] = m[q];       /* symmetric: row- and column-major agree */
    for (int q = 0; q < 3; ++q) row[15 + q] = m[9 + q];
    return SPMD_OK;
}

/* observer workspace: [2][kObsBins][3] profile, 18 Fourier, kObsSpeedBins+1 speed */
static int obs_ws(spmd_sys* s) {
    if (!s->obs)
        if (int rc = dalloc(s, &s->obs, (size_t)6 * kObsBins + 18 + kObsSpeedBins + 1)) return rc;
    return SPMD_OK;
}

static void profile_bins(const RepDev& rp, double width, int axis, int* b0, int* nb) {
    const double h = rp.box.half[axis];
    *b0 = (int)std::floor(-h / width);
    *nb = (int)std::floor(h / width) - *b0 + 1;
}

int spmd_velocity_profile(spmd_sys* s, uint32_t r, double width, uint32_t max_bins, int32_t* b0, uint32_t* nbins,
                          double* sums) {
    if (!valid_rep(s, r) || !(width > 0) || !b0 || !nbins || !sums) return SPMD_ERR_ARG;
    const RepDev& rp = s->reps[r];
    ObsArgs a{};
    a.replica = (int)r;
    a.flags = kObsProfile;
    a.w = width;
    for (int ax = 0; ax < 2; ++ax) {
        profile_bins(rp, width, ax, &a.b0[ax], &a.nb[ax]);
        if (a.nb[ax] > kObsBins) return fail(s, SPMD_ERR_ARG, "profile bin width too small (1024 bins per axis at most)");
        if ((uint32_t)a.nb[ax] > max_bins) return fail(s, SPMD_ERR_ARG, "max_bins smaller than the number of bins");
        b0[ax] = a.b0[ax];
        nbins[ax] = (uint32_t)a.nb[ax];
    }
    if (int rc = obs_ws(s)) return rc;
    if (s->stars[r].n > 0) {
        if (int rc = star_moments(s, r, nullptr)) return rc;
        a.com = s->mom + 12;
    }
    a.prof = s->obs;
    HIPCHK(s, hipMemsetAsync(s->obs, 0, 6ull * kObsBins * 8, s->stream));
    HIPCHK(s, launch_fluid_observe(s->d, s->stream, a, rp.n_fluid));
    std::vector<unsigned long long> h(6ull * kObsBins);
    HIPCHK(s, hipMemcpyAsync(h.data(), s->obs, h.size() * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    const double unfix = 2.3283064365386962890625e-10;
    for (int ax = 0; ax < 2; ++ax)
        for (int b = 0; b < a.nb[ax]; ++b) {
            const unsigned long long* q = h.data() + 3ull * (ax * kObsBins + b);
            double* o = sums + 3ull * ((size_t)ax * max_bins + b);
            o[0] = (double)(long long)q[0] * unfix;
            o[1] = (double)(long long)q[1] * unfix;
            o[2] = (double)q[2];
        }
    return SPMD_OK;
}

int spmd_velocity_profile_noalloc(spmd_sys* s, uint32_t r, double width, uint32_t max_bins, int32_t* b0,
                                  uint32_t* nbins, double* sums_dev) {
    if (!valid_rep(s, r) || !(width > 0) || !b0 || !nbins || !sums_dev) return SPMD_ERR_ARG;
    if (!s->scratch) return fail(s, SPMD_ERR_STATE, "no device scratch on this handle");
    if (s->mono_dirty)
        return fail(s, SPMD_ERR_STATE, "the monomer buffers await a rebuild, which allocates: call spmd_md_init first");
    const RepDev& rp = s->reps[r];
    ObsArgs a{};
    a.replica = (int)r;
    a.flags = kObsProfile;
    a.w = width;
    for (int ax = 0; ax < 2; ++ax) {
        profile_bins(rp, width, ax, &a.b0[ax], &a.nb[ax]);
        if (a.nb[ax] > kObsBins)
            return fail(s, SPMD_ERR_ARG, "profile bin width too small: the scratch holds 1024 bins per axis");
        if ((uint32_t)a.nb[ax] > max_bins) return fail(s, SPMD_ERR_ARG, "max_bins smaller than the number of bins");
    }
    for (int ax = 0; ax < 2; ++ax) { b0[ax] = a.b0[ax]; nbins[ax] = (uint32_t)a.nb[ax]; }
    const DevState& d = s->d;
    /* pending substeps and delrx advances first, as every observer sees them */
    if (s->pending_md || s->pending_delrx) {
        HIPCHK(s, launch_prologue(d, s->stream, 0, s->pending_delrx, s->pending_dt));
        HIPCHK(s, launch_md(d, s->stream, s->pending_md, s->pending_epot, 0, 0, s->md_dt));
        s->pending_md = 0;
        s->pending_delrx = 0;
    }
    unsigned long long* acc = s->scratch;
    double* mom = reinterpret_cast<double*>(s->scratch + 6ull * kObsBins);
    HIPCHK(s, launch_fill_u64(acc, (uint32_t)(6 * kObsBins), s->stream));
    if (s->stars[r].n > 0) {
        HIPCHK(s, launch_star_moments(d, s->stream, (int)r, mom));
        a.com = mom + 12;
    }
    a.prof = acc;
    HIPCHK(s, launch_fluid_observe(d, s->stream, a, rp.n_fluid));
    HIPCHK(s, launch_profile_export(acc, a.nb[0], a.nb[1], max_bins, sums_dev, s->stream));
    return SPMD_OK;
}

int spmd_fourier_modes(spmd_sys* s, uint32_t r, double* modes) {
    if (!valid_rep(s, r) || !modes) return SPMD_ERR_ARG;
    const RepDev& rp = s->reps[r];
    if (rp.n_fluid == 0) return fail(s, SPMD_ERR_STATE, "replica has no solvent");
    if (int rc = obs_ws(s)) return rc;
    ObsArgs a{};
    a.replica = (int)r;
    a.flags = kObsFourier;
    a.nb[0] = a.nb[1] = 1;
    a.four = s->obs + 6ull * kObsBins;
    HIPCHK(s, hipMemsetAsync(a.four, 0, 18 * 8, s->stream));
    HIPCHK(s, launch_fluid_observe(s->d, s->stream, a, rp.n_fluid));
    unsigned long long h[18];
    HIPCHK(s, hipMemcpyAsync(h, a.four, sizeof h, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    const double unfix = 2.3283064365386962890625e-10, n = (double)rp.n_fluid;
    for (int q = 0; q < 18; ++q) modes[q] = (double)(long long)h[q] * unfix / n;
    /* Fourier_Autocorr.h:68: drop the component along k_j */
    for (int j = 0; j < 3; ++j) modes[6 * j + 2 * j] = modes[6 * j + 2 * j + 1] = 0.0;
    return SPMD_OK;
}

int spmd_speed_histogram(spmd_sys* s, uint32_t r, double width, uint32_t nbins, uint64_t* counts) {
    if (!valid_rep(s, r) || !(width > 0) || nbins == 0 || !counts) return SPMD_ERR_ARG;
    if (nbins > (uint32_t)kObsSpeedBins) return fail(s, SPMD_ERR_ARG, "at most 1024 speed bins");
    const RepDev& rp = s->reps[r];
    if (int rc = obs_ws(s)) return rc;
    ObsArgs a{};
    a.replica = (int)r;
    a.flags = kObsSpeed;
    a.vw = width;
    a.nb[0] = a.nb[1] = 1;
    a.nsb = (int)nbins;
    a.speed = s->obs + 6ull * kObsBins + 18;
    HIPCHK(s, hipMemsetAsync(a.speed, 0, (nbins + 1) * 8ull, s->stream));
    HIPCHK(s, launch_fluid_observe(s->d, s->stream, a, rp.n_fluid));
    HIPCHK(s, hipMemcpyAsync(counts, a.speed, (nbins + 1) * 8ull, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return SPMD_OK;
}

int spmd_fluid_vacf_sample(spmd_sys* s, uint32_t r, uint32_t entries) {
    if (!valid_rep(s, r) || entries == 0) return SPMD_ERR_ARG;
    if (entries > (uint32_t)kVacfMaxEntries) return fail(s, SPMD_ERR_ARG, "at most 1024 VACF entries");
    DevState& d = s->d;
    if (!d.ids[d.cur]) return fail(s, SPMD_ERR_STATE, "the fluid VACF follows particles by id: turn tracking on");
    if (s->vacf.size() < s->reps.size()) s->vacf.resize(s->reps.size());
    spmd_sys::Vacf& v = s->vacf[r];
    const RepDev& rp = s->reps[r];
    if (!v.hist) {
        if (int rc = dalloc(s, &v.hist, 3ull * rp.n_fluid * entries)) return rc;
        if (int rc = dalloc(s, &v.sums, entries)) { dfree(v.hist); return rc; }
        HIPCHK(s, hipMemsetAsync(v.sums, 0, entries * 8ull, s->stream));
        v.entries = entries;
        v.calls = v.counter = 0;
    } else if (v.entries != entries) {
        return fail(s, SPMD_ERR_ARG, "entries differ from the running VACF's (reset it first)");
    }
    /* the window after this sample holds samples calls-E+1 .. calls */
    const uint32_t slot = (uint32_t)(v.calls % v.entries);
    const int calc = v.calls + 1 >= v.entries;
    const uint32_t oldest = (uint32_t)((v.calls + 1) % v.entries);
    HIPCHK(s, launch_fluid_vacf(d, s->stream, (int)r, rp.n_fluid, v.hist, v.entries, slot, oldest, calc, v.sums));
    v.calls++;
    if (calc) v.counter++;
    return SPMD_OK;
}

int spmd_fluid_vacf_read(spmd_sys* s, uint32_t r, double* sums, uint64_t* counter) {
    if (!valid_rep(s, r) || !sums || !counter) return SPMD_ERR_ARG;
    if (s->vacf.size() <= r || !s->vacf[r].hist) return fail(s, SPMD_ERR_STATE, "no VACF sample taken");
    const spmd_sys::Vacf& v = s->vacf[r];
    std::vector<unsigned long long> q(v.entries);
    HIPCHK(s, hipMemcpyAsync(q.data(), v.sums, v.entries * 8ull, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    for (uint32_t j = 0; j < v.entries; ++j) sums[j] = (double)(long long)q[j] * (1.0 / 4294967296.0);
    *counter = v.counter;
    return SPMD_OK;
}

int spmd_fluid_vacf_reset(spmd_sys* s, uint32_t r) {
    if (!valid_rep(s, r)) return SPMD_ERR_ARG;
    if (s->vacf.size() <= r) return SPMD_OK;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    dfree(s->vacf[r].hist);
    dfree(s->vacf[r].sums);
    s->vacf[r] = spmd_sys::Vacf{};
    return SPMD_OK;
}

int spmd_get_params(const spmd_sys* s, uint32_t r, spmd_params* p) {
    if (!valid_rep(s, r) || !p) return SPMD_ERR_ARG;
    *p = s->params[r];
    return SPMD_OK;
}

int spmd_star_layout(const spmd_sys* s, uint32_t r, uint32_t* n, uint32_t* arms, uint32_t* A, uint32_t* B) {
    if (!valid_rep(s, r)) return SPMD_ERR_ARG;
    const StarHost& st = s->stars[r];
    if (n) *n = st.n;
    if (arms) *arms = st.arms;
    if (A) *A = st.A;
    if (B) *B = st.B;
    return SPMD_OK;
}

/* ------------------------------------------------------------ snapshots */

}  // extern "C"

namespace {

constexpr char kSnapMagic[8] = {'S', 'P', 'M', 'D', 'S', 'N', 'P', '1'};
constexpr uint32_t kSnapVersion = 1;

struct SnapFile {
    FILE* f;
    bool ok = true;
    explicit SnapFile(FILE* fp) : f(fp) {}
    ~SnapFile() {
        if (f) std::fclose(f);
    }
    void put(const void* p, size_t n) { ok = ok && (n == 0 || std::fwrite(p, 1, n, f) == n); }
    void get(void* p, size_t n) { ok = ok && (n == 0 || std::fread(p, 1, n, f) == n); }
    template <typename T> void put(const T& v) { put(&v, sizeof v); }
    template <typename T> void get(T& v) { get(&v, sizeof v); }
    template <typename T> void put_vec(const std::vector<T>& v) { put(v.data(), v.size() * sizeof(T)); }
    template <typename T> void get_vec(std::vector<T>& v, size_t n) { v.resize(n); get(v.data(), n * sizeof(T)); }
};

}  // namespace

extern "C" {

int spmd_snapshot_save(spmd_sys* s, const char* path, int64_t step) {
    if (!s || !path) return SPMD_ERR_ARG;
    if (int rc = flush_md(s)) return rc;
    if (int rc = ensure_monomers(s)) return rc;
    if (int rc = download_monomers(s)) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (int rc = check_status(s)) return rc;
    SnapFile out(std::fopen(path, "wb"));
    if (!out.f) return fail(s, SPMD_ERR_ARG, std::string("cannot write ") + path);
    const DevState& d = s->d;
    const uint32_t nrep = (uint32_t)d.n_rep, tracking = d.ids[0] ? 1u : 0u;
    out.put(kSnapMagic, sizeof kSnapMagic);
    out.put(kSnapVersion);
    out.put(nrep);
    out.put((uint32_t)sizeof(spmd_params));
    out.put(step);
    out.put(s->md_dt);
    out.put(tracking);
    out.put(s->cl_k);
    std::vector<RepCtl> ctl(nrep);
    HIPCHK(s, hipMemcpy(ctl.data(), d.ctl, nrep * sizeof(RepCtl), hipMemcpyDeviceToHost));
    for (uint32_t r = 0; r < nrep; ++r) {
        const RepDev& rp = s->reps[r];
        out.put(s->params[r]);
        out.put(ctl[r].delrx);
        out.put(ctl[r].k_done);
        const uint64_t n = rp.n_fluid;
        out.put(n);
        std::vector<float> t(n);
        const float* src[6] = {d.x[d.cur], d.y[d.cur], d.z[d.cur], d.vx[d.cur], d.vy[d.cur], d.vz[d.cur]};
        for (int c = 0; c < 6; ++c) {
            HIPCHK(s, hipMemcpy(t.data(), src[c] + rp.part_base, n * 4, hipMemcpyDeviceToHost));
            out.put_vec(t);
        }
        if (tracking) {
            std::vector<uint32_t> id(n);
            HIPCHK(s, hipMemcpy(id.data(), d.ids[d.cur] + rp.part_base, n * 4, hipMemcpyDeviceToHost));
            out.put_vec(id);
        }
        const StarHost& st = s->stars[r];
        out.put(st.n);
        out.put(st.arms);
        out.put(st.A);
        out.put(st.B);
        if (st.n) {
            out.put_vec(st.pos);
            out.put_vec(st.vel);
            out.put_vec(st.force);
            out.put_vec(st.mass);
            out.put_vec(st.btype);
            out.put_vec(st.anchor);
            const uint32_t nb = (uint32_t)st.bonds.size();
            out.put(nb);
            for (auto& b : st.bonds) { out.put(b.first); out.put(b.second); }
        }
    }
    if (!out.ok) return fail(s, SPMD_ERR_ARG, std::string("short write to ") + path);
    return SPMD_OK;
}

int spmd_snapshot_load(const char* path, spmd_sys** res, int64_t* step) {
    if (!path || !res) return SPMD_ERR_ARG;
    *res = nullptr;
    SnapFile in(std::fopen(path, "rb"));
    if (!in.f) return SPMD_ERR_ARG;
    char magic[8];
    uint32_t version = 0, nrep = 0, psize = 0, tracking = 0, cl_k = 0;
    int64_t st0 = 0;
    double md_dt = 0.0;
    in.get(magic, sizeof magic);
    in.get(version);
    in.get(nrep);
    in.get(psize);
    in.get(st0);
    in.get(md_dt);
    in.get(tracking);
    in.get(cl_k);
    if (!in.ok || std::memcmp(magic, kSnapMagic, 8) || version != kSnapVersion || psize != sizeof(spmd_params) ||
        nrep == 0 || nrep > (1u << 20))
        return SPMD_ERR_ARG;
    struct Rep {
        spmd_params p;
        double delrx;
        uint32_t k_done;
        std::vector<float> f[6];
        std::vector<uint32_t> ids;
        StarHost st;
    };
    std::vector<Rep> reps(nrep);
    for (auto& rp : reps) {
        in.get(rp.p);
        in.get(rp.delrx);
        in.get(rp.k_done);
        uint64_t n = 0;
        in.get(n);
        if (!in.ok || n > (1ull << 34)) return SPMD_ERR_ARG;
        for (int c = 0; c < 6; ++c) in.get_vec(rp.f[c], n);
        if (tracking) in.get_vec(rp.ids, n);
        StarHost& st = rp.st;
        in.get(st.n);
        in.get(st.arms);
        in.get(st.A);
        in.get(st.B);
        if (!in.ok || st.n > (uint32_t)kMaxMono) return SPMD_ERR_ARG;
        if (st.n) {
            in.get_vec(st.pos, 3ull * st.n);
            in.get_vec(st.vel, 3ull * st.n);
            in.get_vec(st.force, 3ull * st.n);
            in.get_vec(st.mass, st.n);
            in.get_vec(st.btype, st.n);
            in.get_vec(st.anchor, st.n);
            uint32_t nb = 0;
            in.get(nb);
            if (!in.ok || nb > 4u * st.n) return SPMD_ERR_ARG;
            st.bonds.resize(nb);
            for (auto& b : st.bonds) { in.get(b.first); in.get(b.second); }
        }
        if (!in.ok) return SPMD_ERR_ARG;
    }
    std::vector<spmd_params> ps(nrep);
    for (uint32_t r = 0; r < nrep; ++r) ps[r] = reps[r].p;
    spmd_sys* s = nullptr;
    if (int rc = spmd_create_batch(ps.data(), nrep, &s)) return rc;
    auto bail = [&](int rc) { spmd_destroy(s); return rc; };
    s->md_dt = md_dt;
    s->cl_k = cl_k;
    if (tracking)
        if (int rc = spmd_set_tracking(s, 1)) return bail(rc);
    DevState& d = s->d;
    for (uint32_t r = 0; r < nrep; ++r) {
        const RepDev& rp = s->reps[r];
        if (reps[r].f[0].size() != rp.n_fluid) return bail(SPMD_ERR_ARG);
        float* dst[6] = {d.x[d.cur], d.y[d.cur], d.z[d.cur], d.vx[d.cur], d.vy[d.cur], d.vz[d.cur]};
        for (int c = 0; c < 6; ++c)
            if (hipMemcpy(dst[c] + rp.part_base, reps[r].f[c].data(), rp.n_fluid * 4, hipMemcpyHostToDevice))
                return bail(SPMD_ERR_HIP);
        if (tracking && hipMemcpy(d.ids[d.cur] + rp.part_base, reps[r].ids.data(), rp.n_fluid * 4,
                                  hipMemcpyHostToDevice))
            return bail(SPMD_ERR_HIP);
        if (hipMemcpy(&d.ctl[r].delrx, &reps[r].delrx, 8, hipMemcpyHostToDevice) ||
            hipMemcpy(&d.ctl[r].k_done, &reps[r].k_done, 4, hipMemcpyHostToDevice))
            return bail(SPMD_ERR_HIP);
        s->stars[r] = std::move(reps[r].st);
    }
    s->mono_dirty = true;
    /* Verlet lists and forces from the saved positions (the forces are a
     * function of the positions alone, so they equal the saved ones) */
    if (d.n_rep > 0) {
        bool any = false;
        for (auto& st : s->stars) any = any || st.n > 0;
        if (any)
            if (int rc = spmd_md_init(s)) return bail(rc);
    }
    if (step) *step = st0;
    *res = s;
    return SPMD_OK;
}

int spmd_md_stats(spmd_sys* s, uint32_t r, uint32_t* n_rebuild, uint32_t* vl_entries) {
    if (!valid_rep(s, r)) return SPMD_ERR_ARG;
    if (int rc = flush_md(s)) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    RepCtl c;
    HIPCHK(s, hipMemcpy(&c, s->d.ctl + r, sizeof c, hipMemcpyDeviceToHost));
    if (n_rebuild) *n_rebuild = c.n_rebuild;
    if (vl_entries) *vl_entries = c.vl_count;
    return SPMD_OK;
}

int spmd_profile_enable(spmd_sys* s, int on) {
    if (!s) return SPMD_ERR_ARG;
    s->prof = on != 0;
    return SPMD_OK;
}

int spmd_profile_reset(spmd_sys* s) {
    if (!s) return SPMD_ERR_ARG;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    for (auto& pe : s->pev) {
        for (auto e : pe.ev) if (e) (void)hipEventDestroy(e);
        pe.ev.clear();
    }
    return SPMD_OK;
}

int spmd_profile_read(spmd_sys* s, const char* name, double* avg_ms, int64_t* count) {
    if (!s || !name) return SPMD_ERR_ARG;
    int k = -1;
    for (int i = 0; i < kProfKinds; ++i)
        if (!std::strcmp(name, kProfNames[i])) k = i;
    if (k < 0) return fail(s, SPMD_ERR_ARG, "unknown kernel class");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    double tot = 0.0;
    int64_t n = 0;
    auto& ev = s->pev[k].ev;
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
        float ms = 0.f;
        if (ev[i] && ev[i + 1] && hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) { tot += ms; n++; }
    }
    if (avg_ms) *avg_ms = n ? tot / n : 0.0;
    if (count) *count = n;
    return SPMD_OK;
}

}  // extern "C"
