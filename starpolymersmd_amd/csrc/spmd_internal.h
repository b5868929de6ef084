/*
 * spmd_internal.h -- device-resident state of a handle and the kernel
 * launchers shared by spmd_kernels.hip and spmd_api.hip.
 *
 * Layout in HBM (all replicas of a handle concatenated):
 *   solvent   x,y,z,vx,vy,vz   float[N_tot] each, two copies (cur/alt): the
 *             cell sort writes alt, which becomes cur after the collision
 *   rank      uint16[N_tot]    slot of a particle inside its new cell
 *   counts    uint32[C_tot]    particles per cell (zeroed by the scan)
 *   offs[b]   uint32[C_tot+1]  first particle of each cell in buffer b
 *                              (each collision writes the offsets of the
 *                              buffer it fills)
 *   bucket    uint32[tiles][rt_cap]  row-tile path: indices of the particles
 *                              whose new cell lies in each row tile;
 *                              bcount[b] their numbers at the collision
 *                              that reads buffer b, tbase their scan
 *   monomers  pos/vel/force/vpos double[3*M_tot], mass double[M_tot],
 *             btype uint8[M_tot], Verlet half lists uint32[M_tot*64]
 *             (i | j << 16, per replica from mono_base*64), bonds
 *             uint32[bonds] (i | j << 12 | anchor << 24, each bond once)
 *   mcell     uint32[M_tot]    global cell of each monomer at the last
 *                              collision; each collide tile gathers its own
 *   RepDev    per-replica constants, RepCtl per-replica step state
 */
#ifndef SPMD_INTERNAL_H
#define SPMD_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spmd_math.h"

namespace spmd {

constexpr int kVlPerMono = 64;     /* global Verlet half-list entries reserved per monomer */
constexpr int kMaxMono = 2048;     /* monomers per replica (one MD workgroup) */
constexpr int kMdThreads = 512;
constexpr int kScanItems = 16;     /* counts per thread in the scan */
constexpr int kScanThreads = 256;
constexpr int kScanTile = kScanItems * kScanThreads;

enum : uint32_t {
    ST_CELL_OOB = 1u << 0,
    ST_VERLET_OOB = 1u << 1,
    ST_VERLET_OVERFLOW = 1u << 2,
    ST_RANK_OVERFLOW = 1u << 3,
    ST_SCAN_TIMEOUT = 1u << 4,
    ST_TILE_MONO = 1u << 5,      /* more monomers in a collide tile than its list holds */
    ST_SORT_MISMATCH = 1u << 6,  /* collide recomputed a cell outside its tile (internal error) */
    ST_TILE_OVERFLOW = 1u << 7,  /* more particles arrived in a row tile than it holds */
    ST_GAMMA_MISS = 1u << 8,     /* a cell's Gamma variate was not drawn for its count (internal error) */
};

/* Row-tile path: a row tile is `nr` consecutive y-rows (whole x-rows) of
 * one z-plane. */
constexpr int kRowsMax = 16;        /* rows per row tile */
constexpr int kRowTileCells = 256;  /* cells per row tile */

/* bond entry: partner index | anchor-FENE flag in bit 31 */
constexpr uint32_t kBondAnchor = 0x80000000u;

struct RepDev {
    Box box;               /* delrx field unused here: lives in RepCtl */
    double kT, c, s, lambda;
    uint64_t seed;
    uint32_t replica;
    uint32_t ncells;
    uint64_t cell_base;
    uint64_t part_base;
    uint32_t n_fluid;
    uint32_t mono_base, n_mono;
    uint32_t bond_base, n_bond;  /* this replica's entries of the bond list */
    int vgrid[3];          /* Verlet grid cells per side (Box.cpp:20-22) */
    double vside[3];       /* Verlet grid cell side (Box.cpp:23-25) */
    int amc;               /* angular-momentum-conserving collision (MPC angular_mom) */
    int nr;                /* row-tile path: y-rows per tile */
    int tpp;               /* row tiles per z-plane, ceil(Ly / nr) */
    uint32_t rt_base;      /* first row tile of this replica */
};

struct RepCtl {
    double delrx;
    double shift[3];
    double com[3], com_vel[3];   /* check_bounds recentring of this collision */
    double vcom[3];              /* COM at the last Verlet rebuild */
    double epot;
    double inj_shift[3];
    uint32_t k_cur, k_done;      /* Philox collision counter */
    int32_t recenter;
    int32_t inj_pending, inj_active;
    uint32_t status;
    uint32_t n_rebuild;
    int32_t redo;                /* the speculative sort (no recentring) must be redone */
    uint32_t vl_count;           /* entries of the star's Verlet half list */
    uint32_t closure_oor;        /* monomers whose closure features left the trained range */
    uint32_t cg_k;               /* k_cur + 1 of the drawn cell Gammas (0: none) */
};

struct DevState {
    int n_rep;
    uint64_t n_tot;        /* solvent particles, all replicas */
    uint64_t c_tot;        /* cells, all replicas */
    uint32_t m_tot;        /* monomers, all replicas */
    float *x[2], *y[2], *z[2], *vx[2], *vy[2], *vz[2];
    uint32_t* ids[2];      /* null unless tracking */
    int cur;
    uint16_t* rank;        /* slot of a particle inside its new cell (k_bin) */
    uint32_t* pcell;       /* global new cell of a particle, ~0u if none (k_bin) */
    uint32_t* perm;        /* sorted slot -> particle index in the current buffer (k_perm) */
    uint32_t* counts;
    uint32_t* offs[2];     /* first particle of each cell in buffer 0 / 1 (C_tot + 1) */
    /* row-tile path */
    uint32_t* bcount[2];   /* particles binned per row tile, by the buffer the collision reads */
    uint32_t* bcount_r;    /* rebinning of a recentred replica (redo kernel) */
    uint32_t* bucket;      /* particle indices per row tile, rt_cap each */
    uint32_t* bucket_r;    /* the rebinning's */
    uint32_t* tbase;       /* first slot of each row tile in the new buffer */
    uint32_t* ttag;        /* per row tile: k_cur + 1 of the collision whose monomers it holds */
    uint8_t* blc;          /* beside each bucket entry: the particle's cell in its tile */
    uint8_t* blc_r;        /* the rebinning's */
    /* per cell (C_tot): the Gamma variate of the coming collision drawn
     * ahead of it (k_cell_gamma, the redo kernel) and the member count it
     * was drawn for */
    uint32_t* cg_n;
    double* cg_g;
    uint32_t* rtile_base;  /* row-tile prefix per replica (n_rep+1) */
    uint32_t n_rtiles;
    int rt_cap;            /* particles a row tile holds (1024 / 2048), 0: path unavailable */
    int rt_tcm;            /* cells a row tile holds (64 / 128 / 256) */
    uint64_t* scan_flags;  /* per scan tile: status<<62 | value */
    uint32_t* scan_ticket;
    uint32_t* grid_bar;    /* the redo kernel's grid barrier (reset by the prologue) */
    uint32_t n_scan_tiles;
    double *mpos, *mvel, *mforce, *mvpos, *mmass;
    uint8_t* mbtype;
    uint32_t* vl;          /* Verlet half lists: i | j << 16, kVlPerMono per monomer; then as
                              many entries of rebuild scratch rows */
    uint32_t* bonds;
    uint32_t* mcell;       /* global cell of each monomer at the last collision (~0u: none) */
    double* inj_axis;      /* 2 per cell, all replicas */
    double* inj_gamma;     /* 1 per cell */
    RepDev* reps;          /* device copy */
    RepCtl* ctl;           /* device */
    uint32_t* pblk_base;   /* particle-block prefix per replica (n_rep+1) */
    uint32_t* tile_base;   /* collide-tile prefix per replica (n_rep+1) */
    double* red;           /* scratch for reductions: 4 doubles per block */
    uint32_t n_pblk;       /* total particle blocks */
    uint32_t n_tiles;      /* total collide tiles */
    int md_ns;             /* LDS stride of the MD kernel: max monomers per replica (>= 1) */
    uint32_t md_vl_cap;    /* Verlet entries the MD kernel keeps in LDS */
    int ppt;               /* particles per thread in bin/scatter */
    int tile_cells;        /* cells per collide tile */
    int any_amc;           /* some replica collides with angular-momentum conservation */
};

/* bytes of dynamic LDS the MD kernel needs for stride ns and cap entries */
size_t md_lds_bytes(int ns, uint32_t cap);

/* launchers (spmd_kernels.hip).  Sort modes: kSortNormal applies the
 * recentring flag, kSortSpec assumes none (runs beside the MD kernel),
 * (k_sort_redo redoes the replicas whose MD found a recentring). */
enum { kSortNormal = 0, kSortSpec = 1 };
hipError_t launch_prologue(const DevState& d, hipStream_t st, int do_collision, int n_delrx, double delrx_dt);
hipError_t launch_md(const DevState& d, hipStream_t st, int n_md, int epot_last, int md_init, int do_collision,
                     double md_dt);
hipError_t launch_bin(const DevState& d, hipStream_t st, double h, int mode);
hipError_t launch_scan(const DevState& d, hipStream_t st);
hipError_t launch_sort_redo(const DevState& d, hipStream_t st, double h);
hipError_t launch_perm(const DevState& d, hipStream_t st, int mode);
hipError_t launch_collide(const DevState& d, hipStream_t st, double h);
/* row-tile path: k_advect (stream, new cell, arrivals per row tile), the
 * redo kernel (monomer cells; a recentred replica's advect again) and the
 * row-window collide */
hipError_t launch_advect(const DevState& d, hipStream_t st, double h, int mode);
hipError_t launch_rows_redo(const DevState& d, hipStream_t st, double h);
hipError_t launch_collide_rows(const DevState& d, hipStream_t st, double h);
/* the Gamma variate of every cell for its predicted member count (after
 * k_advect; the redo kernel redraws the cells holding monomers) */
hipError_t launch_cell_gamma(const DevState& d, hipStream_t st);
hipError_t launch_fluid_init(const DevState& d, hipStream_t st);
/* learned closure: 24 features per monomer (all replicas, global monomer
 * order) with an optional range check against [fmin, fmax] (device arrays
 * of 24, null for none); the drift + noise kick from the network output
 * (m_tot x 6: mean and log-variance in units of sdy) */
constexpr int kClosureFeatures = 24;
hipError_t launch_closure_features(const DevState& d, hipStream_t st, float* feat, const float* fmin,
                                   const float* fmax);
/* one replica's star: gyration tensor (out[0..8]), rotation frequency
 * (out[9..11]), COM position and velocity (out[12..17]), kinetic energy
 * relative to the COM velocity (out[18]) */
constexpr int kStarMoments = 19;
hipError_t launch_star_moments(const DevState& d, hipStream_t st, int replica, double* out);
/* patch number and mean patch size of one replica's star (arms <= kPatchArms) */
constexpr int kPatchArms = 128;
hipError_t launch_star_patches(const DevState& d, hipStream_t st, int replica, int arms, int A, int B, double* out);
/* solvent observers of one replica (flags kObs*); sums accumulate into the
 * device arrays, which the caller clears */
constexpr int kObsBins = 1024;        /* bins per axis of the velocity profile */
constexpr int kObsSpeedBins = 1024;   /* speed histogram bins (+1 for the rest) */
constexpr int kObsPerThread = 8;
enum { kObsProfile = 1, kObsFourier = 2, kObsSpeed = 4 };
struct ObsArgs {
    int replica;
    int flags;
    double w, vw;                 /* profile and speed bin widths */
    int b0[2], nb[2];             /* first bin index floor(x/w) and bins, x and y maps */
    int nsb;                      /* speed bins; counts[nsb] takes the speeds beyond */
    const double* com;            /* 6: COM position and velocity subtracted, or null */
    unsigned long long* prof;     /* [2][kObsBins][3]: sum vx, sum vy (2^32 fixed), count */
    unsigned long long* four;     /* [3][3][2] (2^32 fixed) */
    unsigned long long* speed;    /* [nsb + 1] */
};
hipError_t launch_fluid_observe(const DevState& d, hipStream_t st, const ObsArgs& a, uint32_t n_fluid);
/* zero n words; the velocity profile's bins as doubles in the layout of
 * spmd_velocity_profile (sums[(a * max_bins + b) * 3 + c], b < nb[a]) */
hipError_t launch_fill_u64(unsigned long long* p, uint32_t n, hipStream_t st);
hipError_t launch_profile_export(const unsigned long long* acc, int nb0, int nb1, uint32_t max_bins, double* out,
                                 hipStream_t st);
/* Vel_Autocorr_Fluid: window slot `slot` <- velocities by id; with calc,
 * sums[lag] += sum over particles of v(oldest) . v(oldest + lag) */
constexpr int kVacfMaxEntries = 1024;
hipError_t launch_fluid_vacf(const DevState& d, hipStream_t st, int r, uint32_t n_fluid, float* hist, uint32_t E,
                             uint32_t slot, uint32_t oldest, int calc, unsigned long long* sums);
hipError_t launch_closure_kick(const DevState& d, hipStream_t st, const float* out, float sy0, float sy1, float sy2,
                               uint32_t k);
hipError_t launch_temperature(const DevState& d, hipStream_t st, int replica, uint32_t ntiles);
/* mean solvent velocity of cells c0 .. c0+n-1 of a replica at the last
 * collision (3 doubles per cell into out, device) */
hipError_t launch_cell_velocity(const DevState& d, hipStream_t st, int replica, uint32_t c0, uint32_t n, double* out);

}  // namespace spmd

/* the closure network's forward pass enqueued on a caller's stream (spmd_mlp.hip) */
struct spmd_mlp;
int spmd_mlp_forward_on(spmd_mlp* m, const float* x_dev, float* out_dev, uint64_t M, hipStream_t st);

#endif
