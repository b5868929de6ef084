/*
 * spmd.h -- C ABI of the MI355X MPC-MD library (libspmd.so).
 *
 * The reference has no plugin ABI: its hot path is the C++ classes MPC
 * (star-polymers/lib/Hydrodynamics/MPC.h:25-122), Thermostat_None
 * (lib/Thermostats/Thermostat_None.h:14) and Box (lib/Box.h:30-121).  Each
 * entry point below replaces one of their members, cited beside it; the C++
 * facade in starpolymersmd_amd/facade/ puts the reference's class names and
 * signatures back on top of this ABI.
 *
 * Conventions
 *  - Every function returns an int status (SPMD_OK = 0).  No exceptions cross
 *    the ABI; spmd_last_error() gives a message for the last failure.
 *  - Host pointers are caller-owned and copied synchronously.
 *  - One HIP stream per handle; a handle is not thread-safe (the reference is
 *    single-threaded with a global RNG, lib/Rand.cpp:6).
 *  - A handle holds one or more replicas (independent boxes).  The
 *    single-system reference API is replica 0 of a handle made by
 *    spmd_create().  Per-replica calls take the replica index.
 *  - Device state: solvent positions/velocities FP32 (structure of arrays),
 *    monomers FP64, per-cell sums 64-bit fixed point, delrx FP64.
 */
#ifndef SPMD_H
#define SPMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SPMD_OK = 0,
    SPMD_ERR_ARG = 1,            /* invalid argument */
    SPMD_ERR_HIP = 2,            /* HIP runtime failure */
    SPMD_ERR_OOM = 3,            /* device allocation failed */
    SPMD_ERR_CELL_OOB = 4,       /* MPC::sort cell index out of range (MPC.cpp:212-219,228-233 raise SIGSEGV) */
    SPMD_ERR_VERLET_OOB = 5,     /* Box::update_VerletLists grid bounds (Box.cpp:133-142 raise SIGSEGV) */
    SPMD_ERR_VERLET_OVERFLOW = 6,/* a monomer has more Verlet neighbours than the list capacity */
    SPMD_ERR_UNSUPPORTED = 7,    /* an option this library does not provide */
    SPMD_ERR_STATE = 8,          /* call out of order */
    SPMD_ERR_RANK_OVERFLOW = 9,  /* more than 65535 particles in one cell */
    SPMD_ERR_NO_DEVICE = 10,     /* no usable gfx950 device */
    SPMD_ERR_SCAN = 11           /* cell-offset scan did not complete (bounded wait expired) */
};

/* Box ctor (Box.cpp:4-27) + MPC ctor (MPC.cpp:10-32) + Thermostat_None dt
 * (Thermostat_None.cpp:5-12) parameters. */
typedef struct spmd_params {
    int32_t  L[3];            /* box; odd sizes rounded up to even (Box.cpp:14-16) */
    double   kT;              /* Temperature (Box, MPC) */
    int32_t  n_c;             /* solvent per cell = MPC's N_c (star-polymers.cpp:354 passes Mass_mono) */
    double   shear;           /* Lees-Edwards shear rate (MPC aShear) */
    uint32_t step_update;     /* collide on every step_update-th MPC::step call (MPC.cpp:119) */
    double   alpha_deg;       /* SRD rotation angle (MPC.cpp:23-24: 130) */
    double   lambda;          /* B-B attraction (Box Lambda) */
    double   md_dt;           /* velocity-Verlet step (Thermostat DeltaT) */
    uint64_t seed;            /* Philox4x32-10 key (replaces mt19937_64(1), Rand.cpp:6) */
    int32_t  angular_momentum;/* MPC angular_mom flag: angular-momentum-conserving collision (MPC.cpp:258-269) */
} spmd_params;

typedef struct spmd_sys spmd_sys;

const char* spmd_error_string(int code);
const char* spmd_last_error(const spmd_sys* s);
const char* spmd_version(void);

/* MPC::MPC(Box&, T, N_c, shear, step, angular_mom) MPC.h:44 with its Box. */
int spmd_create(const spmd_params* p, spmd_sys** out);
/* Many independent replicas advanced together (no reference counterpart:
 * the reference runs one process per parameter point, runExecuter.py:57-98).
 * step_update and md_dt must agree across replicas. */
int spmd_create_batch(const spmd_params* p, uint32_t n_replicas, spmd_sys** out);
int spmd_destroy(spmd_sys* s);
int spmd_num_replicas(const spmd_sys* s, uint32_t* n);
/* Thermostat::dtime(double&) Thermostat.cpp:11-13 */
int spmd_set_md_dt(spmd_sys* s, double dt);

/* ---- solvent (MPC::Fluid, MPC.h:43) ---- */
/* MPC::initialize() MPC.cpp:34-55 (Philox draws), all replicas. */
int spmd_fluid_initialize(spmd_sys* s);
/* MPC::NumberOfParticles() MPC.h:91 */
int spmd_num_fluid(const spmd_sys* s, uint32_t replica, uint64_t* n);
/* Fluid[i].Position / .Velocity, xyz interleaved per particle.  set: n must
 * equal N_c*L^3.  get: device order (cell order of the last collision);
 * ids (optional) gives each particle's index at spmd_set_fluid time when
 * tracking is on. */
int spmd_set_fluid(spmd_sys* s, uint32_t replica, const float* pos, const float* vel, uint64_t n);
int spmd_get_fluid(spmd_sys* s, uint32_t replica, float* pos, float* vel, uint32_t* ids);
/* carry particle ids through the cell sort (parity hooks; adds 8 B per
 * particle per collision) */
int spmd_set_tracking(spmd_sys* s, int on);

/* ---- star (Box::Molecules, Box.h:50) ---- */
/* Box::add_star(A, B, Arms, Mass, Bond, AnchorBond) Box.cpp:39-46 ->
 * Molecule::initialize_open_star Molecule.cpp:126-176 (Philox draws).
 * Adding a star rebuilds the device monomer buffers of every replica; once
 * spmd_md_init has run, that rebuild also redoes its work (Verlet lists and
 * forces of all replicas), so a star added later interacts at once. */
int spmd_add_star(spmd_sys* s, uint32_t replica, uint32_t A, uint32_t B, uint32_t arms,
                  double mass, double bond, double anchor_bond);
/* Same star with seeded random-coil arms (no reference counterpart: the
 * reference builds straight arms only, which for 30-bead arms leave the
 * Verlet grid of Box.cpp:129-142).  The first beads sit on a Fibonacci
 * sphere of radius `anchor_bond`; bead k>=2 at distance `bond` from bead
 * k-1 with r^2 >= 1 to all other beads (Philox draws; where the core is too
 * crowded for that, the bond is lengthened by up to 30%). */
int spmd_add_star_coil(spmd_sys* s, uint32_t replica, uint32_t A, uint32_t B, uint32_t arms,
                       double mass, double bond, double anchor_bond);
int spmd_num_monomers(const spmd_sys* s, uint32_t replica, uint32_t* n);
/* Molecules[0].Monomers[i].Position/Velocity/Force, xyz interleaved */
int spmd_set_monomers(spmd_sys* s, uint32_t replica, const double* pos, const double* vel);
int spmd_get_monomers(spmd_sys* s, uint32_t replica, double* pos, double* vel, double* force);
/* Molecules[0].Monomers[i].Mass (Particle.h: MDParticle::Mass) */
int spmd_get_monomer_masses(const spmd_sys* s, uint32_t replica, double* mass);

/* ---- stepping ---- */
/* Thermostat_None ctor work (Thermostat_None.cpp:5-12): Box::update_VerletLists
 * + Box::calculate_forces_verlet for every replica. */
int spmd_md_init(spmd_sys* s);
/* Thermostat_None::propagate(calc_epot) Thermostat_None.cpp:21-38 (all
 * replicas).  Enqueued lazily: consecutive calls run as one launch. */
int spmd_md_propagate(spmd_sys* s, int calc_epot);
/* MPC::step(t, dt) MPC.cpp:116-142 (all replicas), including the delrx
 * advance on every call (MPC.cpp:117-118). */
int spmd_mpc_step(spmd_sys* s, int64_t t, double dt);
/* The caller loop of star-polymers.cpp:412-536 without I/O: for t in
 * [t0, t0+n): Thermostat_None::propagate(false); MPC::step(t, mpc_dt).
 * Whole collision intervals are replayed from a captured HIP graph. */
int spmd_run(spmd_sys* s, int64_t t0, int64_t n, double mpc_dt);
/* wait for all enqueued work and report device-side errors */
int spmd_sync(spmd_sys* s);

/* ---- parity hooks ---- */
int spmd_set_delrx(spmd_sys* s, uint32_t replica, double delrx);
int spmd_get_delrx(spmd_sys* s, uint32_t replica, double* delrx);
/* randoms for the next collision of one replica: shift[3] in [-0.5,0.5),
 * axis_u[2*L^3] = (u_phi, u_theta) per cell (MPC.cpp:244-245),
 * gamma[L^3] = Gamma(1.5(n-1), kT) variate per cell (MPC.cpp:289). */
int spmd_inject_collision_randoms(spmd_sys* s, uint32_t replica, const double* shift,
                                  const double* axis_u, const double* gamma);
/* shift, recentring flag (MPC::check_bounds) and COM of the last collision */
int spmd_last_collision(spmd_sys* s, uint32_t replica, double* shift, int* recentered,
                        double* com_pos, double* com_vel);
/* MPCParticle::CellIndex of the last MPC::sort() (MPC.cpp:209,225), by
 * particle id (tracking on) and by monomer index. */
int spmd_get_cell_index(spmd_sys* s, uint32_t replica, uint32_t* fluid_cells, uint32_t* mono_cells);

/* ---- observables ---- */
/* MPC::calculateCurrentTemperature MPC.cpp:405-409 over the cell lists of
 * the last collision. */
int spmd_temperature(spmd_sys* s, uint32_t replica, double* T);
/* MPC::calculateFluidVelocity (MPC.cpp:334-342) of every cell with z in
 * [z_start, z_stop], as MPC::print_fluid (MPC.cpp:533-553) reads it: the
 * mean velocity of the solvent particles the last collision sorted into the
 * cell, 0 for an empty cell (and for every cell before the first collision
 * after the solvent was set).  out[3 * (c - Lx Ly z_start) + k] for cell
 * c = x + Lx y + Lx Ly z, component k. */
int spmd_cell_fluid_velocity(spmd_sys* s, uint32_t replica, int32_t z_start, int32_t z_stop, double* out);
/* Molecule::Epot after the last propagate (Box.cpp:243-286). */
int spmd_epot(spmd_sys* s, uint32_t replica, double* epot);
/* gyration tensor of one replica's star about its centre of mass
 * (Box::calculate_gyration_tensor, Box.cpp:330-352), row-major 3x3 */
int spmd_gyration_tensor(spmd_sys* s, uint32_t replica, double* g9);
/* rotation frequency omega = I^-1 L of one replica's star in its centre-of-
 * mass frame (Box::calculate_rotation_frequency, Box.cpp:354-379) */
int spmd_rotation_frequency(spmd_sys* s, uint32_t replica, double* w3);
/* Box::calculate_patches_new (Box.cpp:434-507): groups of arms joined by a
 * B-B pair of negative energy; number of groups of more than one arm and
 * their mean size in arms (stars of at most 128 arms). */
int spmd_patches(spmd_sys* s, uint32_t replica, double* n_patch, double* patch_size);
/* One row of statistics*.dat (star-polymers.cpp:307, 470-482) without the
 * step: row[0] Epot per monomer (Box::print_Epot Box.cpp:605-613, from the
 * last propagate that computed it), row[1] Ekin and row[2] 2 Ekin/(3(N-1))
 * (Box::print_Ekin_and_Temperature Box.cpp:635-646), row[3..4] patches,
 * row[5] R_gyr = sqrt(tr G), row[6..14] G, row[15..17] omega. */
int spmd_statistics(spmd_sys* s, uint32_t replica, double* row);

/* ---- solvent observers (Analysis functors applied over MPC::Fluid,
 * MPC.h:110-116); one pass over a replica's solvent on the device ---- */
/* VelocityX::operator() Velocity_x.h:76-96: for every solvent particle,
 * position and velocity relative to the star's centre of mass (Box::COM_Pos
 * and COM_Vel as Box::center_of_mass_reference_frame Box.cpp:85-102 sets
 * them; zero without a star), wrapped by MPC::wrap(pos, vel) MPC.cpp:492-495,
 * binned by floor(x/width) (map 0) and floor(y/width) (map 1).  Bin b of
 * map a has key (b0[a] + b) * width; sums[(a * max_bins + b) * 3 + c] =
 * sum vx, sum vy, count.  nbins[a] = floor(L_a/2/width) - b0[a] + 1 <= 1024. */
int spmd_velocity_profile(spmd_sys* s, uint32_t replica, double width, uint32_t max_bins, int32_t* b0,
                          uint32_t* nbins, double* sums);
/* No counterpart in the reference: spmd_velocity_profile for callers that
 * cannot allocate device memory (code under graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same bins and sums, written as doubles
 * to the caller's device buffer sums_dev (2 * max_bins * 3 doubles, layout
 * as above) by work enqueued on the handle's stream; b0 and nbins are set
 * on return, sums_dev is ready after spmd_sync.  Allocates nothing: the bins
 * and the star's COM live in a static device array of 49,408 bytes per GPU
 * (held until the process ends); widths that need more than 1024 bins per
 * axis are rejected (SPMD_ERR_ARG).  Every call on one GPU shares that array
 * and leaves its contents behind; calls that overlap on one GPU (handles on
 * different streams) race.  Returns SPMD_ERR_STATE if a star was added
 * without a rebuild since (call spmd_md_init first). */
int spmd_velocity_profile_noalloc(spmd_sys* s, uint32_t replica, double width, uint32_t max_bins, int32_t* b0,
                                  uint32_t* nbins, double* sums_dev);
/* Fourier_Autocorr::operator() Fourier_Autocorr.h:55-72: V_j = (1/N) sum_i
 * v_i exp(i 2 pi x_ij / L_j) over the solvent for j = x, y, z, with the
 * component along e_j set to zero.  modes[(j * 3 + c) * 2 + {0,1}] =
 * (Re, Im) of component c of V_j. */
int spmd_fourier_modes(spmd_sys* s, uint32_t replica, double* modes);
/* VelocityDistribution::operator() Vel_distribution.h:35-40: counts of
 * floor(|v|/width) over the solvent, bins 0..nbins-1 (nbins <= 1024);
 * counts[nbins] holds the speeds beyond. */
int spmd_speed_histogram(spmd_sys* s, uint32_t replica, double width, uint32_t nbins, uint64_t* counts);
/* Vel_Autocorr_Fluid (Vel_Autocorr_Fluid.h:19-116): one call of its
 * operator() over replica r's solvent -- every particle's velocity enters
 * its window of `entries` samples (particles followed by id: tracking must
 * be on); once the window is full, every call adds v(oldest) . v(oldest +
 * lag) of every particle to the lag sums (calculate(), :68-78).  The
 * windows live on the device; `entries` (at most 1024) is fixed by the
 * first call until a reset. */
int spmd_fluid_vacf_sample(spmd_sys* s, uint32_t r, uint32_t entries);
/* sums[entries]: the lag sums over particles and calls; *counter: the
 * calls that made them (the reference's Counter) */
int spmd_fluid_vacf_read(spmd_sys* s, uint32_t r, double* sums, uint64_t* counter);
int spmd_fluid_vacf_reset(spmd_sys* s, uint32_t r);

/* Verlet rebuilds so far and entries of the current (half) lists of one
 * replica's star -- diagnostics, no reference counterpart. */
int spmd_md_stats(spmd_sys* s, uint32_t replica, uint32_t* n_rebuild, uint32_t* vl_entries);

/* ---- restart and configuration files ---- */
/* spmd_params of a replica as created */
int spmd_get_params(const spmd_sys* s, uint32_t replica, spmd_params* p);
/* a replica's star: monomers, arms and beads per arm (zeros without a star) */
int spmd_star_layout(const spmd_sys* s, uint32_t replica, uint32_t* n, uint32_t* arms, uint32_t* A, uint32_t* B);
/* MPC::print_fluid_complete (MPC.cpp:570-579): "%f %f %f %f %f %f \n" =
 * position, velocity of every solvent particle. */
int spmd_write_fluid(spmd_sys* s, uint32_t replica, const char* path);
/* MPC::initialize(filename, Box_old) (MPC.cpp:57-113): solvent from such a
 * file written for a box old_box[3]; particles outside the present box,
 * and those the file lacks, are placed at random in the region the box
 * gained with U(-1/2,1/2) velocities (Philox, keyed by particle); all are
 * wrapped and the mean velocity removed.  *n_from_file (optional) counts
 * the particles taken from the file. */
int spmd_read_fluid(spmd_sys* s, uint32_t replica, const char* path, const int32_t* old_box, uint64_t* n_from_file);
/* Box::print_PDB_with_velocity (Box.cpp:563-602): one MODEL of the star,
 * appended to the file when append != 0. */
int spmd_write_pdb(spmd_sys* s, uint32_t replica, const char* path, int64_t step, int append);
/* Box::add_star(filename, A, B, Arms, Mass, set_zero) (Box.cpp:48-54,
 * Molecule::star_from_file Molecule.cpp:178-224): the star topology with
 * positions and velocities from the first MODEL of such a file; set_zero
 * moves its centre of mass and COM velocity to zero. */
int spmd_add_star_pdb(spmd_sys* s, uint32_t replica, const char* path, uint32_t A, uint32_t B, uint32_t arms,
                      double mass, int set_zero);
/* Lossless binary snapshot of a handle (no reference counterpart: the
 * reference restarts from the text files above, star-polymers.cpp:381-405):
 * parameters, solvent (and particle ids when tracking), stars, delrx and the
 * Philox collision counters, plus a caller's step number.  A handle loaded
 * from it continues the run bit for bit. */
int spmd_snapshot_save(spmd_sys* s, const char* path, int64_t step);
int spmd_snapshot_load(const char* path, spmd_sys** out, int64_t* step);

/* ---- tuning (no reference counterpart) ---- */
/* particles per thread in the bin/perm kernels (1,2,4,8) and cells per
 * collide tile (64,128,256); chosen by starpolymersmd_amd.autotune. */
int spmd_set_tuning(spmd_sys* s, int ppt, int tile_cells);
int spmd_get_tuning(const spmd_sys* s, int* ppt, int* tile_cells);
/* Collision path (no reference counterpart; both give the same bits).
 * SPMD_PATH_AUTO: once a collision has put the solvent in cell order, the
 * row-tile path (particles binned by row tile into buckets of indices, then
 * one workgroup per tile gathers, sorts by cell in LDS and collides: no
 * cell-level scan or permutation) whenever every replica's x-row fits a
 * tile (L_x <= 256, L_x * N_c <= 1300); the first collision after the
 * fluid was set, and handles whose rows do not fit, take the sort path.  SPMD_PATH_SORT: always the sort path (bin, scan,
 * permutation, collide).  spmd_get_collision_path tells which path the
 * next collision takes. */
enum { SPMD_PATH_AUTO = 0, SPMD_PATH_SORT = 1, SPMD_PATH_ROWS = 2 };
int spmd_set_collision_path(spmd_sys* s, int path);

int spmd_get_collision_path(const spmd_sys* s, int* next_path);

/* ---- timing (device events on the handle's stream) ---- */
/* average duration in ms of each kernel class over the launches since the
 * last reset: names "md", "bin", "scan", "perm", "collide", "prologue",
 * and for the row-tile path "advect", "gamma" (the cells' Gamma variates
 * drawn ahead) and "redo" ("collide" is then the row-window collide). */
int spmd_profile_enable(spmd_sys* s, int on);
int spmd_profile_read(spmd_sys* s, const char* name, double* avg_ms, int64_t* count);
int spmd_profile_reset(spmd_sys* s);

/* ---- learned implicit-solvent closure (no reference counterpart) ----
 * A per-monomer closure of the MPC solvent: a network predicts, from 24
 * features of a monomer's state, the mean and diagonal variance of the
 * velocity change the solvent imparts at one collision; the star then
 * advances by its own velocity-Verlet MD plus that stochastic kick, with no
 * solvent particles.  Explicit MPC stays the fallback and the standard the
 * closure is checked against. */

/* The network: MLP 24-128-128-128-6, SiLU, BF16 matrix-core kernels with
 * FP32 accumulation, Adam.  Pointers named *_dev are device pointers; every
 * call returns after its device work is complete. */
typedef struct spmd_mlp spmd_mlp;
int spmd_mlp_create(uint64_t seed, spmd_mlp** out);
int spmd_mlp_destroy(spmd_mlp* m);
const char* spmd_mlp_last_error(const spmd_mlp* m);
/* 36,998: W0[128][24] b0[128] W1[128][128] b1 W2[128][128] b2 W3[6][128] b3[6] */
uint32_t spmd_mlp_num_params(void);
int spmd_mlp_get_params(spmd_mlp* m, float* host);
int spmd_mlp_set_params(spmd_mlp* m, const float* host);
/* feature standardisation (x - mu_x) / sd_x (24 each) and target scale sd_y (3) */
int spmd_mlp_set_norm(spmd_mlp* m, const float* mu_x, const float* sd_x, const float* sd_y);
int spmd_mlp_get_norm(const spmd_mlp* m, float* mu_x, float* sd_x, float* sd_y);
/* out_dev[M][6] = (mean[3], log-variance[3]) in units of sd_y for x_dev[M][24]
 * (one fused kernel: weights in LDS, activations in registers) */
int spmd_mlp_forward(spmd_mlp* m, const float* x_dev, float* out_dev, uint64_t M);
/* the same outputs through the training path's per-layer products (the
 * activations kept in the workspace for a backward pass) */
int spmd_mlp_forward_unfused(spmd_mlp* m, const float* x_dev, float* out_dev, uint64_t M);
/* gradient of the mean Gaussian negative log-likelihood of y_dev[M][3]
 * (FP32, parameter order as above) and the loss: one fused kernel (forward,
 * loss, backward and weight-gradient products per 64-row chunk, partial
 * gradients in registers) and one that sums the workgroups' partials */
int spmd_mlp_gradient(spmd_mlp* m, const float* x_dev, const float* y_dev, uint64_t M, float* grad_host,
                      double* loss);
/* the same gradient through the per-layer products (a cross-check of the
 * fused pass) */
int spmd_mlp_gradient_unfused(spmd_mlp* m, const float* x_dev, const float* y_dev, uint64_t M, float* grad_host,
                              double* loss);
/* one Adam step (beta 0.9/0.999, eps 1e-8) on the mean loss of the batch
 * (the fused pass; the sum kernel applies the step and updates the BF16
 * weight copies in place) */
int spmd_mlp_train_step(spmd_mlp* m, const float* x_dev, const float* y_dev, uint64_t M, double lr, double* loss);
/* `steps` Adam steps on minibatches of `batch` rows drawn with replacement
 * (Philox keyed by seed, inside the fused pass) from x_dev[N][24],
 * y_dev[N][3]; loss_hist[steps] optional (kept on the device, copied once
 * at the end) */
int spmd_mlp_fit(spmd_mlp* m, const float* x_dev, const float* y_dev, uint64_t N, uint32_t batch, uint32_t steps,
                 double lr, uint64_t seed, double* loss_hist);
/* plain device buffers for callers without their own device allocator */
int spmd_dev_alloc(uint64_t bytes, void** p);
int spmd_dev_free(void* p);
int spmd_dev_upload(void* dst, const void* src, uint64_t bytes);
int spmd_dev_download(void* dst, const void* src, uint64_t bytes);

/* The 24 closure features of every monomer of every replica, global
 * monomer order (replica r from its first monomer; see spmd_num_monomers),
 * of the current state (pending MD substeps are run first).  feat_dev is
 * 16-byte aligned (rows are written as 16-byte stores). */
int spmd_closure_features(spmd_sys* s, float* feat_dev);
/* n_intervals closure intervals for all replicas: step_update velocity-
 * Verlet substeps, check_bounds, then the network's kick on every monomer.
 * With fmin/fmax (host arrays of 24, the trained feature range), monomers
 * whose features leave it (by more than 25% of the range) are counted per
 * replica into oor_host[n_replicas] (optional). */
int spmd_closure_run(spmd_sys* s, spmd_mlp* m, int64_t n_intervals, const float* fmin, const float* fmax,
                     uint32_t* oor_host);

#ifdef __cplusplus
}
#endif
#endif
