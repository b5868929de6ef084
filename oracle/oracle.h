/*
 * oracle.h -- C API of the FP64 CPU oracle (test infrastructure only).
 *
 * The oracle is a plain, single-threaded FP64 restatement of the reference's
 * MPC-MD step (star-polymers/lib/Hydrodynamics/MPC.cpp, lib/Box.cpp,
 * lib/Thermostats/Thermostat_None.cpp, lib/Molecules/{Molecule.cpp,
 * Potentials.h}).  It is the checker the parity tests, smoke() and the CPU
 * leg of bench.py compare the HIP library against.  Nothing in the product
 * links or calls it.
 *
 * Randomness has three sources (orc_params.rng_mode):
 *   ORC_RNG_REFERENCE  std::mt19937_64 seeded 1 with libstdc++'s
 *                      uniform_real_distribution / gamma_distribution, drawn
 *                      in the reference's call order (lib/Rand.cpp:6-43).
 *   ORC_RNG_PHILOX     the counter-based Philox4x32-10 streams the HIP
 *                      library uses (same counters, same Gamma sampler).
 *   ORC_RNG_INJECTED   grid shift, per-cell rotation-axis uniforms and
 *                      per-cell Gamma variates supplied by the caller.
 */
#ifndef SPMD_ORACLE_H
#define SPMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_RNG_REFERENCE = 0, ORC_RNG_PHILOX = 1, ORC_RNG_INJECTED = 2 };

typedef struct orc_params {
    int32_t  L[3];          /* box, rounded up to even like Box.cpp:14-16 */
    double   kT;            /* MPC / star temperature */
    int32_t  n_c;           /* solvent particles per cell (MPC.cpp:17) */
    double   shear;         /* Lees-Edwards shear rate */
    uint32_t step_update;   /* collision every step_update calls (MPC.cpp:119) */
    double   alpha_deg;     /* SRD rotation angle (MPC.cpp:23-24 uses 130) */
    double   lambda;        /* B-B attraction strength (Box ctor) */
    double   md_dt;         /* velocity-Verlet step (Thermostat_None) */
    int32_t  rng_mode;
    uint64_t seed;          /* Philox key (ORC_RNG_PHILOX) */
    uint32_t replica;       /* replica id entering the Philox counter */
    int32_t  amc;           /* angular-momentum-conserving collision (MPC ctor angular_mom) */
} orc_params;

typedef struct orc_sys orc_sys;

orc_sys* orc_create(const orc_params* p);
void     orc_destroy(orc_sys* s);

/* MPC::initialize() (MPC.cpp:34-55). */
void     orc_fluid_initialize(orc_sys* s);
uint64_t orc_num_fluid(const orc_sys* s);
/* OpenMP threads of the solvent loops in counter-based random modes (the
 * results do not depend on the count) */
void orc_set_threads(int n);
int orc_get_threads(void);
void     orc_set_fluid(orc_sys* s, const double* pos, const double* vel, uint64_t n);
void     orc_get_fluid(const orc_sys* s, double* pos, double* vel);

/* Box::add_star -> Molecule::initialize_open_star (Box.cpp:39-46,
 * Molecule.cpp:126-176); builds bonds and draws velocities. */
int      orc_add_star(orc_sys* s, uint32_t A, uint32_t B, uint32_t arms,
                      double mass, double bond, double anchor_bond);
uint32_t orc_num_monomers(const orc_sys* s);
void     orc_set_monomers(orc_sys* s, const double* pos, const double* vel);
void     orc_get_monomers(const orc_sys* s, double* pos, double* vel, double* force);
void     orc_get_monomer_props(const orc_sys* s, double* mass, uint8_t* type, uint8_t* anchor);
/* bonds as directed (owner, other) pairs in the reference's Neighbors order */
uint32_t orc_get_bonds(const orc_sys* s, uint32_t* pairs, uint32_t cap);

/* Thermostat_None ctor work: update_VerletLists + calculate_forces_verlet. */
int      orc_md_init(orc_sys* s);
/* Thermostat_None::propagate(calc_epot) (Thermostat_None.cpp:21-38). */
int      orc_md_propagate(orc_sys* s, int calc_epot);
double   orc_epot(const orc_sys* s);

/* MPC::step(t, dt) (MPC.cpp:116-142).  Returns 0, or -1 when sort() found
 * a cell index out of range (the reference raises SIGSEGV there). */
int      orc_mpc_step(orc_sys* s, int64_t t, double dt);

/* Injected randoms for the next collision (ORC_RNG_INJECTED): shift[3] in
 * [-0.5,0.5), axis_u[2*ncells] = (u_phi, u_theta) per cell, gamma[ncells]. */
void     orc_inject(orc_sys* s, const double* shift, const double* axis_u, const double* gamma);

double   orc_get_delrx(const orc_sys* s);
void     orc_set_delrx(orc_sys* s, double d);
void     orc_last_shift(const orc_sys* s, double* shift);
int      orc_last_recentered(const orc_sys* s);
/* CellIndex of every fluid particle and monomer from the last sort(). */
void     orc_get_cell_index(const orc_sys* s, uint32_t* fluid, uint32_t* mono);

/* MPC::calculateCurrentTemperature (MPC.cpp:384-418). */
double   orc_temperature(orc_sys* s);

/* Box observables (Box.cpp:330-379) on the COM frame of molecule 0. */
void     orc_com_frame(orc_sys* s);
double   orc_gyration_tensor(orc_sys* s, double* G9);
void     orc_rotation_frequency(orc_sys* s, double* omega3);
double   orc_ekin_monomers(orc_sys* s);
/* solvent observers and patches (lib/Analysis/Velocity_x.h:76-96,
 * Fourier_Autocorr.h:55-72, Vel_distribution.h:35-40; Box.cpp:434-507) */
void     orc_velocity_profile(orc_sys* s, double width, uint32_t cap, uint32_t* n_out, double* keys, double* sums);
void     orc_fourier_modes(orc_sys* s, double* modes);
void     orc_speed_histogram(orc_sys* s, double width, uint32_t nbins, uint64_t* counts);
void     orc_patches(orc_sys* s, double* n_patch, double* patch_size);

/* Potentials (Potentials.h:14-120), exposed for known-answer tests. */
double   orc_fene_force(double r2);
double   orc_fene_anchor_force(double r2);
double   orc_aa_force(double r2);
double   orc_bb_force(double r2, double lambda);
double   orc_fene_potential(double r2);
double   orc_fene_anchor_potential(double r2);
double   orc_aa_potential(double r2);
double   orc_bb_potential(double r2, double lambda);

/* RNG building blocks, exposed for known-answer and statistical tests. */
void     orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double   orc_philox_gamma(uint64_t seed, uint32_t replica, uint32_t cell, uint32_t k,
                          double shape, double scale);
/* n draws of Rand::real_uniform / Rand::real_gamma(a,b) in reference mode
 * from a fresh mt19937_64(1), alternating as listed in `ops` (0=uniform,
 * 1=gamma). */
void     orc_reference_rng_sequence(const int* ops, const double* a, const double* b,
                                    int n, double* out);

#ifdef __cplusplus
}
#endif
#endif
