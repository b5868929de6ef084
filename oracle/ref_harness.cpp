/*
 * ref_harness.cpp -- exports the reference's own potentials and RNG, compiled
 * from the reference sources where they lie, as a checker for the oracle.
 * TEST INFRASTRUCTURE ONLY (built into oracle/_ref/, never shipped).
 *
 * Compiled together with star-polymers/lib/Rand.cpp and including
 * star-polymers/lib/Molecules/Potentials.h unmodified.  Those two are the
 * only parts of the hot path that compile without Eigen 3; MPC.cpp, Box.cpp,
 * Molecule.cpp and Thermostat_None.cpp include <../eigen/Eigen/Dense>, which
 * the reference does not vendor and this image does not have.
 */
#include "Potentials.h"
#include "Rand.h"

extern "C" {

double ref_fene_force(double r2) { return Fene_Force(r2); }
double ref_fene_anchor_force(double r2) { return Fene_Anchor_Force(r2); }
double ref_aa_force(double r2) { return TypeAA_Force(r2); }
double ref_bb_force(double r2, double lambda) { return TypeBB_Force(r2, lambda); }
double ref_fene_potential(double r2) { return Fene_Potential(r2); }
double ref_fene_anchor_potential(double r2) { return Fene_Anchor_Potential(r2); }
double ref_aa_potential(double r2) { return TypeAA_Potential(r2); }
double ref_bb_potential(double r2, double lambda) { return TypeBB_Potential(r2, lambda); }

/* Rand's generator is a process-wide static seeded 1 (Rand.cpp:6): this
 * sequence is only meaningful on the first call in a process. */
void ref_rng_sequence(const int* ops, const double* a, const double* b, int n, double* out) {
    for (int i = 0; i < n; ++i) out[i] = ops[i] == 0 ? Rand::real_uniform() : Rand::real_gamma(a[i], b[i]);
}

}
