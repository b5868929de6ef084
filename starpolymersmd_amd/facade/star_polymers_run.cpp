/*
 * The production loop of star-polymers/src/star-polymers.cpp:351-405,
 * 412-536 (star in an MPC solvent under shear) written against the C++
 * facade, without the reference's file output: prints the temperature,
 * Epot and the star's centre of mass every `out` steps.
 *
 *   star_polymers_run L A B Arms Shear Steps Out [dt] [N_c] [coil]
 */
#include <cstdio>
#include <cstdlib>

#include "spmd_facade.h"

int main(int argc, char** argv) {
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s L A B Arms Shear Steps Out [dt] [N_c] [coil]\n", argv[0]);
        return 2;
    }
    const int L = std::atoi(argv[1]);
    const unsigned A = std::atoi(argv[2]), B = std::atoi(argv[3]), Arms = std::atoi(argv[4]);
    const double Shear = std::atof(argv[5]);
    const long Steps = std::atol(argv[6]), Out = std::atol(argv[7]);
    const double dt = argc > 8 ? std::atof(argv[8]) : 0.01;
    const int Nc = argc > 9 ? std::atoi(argv[9]) : 10;
    const bool coil = argc > 10 && std::atoi(argv[10]) != 0;
    const double Temperature = 1.0, Lambda = 1.0, Mass_mono = 10.0;
    try {
        spmd::Box box(L, L, L, Temperature, Lambda);
        box.DeltaT = dt;
        spmd::MPC hydrodynamics{box, Temperature, Nc, Shear, 10, false};
        if (coil) box.add_star_coil(A, B, Arms, Mass_mono);
        else box.add_star(A, B, Arms, Mass_mono);
        spmd::Thermostat_None thermostat{box, dt};
        hydrodynamics.initialize();
        for (long n = 0; n <= Steps; ++n) {
            thermostat.propagate(n % Out == 0);
            hydrodynamics.step(n, 10 * dt);
            if (n % Out == 0 && n > 0) {
                box.sync_molecules();
                double com[3] = {0, 0, 0};
                for (auto& m : box.Molecules[0].Monomers)
                    for (int c = 0; c < 3; ++c) com[c] += m.Position[c];
                const double inv = 1.0 / box.Molecules[0].Monomers.size();
                std::printf("%ld T=%.6f Epot=%.6f com=(%.4f,%.4f,%.4f)\n", n,
                            hydrodynamics.calculateCurrentTemperature(), box.epot(), com[0] * inv, com[1] * inv,
                            com[2] * inv);
            }
        }
        hydrodynamics.sync();
    } catch (const spmd::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
