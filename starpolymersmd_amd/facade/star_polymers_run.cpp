/*
 * The production loop of star-polymers/src/star-polymers.cpp:351-405,
 * 412-536 (star in an MPC solvent under shear) written against the C++
 * facade.  Prints the temperature, Epot and the star's centre of mass every
 * `Out` steps; with an output directory it also writes the reference's
 * files there: statistics.dat (one row per output step,
 * star-polymers.cpp:300-307, 466-482), fluid_profile.dat (VelocityX,
 * :486, 530-534), end_config.pdb and end_fluid.dat (the restart files of
 * :446-458).
 *
 *   star_polymers_run L A B Arms Shear Steps Out [dt] [N_c] [coil] [outdir]
 */
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "spmd_facade.h"

int main(int argc, char** argv) {
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s L A B Arms Shear Steps Out [dt] [N_c] [coil] [outdir]\n", argv[0]);
        return 2;
    }
    const int L = std::atoi(argv[1]);
    const unsigned A = std::atoi(argv[2]), B = std::atoi(argv[3]), Arms = std::atoi(argv[4]);
    const double Shear = std::atof(argv[5]);
    const long Steps = std::atol(argv[6]), Out = std::atol(argv[7]);
    const double dt = argc > 8 ? std::atof(argv[8]) : 0.01;
    const int Nc = argc > 9 ? std::atoi(argv[9]) : 10;
    const bool coil = argc > 10 && std::atoi(argv[10]) != 0;
    const std::string outdir = argc > 11 ? argv[11] : "";
    const double Temperature = 1.0, Lambda = 1.0, Mass_mono = 10.0;
    try {
        spmd::Box box(L, L, L, Temperature, Lambda);
        box.DeltaT = dt;
        spmd::MPC hydrodynamics{box, Temperature, Nc, Shear, 10, false};
        if (coil) box.add_star_coil(A, B, Arms, Mass_mono);
        else box.add_star(A, B, Arms, Mass_mono);
        spmd::Thermostat_None thermostat{box, dt};
        hydrodynamics.initialize();
        spmd::VelocityX velocity_average_x{box, hydrodynamics, 0.5};
        std::ofstream statistic_file;
        if (!outdir.empty()) {
            statistic_file.open(outdir + "/statistics.dat", std::ios::out | std::ios::trunc);
            statistic_file << spmd::statistics_header();
        }
        for (long n = 0; n <= Steps; ++n) {
            if (n > 0 && !(n % Out)) {
                thermostat.propagate(true);
                if (statistic_file.is_open()) {
                    /* star-polymers.cpp:470-482 */
                    box.center_of_mass_reference_frame();
                    std::vector<double> patch_char = box.calculate_patches_new();
                    auto gyration = box.calculate_gyration_tensor();
                    statistic_file << n << " ";
                    box.print_Epot(statistic_file);
                    box.print_Ekin_and_Temperature(statistic_file);
                    statistic_file << patch_char[0] << " " << patch_char[1] << " ";
                    statistic_file << std::get<0>(gyration) << " ";
                    for (double g : std::get<1>(gyration)) statistic_file << g << " ";
                    for (double w : box.calculate_rotation_frequency()) statistic_file << w << " ";
                    statistic_file << "\n";
                    statistic_file.flush();
                    hydrodynamics(velocity_average_x);
                }
                box.sync_molecules();
                double com[3] = {0, 0, 0};
                for (auto& m : box.Molecules[0].Monomers)
                    for (int c = 0; c < 3; ++c) com[c] += m.Position[c];
                const double inv = 1.0 / box.Molecules[0].Monomers.size();
                std::printf("%ld T=%.6f Epot=%.6f com=(%.4f,%.4f,%.4f)\n", n,
                            hydrodynamics.calculateCurrentTemperature(), box.epot(), com[0] * inv, com[1] * inv,
                            com[2] * inv);
            } else {
                thermostat.propagate(false);
            }
            hydrodynamics.step(n, 10 * dt);
        }
        hydrodynamics.sync();
        if (!outdir.empty()) {
            std::ofstream profile(outdir + "/fluid_profile.dat", std::ios::out | std::ios::trunc);
            velocity_average_x.print_result(profile);
            box.print_PDB_with_velocity(outdir + "/end_config.pdb", (int)Steps, false);
            hydrodynamics.print_fluid_complete(outdir + "/end_fluid.dat");
        }
    } catch (const spmd::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
