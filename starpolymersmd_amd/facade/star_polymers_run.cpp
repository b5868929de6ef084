/*
 * The production loop of star-polymers/src/star-polymers.cpp:351-405,
 * 412-536 (star in an MPC solvent under shear) written against the C++
 * facade.  Prints the temperature, Epot and the star's centre of mass every
 * `Out` steps; with an output directory it also writes the reference's
 * files there: statistics (one row per output step, star-polymers.cpp:
 * 300-307, 466-482), fluid_profile (VelocityX, :486, 530-534), output
 * (the step and Box::print_center_of_mass, :488-491), fluid
 * (MPC::print_fluid of the middle z plane, :520-523), end_config (.pdb)
 * and end_fluid (the restart files of :446-458).  As in the reference the
 * loop holds the integrator as a Thermostat* (:101, 382) and steps the
 * solvent through its Hydrodynamics interface.
 *
 *   star_polymers_run L A B Arms Shear Steps Out [dt] [N_c] [coil] [outdir] [tagged]
 *   star_polymers_run <dir>/end_config<tag>.pdb Steps Out [dt] [N_c] [outdir]
 *
 * Without `tagged` the files are statistics.dat, fluid_profile.dat,
 * end_config.pdb and end_fluid.dat.  With it (tagged = 1) every name carries
 * the reference's parameter string (spmd::run_tag, star-polymers.cpp:
 * 270-288), and SIGTERM makes the loop stop at the next step, write the
 * restart files with the step reached in place of the run length, rename
 * the statistics and profile files the same way and exit with the signal's
 * number (:413-463).  The second form continues such a run
 * (star-polymers.cpp:118-133, 199-225, 361-404): the parameters come from
 * the file name, the star from the PDB (velocities included; run 0 in the
 * name zeroes them), the solvent from the end_fluid file of the same name
 * when it exists (otherwise a fresh one), and the steps run from the one
 * after the name's up to Steps.  The reference's SIGSEGV dump
 * (core_dump.pdb) has no counterpart: the library reports those states as
 * status codes, which end the run with an error message.
 */
#include <cctype>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "spmd_facade.h"

namespace {

volatile std::sig_atomic_t g_signal = 0;

extern "C" void on_signal(int signum) { g_signal = signum; }

bool file_exists(const std::string& p) {
    std::ifstream f(p);
    return f.good();
}

bool is_number(const char* s) { return std::strspn(s, ".eE-0123456789") == std::strlen(s); }

}  // namespace

int main(int argc, char** argv) {
    const bool cont = argc > 1 && !is_number(argv[1]);
    if ((cont && argc < 4) || (!cont && argc < 8)) {
        std::fprintf(stderr,
                     "usage: %s L A B Arms Shear Steps Out [dt] [N_c] [coil] [outdir] [tagged]\n"
                     "       %s <dir>/end_config<tag>.pdb Steps Out [dt] [N_c] [outdir]\n",
                     argv[0], argv[0]);
        return 2;
    }
    spmd::RunParams rp;
    long Steps = 0, Out = 0, n0 = 0;
    double dt = 0.01;
    int Nc = 10;
    bool coil = false, tagged = false;
    std::string outdir, restart;
    try {
        if (cont) {
            restart = argv[1];
            if (!file_exists(restart)) {
                std::fprintf(stderr, "input file does not exist: %s\n", restart.c_str());
                return 1;
            }
            rp = spmd::parse_restart_name(restart);
            Steps = std::atol(argv[2]);
            Out = std::atol(argv[3]);
            dt = argc > 4 ? std::atof(argv[4]) : 0.01;
            Nc = argc > 5 ? std::atoi(argv[5]) : 10;
            outdir = argc > 6 ? argv[6] : "";
            tagged = true;
            n0 = rp.Steps + 1;
        } else {
            const int L = std::atoi(argv[1]);
            rp.A = std::atoi(argv[2]);
            rp.B = std::atoi(argv[3]);
            rp.Arms = std::atoi(argv[4]);
            rp.Shear = std::atof(argv[5]);
            Steps = std::atol(argv[6]);
            Out = std::atol(argv[7]);
            dt = argc > 8 ? std::atof(argv[8]) : 0.01;
            Nc = argc > 9 ? std::atoi(argv[9]) : 10;
            coil = argc > 10 && std::atoi(argv[10]) != 0;
            outdir = argc > 11 ? argv[11] : "";
            tagged = argc > 12 && std::atoi(argv[12]) != 0;
            rp.Lx = rp.Ly = rp.Lz = L;
            rp.Mass = 10;
            rp.Lambda = 1.0;
            rp.T = 1.0;
            rp.MPC_on = true;
        }
    } catch (const spmd::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    if (Out <= 0) {
        std::fprintf(stderr, "Out must be positive\n");
        return 2;
    }
    rp.Steps = Steps;
    const std::string tag = tagged ? spmd::run_tag(rp) : "";
    auto name = [&](const std::string& type, const std::string& t, const char* ext) {
        return outdir + "/" + type + t + ext;
    };
    std::signal(SIGTERM, on_signal);
    try {
        spmd::Box box(rp.Lx, rp.Ly, rp.Lz, rp.T, rp.Lambda);
        box.DeltaT = dt;
        spmd::MPC hydrodynamics{box, rp.T, Nc, rp.Shear, 10, false};
        if (cont) box.add_star(restart, rp.A, rp.B, rp.Arms, (double)rp.Mass, n0 == 1);
        else if (coil) box.add_star_coil(rp.A, rp.B, rp.Arms, (double)rp.Mass);
        else box.add_star(rp.A, rp.B, rp.Arms, (double)rp.Mass);
        std::cout << '\n';
        box.print_center_of_mass(std::cout);  /* star-polymers.cpp:371-372 */
        spmd::Thermostat* thermostat = new spmd::Thermostat_None{box, dt};
        spmd::Hydrodynamics& hydro = hydrodynamics;
        std::string fluid_in;
        if (cont) {
            fluid_in = restart;
            fluid_in.replace(fluid_in.find("end_config"), 10, "end_fluid");
            const std::string::size_type k = fluid_in.rfind(".pdb");
            if (k != std::string::npos) fluid_in.replace(k, 4, ".dat");
        }
        if (!fluid_in.empty() && file_exists(fluid_in))
            hydrodynamics.initialize(fluid_in, std::array<int, 3>{rp.Lx, rp.Ly, rp.Lz});
        else
            hydrodynamics.initialize();
        spmd::VelocityX velocity_average_x{box, hydrodynamics, 0.5};
        std::ofstream statistic_file, output_file;
        FILE* fluid_file = nullptr;
        if (!outdir.empty()) {
            statistic_file.open(name("statistics", tag, ".dat"), std::ios::out | std::ios::trunc);
            statistic_file << spmd::statistics_header();
            output_file.open(name("output", tag, ".dat"), std::ios::out | std::ios::trunc);
            fluid_file = std::fopen(name("fluid", tag, ".dat").c_str(), "w");
        }
        for (long n = n0; n <= Steps; ++n) {
            if (g_signal) {
                /* star-polymers.cpp:413-463: restart files named with the
                 * step reached, the open files renamed the same way */
                const int sig = g_signal;
                spmd::RunParams now = rp;
                now.Steps = n;
                const std::string tag_n = spmd::run_tag(now);
                if (!outdir.empty()) {
                    statistic_file.close();
                    std::rename(name("statistics", tag, ".dat").c_str(), name("statistics", tag_n, ".dat").c_str());
                    std::ofstream profile(name("fluid_profile", tag_n, ".dat"), std::ios::out | std::ios::trunc);
                    velocity_average_x.print_result(profile);
                    box.print_PDB_with_velocity(name("end_config", tag_n, ".pdb"), (int)n, false);
                    hydrodynamics.print_fluid_complete(name("end_fluid", tag_n, ".dat"));
                }
                std::printf("signal %d caught, data saved\n", sig);
                std::fflush(stdout);
                return sig;
            }
            if (n > 0 && !(n % Out)) {
                thermostat->propagate(true);
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
                    output_file << n << " ";
                    box.print_center_of_mass(output_file);
                    output_file << '\n';
                    output_file.flush();
                    hydro.print_fluid(fluid_file, (int)n, rp.Lz / 2, rp.Lz / 2);
                    std::fflush(fluid_file);
                }
                const spmd::Mirror<spmd::Molecule>& mols = box.Molecules;  /* read only: nothing to write back */
                const spmd::Vector3d com = mols[0].calculate_center_of_mass();
                std::printf("%ld T=%.6f Epot=%.6f com=(%.4f,%.4f,%.4f)\n", n, hydro.calculateCurrentTemperature(),
                            box.epot(), com[0], com[1], com[2]);
                std::fflush(stdout);
            } else {
                thermostat->propagate(false);
            }
            hydro.step(n, 10 * dt);
        }
        hydrodynamics.sync();
        if (fluid_file) std::fclose(fluid_file);
        delete static_cast<spmd::Thermostat_None*>(thermostat);
        if (!outdir.empty()) {
            std::ofstream profile(name("fluid_profile", tag, ".dat"), std::ios::out | std::ios::trunc);
            velocity_average_x.print_result(profile);
            /* the restart files through FILE*, as star-polymers.cpp:449-458 */
            FILE* end_config_file = std::fopen(name("end_config", tag, ".pdb").c_str(), "w");
            box.print_PDB_with_velocity(end_config_file, (int)Steps);
            std::fclose(end_config_file);
            FILE* end_fluid_file = std::fopen(name("end_fluid", tag, ".dat").c_str(), "w");
            hydrodynamics.print_fluid_complete(end_fluid_file);
            std::fclose(end_fluid_file);
        }
    } catch (const spmd::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
