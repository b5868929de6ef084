/*
 * The solvent test driver star-polymers/src/mpc_test.cpp:42-146 written
 * against the C++ facade: the same numeric arguments and defaults
 * (N, Temperature, BoxX, BoxY, BoxZ, Shear, StepSize, Steps_Equil,
 * Steps_Total, Steps_Output, AMC), the same observers and sampling steps,
 * and the same five files in the working directory: temperature*, fluid*
 * (VelocityX), maxwell* (VelocityDistribution), velocity*
 * (Vel_Autocorr_Fluid) and fourier* (Fourier_Autocorr), named with the
 * reference's parameter suffix.
 *
 *   mpc_test_run [N] [T] [BoxX] [BoxY] [BoxZ] [Shear] [StepSize] [Equil] [Total] [Output] [AMC]
 */
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "spmd_facade.h"

/* mpc_test.cpp:27-30 */
static bool is_number(const std::string& str) { return str.find_first_not_of(".eE-0123456789") == std::string::npos; }

int main(int argc, char* argv[]) {
    /* defaults and parsing of mpc_test.cpp:49-60 */
    double a_para[]{10, 0.5, 5, 5, 5, 0.5, 0.1, 1E3, 1E4, 100, 0};
    const int a_para_size = sizeof(a_para) / sizeof(*a_para);
    for (int i = 1; i < std::min(a_para_size + 1, argc); ++i) {
        if (is_number(argv[i])) a_para[i - 1] = std::stod(argv[i]);
        else break;
    }
    const int N = (int)a_para[0];
    const double Temperature = a_para[1];
    const int BoxX = (int)a_para[2], BoxY = (int)a_para[3], BoxZ = (int)a_para[4];
    const double Shear = a_para[5], StepSize = a_para[6];
    const int Steps_Equil = (int)a_para[7], Steps_Total = (int)a_para[8], Steps_Output = (int)a_para[9];
    const bool AngularMomentumConservation = (bool)a_para[10];
    try {
        spmd::Box box(BoxX, BoxY, BoxZ, Temperature, 1.0);
        spmd::MPC MPCroutine{box, Temperature, N, Shear, 1, AngularMomentumConservation};
        std::stringstream ss_para;  /* mpc_test.cpp:79-82 */
        ss_para << "_N" << BoxX * BoxY * BoxZ * N;
        ss_para << "_Shear" << Shear;
        ss_para << "_h" << StepSize;
        if (AngularMomentumConservation) ss_para << "_AMC";
        std::ofstream fluid_file("fluid" + ss_para.str() + ".dat", std::ios::out | std::ios::trunc);
        std::ofstream temperature_file("temperature" + ss_para.str() + ".dat", std::ios::out | std::ios::trunc);
        std::ofstream autocorr_file("velocity" + ss_para.str() + ".dat", std::ios::out | std::ios::trunc);
        std::ofstream fourier_file("fourier" + ss_para.str() + ".dat", std::ios::out | std::ios::trunc);
        std::ofstream maxwell_file("maxwell" + ss_para.str() + ".dat", std::ios::out | std::ios::trunc);
        std::cout << "Temperature: " << Temperature << std::endl;
        std::cout << "MPC is turned ON with shear rate: " << Shear << std::endl;
        std::cout << "Box Size: " << BoxX << " " << BoxY << " " << BoxZ << std::endl;
        std::cout << "Step Size: " << StepSize << std::endl;
        std::cout << "Total Steps: " << Steps_Total << " Equilibration: " << Steps_Equil
                  << " Output every: " << Steps_Output << std::endl;
        MPCroutine.initialize();
        spmd::VelocityX velocity_average{box, MPCroutine, 0.2};
        spmd::Vel_Autocorr_Fluid autocorr{MPCroutine.NumberOfParticles(), 50, 1.0};
        spmd::VelocityDistribution maxwell{0.01};
        spmd::Fourier_Autocorr fourier{MPCroutine.NumberOfParticles(), 50, 1.0, (double)BoxX};
        const std::clock_t begin = std::clock();
        for (int n = 0; n < Steps_Total; ++n) {  /* mpc_test.cpp:113-134 */
            MPCroutine.step(n, StepSize);
            if (n > Steps_Equil && !(n % Steps_Output)) {
                const double T = MPCroutine.calculateCurrentTemperature();
                std::cout << n << " " << T << std::endl;
                MPCroutine(velocity_average);
                MPCroutine(maxwell);
                temperature_file << n << " " << T << std::endl;
            }
            if (n > Steps_Equil && !(n % 10)) {
                MPCroutine(autocorr);
                MPCroutine(fourier);
            }
        }
        velocity_average.print_result(fluid_file);
        maxwell.print_result(maxwell_file);
        autocorr.print_result(autocorr_file);
        fourier.print_result(fourier_file);
        const std::clock_t end = std::clock();
        std::cout << "time: " << double(end - begin) / CLOCKS_PER_SEC << std::endl;
    } catch (const spmd::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
