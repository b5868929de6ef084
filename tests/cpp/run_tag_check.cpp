/* Prints spmd::run_tag for parameter sets read from stdin, then the fields
 * spmd::parse_restart_name reads back from "./results/end_config<tag>.pdb";
 * last, find_parameter of an absent name and the error of a name without
 * end_config. */
#include <iostream>

#include "../../starpolymersmd_amd/facade/spmd_facade.h"

int main() {
    spmd::RunParams p;
    int chain = 0, mpc = 0;
    while (std::cin >> chain >> p.A >> p.B >> p.Arms >> p.Mass >> p.Lx >> p.Ly >> p.Lz >> p.Lambda >> p.T >> p.Steps >>
           mpc >> p.Shear) {
        p.chain = chain != 0;
        p.MPC_on = mpc != 0;
        const std::string tag = spmd::run_tag(p);
        const spmd::RunParams q = spmd::parse_restart_name("./results/end_config" + tag + ".pdb");
        std::cout << tag << " " << q.chain << " " << q.A << " " << q.B << " " << q.Arms << " " << q.Mass << " " << q.Lx
                  << " " << q.Ly << " " << q.Lz << " " << q.Lambda << " " << q.T << " " << q.Steps << " " << q.MPC_on
                  << " " << q.Shear << "\n";
    }
    std::cout << spmd::find_parameter("_Star_A3_B3", "Shear") << "\n";
    try {
        spmd::parse_restart_name("statistics_Star_A3.dat");
        std::cout << "no error\n";
    } catch (const spmd::Error& e) {
        std::cout << "error " << e.code << "\n";
    }
    return 0;
}
