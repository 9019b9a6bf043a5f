// Harness for the double-shear-layer example of the integration guide.  The
// guide's code block is split in two: its #include / #define lines go to
// example_prelude.inc, the rest (the time loop) to example_body.inc, which is
// pasted into main() below after the setup of src/main_double-shear-layer.cpp
// (lines 70-154).  The setup gives the block the names it uses there:
// lattice.inif, parameters.tau / tau_nh, col, L, n_steps, interval, U_0,
// U_LIMITE, n_fluid, file_energy (precision 10) and vel (3 L^2 doubles).
//
//   integration_dsl_example [--L 256] [--t-final 1.5] [--dt-out 0.0625]
//
// Writes k_energy_lbm.dat and vel_<t*>.vtk in the working directory.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

using namespace std;

#include "example_prelude.inc"

namespace {

const double U_0 = 1. / 32.;
const double U_LIMITE = 0.4;

struct Lattice {
    double *inif;
};

struct Parameters {
    double tau;
    double tau_nh;
};

double option(int argc, char **argv, const char *name, double dflt)
{
    for (int i = 1; i + 1 < argc; ++i)
        if (std::strcmp(argv[i], name) == 0) return std::atof(argv[i + 1]);
    return dflt;
}

}  // namespace

int main(int argc, char **argv)
{
    const int Li = (int)option(argc, argv, "--L", 256);
    const double T_FINAL = option(argc, argv, "--t-final", 1.5);
    const double DT_OUT = option(argc, argv, "--dt-out", 1. / 16.);

    const double L = Li;
    const int col = ERLBM_COL_EXT_REG;
    const Parameters parameters{0.5 + 3. * U_0 * L / 80000., 0.65};
    const int n_fluid = Li * Li;
    const int steps_per_t = (int)round(L / U_0);
    const int n_steps = (int)round(T_FINAL * steps_per_t) + 1;
    const int interval = max(1, (int)round(DT_OUT * steps_per_t));

    vector<double> inif((size_t)n_fluid * 9);
    if (erlbm_init_double_shear(Li, 1.0, inif.data()) != ERLBM_OK) {
        cerr << "erlbm_init_double_shear: " << erlbm_last_error() << endl;
        return 1;
    }
    Lattice lattice{inif.data()};
    ofstream file_energy("k_energy_lbm.dat");
    file_energy << setprecision(10);
    vector<double> vel((size_t)3 * n_fluid);

#include "example_body.inc"

    return 0;
}
