// taylor_green -- the Taylor-Green vortex driver (src/main_taylor-green.cpp)
// on the GPU: same options, same printed result line
//   L tau tau_nh U0 steps E_vector E_magnitude (nu_eff - nu)/nu
//
//   taylor_green [--L 64] [--collision erlbm] [--tau-nh 0.65] [--Re 50]
//                [--u0 2/L] [--tau T] [--init neq|eq] [--density poisson|uniform]
//                [--precision fp64]
//
// Without --tau: nu = U0 L / Re, run to t = T = L / U0; with --tau: run to
// 2 nu k^2 t = 1 (the kappa4 check).  The error is measured on the
// post-collision populations of step nT, as the reference measures it.
#include <cmath>
#include <cstdio>
#include <vector>

#include "args.hpp"

int main(int argc, char **argv)
{
    drv::Args args{argc, argv};
    const int Li = args.integer("--L", 64);
    const int col = drv::collision(args.str("--collision", "erlbm"));
    const double tau_nh_arg = args.real("--tau-nh", 0.65);
    const double Re = args.real("--Re", 50.);
    const double u_0 = args.real("--u0", 2. / Li);
    const bool tau_mode = args.has("--tau");
    const bool neq = args.str("--init", "neq") == "neq";
    const bool poisson = args.str("--density", "poisson") == "poisson";
    const int precision = drv::precision(args.str("--precision", "fp64"));

    double tau, visc;
    if (tau_mode) {
        tau = args.real("--tau", 0.6);
        visc = (tau - 0.5) / 3.;
    } else {
        visc = u_0 * Li / Re;
        tau = 0.5 + 3. * visc;
    }
    const double tau_nh = (col == ERLBM_COL_REG) ? 1.0 : (col == ERLBM_COL_BGK ? tau : tau_nh_arg);

    const double L = Li;
    const double k = 2. * M_PI / L;
    const int nT = tau_mode ? drv::iround(1. / (2. * visc * k * k)) : drv::iround(L / u_0);

    std::vector<double> f((size_t)Li * Li * 9);
    drv::check(erlbm_init_taylor_green(9, Li, u_0, tau, poisson ? 1 : 0, neq ? 1 : 0, f.data()),
               "erlbm_init_taylor_green");
    erlbm_lattice *h = nullptr;
    drv::check(erlbm_lattice_create(9, Li, Li, 1, nullptr, precision, nullptr, &h), "erlbm_lattice_create");
    drv::check(erlbm_set_params(h, tau, tau_nh, 1.0, col), "erlbm_set_params");
    drv::check(erlbm_upload_aos(h, f.data()), "erlbm_upload_aos");
    drv::check(erlbm_step(h, nT + 1, 0, 0.0), "erlbm_step");
    drv::check(erlbm_download_aos(h, ERLBM_POST_COLLISION, f.data()), "erlbm_download_aos");
    erlbm_lattice_destroy(h);

    double e[3];
    drv::check(erlbm_taylor_green_error(9, Li, u_0, visc, nT, f.data(), e), "erlbm_taylor_green_error");
    std::printf("%d %.6g %.6g %.6g %d %.6g %.6g %.6g\n", Li, tau, tau_nh, u_0, nT, e[0], e[1], e[2]);
    return 0;
}
