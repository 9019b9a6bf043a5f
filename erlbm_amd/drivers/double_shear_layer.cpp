// double_shear_layer -- the double shear layer driver
// (src/main_double-shear-layer.cpp) on the GPU: same options, same output
// files in the working directory:
//   k_energy_lbm.dat   t* and kinetic energy at every output step
//   vel_<t*>.vtk       velocity field (rec_velocity's format and name)
//   vor_<step>.vtk     vorticity field (unless --vorticity 0)
//
//   double_shear_layer [--L 256] [--collision erlbm] [--tau-nh 0.65]
//                      [--Re 80000] [--tau T] [--t-final 1.5] [--dt-out 1/16]
//                      [--vorticity 1] [--precision fp64]
//
// An output step is measured after its collision, on the post-collision
// populations, as the reference does; the run stops at an output step whose
// rms velocity exceeds 0.4 or whose energy is not finite.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "args.hpp"

namespace {
const double U_0 = 1. / 32.;
const double U_LIMIT = 0.4;
}  // namespace

int main(int argc, char **argv)
{
    drv::Args args{argc, argv};
    const int Li = args.integer("--L", 256);
    const int col = drv::collision(args.str("--collision", "erlbm"));
    const double tau_nh_arg = args.real("--tau-nh", 0.65);
    const double Re_in = args.real("--Re", 80000.);
    const double t_final = args.real("--t-final", 1.5);
    const double dt_out = args.real("--dt-out", 1. / 16.);
    const bool write_vor = args.integer("--vorticity", 1) != 0;
    const int precision = drv::precision(args.str("--precision", "fp64"));

    const double L = Li;
    const double tau = args.has("--tau") ? args.real("--tau", 0.5) : 0.5 + 3. * U_0 * L / Re_in;
    const double tau_nh = (col == ERLBM_COL_REG) ? 1.0 : tau_nh_arg;

    const int steps_per_t = drv::iround(L / U_0);
    const int n_steps = drv::iround(t_final * steps_per_t) + 1;
    const int interval = std::max(1, drv::iround(dt_out * steps_per_t));

    std::printf("double shear layer (D2Q9, MI355X)\n   collision %s, L %d, tau / tau_nh %g / %g\n"
                "   steps %d (t* final %g), output every %d steps\n",
                drv::collision_name(col), Li, tau, tau_nh, n_steps, t_final, interval);

    std::vector<double> f((size_t)Li * Li * 9);
    drv::check(erlbm_init_double_shear(Li, 1.0, f.data()), "erlbm_init_double_shear");
    erlbm_lattice *h = nullptr;
    drv::check(erlbm_lattice_create(9, Li, Li, 1, nullptr, precision, nullptr, &h), "erlbm_lattice_create");
    drv::check(erlbm_set_params(h, tau, tau_nh, 1.0, col), "erlbm_set_params");
    drv::check(erlbm_upload_aos(h, f.data()), "erlbm_upload_aos");
    std::vector<double>().swap(f);
    std::vector<double> field((size_t)Li * Li * 3);

    std::FILE *fe = std::fopen("k_energy_lbm.dat", "w");
    if (!fe) {
        std::perror("k_energy_lbm.dat");
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    int step = 0;
    bool unstable = false;
    while (step < n_steps && !unstable) {
        // run through the next output step (or the last step)
        int next = ((step + interval - 1) / interval) * interval;
        if (next > n_steps - 1) next = n_steps - 1;
        drv::check(erlbm_step(h, next - step + 1, 0, 0.0), "erlbm_step");
        step = next;
        if (step % interval == 0) {
            const double t_adm = (double)step * U_0 / L;
            double e = 0.0;
            drv::check(erlbm_kinetic_energy(h, ERLBM_POST_COLLISION, 1, &e), "erlbm_kinetic_energy");
            std::printf("t* = %g   kinetic energy = %g\n", t_adm, e);
            std::fprintf(fe, "%.10g %.10g\n", t_adm, e);
            if (!std::isfinite(e) || std::sqrt(2. * e / ((double)Li * Li)) > U_LIMIT) {
                std::printf("unstable at step %d (t* = %g)\n", step, t_adm);
                unstable = true;
                break;
            }
            drv::check(erlbm_velocity_field(h, ERLBM_POST_COLLISION, field.data()), "erlbm_velocity_field");
            const std::string vel = "vel_" + std::to_string(t_adm) + ".vtk";
            drv::check(erlbm_write_velocity_vtk(vel.c_str(), Li, Li, 1, field.data()), "erlbm_write_velocity_vtk");
            if (write_vor) {
                drv::check(erlbm_vorticity_field(h, ERLBM_POST_COLLISION, field.data()), "erlbm_vorticity_field");
                const std::string vor = "vor_" + std::to_string((unsigned)step) + ".vtk";
                drv::check(erlbm_write_vorticity_vtk(vor.c_str(), Li, Li, 1, field.data()),
                           "erlbm_write_vorticity_vtk");
            }
        }
        step++;
    }
    std::fclose(fe);
    drv::check(erlbm_synchronize(h), "erlbm_synchronize");
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    long long done = 0;
    int reason = 0;
    drv::check(erlbm_status(h, &done, &reason), "erlbm_status");
    std::printf("steps run %lld, %.3f s, %.1f MLUPS (including output)\n", done, s,
                s > 0 ? (double)Li * Li * done / (s * 1e6) : 0.);
    erlbm_lattice_destroy(h);
    return 0;
}
