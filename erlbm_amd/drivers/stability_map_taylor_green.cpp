// stability_map_taylor_green -- the Taylor-Green stability map driver
// (src/main_stability-map_taylor-green.cpp) on the GPU: every map point is a
// job of one batch of 16 x 16 x 1 D3Q19 lattices; same options, same map
// file (1 = unstable: E > 2 E0 or non-finite at a check every 100 steps,
// within 100 L / U0 steps).
//
//   stability_map_taylor_green [--collision erlbm] [--tau-nh 0.65] [--L 16]
//                              [--n-ma 400] [--n-nu 400]
//                              [--out stability_map_taylor_green.vtk]
//                              [--precision fp64] [--autotune 0]
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "args.hpp"

int main(int argc, char **argv)
{
    drv::Args args{argc, argv};
    const int col = drv::collision(args.str("--collision", "erlbm"));
    const double tau_nh = args.real("--tau-nh", 0.65);
    const int Li = args.integer("--L", 16);
    const int n_ma = args.integer("--n-ma", 400);
    const int n_nu = args.integer("--n-nu", 400);
    const std::string out = args.str("--out", "stability_map_taylor_green.vtk");
    const int precision = drv::precision(args.str("--precision", "fp64"));
    const bool tune = args.integer("--autotune", 0) != 0;
    if (Li != 16) {
        std::fprintf(stderr, "the batched map runs the published 16 x 16 lattice (--L 16)\n");
        return 1;
    }
    std::printf("stability map, Taylor-Green (D3Q19, L = %d), %s", Li, drv::collision_name(col));
    if (col == ERLBM_COL_EXT_REG) std::printf(" (tau_nh = %g)", tau_nh);
    std::printf(", %d x %d points (Ma x log10 nu)\n", n_ma, n_nu);

    const size_t n = (size_t)n_ma * n_nu;
    std::vector<erlbm_job> jobs(n);
    std::vector<erlbm_job_result> res(n);
    drv::check(erlbm_tg_map_jobs(n_ma, n_nu, col, tau_nh, precision, 0, n_ma, 0, n_nu, jobs.data()),
               "erlbm_tg_map_jobs");
    erlbm_batch *b = nullptr;
    drv::check(erlbm_batch_create(Li, Li, nullptr, -1, -1, nullptr, &b), "erlbm_batch_create");
    if (tune) {
        int w64 = 0, w32 = 0;
        double pps = 0.0;
        drv::check(erlbm_batch_autotune(b, jobs.data(), (int)std::min<size_t>(n, 4096), 2000, &w64, &w32, &pps),
                   "erlbm_batch_autotune");
        std::printf("autotune: %d / %d workgroups per CU (FP64 / FP32)\n", w64, w32);
    }
    const auto t0 = std::chrono::steady_clock::now();
    drv::check(erlbm_batch_run(b, jobs.data(), (int)n, res.data()), "erlbm_batch_run");
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    erlbm_batch_destroy(b);

    std::vector<int> map(n);
    int n_unstable = 0;
    for (size_t k = 0; k < n; k++) {
        map[k] = res[k].unstable;
        n_unstable += map[k];
    }
    drv::check(erlbm_write_tg_map_vtk(out.c_str(), n_ma, n_nu, col, tau_nh, map.data()), "erlbm_write_tg_map_vtk");
    std::printf("%s written: %d of %zu points unstable (%.2f s, %.0f points/s)\n", out.c_str(), n_unstable, n, s,
                n / s);
    return 0;
}
