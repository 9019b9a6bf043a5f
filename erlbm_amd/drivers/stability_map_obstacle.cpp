// stability_map_obstacle -- the obstacle-channel stability map driver
// (src/main_stability-map_obstacle.cpp) on the GPU: every map point is a job
// of one batch of 22 x 12 x 1 D3Q19 channels (Zou-He inlet at x = 1,
// outlet rescale at x = 20); same options, same map file (1 = unstable:
// |rho| > 10 or non-finite at some site after some step of 200 ny / U0).
//
//   stability_map_obstacle [--n-tau 400] [--n-nu 400] [--ma 0.1]
//                          [--out stability_map_obstacle.vtk]
//                          [--precision fp64] [--autotune 0]
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "args.hpp"

int main(int argc, char **argv)
{
    drv::Args args{argc, argv};
    const int n_tau = args.integer("--n-tau", 400);
    const int n_nu = args.integer("--n-nu", 400);
    const double Ma = args.real("--ma", 0.1);
    const std::string out = args.str("--out", "stability_map_obstacle.vtk");
    const int precision = drv::precision(args.str("--precision", "fp64"));
    const bool tune = args.integer("--autotune", 0) != 0;
    std::printf("stability map, obstacle (D3Q19, 20 x 12, Ma = %g), E-RLBM, %d x %d points (tau_nh x log10 nu)\n",
                Ma, n_tau, n_nu);

    const size_t n = (size_t)n_tau * n_nu;
    std::vector<erlbm_job> jobs(n);
    std::vector<erlbm_job_result> res(n);
    drv::check(erlbm_obstacle_map_jobs(n_tau, n_nu, Ma, precision, 0, n_tau, 0, n_nu, jobs.data()),
               "erlbm_obstacle_map_jobs");
    int mask[22 * 12];
    drv::check(erlbm_obstacle_mask(mask), "erlbm_obstacle_mask");
    erlbm_batch *b = nullptr;
    drv::check(erlbm_batch_create(22, 12, mask, 1, 20, nullptr, &b), "erlbm_batch_create");
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
    drv::check(erlbm_write_obstacle_map_vtk(out.c_str(), n_tau, n_nu, Ma, map.data()),
               "erlbm_write_obstacle_map_vtk");
    std::printf("%s written: %d of %zu points unstable (%.2f s, %.0f points/s)\n", out.c_str(), n_unstable, n, s,
                n / s);
    return 0;
}
