// maps.hip -- batched tiny-lattice runs (stability maps and small requests).
//
// Thousands of independent tiny D3Q19 lattices (the Taylor-Green map's
// 16x16x1 and the obstacle map's 22x12x1 channel, src/main_stability-map_*.cpp)
// run in one persistent kernel.  Each workgroup owns one lattice at a time,
// keeps its populations resident in LDS (19 x 256 x 8 B = 38.9 KB in FP64,
// half that in FP32), runs it step after step and, on retirement (stop test
// fired or step budget used), takes the next job from a device-side queue:
// one atomic per job, no host round trip, early exit frees the workgroup at
// once.  A launch may also stop after a budget of steps per job and save the
// job to HBM, so a host loop can admit new work between chunks.
//
// Per step (the plain path's operation order): pull-gather from LDS with
// periodic wrap and halfway bounce-back, inlet / outlet columns, collision,
// the map driver's stop test -- the energy test every check_every steps
// (ordered FP64 sum over the sites, main_stability-map_taylor-green.cpp:71-81)
// or the post-collision density test every step
// (main_stability-map_obstacle.cpp:104-106).  FP64 jobs are bitwise those of
// the reference; FP32 jobs store shifted deviations f_i - w_i rho0 in LDS
// (half the footprint) and collide in FP64 or, on request, in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "collide.hpp"
#include "collide_f32.hpp"
#include "common.hpp"
#include "maps.hpp"

using namespace erlbm;

namespace {

constexpr int Q = 19;
constexpr int TPB = 256;  // threads per workgroup = max fluid sites of a tiny lattice
constexpr uint32_t FLUID = 1u << 31;

struct TinyGeom {
    int nx, ny;                  // nz = 1
    int n_fluid;                 // <= TPB
    int nvp;                     // LDS plane stride (voxels, padded)
    int general;                 // solids or boundary columns present
    int inlet_x, outlet_x, outlet_dx;
    const int *vox_of_fluid;     // [n_fluid]: voxel of thread p (boundary columns last)
    const int *fluid_of_vox;     // [nx*ny]: index in fluid order (ascending voxels), -1 = solid
    const uint32_t *nbr;         // [nx*ny]: bit i = x - c_i solid, FLUID
    const double *sin_tab;       // sin(2 pi x / L), x < max(nx, ny)
    const double *cos_tab;
};

struct RunArgs {
    TinyGeom g;
    const MapJob *jobs;
    MapResult *results;          // in/out: a job with 0 < steps and !done resumes
    const int *job_idx;          // jobs of this launch
    int n_jobs;
    const int *n_jobs_dev;       // if set, the job count is read from device memory
    unsigned int *next_job;      // queue head
    int max_steps;               // steps per job in this launch (<= 0: to retirement)
    void *saved;                 // per job state between launches, Q * nvp values each
};

// ---- storage policies -------------------------------------------------------
struct F64 {
    using T = double;
    static constexpr int code = ERLBM_FP64;
    struct K {
        CollideConsts c;
    };
    __device__ static K consts(const MapJob &J)
    {
        K k;
        k.c.omega1 = 1.0 - 1.0 / J.tau;
        k.c.omega2 = 1.0 - 1.0 / J.tau_nh;
        k.c.oot = 1.0 / J.tau;
        return k;
    }
    __device__ static void init(double (&f)[Q], double rho0, double ux, double uy)
    {
        feq_exact<Q>(rho0, ux, uy, 0.0, f);
    }
    __device__ static double ke(const double (&f)[Q], double)
    {
        double rho, vx, vy, vz;
        moments_exact<Q, true, false>(f, rho, vx, vy, vz);
        return ke_term<Q>(vx, vy, vz);
    }
    template <int COL>
    __device__ static void collide(double (&f)[Q], const K &k, double)
    {
        double vx, vy, vz;
        collide_exact<Q, COL, false, false>(f, k.c, vx, vy, vz);
    }
    __device__ static double density(const double (&f)[Q], double) { return density_exact<Q>(f); }
    __device__ static void to_abs(const double (&f)[Q], double (&a)[Q], double)
    {
        for (int i = 0; i < Q; i++) a[i] = f[i];
    }
    __device__ static void from_abs(const double (&a)[Q], double (&f)[Q], double)
    {
        for (int i = 0; i < Q; i++) f[i] = a[i];
    }
};

// FP32-shifted storage d_i = f_i - w_i rho0; FP64 arithmetic on the
// absolute populations (the exact collision of collide.hpp).
struct F32 {
    using T = float;
    static constexpr int code = ERLBM_FP32_SHIFTED;
    struct K {
        CollideConsts c;
    };
    __device__ static K consts(const MapJob &J) { return K{F64::consts(J).c}; }
    __device__ static void to_abs(const float (&d)[Q], double (&a)[Q], double rho0)
    {
        for (int i = 0; i < Q; i++) a[i] = (double)d[i] + VSet<Q>::w[i] * rho0;
    }
    __device__ static void from_abs(const double (&a)[Q], float (&d)[Q], double rho0)
    {
        for (int i = 0; i < Q; i++) d[i] = (float)(a[i] - VSet<Q>::w[i] * rho0);
    }
    __device__ static void init(float (&d)[Q], double rho0, double ux, double uy)
    {
        double f[Q];
        feq_exact<Q>(rho0, ux, uy, 0.0, f);
        from_abs(f, d, rho0);
    }
    __device__ static double ke(const float (&d)[Q], double rho0)
    {
        double f[Q];
        to_abs(d, f, rho0);
        return F64::ke(f, rho0);
    }
    template <int COL>
    __device__ static void collide(float (&d)[Q], const K &k, double rho0)
    {
        double f[Q];
        to_abs(d, f, rho0);
        double vx, vy, vz;
        collide_exact<Q, COL, false, false>(f, k.c, vx, vy, vz);
        from_abs(f, d, rho0);
    }
    __device__ static double density(const float (&d)[Q], double rho0)
    {
        double f[Q];
        to_abs(d, f, rho0);
        return density_exact<Q>(f);
    }
};

// FP32-shifted storage with FP32 arithmetic (collide_f32.hpp).
struct F32A {
    using T = float;
    static constexpr int code = ERLBM_FP32_SHIFTED_F32;
    struct K {
        CollideConstsF32 c;
    };
    __device__ static K consts(const MapJob &J)
    {
        K k;
        k.c.omega1 = (float)(1.0 - 1.0 / J.tau);
        k.c.omega2 = (float)(1.0 - 1.0 / J.tau_nh);
        k.c.oot = (float)(1.0 / J.tau);
        k.c.rho0 = (float)J.rho0;
        return k;
    }
    __device__ static void init(float (&d)[Q], double rho0, double ux, double uy) { F32::init(d, rho0, ux, uy); }
    __device__ static double ke(const float (&d)[Q], double rho0)
    {
        float drho, rho, vx, vy, vz, jx, jy, jz;
        moments_shifted<Q>(d, (float)rho0, drho, rho, vx, vy, vz, jx, jy, jz);
        const double x = vx, y = vy, z = vz;
        return 0.5 * (x * x + y * y + z * z);
    }
    template <int COL>
    __device__ static void collide(float (&d)[Q], const K &k, double)
    {
        float drho, vx, vy, vz;
        collide_shifted_f32<Q, COL>(d, k.c, drho, vx, vy, vz);
    }
    __device__ static double density(const float (&d)[Q], double rho0) { return density_shifted_f64<Q>(d, rho0); }
    __device__ static void to_abs(const float (&d)[Q], double (&a)[Q], double rho0) { F32::to_abs(d, a, rho0); }
    __device__ static void from_abs(const double (&a)[Q], float (&d)[Q], double rho0) { F32::from_abs(a, d, rho0); }
};

// ZouHeD3Q19_vx_x (Boundary_conditions.cpp:350-442), west branch, on
// absolute FP64 populations.
__device__ __forceinline__ void zouhe_west(double (&f)[Q], double vx)
{
    double sum_f_xminus = f[1] + f[7] + f[9] + f[11] + f[13];
    double sum_f_xzero = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18];
    double rho = (2.0 * sum_f_xminus + sum_f_xzero) / (1.0 - vx);
    double vy = 0.0, vz = 0.0;
    f[1] = f[1] + rho * vx / 3.0;
    f[7] = f[7] + rho * vx / 6.0 + 0.5 * rho * vy - 0.5 * (f[18] + f[3] + f[16] - f[15] - f[4] - f[17]);
    f[9] = f[9] + rho * vx / 6.0 - 0.5 * rho * vy + 0.5 * (f[18] + f[3] + f[16] - f[15] - f[4] - f[17]);
    f[11] = f[11] + rho * vx / 6.0 + 0.5 * rho * vz - 0.5 * (f[17] + f[5] + f[16] - f[15] - f[6] - f[18]);
    f[13] = f[13] + rho * vx / 6.0 - 0.5 * rho * vz + 0.5 * (f[17] + f[5] + f[16] - f[15] - f[6] - f[18]);
}

__device__ __forceinline__ void zouhe_east(double (&f)[Q], double vx)
{
    double sum_f_xplus = f[2] + f[8] + f[10] + f[12] + f[14];
    double sum_f_xzero = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18];
    double rho = (2.0 * sum_f_xplus + sum_f_xzero) / (1.0 - vx);
    double vy = 0.0, vz = 0.0;
    f[2] = f[2] - rho * vx / 3.0;
    f[10] = f[10] - rho * vx / 6.0 + 0.5 * rho * vy - 0.5 * (f[18] + f[3] + f[16] - f[15] - f[4] - f[17]);
    f[8] = f[8] - rho * vx / 6.0 - 0.5 * rho * vy + 0.5 * (f[18] + f[3] + f[16] - f[15] - f[4] - f[17]);
    f[14] = f[14] - rho * vx / 6.0 + 0.5 * rho * vz - 0.5 * (f[17] + f[5] + f[16] - f[15] - f[6] - f[18]);
    f[12] = f[12] - rho * vx / 6.0 - 0.5 * rho * vz + 0.5 * (f[17] + f[5] + f[16] - f[15] - f[6] - f[18]);
}

// Pull-gather of voxel (x, y) from the voxel-indexed LDS populations
// s[i * nvp + v]: x - c_i with periodic wrap, or the opposite population of
// the voxel itself when x - c_i is solid (bit i of `bounce`).  stream =
// false reads the voxel's own populations (the first step after init).
template <typename T>
__device__ __forceinline__ void gather(const T *s, int nvp, int x, int y, int nx, int ny, uint32_t bounce,
                                       bool stream, T (&f)[Q])
{
    using V = VSet<Q>;
    const int v = x + nx * y;
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        if (!stream || (V::cx[i] == 0 && V::cy[i] == 0)) {
            f[i] = s[i * nvp + v];
        } else if (bounce & (1u << i)) {
            f[i] = s[V::opp[i] * nvp + v];
        } else {
            const int xs = V::cx[i] == 0 ? x : (V::cx[i] > 0 ? xm : xp);
            const int ys = V::cy[i] == 0 ? y : (V::cy[i] > 0 ? ym : yp);
            f[i] = s[i * nvp + xs + nx * ys];
        }
    });
}

// The LDS element offsets gather() reads for a streaming step, computed once
// per job: the pulls of a voxel never change during a run.
__device__ __forceinline__ void pull_offsets(int nvp, int x, int y, int nx, int ny, uint32_t bounce, int (&off)[Q])
{
    using V = VSet<Q>;
    const int v = x + nx * y;
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        if (V::cx[i] == 0 && V::cy[i] == 0) {
            off[i] = i * nvp + v;
        } else if (bounce & (1u << i)) {
            off[i] = V::opp[i] * nvp + v;
        } else {
            const int xs = V::cx[i] == 0 ? x : (V::cx[i] > 0 ? xm : xp);
            const int ys = V::cy[i] == 0 ? y : (V::cy[i] > 0 ? ym : yp);
            off[i] = i * nvp + xs + nx * ys;
        }
    });
}

// Ordered sum of n doubles held in LDS (the reference's sequential loop).
__device__ __forceinline__ double ordered_sum(const double *v, int n)
{
    double s = 0;
    for (int p = 0; p < n; p++) s = s + v[p];
    return s;
}

// Energy of the gathered state, ordered over the sites; `scratch` aliases the
// population LDS, so every gather of this step must be complete (barrier).
template <class P>
__device__ __forceinline__ double ordered_energy(typename P::T (&f)[Q], bool active, int rank, double rho0,
                                                 double *scratch, int n_fluid, double *bcast)
{
    __syncthreads();
    if (active) scratch[rank] = P::ke(f, rho0);   // rank: this site's index in fluid order
    __syncthreads();
    if (threadIdx.x == 0) bcast[0] = ordered_sum(scratch, n_fluid);
    __syncthreads();
    return bcast[0];
}

// One job in one workgroup; thread p owns voxel vox_of_fluid[p] (sites of the
// boundary columns on the last threads), whose index in fluid order is
// fluid_of_vox of that voxel.
// GENERAL: solids and/or boundary columns.  NX, NY > 0: the lattice sizes
// known at compile time (the Taylor-Green maps' 16 x 16, the obstacle map's
// 22 x 12 channel), so a pull's LDS address is one of nine neighbour offsets
// plus a constant population offset the LDS instruction carries; halfway
// bounce-back then goes through ghost values (below) instead of per-site
// pull offsets.
template <class P, int COL, bool GENERAL, int NX = 0, int NY = 0>
__device__ void run_job(const RunArgs &a, const MapJob &J, MapResult *res, typename P::T *sf, double *bcast)
{
    using T = typename P::T;
    using V = VSet<Q>;
    static_assert((NX == 0) == (NY == 0) && NX * NY <= 400, "compile-time sizes: both or neither");
    constexpr int NVP = (NX * NY + 7) / 8 * 8;
    const int tid = threadIdx.x;
    const int nf = a.g.n_fluid;
    const int nx = NX ? NX : a.g.nx, ny = NY ? NY : a.g.ny, nvp = NX ? NVP : a.g.nvp;
    const bool active = tid < nf;
    const auto k = P::consts(J);
    const double rho0 = J.rho0;

    int my_v = 0, my_x = 0, my_y = 0, my_rank = tid;
    uint32_t my_bounce = 0, out_bounce = 0;
    bool out_fluid = false;
    if (active) {
        my_v = GENERAL ? a.g.vox_of_fluid[tid] : tid;
        if constexpr (GENERAL) my_rank = a.g.fluid_of_vox[my_v];
        my_x = my_v % nx;
        my_y = my_v / nx;
        if constexpr (GENERAL) {
            my_bounce = a.g.nbr[my_v] & ~FLUID;
            if (my_x == a.g.outlet_x) {
                const uint32_t nb = a.g.nbr[my_v + a.g.outlet_dx];
                out_fluid = (nb & FLUID) != 0;
                out_bounce = nb & ~FLUID;
            }
        }
    }
    // scratch for the ordered energy sums: the population LDS itself (only
    // used between a completed gather and the next write)
    double *scratch = reinterpret_cast<double *>(sf);

    // ---- initial state and E0, or the state saved by the previous launch
    T f[Q];
    const MapResult prev = *res;
    const bool resume = a.saved != nullptr && prev.steps > 0 && !prev.done;
    T *slot = a.saved ? static_cast<T *>(a.saved) + (size_t)(res - a.results) * Q * nvp : nullptr;
    double e0 = 0.0;
    if (resume) {
        e0 = prev.e0;
        if (active)
            for (int i = 0; i < Q; i++) sf[i * nvp + my_v] = slot[i * nvp + my_v];
    } else {
        if (active) {
            double ux = 0.0, uy = 0.0;
            if (J.init == INIT_TAYLOR_GREEN) {
                ux = J.u0 * a.g.sin_tab[my_x] * a.g.cos_tab[my_y];
                uy = -J.u0 * a.g.cos_tab[my_x] * a.g.sin_tab[my_y];
            }
            P::init(f, rho0, ux, uy);
        }
        if (J.check_every > 0) e0 = ordered_energy<P>(f, active, my_rank, rho0, scratch, nf, bcast);
        __syncthreads();
        if (active)
            for (int i = 0; i < Q; i++) sf[i * nvp + my_v] = f[i];
    }
    __syncthreads();

    long long s = resume ? prev.steps : 0;
    long long end = J.n_steps;
    if (a.max_steps > 0 && s + a.max_steps < end) end = s + a.max_steps;
    int unstable = 0;
    double e_last = resume ? prev.energy : e0;
    double e_max = e0;      // largest energy seen by the energy test
    double r_max = 0.0;     // largest |rho| seen by this thread's density test
    // The voxel this thread gathers: its own, or on the outlet column the
    // upstream neighbour, whose populations devnull_rho_x copies
    // (Boundary_conditions.cpp:117-126).  Its pull offsets are fixed for the
    // run, so they are computed once.
    const bool outlet_copy = GENERAL && active && my_x == a.g.outlet_x && out_fluid;
    const int gx = outlet_copy ? my_x + a.g.outlet_dx : my_x;
    int off[NX ? 1 : Q];
    int nb9[9];   // NX > 0: voxel of x - c_x, y - c_y for (c_x + 1) * 3 + c_y + 1
    if constexpr (NX > 0) {
        const int xm = gx == 0 ? NX - 1 : gx - 1, xp = gx == NX - 1 ? 0 : gx + 1;
        const int ym = my_y == 0 ? NY - 1 : my_y - 1, yp = my_y == NY - 1 ? 0 : my_y + 1;
        const int xs[3] = {xp, gx, xm}, ys[3] = {yp, my_y, ym};
        for (int p = 0; p < 9; p++) nb9[p] = xs[p / 3] + NX * ys[p % 3];
    } else if (active) {
        pull_offsets(nvp, gx, my_y, nx, ny, outlet_copy ? out_bounce : my_bounce, off);
    }
    // Ghost values (GENERAL, NX > 0): a site whose pull of population i
    // would come from a solid voxel x - c_i writes its own post-collision
    // f_opp(i) into that voxel's slot of population i -- which only this
    // site's pull reads -- so every pull reads x - c_i alike.  The outlet
    // column's sites gather at their upstream neighbour, whose own thread
    // writes the ghosts that gather reads.
    const bool ghosts = GENERAL && NX > 0 && active && gx == my_x && my_bounce != 0;
    auto write_ghosts = [&](auto own) {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            if constexpr (V::cx[i] != 0 || V::cy[i] != 0)
                if (my_bounce & (1u << i)) sf[i * NVP + nb9[(V::cx[i] + 1) * 3 + V::cy[i] + 1]] = own(V::opp[i]);
        });
    };
    if constexpr (GENERAL && NX > 0) {
        if (resume) {   // the ghosts of the saved state
            if (ghosts) write_ghosts([&](int i) { return sf[i * NVP + my_v]; });
            __syncthreads();
        }
    }
    for (; s < end; s++) {
        double fator = 0.0;   // outlet column: rho0 / rho(upstream neighbour)
        if (active) {
            if (s > 0) {
                if constexpr (NX > 0)
                    static_for<Q>([&](auto I) {
                        constexpr int i = I;
                        f[i] = sf[i * NVP + nb9[(V::cx[i] + 1) * 3 + V::cy[i] + 1]];
                    });
                else
                    static_for<Q>([&](auto I) { f[(int)I] = sf[off[(int)I]]; });
            } else {                 // first step after the initial state: no stream
                static_for<Q>([&](auto I) { f[(int)I] = sf[(int)I * nvp + gx + nx * my_y]; });
            }
            if (outlet_copy) {
                double fn[Q];
                P::to_abs(f, fn, rho0);
                fator = rho0 / density_exact<Q>(fn);
            }
        }
        if (J.check_every > 0 && s > 0 && s % J.check_every == 0) {
            // (energy tests run on lattices without an outlet column; checked on the host)
            const double e = ordered_energy<P>(f, active, my_rank, rho0, scratch, nf, bcast);
            e_last = e;
            e_max = (e > e_max || !isfinite(e)) ? e : e_max;
            if (e > J.energy_factor * e0 || !isfinite(e)) {
                unstable = 1;
                break;
            }
        }
        bool bad = false;
        if (active) {
            if constexpr (GENERAL) {
                if (my_x == a.g.outlet_x) {
                    double fd[Q];
                    if (out_fluid) {
                        P::to_abs(f, fd, rho0);
                        for (int i = 0; i < Q; i++) fd[i] = fd[i] * fator;
                    } else {
                        for (int i = 0; i < Q; i++) fd[i] = V::w[i] * rho0;   // dist_eq(f, rho)
                    }
                    P::from_abs(fd, f, rho0);
                } else if (my_x == a.g.inlet_x) {
                    double fd[Q];
                    P::to_abs(f, fd, rho0);
                    if (a.g.inlet_x < nx / 2) zouhe_west(fd, J.u0);
                    else zouhe_east(fd, J.u0);
                    P::from_abs(fd, f, rho0);
                }
            }
            P::template collide<COL>(f, k, rho0);
            if (J.dens_limit > 0.0) {
                const double d = P::density(f, rho0);
                bad = !(fabs(d) <= J.dens_limit);
                r_max = (fabs(d) > r_max || !isfinite(d)) ? fabs(d) : r_max;
            }
        }
        __syncthreads();  // every gather of this step is done
        if (active)
            for (int i = 0; i < Q; i++) sf[i * nvp + my_v] = f[i];
        if constexpr (GENERAL && NX > 0)
            if (ghosts) write_ghosts([&](int i) { return f[i]; });
        const int any_bad = __syncthreads_or(bad ? 1 : 0);
        if (any_bad) {
            unstable = 1;
            s++;
            break;
        }
    }
    const bool finished = unstable || s >= J.n_steps;
    if (!finished && slot && active)
        for (int i = 0; i < Q; i++) slot[i * nvp + my_v] = sf[i * nvp + my_v];
    double margin = 0.0;
    if (J.dens_limit > 0.0) {
        // block max of the per-site |rho| maxima in fluid order (NaN
        // propagates; threads without a site hold 0 beyond the sites)
        __syncthreads();
        scratch[active ? my_rank : tid] = r_max;
        __syncthreads();
        if (tid == 0) {
            double m = 0.0;
            for (int p = 0; p < TPB; p++) m = (scratch[p] > m || !isfinite(scratch[p])) ? scratch[p] : m;
            bcast[1] = m / J.dens_limit;
        }
        __syncthreads();
        margin = bcast[1];
    } else if (J.check_every > 0) {
        margin = e_max / e0;
    }
    if (resume) {   // running maximum over launches; a non-finite value sticks
        if (!isfinite(prev.margin)) margin = prev.margin;
        else if (isfinite(margin) && prev.margin > margin) margin = prev.margin;
    }
    if (tid == 0) {
        MapResult r;
        r.margin = margin;
        r.unstable = unstable;
        r.steps = (int)s;
        r.done = finished ? 1 : 0;
        r.precision = P::code;
        r.e0 = e0;
        r.energy = e_last;
        *res = r;
    }
    __syncthreads();
}

// Wavefronts per SIMD the compiler must leave room for: three for the FP64
// periodic lattices -- at run-time sizes 168 VGPRs instead of 192, a few
// spills, 9% faster over the full E-RLBM map -- and for the 16 x 16 ones
// four under RLBM (128 VGPRs, 40 bytes spilled: TG RLBM map 55.7k -> 62.4k
// points/s; E-RLBM at four spills 148 bytes and runs 2% slower than at
// three, 162 VGPRs, no spill); three for FP32 storage of the 16 x 16 ones
// (E-RLBM 172 -> 168 VGPRs: its TG map 35.5k -> 40.8k points/s); the
// natural two otherwise (the boundary-column code of the obstacle channel
// spills 51 VGPRs at three and runs 14% slower).
template <class P, int COL, bool GENERAL, int NX>
constexpr int tiny_waves_per_eu()
{
    if (GENERAL) return NX > 0 && !std::is_same<P, F32A>::value ? 3 : 2;
    if (std::is_same<P, F64>::value) return NX > 0 && COL == COL_REG ? 4 : 3;
    return std::is_same<P, F32>::value && NX > 0 ? 3 : 2;
}

template <class P, int COL, bool GENERAL, int NX = 0, int NY = 0>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(tiny_waves_per_eu<P, COL, GENERAL, NX>())))
void k_tiny(RunArgs a)
{
    extern __shared__ double lds_raw[];   // Q * nvp populations of type P::T
    __shared__ double bcast[2];
    __shared__ int sjob;
    typename P::T *sf = reinterpret_cast<typename P::T *>(lds_raw);
    const int n_jobs = a.n_jobs_dev ? *a.n_jobs_dev : a.n_jobs;
    for (;;) {
        if (threadIdx.x == 0) sjob = (int)atomicAdd(a.next_job, 1u);
        __syncthreads();
        const int k = sjob;
        __syncthreads();
        if (k >= n_jobs) return;
        const int j = a.job_idx[k];
        const MapJob J = a.jobs[j];
        run_job<P, COL, GENERAL, NX, NY>(a, J, &a.results[j], sf, bcast);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
struct erlbm_batch {
    int nx = 0, ny = 0, n_fluid = 0;
    bool general = false;
    int inlet_x = -1, outlet_x = -1, outlet_dx = -1;
    hipStream_t stream = nullptr;
    int *vox_of_fluid = nullptr, *fluid_of_vox = nullptr;
    uint32_t *nbr = nullptr;
    double *tabs = nullptr;         // sin | cos tables
    int tab_n = 0;
    unsigned int *counters = nullptr;
    int wg_per_cu[2] = {0, 0};      // FP64, FP32 (0 = occupancy max)
    int n_cu = 0;
    int chunk_steps = 0;            // > 0: graph-replayed chunks of this many steps
    hipStream_t cstream = nullptr;  // capture stream of the chunked mode (created on first use)
    void *d_saved = nullptr;        // per job state between chunks
    size_t saved_cap = 0;
    int *d_n = nullptr;             // live job count of the chunked mode
    // job buffers (grown on demand)
    MapJob *d_jobs = nullptr;
    MapResult *d_results = nullptr;
    int *d_idx = nullptr;
    int cap = 0;
};

namespace {

template <class P, int COL, bool GENERAL, int NX = 0, int NY = 0>
int launch_group(erlbm_batch *b, const RunArgs &a, unsigned int *counter)
{
    if constexpr (NX == 0) {   // the maps' lattices: sizes compiled in
        if (!GENERAL && a.g.nx == 16 && a.g.ny == 16) return launch_group<P, COL, false, 16, 16>(b, a, counter);
        if (GENERAL && a.g.nx == 22 && a.g.ny == 12) return launch_group<P, COL, true, 22, 12>(b, a, counter);
    }
    int per_cu = 0;
    const size_t lds = sizeof(typename P::T) * Q * (size_t)a.g.nvp;
    ERLBM_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tiny<P, COL, GENERAL, NX, NY>, TPB, lds));
    const int want = b->wg_per_cu[P::code == ERLBM_FP64 ? 0 : 1];
    if (want > 0 && want < per_cu) per_cu = want;
    if (per_cu < 1) per_cu = 1;
    int grid = b->n_cu * per_cu;
    if (grid > a.n_jobs) grid = a.n_jobs;
    if (grid < 1) return ERLBM_OK;
    RunArgs ra = a;
    ra.next_job = counter;
    hipLaunchKernelGGL((k_tiny<P, COL, GENERAL, NX, NY>), dim3(grid), dim3(TPB), lds, b->stream, ra);
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

template <class P>
int launch_col(erlbm_batch *b, int col, const RunArgs &a, unsigned int *counter)
{
    const bool g = b->general;
    switch (col) {
    case ERLBM_COL_BGK: return g ? launch_group<P, COL_BGK, true>(b, a, counter) : launch_group<P, COL_BGK, false>(b, a, counter);
    case ERLBM_COL_REG: return g ? launch_group<P, COL_REG, true>(b, a, counter) : launch_group<P, COL_REG, false>(b, a, counter);
    case ERLBM_COL_EXT_REG:
        return g ? launch_group<P, COL_EXT_REG, true>(b, a, counter) : launch_group<P, COL_EXT_REG, false>(b, a, counter);
    default: return ERLBM_ERR_UNSUPPORTED;
    }
}

int resolve_precision(const MapJob &j)
{
    if (j.precision == ERLBM_FP_AUTO) return (j.tau - 0.5 < 1e-3) ? ERLBM_FP64 : ERLBM_FP32_SHIFTED;
    return j.precision;
}

}  // namespace

namespace {

void fill_geom(const erlbm_batch *b, TinyGeom &g)
{
    g.nx = b->nx;
    g.ny = b->ny;
    g.n_fluid = b->n_fluid;
    g.nvp = (b->nx * b->ny + 7) / 8 * 8;
    g.general = b->general ? 1 : 0;
    g.inlet_x = b->inlet_x;
    g.outlet_x = b->outlet_x;
    g.outlet_dx = b->outlet_dx;
    g.vox_of_fluid = b->vox_of_fluid;
    g.fluid_of_vox = b->fluid_of_vox;
    g.nbr = b->nbr;
    g.sin_tab = b->tabs;
    g.cos_tab = b->tabs + b->tab_n;
}

__global__ void k_zero_counter(unsigned int *c)
{
    if (threadIdx.x == 0) *c = 0u;
}

// Early-exit compaction between chunks: idx[0..*n) keeps, in order, the jobs
// whose result is not done, and *n becomes their count.  One workgroup walks
// the list in slices of CB; a kept entry moves to a position no later than
// its own, inside the slice every thread has already read.
constexpr int CB = 1024;
__global__ __launch_bounds__(CB) void k_compact(int *idx, int *n, const MapResult *results)
{
    __shared__ int wsum[CB / 64];
    const int n0 = *n;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int base = 0;
    for (int start = 0; start < n0; start += CB) {
        const int k = start + (int)threadIdx.x;
        int j = -1;
        bool keep = false;
        if (k < n0) {
            j = idx[k];
            keep = results[j].done == 0;
        }
        const unsigned long long m = __ballot(keep);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int off = 0, total = 0;
        for (int q = 0; q < CB / 64; q++) {
            if (q < w) off += wsum[q];
            total += wsum[q];
        }
        if (keep) idx[base + off + before] = j;
        base += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *n = base;
}

}  // namespace

namespace erlbm {
namespace tiny {

size_t state_bytes(const erlbm_batch *b, int precision)
{
    const size_t nvp = (size_t)(b->nx * b->ny + 7) / 8 * 8;
    return (size_t)Q * nvp * (precision == ERLBM_FP64 ? sizeof(double) : sizeof(float));
}

int fluid_sites(const erlbm_batch *b) { return b->n_fluid; }

int launch(erlbm_batch *b, int precision, int collision, const MapJob *jobs, MapResult *results, const int *job_idx,
           const int *n_jobs_dev, int n_jobs_max, unsigned int *counter, int max_steps, void *saved,
           hipStream_t stream)
{
    RunArgs a;
    fill_geom(b, a.g);
    a.jobs = jobs;
    a.results = results;
    a.job_idx = job_idx;
    a.n_jobs = n_jobs_max;
    a.n_jobs_dev = n_jobs_dev;
    a.next_job = counter;
    a.max_steps = max_steps;
    a.saved = saved;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, stream, counter);
    const hipStream_t keep = b->stream;
    b->stream = stream;
    int rc = (precision == ERLBM_FP64)            ? launch_col<F64>(b, collision, a, counter)
             : (precision == ERLBM_FP32_SHIFTED) ? launch_col<F32>(b, collision, a, counter)
                                                 : launch_col<F32A>(b, collision, a, counter);
    b->stream = keep;
    return rc;
}

}  // namespace tiny
}  // namespace erlbm

extern "C" int erlbm_batch_create(int nx, int ny, const int *mask, int inlet_x, int outlet_x, void *stream,
                                  erlbm_batch **out)
{
    if (!out || nx < 1 || ny < 1 || inlet_x >= nx || outlet_x >= nx) return ERLBM_ERR_ARG;
    const int nv = nx * ny;
    std::vector<int> fov(nv, -1), vof;
    std::vector<uint32_t> nbr(nv, 0);
    for (int v = 0; v < nv; v++)
        if (!mask || mask[v]) {
            fov[v] = (int)vof.size();
            vof.push_back(v);
        }
    const int nf = (int)vof.size();
    if (nf < 1 || nf > TPB || nv > 400) {
        set_last_error_msg("erlbm_batch_create: a tiny lattice needs 1..256 fluid sites in at most 400 voxels");
        return ERLBM_ERR_ARG;
    }
    for (int v : vof) {
        const int x = v % nx, y = v / nx;
        uint32_t bits = FLUID;
        for (int i = 1; i < Q; i++) {
            const int xs = ((x - VSet<Q>::cx[i]) % nx + nx) % nx;
            const int ys = ((y - VSet<Q>::cy[i]) % ny + ny) % ny;
            if (fov[xs + nx * ys] < 0) bits |= 1u << i;
        }
        nbr[v] = bits;
    }
    // Thread order: the sites of the inlet and outlet columns and those next
    // to a solid last, so that only the last wavefronts hold the lanes that
    // run the boundary conditions and write ghost values (a divergent branch
    // costs every wavefront holding one of its lanes); fov keeps the fluid
    // order, which the ordered sums follow.
    std::stable_partition(vof.begin(), vof.end(), [&](int v) {
        return v % nx != inlet_x && v % nx != outlet_x && (nbr[v] & ~FLUID) == 0;
    });
    auto *b = new erlbm_batch();
    b->nx = nx;
    b->ny = ny;
    b->n_fluid = nf;
    b->inlet_x = inlet_x;
    b->outlet_x = outlet_x;
    b->outlet_dx = (outlet_x >= 0 && outlet_x < nx / 2) ? 1 : -1;
    if (outlet_x >= 0 && (outlet_x + b->outlet_dx < 0 || outlet_x + b->outlet_dx >= nx)) {
        delete b;
        set_last_error_msg("erlbm_batch_create: outlet neighbour outside the lattice");
        return ERLBM_ERR_ARG;
    }
    b->general = (nf != nv) || inlet_x >= 0 || outlet_x >= 0;
    b->stream = (hipStream_t)stream;
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (b->n_cu < 1) b->n_cu = 1;
    b->tab_n = std::max(nx, ny);
    std::vector<double> tabs(2 * b->tab_n);
    // sin / cos (2 pi x / L) of the Taylor-Green field (L = nx; the map is square)
    erlbm_tg_map_tables(nx, tabs.data(), tabs.data() + b->tab_n);
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t r) {
        if (r != hipSuccess && e == hipSuccess) e = r;
    };
    ok(hipMalloc(&b->vox_of_fluid, sizeof(int) * nf));
    ok(hipMalloc(&b->fluid_of_vox, sizeof(int) * nv));
    ok(hipMalloc(&b->nbr, sizeof(uint32_t) * nv));
    ok(hipMalloc(&b->tabs, sizeof(double) * 2 * b->tab_n));
    ok(hipMalloc(&b->counters, sizeof(unsigned int) * 16));
    if (e == hipSuccess) {
        ok(hipMemcpy(b->vox_of_fluid, vof.data(), sizeof(int) * nf, hipMemcpyHostToDevice));
        ok(hipMemcpy(b->fluid_of_vox, fov.data(), sizeof(int) * nv, hipMemcpyHostToDevice));
        ok(hipMemcpy(b->nbr, nbr.data(), sizeof(uint32_t) * nv, hipMemcpyHostToDevice));
        ok(hipMemcpy(b->tabs, tabs.data(), sizeof(double) * 2 * b->tab_n, hipMemcpyHostToDevice));
    }
    if (e != hipSuccess) {
        set_last_error("erlbm_batch_create", e);
        erlbm_batch_destroy(b);
        return ERLBM_ERR_NOMEM;
    }
    *out = b;
    return ERLBM_OK;
}

extern "C" int erlbm_batch_destroy(erlbm_batch *b)
{
    if (!b) return ERLBM_OK;
    (void)hipStreamSynchronize(b->stream);
    (void)hipFree(b->vox_of_fluid);
    (void)hipFree(b->fluid_of_vox);
    (void)hipFree(b->nbr);
    (void)hipFree(b->tabs);
    (void)hipFree(b->counters);
    (void)hipFree(b->d_jobs);
    (void)hipFree(b->d_results);
    (void)hipFree(b->d_idx);
    (void)hipFree(b->d_saved);
    (void)hipFree(b->d_n);
    if (b->cstream) (void)hipStreamDestroy(b->cstream);
    delete b;
    return ERLBM_OK;
}

extern "C" int erlbm_batch_set_launch(erlbm_batch *b, int wg_per_cu_f64, int wg_per_cu_f32)
{
    if (!b || wg_per_cu_f64 < 0 || wg_per_cu_f32 < 0) return ERLBM_ERR_ARG;
    b->wg_per_cu[0] = wg_per_cu_f64;
    b->wg_per_cu[1] = wg_per_cu_f32;
    return ERLBM_OK;
}

extern "C" int erlbm_batch_set_chunk(erlbm_batch *b, int chunk_steps)
{
    if (!b || chunk_steps < 0) return ERLBM_ERR_ARG;
    b->chunk_steps = chunk_steps;
    return ERLBM_OK;
}

namespace {

// Chunked mode of erlbm_batch_run: per (precision, collision) group, one
// captured graph = [queue reset, persistent kernel advancing every live job
// by at most chunk_steps steps (state saved per job), compaction of the live
// list]; the graph is replayed until no job is live.
int run_chunked(erlbm_batch *b, const erlbm_job *jobs, int n_jobs, const std::vector<int> &idx,
                const std::vector<int> &prec)
{
    ERLBM_HIP_TRY(hipStreamSynchronize(b->stream));
    if (!b->cstream) ERLBM_HIP_TRY(hipStreamCreateWithFlags(&b->cstream, hipStreamNonBlocking));
    if (!b->d_n) ERLBM_HIP_TRY(hipMalloc(&b->d_n, sizeof(int) * 16));
    const size_t need = (size_t)n_jobs * tiny::state_bytes(b, ERLBM_FP64);
    if (need > b->saved_cap) {
        (void)hipFree(b->d_saved);
        b->d_saved = nullptr;
        b->saved_cap = 0;
        ERLBM_HIP_TRY(hipMalloc(&b->d_saved, need));
        b->saved_cap = need;
    }
    auto key = [&](int k) { return prec[k] * 8 + jobs[k].collision; };
    int start = 0, group = 0;
    std::vector<int> counts(16, 0);
    while (start < n_jobs) {
        int end = start;
        while (end < n_jobs && key(idx[end]) == key(idx[start])) end++;
        if (group >= 16) return ERLBM_ERR_UNSUPPORTED;
        const int p = prec[idx[start]];
        const int col = jobs[idx[start]].collision;
        counts[group] = end - start;
        int *d_idx = b->d_idx + start;
        int *d_n = b->d_n + group;
        ERLBM_HIP_TRY(hipMemcpyAsync(d_n, &counts[group], sizeof(int), hipMemcpyHostToDevice, b->cstream));
        ERLBM_HIP_TRY(hipStreamSynchronize(b->cstream));
        ERLBM_HIP_TRY(hipStreamBeginCapture(b->cstream, hipStreamCaptureModeThreadLocal));
        int rc = tiny::launch(b, p, col, b->d_jobs, b->d_results, d_idx, d_n, end - start, b->counters + group,
                              b->chunk_steps, b->d_saved, b->cstream);
        hipLaunchKernelGGL(k_compact, dim3(1), dim3(CB), 0, b->cstream, d_idx, d_n, (const MapResult *)b->d_results);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->cstream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) {
            set_last_error("erlbm_batch_run: graph capture", e);
            return ERLBM_ERR_HIP;
        }
        hipGraphExec_t ge = nullptr;
        hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        int live = end - start;
        while (ei == hipSuccess && live > 0) {
            ei = hipGraphLaunch(ge, b->cstream);
            if (ei == hipSuccess) ei = hipMemcpyAsync(&live, d_n, sizeof(int), hipMemcpyDeviceToHost, b->cstream);
            if (ei == hipSuccess) ei = hipStreamSynchronize(b->cstream);
        }
        if (ge) (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            set_last_error("erlbm_batch_run: chunk", ei);
            return ERLBM_ERR_HIP;
        }
        start = end;
        group++;
    }
    return ERLBM_OK;
}

}  // namespace

namespace {

// The checks erlbm_batch_run applies to a job list (ERLBM_OK or the error).
int validate_jobs(const erlbm_batch *b, const erlbm_job *jobs, int n_jobs)
{
    for (int k = 0; k < n_jobs; k++) {
        const MapJob &j = jobs[k];
        if (!(j.tau > 0) || !(j.tau_nh > 0) || j.n_steps < 0 || j.n_steps > 2000000000LL ||
            (j.collision != ERLBM_COL_BGK && j.collision != ERLBM_COL_REG && j.collision != ERLBM_COL_EXT_REG) ||
            (j.init != ERLBM_INIT_REST && j.init != ERLBM_INIT_TAYLOR_GREEN) ||
            (j.precision != ERLBM_FP64 && j.precision != ERLBM_FP32_SHIFTED && j.precision != ERLBM_FP_AUTO &&
             j.precision != ERLBM_FP32_SHIFTED_F32)) {
            set_last_error_msg("erlbm_batch_run: invalid job");
            return ERLBM_ERR_ARG;
        }
        if (j.check_every > 0 && b->outlet_x >= 0) {
            set_last_error_msg("erlbm_batch_run: the energy test is not defined on a lattice with an outlet column");
            return ERLBM_ERR_UNSUPPORTED;
        }
        if (j.init == ERLBM_INIT_TAYLOR_GREEN && b->nx != b->ny) {
            set_last_error_msg("erlbm_batch_run: the Taylor-Green field needs a square lattice");
            return ERLBM_ERR_ARG;
        }
    }
    return ERLBM_OK;
}

}  // namespace

extern "C" int erlbm_batch_run(erlbm_batch *b, const erlbm_job *jobs, int n_jobs, erlbm_job_result *results)
{
    if (!b || (n_jobs > 0 && (!jobs || !results)) || n_jobs < 0) return ERLBM_ERR_ARG;
    if (n_jobs == 0) return ERLBM_OK;
    if (int rc = validate_jobs(b, jobs, n_jobs)) return rc;
    if (n_jobs > b->cap) {
        (void)hipFree(b->d_jobs);
        (void)hipFree(b->d_results);
        (void)hipFree(b->d_idx);
        b->d_jobs = nullptr;
        b->d_results = nullptr;
        b->d_idx = nullptr;
        b->cap = 0;
        ERLBM_HIP_TRY(hipMalloc(&b->d_jobs, sizeof(MapJob) * n_jobs));
        ERLBM_HIP_TRY(hipMalloc(&b->d_results, sizeof(MapResult) * n_jobs));
        ERLBM_HIP_TRY(hipMalloc(&b->d_idx, sizeof(int) * n_jobs));
        b->cap = n_jobs;
    }
    // group jobs by (precision, collision); longest budgets first within a
    // group so the queue tail is made of short runs
    std::vector<int> idx(n_jobs);
    std::vector<int> prec(n_jobs);
    for (int k = 0; k < n_jobs; k++) {
        idx[k] = k;
        prec[k] = resolve_precision(jobs[k]);
    }
    auto key = [&](int k) { return prec[k] * 8 + jobs[k].collision; };
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
        if (key(x) != key(y)) return key(x) < key(y);
        return jobs[x].n_steps > jobs[y].n_steps;
    });
    ERLBM_HIP_TRY(hipMemcpyAsync(b->d_jobs, jobs, sizeof(MapJob) * n_jobs, hipMemcpyHostToDevice, b->stream));
    ERLBM_HIP_TRY(hipMemcpyAsync(b->d_idx, idx.data(), sizeof(int) * n_jobs, hipMemcpyHostToDevice, b->stream));
    ERLBM_HIP_TRY(hipMemsetAsync(b->counters, 0, sizeof(unsigned int) * 16, b->stream));
    ERLBM_HIP_TRY(hipMemsetAsync(b->d_results, 0, sizeof(MapResult) * n_jobs, b->stream));
    if (b->chunk_steps > 0) {
        int rc = run_chunked(b, jobs, n_jobs, idx, prec);
        if (rc) return rc;
    } else {
        RunArgs a;
        fill_geom(b, a.g);
        a.jobs = b->d_jobs;
        a.results = b->d_results;
        a.n_jobs_dev = nullptr;
        a.max_steps = 0;
        a.saved = nullptr;
        int start = 0, group = 0;
        while (start < n_jobs) {
            int end = start;
            while (end < n_jobs && key(idx[end]) == key(idx[start])) end++;
            if (group >= 16) return ERLBM_ERR_UNSUPPORTED;
            a.job_idx = b->d_idx + start;
            a.n_jobs = end - start;
            const int p = prec[idx[start]];
            const int col = jobs[idx[start]].collision;
            int rc = (p == ERLBM_FP64)            ? launch_col<F64>(b, col, a, b->counters + group)
                     : (p == ERLBM_FP32_SHIFTED) ? launch_col<F32>(b, col, a, b->counters + group)
                                                 : launch_col<F32A>(b, col, a, b->counters + group);
            if (rc) return rc;
            start = end;
            group++;
        }
    }
    ERLBM_HIP_TRY(hipMemcpyAsync(results, b->d_results, sizeof(MapResult) * n_jobs, hipMemcpyDeviceToHost,
                                 b->stream));
    ERLBM_HIP_TRY(hipStreamSynchronize(b->stream));
    return ERLBM_OK;
}

// ---- batch run without allocation ------------------------------------------------
// erlbm_batch_run's device-side job records, results, queue order and queue
// counters in one static device array (`scratch`) instead of buffers the batch
// allocates: NOALLOC_MAX_JOBS jobs at most.  Host code uses the array through
// the address hipGetSymbolAddress gives for the current device (each GPU holds
// its own copy), never through the symbol's name.
namespace {

constexpr int NOALLOC_MAX_JOBS = ERLBM_NOALLOC_MAX_JOBS;
constexpr int NOALLOC_GROUPS = 16;   // (precision, collision) groups of one call
constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t SCR_JOBS = 0;
constexpr size_t SCR_RESULTS = align256(SCR_JOBS + sizeof(MapJob) * NOALLOC_MAX_JOBS);
constexpr size_t SCR_IDX = align256(SCR_RESULTS + sizeof(MapResult) * NOALLOC_MAX_JOBS);
constexpr size_t SCR_COUNTERS = align256(SCR_IDX + sizeof(int) * NOALLOC_MAX_JOBS);
constexpr size_t SCRATCH_BYTES = align256(SCR_COUNTERS + sizeof(unsigned int) * NOALLOC_GROUPS);

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, taken on the first call
// per device.
int scratch_address(unsigned char **out)
{
    int dev = 0;
    ERLBM_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_last_error_msg("erlbm_batch_run_noalloc: device index out of range");
        return ERLBM_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_last_error("erlbm_batch_run_noalloc: hipGetSymbolAddress(scratch)", e);
            return ERLBM_ERR_HIP;
        }
        g_scratch_dev[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_dev[dev];
    return ERLBM_OK;
}

// Clears the results of n jobs and the queue counters (a kernel, so that a
// caller may capture the call in a graph).
__global__ void k_clear_noalloc(MapResult *results, int n, unsigned int *counters)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) results[i] = MapResult{};
    if (i < NOALLOC_GROUPS) counters[i] = 0u;
}

}  // namespace

extern "C" int erlbm_batch_run_noalloc(erlbm_batch *b, const erlbm_job *jobs, int n_jobs,
                                       erlbm_job_result *results)
{
    if (!b || (n_jobs > 0 && (!jobs || !results)) || n_jobs < 0) return ERLBM_ERR_ARG;
    if (n_jobs > NOALLOC_MAX_JOBS) {
        set_last_error_msg("erlbm_batch_run_noalloc: more than ERLBM_NOALLOC_MAX_JOBS (16384) jobs");
        return ERLBM_ERR_ARG;
    }
    if (n_jobs == 0) return ERLBM_OK;
    if (int rc = validate_jobs(b, jobs, n_jobs)) return rc;
    unsigned char *base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    MapJob *d_jobs = reinterpret_cast<MapJob *>(base + SCR_JOBS);
    MapResult *d_results = reinterpret_cast<MapResult *>(base + SCR_RESULTS);
    int *d_idx = reinterpret_cast<int *>(base + SCR_IDX);
    unsigned int *d_counters = reinterpret_cast<unsigned int *>(base + SCR_COUNTERS);
    // the queue order of erlbm_batch_run -- grouped by (precision,
    // collision), longest budgets first, ties by job index -- formed in a
    // stack array (no heap allocation)
    int idx[NOALLOC_MAX_JOBS];
    auto key = [&](int k) { return resolve_precision(jobs[k]) * 8 + jobs[k].collision; };
    for (int k = 0; k < n_jobs; k++) idx[k] = k;
    std::sort(idx, idx + n_jobs, [&](int x, int y) {
        if (key(x) != key(y)) return key(x) < key(y);
        if (jobs[x].n_steps != jobs[y].n_steps) return jobs[x].n_steps > jobs[y].n_steps;
        return x < y;
    });
    ERLBM_HIP_TRY(hipMemcpyAsync(d_jobs, jobs, sizeof(MapJob) * n_jobs, hipMemcpyHostToDevice, b->stream));
    ERLBM_HIP_TRY(hipMemcpyAsync(d_idx, idx, sizeof(int) * n_jobs, hipMemcpyHostToDevice, b->stream));
    const int clear = std::max(n_jobs, NOALLOC_GROUPS);
    hipLaunchKernelGGL(k_clear_noalloc, dim3((clear + 255) / 256), dim3(256), 0, b->stream, d_results, n_jobs,
                       d_counters);
    ERLBM_HIP_TRY(hipGetLastError());
    RunArgs a;
    fill_geom(b, a.g);
    a.jobs = d_jobs;
    a.results = d_results;
    a.n_jobs_dev = nullptr;
    a.next_job = nullptr;
    a.max_steps = 0;
    a.saved = nullptr;
    int start = 0, group = 0;
    while (start < n_jobs) {
        int end = start;
        while (end < n_jobs && key(idx[end]) == key(idx[start])) end++;
        if (group >= NOALLOC_GROUPS) return ERLBM_ERR_UNSUPPORTED;
        a.job_idx = d_idx + start;
        a.n_jobs = end - start;
        const int p = resolve_precision(jobs[idx[start]]);
        const int col = jobs[idx[start]].collision;
        int rc = (p == ERLBM_FP64)            ? launch_col<F64>(b, col, a, d_counters + group)
                 : (p == ERLBM_FP32_SHIFTED) ? launch_col<F32>(b, col, a, d_counters + group)
                                             : launch_col<F32A>(b, col, a, d_counters + group);
        if (rc) return rc;
        start = end;
        group++;
    }
    ERLBM_HIP_TRY(hipMemcpyAsync(results, d_results, sizeof(MapResult) * n_jobs, hipMemcpyDeviceToHost, b->stream));
    ERLBM_HIP_TRY(hipStreamSynchronize(b->stream));
    return ERLBM_OK;
}

// ---- autotuner --------------------------------------------------------------------
namespace {

template <class P, int COL>
int occupancy_of(const erlbm_batch *b)
{
    int per_cu = 0;
    const size_t nvp = (size_t)(b->nx * b->ny + 7) / 8 * 8;
    const size_t lds = sizeof(typename P::T) * Q * nvp;
    const bool tg = !b->general && b->nx == 16 && b->ny == 16, ch = b->general && b->nx == 22 && b->ny == 12;
    hipError_t e = tg   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tiny<P, COL, false, 16, 16>, TPB, lds)
                   : ch ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tiny<P, COL, true, 22, 12>, TPB, lds)
                   : b->general ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tiny<P, COL, true>, TPB, lds)
                                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tiny<P, COL, false>, TPB, lds);
    return e == hipSuccess && per_cu > 0 ? per_cu : 1;
}

template <class P>
int occupancy_col(const erlbm_batch *b, int col)
{
    switch (col) {
    case ERLBM_COL_BGK: return occupancy_of<P, COL_BGK>(b);
    case ERLBM_COL_REG: return occupancy_of<P, COL_REG>(b);
    default: return occupancy_of<P, COL_EXT_REG>(b);
    }
}

int occupancy_max(const erlbm_batch *b, int precision, int col)
{
    if (precision == ERLBM_FP64) return occupancy_col<F64>(b, col);
    if (precision == ERLBM_FP32_SHIFTED) return occupancy_col<F32>(b, col);
    return occupancy_col<F32A>(b, col);
}

}  // namespace

extern "C" int erlbm_batch_autotune(erlbm_batch *b, const erlbm_job *sample, int n_sample, long long max_steps,
                                    int *wg_per_cu_f64, int *wg_per_cu_f32, double *points_per_s)
{
    if (!b || !sample || n_sample < 1 || max_steps < 1) return ERLBM_ERR_ARG;
    // Every job of the sample gets the same budget (max_steps, or less if
    // its own is shorter), so the time measures the per-step throughput of
    // a launch shape rather than the sample's longest job.
    std::vector<erlbm_job> cls[2];
    for (int k = 0; k < n_sample; k++) {
        erlbm_job j = sample[k];
        if (j.n_steps > max_steps) j.n_steps = max_steps;
        cls[resolve_precision(j) == ERLBM_FP64 ? 0 : 1].push_back(j);
    }
    std::vector<MapResult> res(n_sample);
    int rc = ERLBM_OK;
    auto timed = [&](std::vector<erlbm_job> &v) -> double {
        const auto t0 = std::chrono::steady_clock::now();
        rc = erlbm_batch_run(b, v.data(), (int)v.size(), res.data());
        const auto t1 = std::chrono::steady_clock::now();
        return rc ? -1.0 : std::chrono::duration<double>(t1 - t0).count();
    };
    const int chunk0 = b->chunk_steps;
    b->chunk_steps = 0;
    double total = 0.0;
    // workgroups per CU, per precision class: 1 .. the occupancy limit
    for (int c = 0; c < 2 && rc == ERLBM_OK; c++) {
        if (cls[c].empty()) continue;
        const int maxpc = occupancy_max(b, resolve_precision(cls[c][0]), cls[c][0].collision);
        if (timed(cls[c]) < 0) break;   // warm-up (buffer growth)
        double best = 1e300;
        int best_v = 0;
        for (int v = 1; v <= maxpc; v++) {
            b->wg_per_cu[c] = v;
            const double t = std::min(timed(cls[c]), timed(cls[c]));
            if (rc) break;
            if (t < best) {
                best = t;
                best_v = v;
            }
        }
        b->wg_per_cu[c] = best_v;
        total += best;
    }
    b->chunk_steps = chunk0;
    if (rc) return rc;
    if (wg_per_cu_f64) *wg_per_cu_f64 = b->wg_per_cu[0];
    if (wg_per_cu_f32) *wg_per_cu_f32 = b->wg_per_cu[1];
    if (points_per_s) *points_per_s = total > 0 ? n_sample / total : 0.0;
    return ERLBM_OK;
}

// ---- job lists of the two published maps ----------------------------------------
extern "C" int erlbm_tg_map_jobs(int n_ma, int n_nu, int collision, double tau_nh, int precision, int i0, int i1,
                                 int j0, int j1, erlbm_job *jobs)
{
    if (!jobs || i0 < 0 || j0 < 0 || i1 > n_ma || j1 > n_nu || i0 > i1 || j0 > j1) return ERLBM_ERR_ARG;
    int k = 0;
    for (int j = j0; j < j1; j++)
        for (int i = i0; i < i1; i++) {
            erlbm_job J;
            int n = 0;
            double u0, tau, tnh;
            int rc = erlbm_tg_map_point_params(i, j, n_ma, n_nu, 16, collision, tau_nh, &u0, &tau, &tnh, &n);
            if (rc) return rc;
            J.tau = tau;
            J.tau_nh = tnh;
            J.u0 = u0;
            J.rho0 = 1.0;
            J.energy_factor = 2.0;
            J.dens_limit = 0.0;
            J.n_steps = n;
            J.check_every = 100;
            J.init = ERLBM_INIT_TAYLOR_GREEN;
            J.collision = collision;
            J.precision = precision;
            jobs[k++] = J;
        }
    return ERLBM_OK;
}

extern "C" int erlbm_obstacle_map_jobs(int n_tau, int n_nu, double Ma, int precision, int i0, int i1, int j0,
                                       int j1, erlbm_job *jobs)
{
    if (!jobs || i0 < 0 || j0 < 0 || i1 > n_tau || j1 > n_nu || i0 > i1 || j0 > j1) return ERLBM_ERR_ARG;
    int k = 0;
    for (int j = j0; j < j1; j++)
        for (int i = i0; i < i1; i++) {
            erlbm_job J;
            int n = 0;
            double u, tau, tnh;
            int rc = erlbm_obstacle_map_point_params(i, j, n_tau, n_nu, Ma, &u, &tau, &tnh, &n);
            if (rc) return rc;
            J.tau = tau;
            J.tau_nh = tnh;
            J.u0 = u;
            J.rho0 = 1.0;
            J.energy_factor = 0.0;
            J.dens_limit = 10. * 1.0;
            J.n_steps = n;
            J.check_every = 0;
            J.init = ERLBM_INIT_REST;
            J.collision = ERLBM_COL_EXT_REG;
            J.precision = precision;
            jobs[k++] = J;
        }
    return ERLBM_OK;
}
