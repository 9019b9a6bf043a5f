// lattice_impl.hpp -- definitions shared by the plain lattice path's
// translation units (FP64 kernels in lattice.hip, shifted low-precision
// kernels in lattice_lowp.hip): launch geometry, boundary-column helpers and
// the lattice handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "collide_moment.hpp"
#include "common.hpp"

namespace erlbm {
namespace lat {

constexpr int BLOCK = 256;
constexpr uint32_t FLUID_BIT = 1u << 31;

struct Geometry {
    int nx, ny, nz;
    long long si;   // population stride
    long long sr;   // row stride
};

__host__ __device__ inline long long at(const Geometry &g, int i, long long row, int x)
{
    return i * g.si + row * g.sr + x;
}

struct BCArgs {
    int inlet_x;    // < 0: none
    int inlet_west; // pos_x < nx/2 branch of ZouHeD3Q19_vx_x
    double u_in;
    int outlet_x;   // < 0: none
    int outlet_dx;  // +1 west half, -1 east half (devnull_rho_x :90-92)
    double rho_out;
};

struct StepArgs {
    const double *__restrict__ src;
    double *__restrict__ dst;
    Geometry g;
    const uint32_t *__restrict__ nbr;  // general kernel only
    CollideConsts k;
    BodyForce bf;                      // forced lattices (general kernel) only
    BCArgs bc;
    double dens_limit;                 // <= 0: no test
    unsigned long long *status;        // written (atomics) by this launch
    const unsigned long long *__restrict__ stop_ro;  // the same word, read as of the launch
    unsigned long long step;
};

__device__ inline bool stopped(const unsigned long long *status, unsigned long long step)
{
    const unsigned long long s =
        __hip_atomic_load(&status[ST_STOP_STEP], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return step >= s;
}

__device__ inline void flag_density(unsigned long long *status, unsigned long long step, bool bad)
{
    // one atomic per wavefront
    if (__any(bad)) {
        if ((threadIdx.x & 63) == 0) {
            atomicMin(&status[ST_STOP_STEP], step + 1);
            atomicOr(&status[ST_REASON], (unsigned long long)ERLBM_STOP_DENSITY);
        }
    }
}

__device__ inline int wrap(int a, int n)
{
    return a < 0 ? a + n : (a >= n ? a - n : a);
}

// ZouHeD3Q19_vx_x (Boundary_conditions.cpp:350-442) on one site, in registers.
__device__ inline void zouhe_d3q19(double (&f)[19], double vx, bool west)
{
    if (west) {
        double sum_f_xminus = f[1] + f[7] + f[9] + f[11] + f[13];
        double sum_f_xzero = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18];
        double rho = (2.0 * sum_f_xminus + sum_f_xzero) / (1.0 - vx);
        double vy = 0.0, vz = 0.0;
        f[1] = f[1] + rho * vx / 3.0;
        f[7] = f[7] + rho * vx / 6.0 + 0.5 * rho * vy - 0.5 * (f[18] + f[3] + f[16] - f[15] - f[4] - f[17]);
        f[9] = f[9] + rho * vx / 6.0 - 0.5 * rho * vy + 0.5 * (f[18] + f[3] + f[16] - f[15] - f[4] - f[17]);
        f[11] = f[11] + rho * vx / 6.0 + 0.5 * rho * vz - 0.5 * (f[17] + f[5] + f[16] - f[15] - f[6] - f[18]);
        f[13] = f[13] + rho * vx / 6.0 - 0.5 * rho * vz + 0.5 * (f[17] + f[5] + f[16] - f[15] - f[6] - f[18]);
    } else {
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
}

// ZouHeD2Q9_vx_x (Boundary_conditions.cpp:710-735) on one site.
__device__ inline void zouhe_d2q9(double (&f)[9], double vx)
{
    double rho = (f[0] + f[3] + f[4] + 2 * (f[2] + f[8] + f[6])) / (1. - vx);
    f[1] = (rho / 9.0) * (1.0 + 3.0 * vx + 3.0 * vx * vx);
    f[5] = (rho / 36.0) * (1.0 + 3.0 * vx + 3.0 * vx * vx);
    f[7] = (rho / 36.0) * (1.0 + 3.0 * vx + 3.0 * vx * vx);
}

template <int Q>
__device__ inline void zouhe(double (&f)[Q], double vx, bool west)
{
    if constexpr (Q == 19) zouhe_d3q19(f, vx, west);
    else zouhe_d2q9(f, vx);
}


}  // namespace lat
}  // namespace erlbm

// Multi-step shapes (ERLBM_FUSE values) that the FP32 shifted-storage path
// also has.
constexpr int FUSE2_STRIP_256x32 = 11, FUSE2_STRIP_128x32 = 13;
constexpr int FUSE3_STRIP_256x32 = 22, FUSE3_STRIP_128x32 = 24, FUSE3_STRIP_SIZED = 26;
// D3Q19 two-step z-walks (x-y tiles)
constexpr int FUSE2_ZWALK_32x8 = 30, FUSE2_ZWALK_64x4 = 31, FUSE2_ZWALK_32x4 = 32, FUSE2_ZWALK_64x8 = 33,
              FUSE2_ZWALK_128x4 = 34, FUSE2_ZWALK_32x16 = 35;
// D2Q9 three-step walks on strips of any width, 1 .. 4 wavefronts per workgroup
constexpr int FUSE3_GSTRIP_1W = 36, FUSE3_GSTRIP_2W = 37, FUSE3_GSTRIP_3W = 38, FUSE3_GSTRIP_4W = 39;
// the same with the lanes aligned across the walk's phases, 2, 3, 4, 8 wavefronts
constexpr int FUSE3_ASTRIP_2W = 40, FUSE3_ASTRIP_3W = 41, FUSE3_ASTRIP_4W = 42, FUSE3_ASTRIP_8W = 43,
              FUSE3_ASTRIP_5W = 44, FUSE3_ASTRIP_6W = 45;
// D2Q9 four, five and six steps per launch, lane-aligned strips of any width, 4 wavefronts
constexpr int FUSE4_ASTRIP_4W = 46, FUSE5_ASTRIP_4W = 47, FUSE6_ASTRIP_4W = 48;

struct erlbm_lattice {
    int q = 0, nx = 0, ny = 0, nz = 0;
    long long nvox = 0, si = 0, sr = 0, extent = 0;
    int layout = 1;
    int n_fluid = 0;
    bool all_fluid = true;
    int precision = ERLBM_FP64;
    hipStream_t stream = nullptr;
    int device = 0;

    void *buf[2] = {nullptr, nullptr};   // population buffers, elem bytes per value
    int elem = 8;                        // 8 FP64, 4 FP32-shifted, 2 FP16-shifted
    bool f32_arith = false;              // shifted storage with FP32 arithmetic
    double shift = 1.0;                  // rho0 of the shifted storage (set at upload)
    int base = 0;             // buffer holding the uploaded state
    bool uploaded = false;
    long long launched = 0;   // steps enqueued since upload

    uint32_t *nbr = nullptr;          // per voxel (always allocated)
    int *vox_of_fluid = nullptr;      // null when all fluid
    double *ke_node = nullptr, *partial_ke = nullptr, *partial_rho = nullptr;
    double *scalars = nullptr;        // [0] E, [1] max|rho|, [2] E0
    int *nonfinite = nullptr;
    unsigned long long *status = nullptr;

    double tau = 1.0, tau_nh = 1.0, rho0 = 1.0;
    int col = ERLBM_COL_EXT_REG;
    int inlet_x = -1;
    double u_in = 0.0;
    int outlet_x = -1;
    double rho_out = 1.0;
    double dens_limit = 0.0;
    bool e0_valid = false;
    // body force (erlbm_set_body_force): forced lattices step on the general
    // one-step kernel
    bool forced = false;
    double acc[3] = {0.0, 0.0, 0.0};
    // collision arithmetic of the multi-step kernels (erlbm_set_arith)
    int arith = ERLBM_ARITH_EXACT;

    // launch tuning (environment overrides read at creation):
    //   ERLBM_PLANE_PAD  doubles added to each population plane's stride
    //   ERLBM_NPT        nodes per thread of the periodic D2Q9 kernel (1, 2, 4)
    //   ERLBM_NT_LOADS   1: nontemporal loads in that kernel
    //   ERLBM_XCD_REMAP  1: renumber blocks so each XCD walks a band of rows
    //   ERLBM_PAIR       1: D2Q9 kernel with two adjacent nodes per thread
    long long tune_pad = 0;
    //   ERLBM_BLOCK      threads per block of that kernel (128, 256, 512)
    //   ERLBM_VEC        vectorised 16-byte kernel for periodic lattices (lattice_vec.hip):
    //                    -1 auto, 0 off, 1 edge elements by wavefront shuffle, 2 edge elements loaded
    int tune_npt = 1, tune_ntl = 2, tune_remap = 0, tune_pair = 0, tune_block = 256, tune_vec = -1;
    //   ERLBM_FUSE       several steps per launch for periodic D2Q9 runs without stop
    //                    tests: 0 off; two steps: 1..10 tile shape (64x8, 32x16, 32x8,
    //                    128x4, 16x16, 32x32, 64x4, 128x8, 256x4, 64x16), 11..15 and
    //                    19..21 strip walks (256x32, 256x16, 128x32, 256x64, 128x64;
    //                    256x8, 128x16, 128x8), 16..18 wavefront strips; three steps:
    //                    22..25 strip walks (256x32, 256x16, 128x32, 256x64), 26..29 with
    //                    segments sized to the resident workgroups (256- and 128-wide,
    //                    lean rings; 256x32 lean; 256-wide 18-row rings); D3Q19 (two steps,
    //                    x-y tiles walked along z): 30..35 (32x8, 64x4, 32x4, 64x8, 128x4, 32x16);
    //                    D2Q9 three steps on strips of any width: 36..39 (1..4 wavefronts)
    int tune_fuse = 26;
    int tune_fuse_remap = 1;  //   ERLBM_FUSE_REMAP 1: XCD-contiguous runs of tiles (0: dispatch order)
    int tune_fuse_slots = 0;  //   ERLBM_FUSE_SLOTS workgroups per CU the walks' segments are sized for (0: all that fit)
    int tune_fuse_prio = -1;  //   ERLBM_FUSE_PRIO issue priority by progress in the aligned walks (-1 auto, 0 off,
                              //                   1 equal quarters, 2 bands ending at 50 / 75 / 90%)
    int tune_zwalk_nt = 0;    //   ERLBM_ZWALK_NT   nontemporal step-k loads of the 64x8 z-walk: 0 none (default),
                              //                    1 c_x = 0, 2 c_x = c_y = 0
    int tune_walk_nt = 0;     //   ERLBM_WALK_NT    nontemporal step-k loads of the c_x = 0 populations in the
                              //                    four-wavefront lane-aligned walks, FP64 and FP32 (1), or none (0, default)
    bool checked = false;     // a stop test ran and no later status read cleared it (no fused steps then)
    bool fuse_requested = false;   // ERLBM_FUSE set explicitly
    int fuse_active = 0;      // multi-step shape of the current erlbm_step call (0: none)
    int n_cu = 256;           // compute units of the device
};


namespace erlbm {
namespace lat {

// Constants of the moment-space collision (collide_moment.hpp) for a
// velocity set, operator and relaxation times: (W, K0, K1) = (1 - 1/tau_nh,
// 1/tau_nh, 1/tau_nh - 1/tau) for E-RLBM, (0, 1, 1 - 1/tau) for the
// regularized operator, (1 - 1/tau, 1/tau, 0) for BGK; T = 1/tau.
inline MomentConsts moment_consts(int q, int col, double tau, double tau_nh)
{
    double W, K0, K1;
    if (col == ERLBM_COL_BGK) {
        W = 1.0 - 1.0 / tau;
        K0 = 1.0 / tau;
        K1 = 0.0;
    } else if (col == ERLBM_COL_REG) {
        W = 0.0;
        K0 = 1.0;
        K1 = 1.0 - 1.0 / tau;
    } else {
        W = 1.0 - 1.0 / tau_nh;
        K0 = 1.0 / tau_nh;
        K1 = 1.0 / tau_nh - 1.0 / tau;
    }
    const double T = 1.0 / tau;
    const double D = q == 9 ? 2.0 : 3.0;
    const double wa = q == 9 ? 1.0 / 9.0 : 1.0 / 18.0, wd = 1.0 / 36.0;
    MomentConsts m;
    m.W = W;
    m.cr = K0 + 0.5 * D * K1;
    m.t15 = 1.5 * T;
    m.k15 = 1.5 * K1;
    m.t45 = 4.5 * T;
    m.k45 = 4.5 * K1;
    m.b_axis = (3.0 * wa) * K0;
    m.b_diag = (3.0 * wd) * K0;
    return m;
}

inline int mon_blocks(const erlbm_lattice *h) { return (h->n_fluid + BLOCK - 1) / BLOCK; }

inline Geometry geom(const erlbm_lattice *h) { return Geometry{h->nx, h->ny, h->nz, h->si, h->sr}; }

// Shifted low-precision storage (lattice_lowp.hip).  `src`/`dst` are the
// handle's population buffers of element size h->elem.
int lowp_step(erlbm_lattice *h, const StepArgs &a, bool stream);
// Two steps per launch for periodic D2Q9 shifted storage (FP64 arithmetic),
// strip walk of h->tune_fuse (13: 128-wide, otherwise 256-wide), 32 rows.
int lowp_step2(erlbm_lattice *h, const StepArgs &a);
// AoS FP64 (device, reference fluid order) <-> shifted storage.
int lowp_from_aos(erlbm_lattice *h, const double *aos, void *dst);
int lowp_to_aos(erlbm_lattice *h, const void *src, bool stream, double *aos);
// Per-site 0.5 |u|^2 and |rho| partials (FP64) of the state in `src`.
int lowp_monitor(erlbm_lattice *h, const void *src, bool stream, bool ordered);

// Device AoS copy of the populations `which` names (lattice.hip); *owned
// (to be hipFree'd) is set when the copy was allocated for the call.
int to_aos_device(erlbm_lattice *h, int which, double **aos, void **owned);

// Vectorised periodic kernel (lattice_vec.hip), any storage width.
bool vec_step_applies(const erlbm_lattice *h);
int vec_step(erlbm_lattice *h, const StepArgs &a, bool stream);

}  // namespace lat
}  // namespace erlbm
