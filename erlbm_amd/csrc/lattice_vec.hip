// lattice_vec.hip -- vectorised collide-stream kernel for periodic all-fluid
// lattices, any storage width.
//
// Each thread owns V x-consecutive nodes of one row, V * sizeof(T) = 16 B, so
// every population row segment is read with one aligned 16-byte load per
// thread and written with one aligned 16-byte store.  The x-shift of the
// pull (c_x = +-1) is done in registers: a lane takes the missing neighbour
// from the adjacent lane with a wavefront shuffle, and only the two lanes at
// the wave's edges load that element from memory (with the periodic wrap).
// y / z shifts are whole-row offsets.  Collision in FP64 (bitwise that of
// collide.hpp for T = double; on f = d + w rho0 for shifted storage).
#include <hip/hip_runtime.h>

#include "lattice_impl.hpp"

namespace erlbm {
namespace lat {

namespace {

template <typename T, int V>
using VecT = T __attribute__((ext_vector_type(V)));

template <typename T>
__device__ __forceinline__ T shfl_up1(T x)
{
    if constexpr (sizeof(T) == 8) {
        return __shfl_up(x, 1, 64);
    } else if constexpr (sizeof(T) == 4) {
        return __shfl_up(x, 1, 64);
    } else {
        const unsigned short b = __builtin_bit_cast(unsigned short, x);
        const int r = __shfl_up((int)b, 1, 64);
        return __builtin_bit_cast(T, (unsigned short)r);
    }
}

template <typename T>
__device__ __forceinline__ T shfl_down1(T x)
{
    if constexpr (sizeof(T) == 8) {
        return __shfl_down(x, 1, 64);
    } else if constexpr (sizeof(T) == 4) {
        return __shfl_down(x, 1, 64);
    } else {
        const unsigned short b = __builtin_bit_cast(unsigned short, x);
        const int r = __shfl_down((int)b, 1, 64);
        return __builtin_bit_cast(T, (unsigned short)r);
    }
}

struct VecArgs {
    const void *src;
    void *dst;
    Geometry g;
    CollideConsts k;
    double shift;        // rho0 of shifted storage (unused for T = double)
    double dens_limit;
    unsigned long long *status;
    const unsigned long long *__restrict__ stop_ro;
    unsigned long long step;
};

template <int Q, int COL, typename T, int V, bool STREAM, bool SHUF>
__global__ __launch_bounds__(BLOCK) void k_vec(VecArgs a)
{
    using Vv = VSet<Q>;
    const T *__restrict__ src = static_cast<const T *>(a.src);
    T *__restrict__ dst = static_cast<T *>(a.dst);
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const int row = blockIdx.y;
    const int y = row % ny, z = row / ny;
    const int x0 = (blockIdx.x * BLOCK + (int)threadIdx.x) * V;
    const int lane = threadIdx.x & 63;
    const bool in = x0 < nx;   // nx % V == 0, so a lane is all in or all out
    const int xl = x0 == 0 ? nx - 1 : x0 - 1;      // left neighbour column (wrapped)
    const int xr = x0 + V >= nx ? 0 : x0 + V;      // right neighbour column (wrapped)

    // 1. all row-segment loads, then the single elements across the
    //    segment edges (SHUF: only the wave's edge lanes load them, the
    //    others take them from the adjacent lane below)
    VecT<T, V> raw[Q];
    T edge[Q];
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        int ys = y, zs = z;
        if constexpr (STREAM && Vv::cy[i] != 0) ys = wrap(y - Vv::cy[i], ny);
        if constexpr (STREAM && Vv::cz[i] != 0) zs = wrap(z - Vv::cz[i], nz);
        const T *rowp = src + at(a.g, i, ys + (long long)ny * zs, 0);
        raw[i] = in ? __builtin_nontemporal_load(reinterpret_cast<const VecT<T, V> *>(rowp + x0)) : VecT<T, V>(0);
        edge[i] = T(0);
        if constexpr (STREAM && Vv::cx[i] != 0) {
            const int xe = Vv::cx[i] > 0 ? xl : xr;
            const bool need = SHUF ? (Vv::cx[i] > 0 ? lane == 0 : (lane == 63 || x0 + V >= nx)) : true;
            if (in && need) edge[i] = rowp[xe];
        }
    });
    // 2. x-shift in registers
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        if constexpr (STREAM && Vv::cx[i] != 0) {
            T e = edge[i];
            if constexpr (SHUF) {
                const T nb = Vv::cx[i] > 0 ? shfl_up1((T)raw[i][V - 1]) : shfl_down1((T)raw[i][0]);
                const bool own = Vv::cx[i] > 0 ? lane == 0 : (lane == 63 || x0 + V >= nx);
                e = own ? e : nb;
            }
            if constexpr (Vv::cx[i] > 0) {
#pragma unroll
                for (int k = V - 1; k > 0; k--) raw[i][k] = raw[i][k - 1];
                raw[i][0] = e;
            } else {
#pragma unroll
                for (int k = 0; k < V - 1; k++) raw[i][k] = raw[i][k + 1];
                raw[i][V - 1] = e;
            }
        }
    });
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    bool bad = false;
    if (in) {
#pragma unroll
        for (int k = 0; k < V; k++) {
            double f[Q];
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                if constexpr (std::is_same<T, double>::value) f[i] = raw[i][k];
                else f[i] = (double)(float)raw[i][k] + Vv::w[i] * a.shift;
            });
            double vx, vy, vz;
            collide_exact<Q, COL>(f, a.k, vx, vy, vz);
            if (a.dens_limit > 0.0) {
                const double d = density_exact<Q>(f);
                bad = bad || !(fabs(d) <= a.dens_limit);
            }
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                if constexpr (std::is_same<T, double>::value) raw[i][k] = f[i];
                else raw[i][k] = (T)(float)(f[i] - Vv::w[i] * a.shift);
            });
        }
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            __builtin_nontemporal_store(raw[i], reinterpret_cast<VecT<T, V> *>(dst + at(a.g, i, row, x0)));
        });
    }
    flag_density(a.status, a.step, bad);
}

template <int Q, int COL, typename T>
void launch_qct(erlbm_lattice *h, const VecArgs &v, bool stream)
{
    constexpr int V = 16 / (int)sizeof(T);
    const dim3 grid((h->nx / V + BLOCK - 1) / BLOCK, h->ny * h->nz);
    if (h->tune_vec == 2) {   // edge elements loaded by every lane
        if (stream) hipLaunchKernelGGL((k_vec<Q, COL, T, V, true, false>), grid, dim3(BLOCK), 0, h->stream, v);
        else hipLaunchKernelGGL((k_vec<Q, COL, T, V, false, false>), grid, dim3(BLOCK), 0, h->stream, v);
    } else {
        if (stream) hipLaunchKernelGGL((k_vec<Q, COL, T, V, true, true>), grid, dim3(BLOCK), 0, h->stream, v);
        else hipLaunchKernelGGL((k_vec<Q, COL, T, V, false, true>), grid, dim3(BLOCK), 0, h->stream, v);
    }
}

template <int Q, typename T>
int launch_qt(erlbm_lattice *h, const VecArgs &v, bool stream)
{
    switch (h->col) {
    case ERLBM_COL_BGK: launch_qct<Q, COL_BGK, T>(h, v, stream); break;
    case ERLBM_COL_REG: launch_qct<Q, COL_REG, T>(h, v, stream); break;
    case ERLBM_COL_EXT_REG: launch_qct<Q, COL_EXT_REG, T>(h, v, stream); break;
    case ERLBM_COL_MRT:
        if constexpr (Q == 9) {
            launch_qct<Q, COL_MRT, T>(h, v, stream);
            break;
        }
        return ERLBM_ERR_UNSUPPORTED;
    default: return ERLBM_ERR_UNSUPPORTED;
    }
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

}  // namespace

bool vec_step_applies(const erlbm_lattice *h)
{
    if (h->tune_vec == 0 || h->f32_arith) return false;
    // automatic choice (-1): measured on MI355X, the 16-byte kernel pays off
    // for FP16 storage only (FP64 and FP32 are as fast or faster per node
    // with one element per thread)
    if (h->tune_vec < 0 && h->elem != 2) return false;
    if (!(h->all_fluid && !h->forced && h->inlet_x < 0 && h->outlet_x < 0)) return false;
    const int V = 16 / h->elem;
    // whole 16-byte vectors per row and 16-byte aligned row starts
    return (h->nx % V) == 0 && (h->si % V) == 0 && (h->sr % V) == 0;
}

int vec_step(erlbm_lattice *h, const StepArgs &a, bool stream)
{
    VecArgs v;
    v.src = a.src;
    v.dst = a.dst;
    v.g = a.g;
    v.k = a.k;
    v.shift = h->shift;
    v.dens_limit = a.dens_limit;
    v.status = a.status;
    v.stop_ro = a.stop_ro;
    v.step = a.step;
    if (h->q == 9) {
        if (h->elem == 8) return launch_qt<9, double>(h, v, stream);
        if (h->elem == 4) return launch_qt<9, float>(h, v, stream);
        return launch_qt<9, _Float16>(h, v, stream);
    }
    if (h->elem == 8) return launch_qt<19, double>(h, v, stream);
    if (h->elem == 4) return launch_qt<19, float>(h, v, stream);
    return launch_qt<19, _Float16>(h, v, stream);
}

}  // namespace lat
}  // namespace erlbm
