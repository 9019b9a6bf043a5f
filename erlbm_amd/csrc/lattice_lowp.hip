// lattice_lowp.hip -- the plain lattice path with shifted low-precision
// population storage.
//
// Populations are stored as deviations d_i = f_i - w_i rho0 (rho0 = the
// lattice's reference density at upload) in FP32 or FP16.  The collision
// runs in FP64 on f_i = d_i + w_i rho0 (collide.hpp), or, with F32A, in
// FP32 on the deviations (collide_f32.hpp: compensated moment sums, FMA).  The stream, the
// boundary columns (Zou-He inlet and outlet rescale, evaluated on absolute
// FP64 populations), the stop tests and the monitors follow the FP64 path
// (lattice.hip) step for step; energies and densities are reduced in FP64.
// Storage traffic per D2Q9 node-step: 72 B (FP32) or 36 B (FP16) instead of
// 144 B.
#include <hip/hip_runtime.h>

#include "collide_f32.hpp"
#include "lattice_impl.hpp"
#include "walk.hpp"

namespace erlbm {
namespace lat {

namespace {

template <typename T>
__device__ __forceinline__ float ld(const T *p, bool nt)
{
    if (nt) return (float)__builtin_nontemporal_load(p);
    return (float)*p;
}

template <typename T>
__device__ __forceinline__ void st(T *p, float v)
{
    __builtin_nontemporal_store((T)v, p);
}

struct LPArgs {
    const void *src;
    void *dst;
    Geometry g;
    const uint32_t *__restrict__ nbr;
    CollideConstsF32 k;
    CollideConsts k64;
    BCArgs bc;
    double shift;                      // rho0 of the storage
    double dens_limit;
    unsigned long long *status;
    const unsigned long long *__restrict__ stop_ro;
    unsigned long long step;
};

template <int Q, typename T, bool STREAM, bool SOLID>
__device__ __forceinline__ void gather_lp(const T *__restrict__ src, const Geometry &g, uint32_t nb, int x, int y,
                                          int z, float (&d)[Q])
{
    using V = VSet<Q>;
    const long long row = y + (long long)g.ny * z;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        if constexpr (!STREAM || i == 0) {
            d[i] = ld(&src[at(g, i, row, x)], true);
        } else {
            if (SOLID && (nb & (1u << i))) {
                d[i] = ld(&src[at(g, V::opp[i], row, x)], true);
            } else {
                const int xs = V::cx[i] == 0 ? x : wrap(x - V::cx[i], g.nx);
                const int ys = V::cy[i] == 0 ? y : wrap(y - V::cy[i], g.ny);
                const int zs = V::cz[i] == 0 ? z : wrap(z - V::cz[i], g.nz);
                d[i] = ld(&src[at(g, i, ys + (long long)g.ny * zs, xs)], true);
            }
        }
    });
}

template <int Q>
__device__ __forceinline__ void to_abs(const float (&d)[Q], double (&f)[Q], double shift)
{
    static_for<Q>([&](auto I) { f[(int)I] = (double)d[(int)I] + VSet<Q>::w[(int)I] * shift; });
}

template <int Q>
__device__ __forceinline__ void from_abs(const double (&f)[Q], float (&d)[Q], double shift)
{
    static_for<Q>([&](auto I) { d[(int)I] = (float)(f[(int)I] - VSet<Q>::w[(int)I] * shift); });
}

// Periodic all-fluid lattice.
// Collision of shifted populations d (and the density for the stop test).
template <int Q, int COL, bool F32A>
__device__ __forceinline__ double collide_lp(float (&d)[Q], const LPArgs &a)
{
    if constexpr (F32A) {
        float drho, vx, vy, vz;
        collide_shifted_f32<Q, COL>(d, a.k, drho, vx, vy, vz);
        return density_shifted_f64<Q>(d, a.shift);
    } else {
        double f[Q];
        to_abs<Q>(d, f, a.shift);
        double vx, vy, vz;
        collide_sel<Q, COL>(f, a.k64, vx, vy, vz);
        from_abs<Q>(f, d, a.shift);
        return density_exact<Q>(f);
    }
}

template <int Q, int COL, typename T, bool STREAM, bool F32A>
__global__ __launch_bounds__(BLOCK) void k_lp_periodic(LPArgs a)
{
    const T *__restrict__ src = static_cast<const T *>(a.src);
    T *__restrict__ dst = static_cast<T *>(a.dst);
    const int row = blockIdx.y;
    const int x = blockIdx.x * BLOCK + (int)threadIdx.x;
    const int y = row % a.g.ny, z = row / a.g.ny;
    float d[Q];
    if (x < a.g.nx) gather_lp<Q, T, STREAM, false>(src, a.g, 0u, x, y, z, d);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    bool bad = false;
    if (x < a.g.nx) {
        const double r = collide_lp<Q, COL, F32A>(d, a);
        if (a.dens_limit > 0.0) bad = !(fabs(r) <= a.dens_limit);
        static_for<Q>([&](auto I) { st(&dst[at(a.g, (int)I, row, x)], d[(int)I]); });
    }
    flag_density(a.status, a.step, bad);
}

// Any geometry, boundary columns on absolute FP64 populations.
template <int Q, int COL, typename T, bool STREAM, bool F32A>
__global__ __launch_bounds__(BLOCK) void k_lp_general(LPArgs a)
{
    using V = VSet<Q>;
    const T *__restrict__ src = static_cast<const T *>(a.src);
    T *__restrict__ dst = static_cast<T *>(a.dst);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    const long long nvox = (long long)a.g.nx * a.g.ny * a.g.nz;
    const long long v = (long long)blockIdx.x * BLOCK + threadIdx.x;
    bool bad = false;
    if (v < nvox) {
        const uint32_t nb = a.nbr[v];
        if (nb & FLUID_BIT) {
            const int x = (int)(v % a.g.nx);
            const long long yz = v / a.g.nx;
            const int y = (int)(yz % a.g.ny);
            const int z = (int)(yz / a.g.ny);
            float d[Q];
            gather_lp<Q, T, STREAM, true>(src, a.g, nb, x, y, z, d);
            if (x == a.bc.inlet_x || x == a.bc.outlet_x) {
                double f[Q];
                to_abs<Q>(d, f, a.shift);
                if (x == a.bc.inlet_x) zouhe<Q>(f, a.bc.u_in, a.bc.inlet_west);
                if (x == a.bc.outlet_x) {
                    const int xn = x + a.bc.outlet_dx;
                    const long long vn = xn + (long long)a.g.nx * yz;
                    const uint32_t nbn = a.nbr[vn];
                    if (nbn & FLUID_BIT) {
                        float dn[Q];
                        gather_lp<Q, T, STREAM, true>(src, a.g, nbn, xn, y, z, dn);
                        double fa[Q];
                        to_abs<Q>(dn, fa, a.shift);
                        if (xn == a.bc.inlet_x) zouhe<Q>(fa, a.bc.u_in, a.bc.inlet_west);
                        const double fator = a.bc.rho_out / density_exact<Q>(fa);
                        static_for<Q>([&](auto I) { f[(int)I] = fa[(int)I] * fator; });
                    } else {
                        static_for<Q>([&](auto I) { f[(int)I] = V::w[(int)I] * a.bc.rho_out; });
                    }
                }
                from_abs<Q>(f, d, a.shift);
            }
            const double r = collide_lp<Q, COL, F32A>(d, a);
            if (a.dens_limit > 0.0) bad = !(fabs(r) <= a.dens_limit);
            static_for<Q>([&](auto I) { st(&dst[at(a.g, (int)I, yz, x)], d[(int)I]); });
        }
    }
    flag_density(a.status, a.step, bad);
}

template <int Q, typename T>
__global__ __launch_bounds__(BLOCK) void k_lp_from_aos(const double *__restrict__ aos, Geometry g,
                                                       const int *__restrict__ vof, int n_fluid, double shift,
                                                       T *__restrict__ dst)
{
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n_fluid) return;
    const long long v = vof ? vof[p] : p;
    const int x = (int)(v % g.nx);
    const long long row = v / g.nx;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        dst[at(g, i, row, x)] = (T)(float)(aos[(long long)p * Q + i] - VSet<Q>::w[i] * shift);
    });
}

template <int Q, typename T, bool STREAM>
__global__ __launch_bounds__(BLOCK) void k_lp_to_aos(const T *__restrict__ src, Geometry g,
                                                     const uint32_t *__restrict__ nbr, const int *__restrict__ vof,
                                                     int n_fluid, double shift, double *__restrict__ aos)
{
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n_fluid) return;
    const long long v = vof ? vof[p] : p;
    const int x = (int)(v % g.nx);
    const long long yz = v / g.nx;
    const int y = (int)(yz % g.ny), z = (int)(yz / g.ny);
    float d[Q];
    if (nbr) gather_lp<Q, T, STREAM, true>(src, g, nbr[v], x, y, z, d);
    else gather_lp<Q, T, STREAM, false>(src, g, 0u, x, y, z, d);
    double f[Q];
    to_abs<Q>(d, f, shift);
    static_for<Q>([&](auto I) { aos[(long long)p * Q + (int)I] = f[(int)I]; });
}

// Per-site FP64 energy and density of the absolute populations.
template <int Q, typename T, bool STREAM>
__global__ __launch_bounds__(BLOCK) void k_lp_monitor(const T *__restrict__ src, Geometry g,
                                                      const uint32_t *__restrict__ nbr, const int *__restrict__ vof,
                                                      int n_fluid, double shift, double *ke_node,
                                                      double *partial_ke, double *partial_rho, int *nonfinite)
{
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    double e = 0.0, r_abs = 0.0;
    bool nf = false;
    if (p < n_fluid) {
        const long long v = vof ? vof[p] : p;
        const int x = (int)(v % g.nx);
        const long long yz = v / g.nx;
        const int y = (int)(yz % g.ny), z = (int)(yz / g.ny);
        float d[Q];
        if (nbr) gather_lp<Q, T, STREAM, true>(src, g, nbr[v], x, y, z, d);
        else gather_lp<Q, T, STREAM, false>(src, g, 0u, x, y, z, d);
        double f[Q];
        to_abs<Q>(d, f, shift);
        double rho, vx, vy, vz;
        moments_exact<Q>(f, rho, vx, vy, vz);
        e = ke_term<Q>(vx, vy, vz);
        r_abs = fabs(rho);
        nf = !isfinite(rho);
        if (ke_node) ke_node[p] = e;
    }
    if (__any(nf) && (threadIdx.x & 63) == 0) atomicOr(nonfinite, 1);
    __shared__ double s_e[BLOCK / 64], s_r[BLOCK / 64];
    e = wave_sum(e);
    r_abs = wave_max(r_abs);
    if ((threadIdx.x & 63) == 0) {
        s_e[threadIdx.x >> 6] = e;
        s_r[threadIdx.x >> 6] = r_abs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double te = 0.0, tr = 0.0;
        for (int w = 0; w < BLOCK / 64; w++) {
            te += s_e[w];
            tr = fmax(tr, s_r[w]);
        }
        partial_ke[blockIdx.x] = te;
        partial_rho[blockIdx.x] = tr;
    }
}

// Two steps per launch for periodic all-fluid D2Q9 lattices with shifted
// storage and FP64 arithmetic: the strip walk of lattice.hip's
// k_step2_strip (three-row LDS ring, one halo column each side), with the
// step-k populations kept in the ring in the storage type T, exactly the
// values a one-step launch would have written to HBM and read back.

template <int COL, typename T, int TX, int SEG>
__global__ __launch_bounds__(strip_threads<TX>()) void k_lp_step2_strip(LPArgs a, int remap)
{
    using V = VSet<9>;
    constexpr int EX = TX + 2;
    extern __shared__ double ring_raw[];
    T *ring = reinterpret_cast<T *>(ring_raw);   // [3][9][EX]
    const T *__restrict__ src = static_cast<const T *>(a.src);
    T *__restrict__ dst = static_cast<T *>(a.dst);
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if (remap) b = xcd_runs(b, nb);
    const int nstrip = a.g.nx / TX;
    const int x0 = (int)(b % (unsigned)nstrip) * TX;
    const int ya = (int)(b / (unsigned)nstrip) * SEG;
    const int t = (int)threadIdx.x;
    const int nx = a.g.nx, ny = a.g.ny;
    const bool act1 = t < EX;
    const int x1 = act1 ? wrap(x0 - 1 + t, nx) : 0;
    auto load_row = [&](int y, float (&d)[9]) {
        if (!act1) return;
        const int yw = wrap(y, ny);
        static_for<9>([&](auto I) {
            constexpr int i = I;
            const int xs = V::cx[i] == 0 ? x1 : wrap(x1 - V::cx[i], nx);
            const int ys = V::cy[i] == 0 ? yw : wrap(yw - V::cy[i], ny);
            d[i] = ld(&src[at(a.g, i, ys, xs)], V::cx[i] == 0);
        });
    };
    auto slot = [&](int y) { return (y - ya + 1) % 3; };
    auto collide_to_ring = [&](int y, float (&d)[9]) {
        if (!act1) return;
        (void)collide_lp<9, COL, false>(d, a);
        T *r = ring + slot(y) * 9 * EX;
        static_for<9>([&](auto I) { r[(int)I * EX + t] = (T)d[(int)I]; });
    };
    auto output_row = [&](int y) {
        if (t < TX) {
            float d[9];
            static_for<9>([&](auto I) {
                constexpr int i = I;
                d[i] = (float)ring[slot(y - V::cy[i]) * 9 * EX + i * EX + (t + 1 - V::cx[i])];
            });
            (void)collide_lp<9, COL, false>(d, a);
            static_for<9>([&](auto I) { st(&dst[at(a.g, (int)I, y, x0 + t)], d[(int)I]); });
        }
    };
    float da[9], db[9];
    load_row(ya - 1, da);
    load_row(ya, db);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    collide_to_ring(ya - 1, da);
    collide_to_ring(ya, db);
    load_row(ya + 1, da);
    for (int r = 0; r < SEG; r++) {
        const int y = ya + r;
        collide_to_ring(y + 1, da);
        if (r + 1 < SEG) load_row(y + 2, da);
        lds_barrier();
        output_row(y);
        lds_barrier();
    }
}

// Three steps per launch for FP32 shifted storage with FP64 arithmetic:
// the three-step strip walk of lattice.hip's k_step3_strip (two LDS rings,
// odd segments walking upwards), with the step-k and step-(k+1)
// populations kept in the rings in FP32, exactly the values one-step
// launches would have written to HBM and read back.
template <int COL, int TX, int SEG, int DIR>
__device__ __forceinline__ void lp_strip3_walk(const LPArgs &a, int x0, int ya, float *ring1, float *ring2)
{
    using V = VSet<9>;
    constexpr int E1 = TX + 4, E2 = TX + 2;
    const int t = (int)threadIdx.x;
    const int nx = a.g.nx, ny = a.g.ny;
    const bool act1 = t < E1, act2 = t < E2, act3 = t < TX;
    const int x1 = act1 ? wrap(x0 - 2 + t, nx) : 0;
    const __amdgpu_buffer_rsrc_t rs = pop_buffer<4>(a.src, a.g), rd = pop_buffer<4>(a.dst, a.g);
    const unsigned oob1 = act1 ? 0u : BUF_OOB;
    const unsigned vo[3] = {4u * wrap(x1 + 1, nx) | oob1, 4u * x1 | oob1, 4u * wrap(x1 - 1, nx) | oob1};
    const unsigned vst = act3 ? 4u * (x0 + t) : 0u;
    auto load_row = [&](int y, float (&d)[9], bool skip) {   // pulls of step-k row y (-3 .. ny+2)
        const unsigned off = skip ? BUF_OOB : 0u;
        const int yw = wrap(y, ny);
        static_for<9>([&](auto I) {
            constexpr int i = I;
            const int ys = V::cy[i] == 0 ? yw : wrap(yw - V::cy[i], ny);
            d[i] = buf_load_f32<V::cx[i] == 0 ? BUF_NT : 0>(rs, vo[V::cx[i] + 1] | off, (int)(4 * at(a.g, i, ys, 0)));
        });
    };
    auto rrow = [&](int i, int y) { return ring_base<DIR>(i) + (y - ya + 2) % ring_depth<DIR>(i); };
    const int r0 = DIR > 0 ? ya - 2 : ya + SEG + 1;
    float fq[9];
    load_row(r0, fq, false);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    nine_dropped_stores<float>(rd);
    for (int n = 0; n < SEG + 4; n++) {
        const int r = r0 + DIR * n;
        float d[9];
        static_for<9>([&](auto I) { d[(int)I] = fq[(int)I]; });
        load_row(r + DIR, fq, n + 1 >= SEG + 4);                     // in flight meanwhile
        if (act1) {                                                   // step k, row r
            (void)collide_lp<9, COL, false>(d, a);
            static_for<9>([&](auto I) { ring1[rrow((int)I, r) * E1 + t] = d[(int)I]; });
        }
        lds_barrier();
        if (act2 && n >= 2) {                                         // step k+1, row r-DIR
            const int y = r - DIR;
            float g[9];
            static_for<9>([&](auto I) {
                constexpr int i = I;
                g[i] = ring1[rrow(i, y - V::cy[i]) * E1 + (t + 1 - V::cx[i])];
            });
            (void)collide_lp<9, COL, false>(g, a);
            static_for<9>([&](auto I) { ring2[rrow((int)I, y) * E2 + t] = g[(int)I]; });
        }
        lds_barrier();
        const bool out = n >= 4;                                      // step k+2, row r-2 DIR
        const int y = out ? r - 2 * DIR : ya;
        float g[9] = {};
        if (act3 && out) {
            static_for<9>([&](auto I) {
                constexpr int i = I;
                g[i] = ring2[rrow(i, y - V::cy[i]) * E2 + (t + 1 - V::cx[i])];
            });
            (void)collide_lp<9, COL, false>(g, a);
        }
        const unsigned sto = act3 && out ? 0u : BUF_OOB;
        static_for<9>([&](auto I) { buf_store_nt(g[(int)I], rd, vst | sto, (int)(4 * at(a.g, (int)I, y, 0))); });
    }
}

template <int COL, int TX, int SEG>
__global__ __launch_bounds__(strip_threads<TX>()) void k_lp_step3_strip(LPArgs a, int remap)
{
    constexpr int E1 = TX + 4;
    extern __shared__ float lp_rings[];   // ring 1 [RING_ROWS][TX+4], ring 2 [RING_ROWS][TX+2]
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if (remap) b = xcd_runs(b, nb);
    const int nstrip = a.g.nx / TX;
    const int x0 = (int)(b % (unsigned)nstrip) * TX;
    const int seg = (int)(b / (unsigned)nstrip);
    if (seg & 1) lp_strip3_walk<COL, TX, SEG, -1>(a, x0, seg * SEG, lp_rings, lp_rings + RING_ROWS * E1);
    else lp_strip3_walk<COL, TX, SEG, 1>(a, x0, seg * SEG, lp_rings, lp_rings + RING_ROWS * E1);
}

// The lane-aligned three-step walk of lattice.hip's k_step3_astrip for FP32
// shifted storage (FP64 arithmetic): lane t holds position x0-2+t in all
// three phases, the c_x = 0 populations stay in registers, the 8-row lean
// FP32 rings hold the other six (16 KB for four wavefronts); the values
// kept, in LDS or registers, are the FP32 ones one-step launches would
// store and read back.
// NTP 1: the c_x = 0 step-k loads carry the nontemporal hint; 0 (default):
// none, so the halo rows neighbouring segments share stay in L2.
template <int COL, int NW, int DIR, int NTP = 0>
__device__ __forceinline__ void lp_strip3_walk_aligned(const LPArgs &a, int x0, int w, int ya, int rows, float *ring1,
                                                       float *ring2, int prio)
{
    using V = VSet<9>;
    constexpr int E = 64 * NW;
    const int t = (int)threadIdx.x;
    const int nx = a.g.nx, ny = a.g.ny;
    const bool act1 = t < w + 4, act2 = t >= 1 && t < w + 3, act3 = t >= 2 && t < w + 2;
    const int x1 = act1 ? wrap(x0 - 2 + t, nx) : 0;
    const __amdgpu_buffer_rsrc_t rs = pop_buffer<4>(a.src, a.g), rd = pop_buffer<4>(a.dst, a.g);
    const unsigned oob1 = act1 ? 0u : BUF_OOB;
    const unsigned vo[3] = {4u * wrap(x1 + 1, nx) | oob1, 4u * x1 | oob1, 4u * wrap(x1 - 1, nx) | oob1};
    const unsigned vst = act3 ? 4u * (x0 - 2 + t) : 0u;
    auto load_row = [&](int y, float (&d)[9], bool skip) {   // pulls of step-k row y (-3 .. ny+2)
        const unsigned off = skip ? BUF_OOB : 0u;
        const int yw = wrap(y, ny);
        static_for<9>([&](auto I) {
            constexpr int i = I;
            const int ys = V::cy[i] == 0 ? yw : wrap(yw - V::cy[i], ny);
            d[i] = buf_load_f32<NTP == 1 && V::cx[i] == 0 ? BUF_NT : 0>(rs, vo[V::cx[i] + 1] | off, (int)(4 * at(a.g, i, ys, 0)));
        });
    };
    auto rrow = [&](int i, int y) { return xring_base<DIR>(i) + (y - ya + 2) % xring_depth<DIR>(i); };
    auto inring = [](int i) { return V::cx[i] != 0; };
    auto now = [](int i) { return V::cy[i] * DIR < 0; };
    auto age = [](int i) { return 1 + V::cy[i] * DIR; };
    const int r0 = DIR > 0 ? ya - 2 : ya + rows + 1;
    float fq[9];
    float h2[9];
    float k1a[9], k1b[9];
    float k2a[9], k2b[9];
    load_row(r0, fq, false);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    nine_dropped_stores<float>(rd);
    for (int n = 0; n < rows + 4; n++) {
        if (prio) progress_prio(n, rows + 4, prio);
        const int r = r0 + DIR * n;
        float f[9];
        static_for<9>([&](auto I) { f[(int)I] = fq[(int)I]; });
        load_row(r + DIR, fq, n + 1 >= rows + 4);                    // in flight meanwhile
        if (act1) {                                                   // phase 1: step k, row r
            (void)collide_lp<9, COL, false>(f, a);
            static_for<9>([&](auto I) {
                if (inring(I) && now(I)) ring1[rrow((int)I, r) * E + t] = f[(int)I];
            });
        }
        lds_barrier();
        if (act2 && n >= 3) {                                         // the rest of row r-2 DIR
            static_for<9>([&](auto I) {
                if (inring(I) && !now(I)) ring2[rrow((int)I, r - 2 * DIR) * E + t] = h2[(int)I];
            });
        }
        if (act2 && n >= 2) {                                         // phase 2: step k+1, row r-DIR
            const int y = r - DIR;
            float g[9];
            static_for<9>([&](auto I) {
                constexpr int i = I;
                if constexpr (V::cx[i] != 0) g[i] = ring1[rrow(i, y - V::cy[i]) * E + (t - V::cx[i])];
                else g[i] = age(i) == 0 ? f[i] : (age(i) == 1 ? k1a[i] : k1b[i]);
            });
            (void)collide_lp<9, COL, false>(g, a);
            static_for<9>([&](auto I) {
                if (inring(I) && now(I)) ring2[rrow((int)I, y) * E + t] = g[(int)I];
                h2[(int)I] = g[(int)I];
            });
        }
        static_for<9>([&](auto I) {
            if constexpr (V::cx[(int)I] == 0) {
                k1b[(int)I] = k1a[(int)I];
                k1a[(int)I] = f[(int)I];
            }
        });
        lds_barrier();
        if (act1) {                                                   // the rest of row r
            static_for<9>([&](auto I) {
                if (inring(I) && !now(I)) ring1[rrow((int)I, r) * E + t] = f[(int)I];
            });
        }
        const bool out = n >= 4;                                      // phase 3: step k+2, row r-2 DIR
        const int y = out ? r - 2 * DIR : ya;
        float o[9] = {};
        if (act3 && out) {
            static_for<9>([&](auto I) {
                constexpr int i = I;
                if constexpr (V::cx[i] != 0) o[i] = ring2[rrow(i, y - V::cy[i]) * E + (t - V::cx[i])];
                else o[i] = age(i) == 0 ? h2[i] : (age(i) == 1 ? k2a[i] : k2b[i]);
            });
            (void)collide_lp<9, COL, false>(o, a);
        }
        static_for<9>([&](auto I) {
            if constexpr (V::cx[(int)I] == 0) {
                k2b[(int)I] = k2a[(int)I];
                k2a[(int)I] = h2[(int)I];
            }
        });
        const unsigned sto = act3 && out ? 0u : BUF_OOB;
        static_for<9>([&](auto I) { buf_store_nt(o[(int)I], rd, vst | sto, (int)(4 * at(a.g, (int)I, y, 0))); });
    }
}

template <int COL, int NW, int NTP = 0>
__global__ __launch_bounds__(64 * NW) void k_lp_step3_astrip(LPArgs a, int remap, int nstrip, int nseg, int unit,
                                                            int prio)
{
    constexpr int E = 64 * NW;
    extern __shared__ float lp_rings[];   // ring 1 [XRING_ROWS][E], ring 2 [XRING_ROWS][E]
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if (remap) b = xcd_runs(b, nb);
    const int s = (int)(b % (unsigned)nstrip);
    const int units = a.g.nx / unit;
    const int x0 = (int)((long long)s * units / nstrip) * unit;
    const int w = (int)((long long)(s + 1) * units / nstrip) * unit - x0;
    const int seg = (int)(b / (unsigned)nstrip);
    const int ya = (int)((long long)seg * a.g.ny / nseg);
    const int rows = (int)((long long)(seg + 1) * a.g.ny / nseg) - ya;
    if (seg & 1) lp_strip3_walk_aligned<COL, NW, -1, NTP>(a, x0, w, ya, rows, lp_rings, lp_rings + XRING_ROWS * E, prio);
    else lp_strip3_walk_aligned<COL, NW, 1, NTP>(a, x0, w, ya, rows, lp_rings, lp_rings + XRING_ROWS * E, prio);
}

// Four steps per launch on FP32 shifted storage: the FP64 kernel's walk
// (stripP_walk_aligned) with FP32 rings (18 KB for NW = 4), the kept values
// the FP32 ones one-step launches would store, so bitwise the one-step FP32
// path.
template <int COL, int NW>
__global__ __launch_bounds__(64 * NW) void k_lp_step4_astrip(LPArgs a, int remap, int nstrip, int nseg, int unit,
                                                            int prio)
{
    extern __shared__ float lp_rings[];
    auto collide = [&](float (&d)[9]) { (void)collide_lp<9, COL, false>(d, a); };
    astripP_workgroup<float, NW, 4>(a, remap, nstrip, nseg, unit, prio, lp_rings, collide);
}

template <int COL, int NW>
__global__ __launch_bounds__(64 * NW) void k_lp_step5_astrip(LPArgs a, int remap, int nstrip, int nseg, int unit,
                                                            int prio)
{
    extern __shared__ float lp_rings[];
    auto collide = [&](float (&d)[9]) { (void)collide_lp<9, COL, false>(d, a); };
    astripP_workgroup<float, NW, 5>(a, remap, nstrip, nseg, unit, prio, lp_rings, collide);
}

// P steps per launch (4 or 5): strips of at most 64 NW - 2 (P-1) positions,
// edges on 16-position units where the width allows; segments sized to the
// resident workgroups.
template <int COL, int P>
void launch_lp_astripP(erlbm_lattice *h, const LPArgs &l)
{
    constexpr int NW = 4;
    constexpr size_t lds = sizeof(float) * (P - 1) * XSLOTS * (64 * NW);
    const void *k = P == 4 ? reinterpret_cast<const void *>(&k_lp_step4_astrip<COL, NW>)
                           : reinterpret_cast<const void *>(&k_lp_step5_astrip<COL, NW>);
    static const int per_cu = [&] {
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, k, 64 * NW, lds) != hipSuccess || nblk < 1) nblk = 1;
        return nblk;
    }();
    const int pc = h->tune_fuse_slots > 0 ? std::min(per_cu, h->tune_fuse_slots) : per_cu;
    const int unit = h->nx % 16 == 0 ? 16 : 1;
    constexpr int TX = 64 * NW - 2 * (P - 1);
    const int nstrip = (h->nx / unit + TX / unit - 1) / (TX / unit);
    const int nseg = (int)std::max(1ll, std::min<long long>(h->ny / 16, (long long)h->n_cu * pc / nstrip));
    const int prio = h->tune_fuse_prio >= 0 ? h->tune_fuse_prio : (nstrip * nseg >= 2 * h->n_cu ? 2 : 0);
    if constexpr (P == 4)
        hipLaunchKernelGGL((k_lp_step4_astrip<COL, NW>), dim3((unsigned)(nstrip * nseg)), dim3(64 * NW), lds, h->stream,
                           l, h->tune_fuse_remap, nstrip, nseg, unit, prio);
    else
        hipLaunchKernelGGL((k_lp_step5_astrip<COL, NW>), dim3((unsigned)(nstrip * nseg)), dim3(64 * NW), lds, h->stream,
                           l, h->tune_fuse_remap, nstrip, nseg, unit, prio);
}

// Strips of at most 252 positions, edges on 16-position units where the
// width allows; segments as many as make one workgroup per resident slot,
// at least 16 rows each.
template <int COL>
void launch_lp_astrip(erlbm_lattice *h, const LPArgs &l)
{
    constexpr int NW = 4;
    constexpr size_t lds = sizeof(float) * XRING_ROWS * 2 * (64 * NW);
    static const int per_cu = [] {
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, reinterpret_cast<const void *>(&k_lp_step3_astrip<COL, NW>),
                                                         64 * NW, lds) != hipSuccess ||
            nblk < 1)
            nblk = 1;
        return nblk;
    }();
    const int unit = h->nx % 16 == 0 ? 16 : 1;
    const int nstrip = (h->nx / unit + (64 * NW - 4) / unit - 1) / ((64 * NW - 4) / unit);
    const int nseg = (int)std::max(1ll, std::min<long long>(h->ny / 16, (long long)h->n_cu * per_cu / nstrip));
    // issue priority by progress where two or more workgroups share a CU
    const int prio = h->tune_fuse_prio >= 0 ? h->tune_fuse_prio : (nstrip * nseg >= 2 * h->n_cu ? 2 : 0);
    if (h->tune_walk_nt == 1)
        hipLaunchKernelGGL((k_lp_step3_astrip<COL, NW, 1>), dim3((unsigned)(nstrip * nseg)), dim3(64 * NW), lds,
                           h->stream, l, h->tune_fuse_remap, nstrip, nseg, unit, prio);
    else
        hipLaunchKernelGGL((k_lp_step3_astrip<COL, NW>), dim3((unsigned)(nstrip * nseg)), dim3(64 * NW), lds,
                           h->stream, l, h->tune_fuse_remap, nstrip, nseg, unit, prio);
}

template <int COL, int TX, int SEG>
void launch_lp_strip3(erlbm_lattice *h, const LPArgs &l)
{
    constexpr size_t lds = sizeof(float) * RING_ROWS * ((TX + 4) + (TX + 2));
    const unsigned nb = (unsigned)((h->nx / TX) * (h->ny / SEG));
    hipLaunchKernelGGL((k_lp_step3_strip<COL, TX, SEG>), dim3(nb), dim3(strip_threads<TX>()), lds, h->stream, l,
                       h->tune_fuse_remap);
}

template <int COL, typename T, int TX, int SEG>
void launch_lp_strip(erlbm_lattice *h, const LPArgs &l)
{
    constexpr size_t lds = sizeof(T) * 3 * 9 * (TX + 2);
    const unsigned nb = (unsigned)((h->nx / TX) * (h->ny / SEG));
    hipLaunchKernelGGL((k_lp_step2_strip<COL, T, TX, SEG>), dim3(nb), dim3(strip_threads<TX>()), lds, h->stream,
                       l, h->tune_fuse_remap);
}

template <int COL, typename T>
void launch_lp_step2_ct(erlbm_lattice *h, const LPArgs &l)
{
    if constexpr ((COL & COL_MOMENT) != 0) {
        // the moment-space collision on FP32 storage: the lane-aligned
        // three-step walk and the two-step remainder walk; any other shape
        // keeps the reference's order
        if constexpr (sizeof(T) == 4) {
            if (h->fuse_active == FUSE5_ASTRIP_4W) return launch_lp_astripP<COL, 5>(h, l);
            if (h->fuse_active == FUSE4_ASTRIP_4W) return launch_lp_astripP<COL, 4>(h, l);
            if (h->fuse_active == FUSE3_ASTRIP_4W) return launch_lp_astrip<COL>(h, l);
            if (h->fuse_active == FUSE2_STRIP_256x32) return launch_lp_strip<COL, T, 256, 32>(h, l);
        }
        return launch_lp_step2_ct<(COL & ~COL_MOMENT), T>(h, l);
    } else if constexpr (sizeof(T) == 4) {
        if (h->fuse_active == FUSE3_STRIP_256x32) return launch_lp_strip3<COL, 256, 32>(h, l);
        if (h->fuse_active == FUSE3_STRIP_128x32) return launch_lp_strip3<COL, 128, 32>(h, l);
        if (h->fuse_active == FUSE3_ASTRIP_4W) return launch_lp_astrip<COL>(h, l);
        if (h->fuse_active == FUSE4_ASTRIP_4W) return launch_lp_astripP<COL, 4>(h, l);
        if (h->fuse_active == FUSE5_ASTRIP_4W) return launch_lp_astripP<COL, 5>(h, l);
    }
    if (h->fuse_active == FUSE2_STRIP_128x32) launch_lp_strip<COL, T, 128, 32>(h, l);
    else launch_lp_strip<COL, T, 256, 32>(h, l);
}

template <typename T>
int launch_lp_step2_t(erlbm_lattice *h, const LPArgs &l)
{
    if (h->arith == ERLBM_ARITH_MOMENT) {
        switch (h->col) {   // coll_MRT keeps the reference's order
        case ERLBM_COL_BGK: launch_lp_step2_ct<COL_BGK | COL_MOMENT, T>(h, l); break;
        case ERLBM_COL_MRT: launch_lp_step2_ct<COL_MRT, T>(h, l); break;
        case ERLBM_COL_REG: launch_lp_step2_ct<COL_REG | COL_MOMENT, T>(h, l); break;
        case ERLBM_COL_EXT_REG: launch_lp_step2_ct<COL_EXT_REG | COL_MOMENT, T>(h, l); break;
        default: return ERLBM_ERR_UNSUPPORTED;
        }
        ERLBM_HIP_TRY(hipGetLastError());
        return ERLBM_OK;
    }
    switch (h->col) {
    case ERLBM_COL_BGK: launch_lp_step2_ct<COL_BGK, T>(h, l); break;
    case ERLBM_COL_MRT: launch_lp_step2_ct<COL_MRT, T>(h, l); break;
    case ERLBM_COL_REG: launch_lp_step2_ct<COL_REG, T>(h, l); break;
    case ERLBM_COL_EXT_REG: launch_lp_step2_ct<COL_EXT_REG, T>(h, l); break;
    default: return ERLBM_ERR_UNSUPPORTED;
    }
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

LPArgs make_args(erlbm_lattice *h, const StepArgs &a)
{
    LPArgs l;
    l.src = a.src;
    l.dst = a.dst;
    l.g = a.g;
    l.nbr = a.nbr;
    l.k.omega1 = (float)(1.0 - 1.0 / h->tau);
    l.k.omega2 = (float)(1.0 - 1.0 / h->tau_nh);
    l.k.oot = (float)(1.0 / h->tau);
    l.k.rho0 = (float)h->shift;
    l.k64.omega1 = 1.0 - 1.0 / h->tau;
    l.k64.omega2 = 1.0 - 1.0 / h->tau_nh;
    l.k64.oot = 1.0 / h->tau;
    l.k64.m = moment_consts(h->q, h->col, h->tau, h->tau_nh);
    l.bc = a.bc;
    l.shift = h->shift;
    l.dens_limit = a.dens_limit;
    l.status = a.status;
    l.stop_ro = a.stop_ro;
    l.step = a.step;
    return l;
}

template <int Q, int COL, typename T, bool F32A>
void step_qcta(erlbm_lattice *h, const LPArgs &l, bool stream)
{
    if (h->all_fluid && h->inlet_x < 0 && h->outlet_x < 0) {
        const dim3 grid((h->nx + BLOCK - 1) / BLOCK, h->ny * h->nz);
        if (stream) hipLaunchKernelGGL((k_lp_periodic<Q, COL, T, true, F32A>), grid, dim3(BLOCK), 0, h->stream, l);
        else hipLaunchKernelGGL((k_lp_periodic<Q, COL, T, false, F32A>), grid, dim3(BLOCK), 0, h->stream, l);
    } else {
        const unsigned nb = (unsigned)((h->nvox + BLOCK - 1) / BLOCK);
        if (stream) hipLaunchKernelGGL((k_lp_general<Q, COL, T, true, F32A>), dim3(nb), dim3(BLOCK), 0, h->stream, l);
        else hipLaunchKernelGGL((k_lp_general<Q, COL, T, false, F32A>), dim3(nb), dim3(BLOCK), 0, h->stream, l);
    }
}

template <int Q, int COL, typename T>
int step_qct(erlbm_lattice *h, const LPArgs &l, bool stream)
{
    if constexpr (COL == COL_MRT) {
        // the FP32-arithmetic variant has no MRT form; FP64 arithmetic only
        if (h->f32_arith) return ERLBM_ERR_UNSUPPORTED;
        step_qcta<Q, COL, T, false>(h, l, stream);
    } else if constexpr (std::is_same<T, float>::value) {
        if (h->f32_arith) step_qcta<Q, COL, T, true>(h, l, stream);
        else step_qcta<Q, COL, T, false>(h, l, stream);
    } else {
        step_qcta<Q, COL, T, false>(h, l, stream);
    }
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

template <int Q, typename T>
int step_qt(erlbm_lattice *h, const LPArgs &l, bool stream)
{
    switch (h->col) {
    case ERLBM_COL_BGK: return step_qct<Q, COL_BGK, T>(h, l, stream);
    case ERLBM_COL_REG: return step_qct<Q, COL_REG, T>(h, l, stream);
    case ERLBM_COL_EXT_REG: return step_qct<Q, COL_EXT_REG, T>(h, l, stream);
    case ERLBM_COL_MRT:
        if constexpr (Q == 9) return step_qct<Q, COL_MRT, T>(h, l, stream);
        return ERLBM_ERR_UNSUPPORTED;
    default: return ERLBM_ERR_UNSUPPORTED;
    }
}

template <int Q, typename T>
int to_aos_qt(erlbm_lattice *h, const void *src, bool stream, double *aos)
{
    const uint32_t *nbr = h->all_fluid ? nullptr : h->nbr;
    const T *s = static_cast<const T *>(src);
    if (stream)
        hipLaunchKernelGGL((k_lp_to_aos<Q, T, true>), dim3(mon_blocks(h)), dim3(BLOCK), 0, h->stream, s, geom(h), nbr,
                           h->vox_of_fluid, h->n_fluid, h->shift, aos);
    else
        hipLaunchKernelGGL((k_lp_to_aos<Q, T, false>), dim3(mon_blocks(h)), dim3(BLOCK), 0, h->stream, s, geom(h),
                           nbr, h->vox_of_fluid, h->n_fluid, h->shift, aos);
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

template <int Q, typename T>
int monitor_qt(erlbm_lattice *h, const void *src, bool stream, bool ordered)
{
    const uint32_t *nbr = h->all_fluid ? nullptr : h->nbr;
    const T *s = static_cast<const T *>(src);
    double *ke = ordered ? h->ke_node : nullptr;
    if (stream)
        hipLaunchKernelGGL((k_lp_monitor<Q, T, true>), dim3(mon_blocks(h)), dim3(BLOCK), 0, h->stream, s, geom(h),
                           nbr, h->vox_of_fluid, h->n_fluid, h->shift, ke, h->partial_ke, h->partial_rho,
                           h->nonfinite);
    else
        hipLaunchKernelGGL((k_lp_monitor<Q, T, false>), dim3(mon_blocks(h)), dim3(BLOCK), 0, h->stream, s, geom(h),
                           nbr, h->vox_of_fluid, h->n_fluid, h->shift, ke, h->partial_ke, h->partial_rho,
                           h->nonfinite);
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

template <int Q, typename T>
int from_aos_qt(erlbm_lattice *h, const double *aos, void *dst)
{
    hipLaunchKernelGGL((k_lp_from_aos<Q, T>), dim3(mon_blocks(h)), dim3(BLOCK), 0, h->stream, aos, geom(h),
                       h->vox_of_fluid, h->n_fluid, h->shift, static_cast<T *>(dst));
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

}  // namespace

#define ERLBM_LP_DISPATCH(fn, ...)                                                          \
    do {                                                                                    \
        if (h->q == 9)                                                                      \
            return h->elem == 4 ? fn<9, float>(h, __VA_ARGS__) : fn<9, _Float16>(h, __VA_ARGS__);   \
        return h->elem == 4 ? fn<19, float>(h, __VA_ARGS__) : fn<19, _Float16>(h, __VA_ARGS__);     \
    } while (0)

int lowp_step(erlbm_lattice *h, const StepArgs &a, bool stream)
{
    const LPArgs l = make_args(h, a);
    ERLBM_LP_DISPATCH(step_qt, l, stream);
}

int lowp_step2(erlbm_lattice *h, const StepArgs &a)
{
    const LPArgs l = make_args(h, a);
    return h->elem == 4 ? launch_lp_step2_t<float>(h, l) : launch_lp_step2_t<_Float16>(h, l);
}

int lowp_from_aos(erlbm_lattice *h, const double *aos, void *dst) { ERLBM_LP_DISPATCH(from_aos_qt, aos, dst); }

int lowp_to_aos(erlbm_lattice *h, const void *src, bool stream, double *aos)
{
    ERLBM_LP_DISPATCH(to_aos_qt, src, stream, aos);
}

int lowp_monitor(erlbm_lattice *h, const void *src, bool stream, bool ordered)
{
    ERLBM_LP_DISPATCH(monitor_qt, src, stream, ordered);
}

}  // namespace lat
}  // namespace erlbm
