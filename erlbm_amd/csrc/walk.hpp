// walk.hpp -- pieces shared by the multi-step strip walks (lattice.hip for
// FP64 populations, lattice_lowp.hip for FP32 shifted storage): raw-buffer
// access to the population arrays and the row layout of the LDS rings.
#pragma once

#include <hip/hip_runtime.h>

#include "lattice_impl.hpp"
#include "vset.hpp"

namespace erlbm {
namespace lat {

typedef unsigned int buf_u2 __attribute__((ext_vector_type(2)));
constexpr int BUF_DWORD3 = 0x00020000;   // raw (stride 0) buffer of dwords
constexpr int BUF_NT = 2;                // nontemporal cache policy
// A VGPR offset at or above the buffer's size is out of range: the load
// returns 0 without touching memory, the store is dropped.  Masking lanes
// and skipped prefetches this way keeps the walks free of branches around
// memory instructions, so the wait for a row's loads never has to wait for
// the stores issued after them.
constexpr unsigned BUF_OOB = 0x80000000u;

// Raw buffer over a population array of E-byte elements (byte size below
// 2^31, checked by the launchers): loads and stores at a per-lane VGPR
// offset plus a wave-uniform SGPR offset; a VGPR offset with BUF_OOB or-ed
// in is out of range whether or not the SGPR offset (below 2^31) is added
// before the range check.
template <int E>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pop_buffer(const void *p, const Geometry &g)
{
    const unsigned nbytes = (unsigned)(E * at(g, 8, (long long)g.ny * g.nz - 1, g.nx));
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)nbytes, BUF_DWORD3);
}
template <int AUX>
__device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
}
template <int AUX>
__device__ __forceinline__ float buf_load_f32(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
}
__device__ __forceinline__ void buf_store_nt(double v, __amdgpu_buffer_rsrc_t r, unsigned voff, int soff = 0)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(buf_u2, v), r, voff, soff, BUF_NT);
}
__device__ __forceinline__ void buf_store_nt(float v, __amdgpu_buffer_rsrc_t r, unsigned voff, int soff = 0)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, BUF_NT);
}
// Nine dropped stores: issued after the loads a loop waits for at its top,
// they make the state entering the loop match the one at its back edge
// (those loads, then one row's nine stores), so the wait at the top can
// leave the previous row's stores in flight instead of draining them.
template <typename T>
__device__ __forceinline__ void nine_dropped_stores(__amdgpu_buffer_rsrc_t r)
{
    static_for<9>([&](auto I) { buf_store_nt(T(0), r, BUF_OOB | ((unsigned)sizeof(T) * (int)I)); });
}

// The LDS rings of the D2Q9 strip walks keep each population only as long
// as it is read.  A pull of row y reads the c_y = +1 populations from row
// y-1, c_y = 0 from row y and c_y = -1 from row y+1, and the newest of the
// three is written just before the pull, so walking in direction DIR
// (+1: rows ascending) the populations whose source row is the oldest
// (c_y = DIR), the middle and the newest need 3, 2 and 1 row slots: 18 rows
// in all instead of 27.
template <int DIR>
__host__ __device__ constexpr int ring_depth(int i)
{
    return VSet<9>::cy[i] * DIR > 0 ? 3 : (VSet<9>::cy[i] == 0 ? 2 : 1);
}
template <int DIR>
__host__ __device__ constexpr int ring_base(int i)
{
    return i == 0 ? 0 : ring_base<DIR>(i - 1) + ring_depth<DIR>(i - 1);
}
constexpr int RING_ROWS = ring_base<1>(8) + ring_depth<1>(8);
static_assert(RING_ROWS == 18 && ring_base<-1>(8) + ring_depth<-1>(8) == RING_ROWS, "D2Q9 ring rows");

// The lean rings of the three-step walk write a population into the ring
// only when its first reader needs it (the newest-row ones at once, the
// others one barrier later, kept in registers meanwhile): 2 / 1 / 1 row
// slots, 12 rows in all.
template <int DIR>
__host__ __device__ constexpr int lean_depth(int i)
{
    return VSet<9>::cy[i] * DIR > 0 ? 2 : 1;
}
template <int DIR>
__host__ __device__ constexpr int lean_base(int i)
{
    return i == 0 ? 0 : lean_base<DIR>(i - 1) + lean_depth<DIR>(i - 1);
}
constexpr int LEAN_ROWS = lean_base<1>(8) + lean_depth<1>(8);
static_assert(LEAN_ROWS == 12 && lean_base<-1>(8) + lean_depth<-1>(8) == LEAN_ROWS, "D2Q9 lean ring rows");

// Lean rings of the lane-aligned three-step walk: the c_x != 0 populations
// only (a c_x = 0 population is pulled by the lane that computed it, from
// registers): 8 rows.
template <int DIR>
__host__ __device__ constexpr int xring_depth(int i)
{
    return VSet<9>::cx[i] == 0 ? 0 : lean_depth<DIR>(i);
}
template <int DIR>
__host__ __device__ constexpr int xring_base(int i)
{
    return i == 0 ? 0 : xring_base<DIR>(i - 1) + xring_depth<DIR>(i - 1);
}
constexpr int XRING_ROWS = xring_base<1>(8) + xring_depth<1>(8);
static_assert(XRING_ROWS == 8 && xring_base<-1>(8) + xring_depth<-1>(8) == XRING_ROWS, "D2Q9 aligned ring rows");

// Lean rings of the D3Q19 z-walk, by c_z: 2 / 1 / 1 plane slots for the
// populations a pull reads from the oldest (c_z = DIR), middle and newest
// of its three planes; 24 planes in all.
template <int DIR>
__host__ __device__ constexpr int zring_depth(int i)
{
    return VSet<19>::cz[i] * DIR > 0 ? 2 : 1;
}
template <int DIR>
__host__ __device__ constexpr int zring_base(int i)
{
    return i == 0 ? 0 : zring_base<DIR>(i - 1) + zring_depth<DIR>(i - 1);
}
constexpr int ZRING_PLANES = zring_base<1>(18) + zring_depth<1>(18);
static_assert(ZRING_PLANES == 24 && zring_base<-1>(18) + zring_depth<-1>(18) == ZRING_PLANES, "D3Q19 ring planes");

// Workgroup b of nb, dispatched to XCD b mod 8, renumbered so that each
// XCD's workgroups take a contiguous run of the logical numbers (the first
// nb mod 8 XCDs one more than the others): neighbouring strips and segments
// then share an L2.  A bijection of 0 .. nb-1 for any nb.
__device__ __forceinline__ unsigned xcd_runs(unsigned b, unsigned nb)
{
    const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Issue priority by progress (walking step n of steps): the SIMD arbiter
// prefers the oldest ready wave, so of the workgroups sharing a CU the first
// started would run ahead and the last finish alone, latency-bound;
// lowering a walk's priority a level per band of its steps lets the ones
// behind catch up.  Scheme 2: bands ending at 50, 75 and 90% of the walk
// (what matters is how far apart the workgroups are when the first ends);
// 1: equal quarters.
__device__ __forceinline__ void progress_prio(int n, int steps, int scheme)
{
    const int q = scheme == 2 ? (20 * n < 10 * steps ? 0 : (20 * n < 15 * steps ? 1 : (20 * n < 18 * steps ? 2 : 3)))
                              : 4 * n / steps;
    if (q == 0) __builtin_amdgcn_s_setprio(3);
    else if (q == 1) __builtin_amdgcn_s_setprio(2);
    else if (q == 2) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// N dropped stores of T (see nine_dropped_stores).
template <typename T, int N>
__device__ __forceinline__ void dropped_stores(__amdgpu_buffer_rsrc_t r)
{
    static_for<N>([&](auto I) { buf_store_nt(T(0), r, BUF_OOB | ((unsigned)sizeof(T) * (int)I)); });
}

// One element of a population array of element type T (raw buffer load).
template <typename T>
__device__ __forceinline__ T buf_load_elem(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff);
template <>
__device__ __forceinline__ double buf_load_elem<double>(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff)
{
    return buf_load<0>(r, voff, soff);
}
template <>
__device__ __forceinline__ float buf_load_elem<float>(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff)
{
    return buf_load_f32<0>(r, voff, soff);
}

// Slot of a c_x != 0 population in the rings of the P-step walks (one row
// per population): 0 .. 5 for populations 1, 2, 5, 6, 7, 8.
__host__ __device__ constexpr int xslot(int i)
{
    return i == 0 ? 0 : xslot(i - 1) + (VSet<9>::cx[i - 1] != 0 ? 1 : 0);
}
constexpr int XSLOTS = 6;
static_assert(xslot(8) == 5 && xslot(5) == 2, "D2Q9 ring slots");

// Four to six steps per launch (P phases): the lane-aligned walk with P-1
// rings.  Lane t holds position x0 - (P-1) + t in all phases (phase p:
// lanes p-1 .. w+2P-p-2; the last phase stores lanes P-1 .. w+P-2).  With two
// barriers per walking step the phases alternate between the two intervals
// they make: walking step n (row r = r0 + DIR n) runs phase 1 on row r
// before the first barrier, phases 2 (row r - DIR) and 4 (row r - 4 DIR)
// between the barriers, phases 3 (row r - 2 DIR) and 5 (row r - 5 DIR)
// after the second; phase 4 so reads the ring-3 rows phase 3 wrote in the
// walking steps before, phase 5 the ring-4 rows phase 4 wrote in the same
// one, phase 6 (between the barriers, row r - 7 DIR) the ring-5 rows phase
// 5 wrote in the walking steps before.  Ring p holds the c_x != 0 results of phase p, one row per
// population (6 x 64 NW values): every ring value has exactly one reader,
// so each is written in the interval just before the one its reader runs
// in, from registers — the populations read from the newest row (c_y DIR <
// 0) at once, those read from the middle row one walking step later, those
// read from the oldest row two walking steps later — and each slot is free
// again when the next row's value arrives.  Rings 1, 3 and 5 are written
// after the second barrier (and ring 1 before the first), rings 2 and 4
// between the barriers: every ring is written in one interval kind and
// read in the other.  The c_x = 0 results stay in the lane's registers.
// Output rows ya .. ya+rows-1 need step-k rows ya-(P-1) .. ya+rows+P-2:
// rows + 7 (P = 4), rows + 9 (P = 5) or rows + 12 (P = 6) walking steps.
// T: the population arrays' element type (FP64 storage, or FP32 shifted
// storage whose collision converts to FP64 and back); Args: the launch's
// arguments (src, dst, g, step, stop_ro); collide(T (&)[9]) collides one
// node in place; ring: P-1 rings of XSLOTS x 64 NW elements.
template <typename T, int NW, int DIR, int P, class Args, class Collide>
__device__ __forceinline__ void stripP_walk_aligned(const Args &a, int x0, int w, int ya, int rows, T *ring, int prio,
                                                    Collide collide)
{
    static_assert(P >= 4 && P <= 6, "four, five or six phases");
    constexpr unsigned SZ = sizeof(T);
    using V = VSet<9>;
    constexpr int E = 64 * NW;
    T *const ring1 = ring, *const ring2 = ring + XSLOTS * E, *const ring3 = ring + 2 * XSLOTS * E;
    T *const ring4 = ring + 3 * XSLOTS * E;   // P >= 5
    T *const ring5 = ring + 4 * XSLOTS * E;   // P = 6
    const int t = (int)threadIdx.x;
    const int nx = a.g.nx, ny = a.g.ny;
    // phase p covers positions x0-(P-p) .. x0+w-1+(P-p): lanes p-1 .. w+2P-p-2
    auto act = [&](int p) { return t >= p - 1 && t < w + 2 * P - p - 1; };
    const bool act1 = act(1), act2 = act(2), act3 = act(3), act4 = act(4), act5 = act(5), act6 = act(6);
    const bool actl = P == 4 ? act4 : (P == 5 ? act5 : act6);   // lanes of the last phase: the stores
    const int x1 = act1 ? wrap(x0 - (P - 1) + t, nx) : 0;
    const __amdgpu_buffer_rsrc_t rs = pop_buffer<SZ>(a.src, a.g), rd = pop_buffer<SZ>(a.dst, a.g);
    const unsigned oob1 = act1 ? 0u : BUF_OOB;
    const unsigned vo[3] = {SZ * wrap(x1 + 1, nx) | oob1, SZ * x1 | oob1, SZ * wrap(x1 - 1, nx) | oob1};
    const unsigned vst = actl ? SZ * (x0 - (P - 1) + t) : 0u;
    auto load_row = [&](int y, T (&f)[9], bool skip) {   // pulls of step-k row y (-P .. ny+P-1)
        const unsigned off = skip ? BUF_OOB : 0u;
        const int yw = wrap(y, ny);
        static_for<9>([&](auto I) {
            constexpr int i = I;
            const int ys = V::cy[i] == 0 ? yw : wrap(yw - V::cy[i], ny);
            f[i] = buf_load_elem<T>(rs, vo[V::cx[i] + 1] | off, (int)(SZ * at(a.g, i, ys, 0)));
        });
    };
    auto age = [](int i) { return 1 + V::cy[i] * DIR; };   // 0: newest source row, 1: middle, 2: oldest
    // pull of a phase from the ring of the phase before (c_x != 0) and from
    // that phase's results in registers (c_x = 0: this lane's own, of the
    // newest, middle and oldest source row)
    auto pull = [&](T (&g)[9], const T *rg, const T (&r0)[9], const T (&r1)[9], const T (&r2)[9]) {
        static_for<9>([&](auto I) {
            constexpr int i = I;
            if constexpr (V::cx[i] != 0) g[i] = rg[xslot(i) * E + (t - V::cx[i])];
            else g[i] = age(i) == 0 ? r0[i] : (age(i) == 1 ? r1[i] : r2[i]);
        });
    };
    // ring writes of the populations whose reader takes them from the
    // newest / middle / oldest of its three source rows
    auto put = [&](T *rg, const T (&g)[9], int which) {
        static_for<9>([&](auto I) {
            constexpr int i = I;
            if constexpr (V::cx[i] != 0) {
                if (age(i) == which) rg[xslot(i) * E + t] = g[i];
            }
        });
    };
    auto keep = [](T (&dst)[9], const T (&src)[9]) { static_for<9>([&](auto I) { dst[(int)I] = src[(int)I]; }); };
    auto shift = [](T (&ka)[9], T (&kb)[9], const T (&g)[9]) {   // c_x = 0 results one walking step older
        static_for<9>([&](auto I) {
            if constexpr (V::cx[(int)I] == 0) {
                kb[(int)I] = ka[(int)I];
                ka[(int)I] = g[(int)I];
            }
        });
    };
    auto store = [&](const T (&o)[9], int y, bool on) {
        const unsigned sto = on ? 0u : BUF_OOB;
        static_for<9>([&](auto I) { buf_store_nt(o[(int)I], rd, vst | sto, (int)(SZ * at(a.g, (int)I, y, 0))); });
    };
    const int r0 = DIR > 0 ? ya - (P - 1) : ya + rows + (P - 2);
    const int nrows1 = rows + 2 * (P - 1);           // rows of phase 1
    const int nsteps = P == 4 ? rows + 7 : (P == 5 ? rows + 9 : rows + 12);  // walking steps
    T fq[9];
    T f1[9];                       // phase-1 row of the walking step before (its oldest-row ring writes)
    T h2[9], h3[9], h4[9], h5[9];  // phase-2 / 3 / 4 / 5 rows of the walking step before
    T h2p[9], h3p[9], h4p[9], h5p[9];   // the same two walking steps before
    T k1a[9], k1b[9];              // c_x = 0: phase-1 results of the rows one and two walking steps back
    T k2a[9], k2b[9];              // phase 2, likewise
    T k3a[9], k3b[9];              // phase 3, one and two walking steps before h3
    T k4a[9], k4b[9];              // phase 4, likewise (P >= 5)
    T k5a[9], k5b[9];              // phase 5, one and two walking steps before h5 (P = 6)
    load_row(r0, fq, false);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    nine_dropped_stores<T>(rd);
    for (int n = 0; n < nsteps; n++) {
        if (prio) progress_prio(n, nsteps, prio);
        const int r = r0 + DIR * n;
        T f[9];
        static_for<9>([&](auto I) { f[(int)I] = fq[(int)I]; });
        load_row(r + DIR, fq, n + 1 >= nrows1);                      // in flight meanwhile
        const bool p1 = act1 && n < nrows1;
        if (p1) {                                                     // phase 1: step k, row r
            collide(f);
            put(ring1, f, 0);
        }
        lds_barrier();
        if (act2 && n >= 3) put(ring2, h2, 1);                        // phase-2 rows of the steps before
        if (act2 && n >= 4) put(ring2, h2p, 2);
        if constexpr (P >= 5) {
            if (act4 && n >= 8) put(ring4, h4, 1);                    // phase-4 rows of the steps before
            if (act4 && n >= 9) put(ring4, h4p, 2);
        }
        if (act2 && n >= 2 && n < nrows1) {                           // phase 2: step k+1, row r-DIR
            T g[9];
            pull(g, ring1, f, k1a, k1b);
            collide(g);
            put(ring2, g, 0);
            keep(h2p, h2);
            keep(h2, g);
        }
        shift(k1a, k1b, f);
        const int y4 = n >= 7 ? r - 4 * DIR : ya;                     // phase 4: step k+3, row r-4 DIR
        if constexpr (P == 4) {
            T o[9] = {};
            if (act4 && n >= 7) {
                pull(o, ring3, h3, k3a, k3b);
                collide(o);
            }
            shift(k3a, k3b, h3);
            store(o, y4, act4 && n >= 7);
        } else {
            if (act4 && n >= 7) {
                T g[9];
                pull(g, ring3, h3, k3a, k3b);
                collide(g);
                put(ring4, g, 0);
                keep(h4p, h4);
                keep(h4, g);
            }
            shift(k3a, k3b, h3);
        }
        if constexpr (P == 6) {                                       // phase 6: step k+5, row r-7 DIR
            const int y6 = n >= 12 ? r - 7 * DIR : ya;
            T o[9] = {};
            if (act6 && n >= 12) {
                pull(o, ring5, h5, k5a, k5b);
                collide(o);
            }
            shift(k5a, k5b, h5);
            store(o, y6, act6 && n >= 12);
        }
        lds_barrier();
        if (p1) put(ring1, f, 1);                                     // phase-1 rows r and r-DIR
        if (act1 && n >= 1 && n <= nrows1) put(ring1, f1, 2);
        keep(f1, f);
        if (act3 && n >= 5) put(ring3, h3, 1);                        // phase-3 rows of the steps before
        if (act3 && n >= 6) put(ring3, h3p, 2);
        if (act3 && n >= 4 && n < nrows1) {                           // phase 3: step k+2, row r-2 DIR
            T g[9];
            pull(g, ring2, h2, k2a, k2b);
            collide(g);
            put(ring3, g, 0);
            keep(h3p, h3);
            keep(h3, g);
        }
        shift(k2a, k2b, h2);
        if constexpr (P == 5) {                                       // phase 5: step k+4, row r-5 DIR
            const int y5 = n >= 9 ? r - 5 * DIR : ya;
            T o[9] = {};
            if (act5 && n >= 9) {
                pull(o, ring4, h4, k4a, k4b);
                collide(o);
            }
            shift(k4a, k4b, h4);
            store(o, y5, act5 && n >= 9);
        } else if constexpr (P == 6) {
            if (act5 && n >= 10) put(ring5, h5, 1);                   // phase-5 rows of the steps before
            if (act5 && n >= 11) put(ring5, h5p, 2);
            if (act5 && n >= 9) {                                     // phase 5: step k+4, row r-5 DIR
                T g[9];
                pull(g, ring4, h4, k4a, k4b);
                collide(g);
                put(ring5, g, 0);
                keep(h5p, h5);
                keep(h5, g);
            }
            shift(k4a, k4b, h4);
        }
    }
}

// One workgroup of the P-step lane-aligned walk: strip s = b mod nstrip
// (edges on multiples of `unit` positions), segment b / nstrip (rows split
// evenly, odd segments walking upwards), b the workgroup number after the
// XCD-contiguous renumbering.
template <typename T, int NW, int P, class Args, class Collide>
__device__ __forceinline__ void astripP_workgroup(const Args &a, int remap, int nstrip, int nseg, int unit, int prio,
                                                  T *rings, Collide collide)
{
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
    if (seg & 1) stripP_walk_aligned<T, NW, -1, P>(a, x0, w, ya, rows, rings, prio, collide);
    else stripP_walk_aligned<T, NW, 1, P>(a, x0, w, ya, rows, rings, prio, collide);
}

// Threads of a strip-walk workgroup for TX-wide strips (the halo positions
// of the first steps spill into a fifth wavefront at TX = 256).
template <int TX>
constexpr int strip_threads() { return TX + 64; }

}  // namespace lat
}  // namespace erlbm
