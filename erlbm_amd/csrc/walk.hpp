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

// Threads of a strip-walk workgroup for TX-wide strips (the halo positions
// of the first steps spill into a fifth wavefront at TX = 256).
template <int TX>
constexpr int strip_threads() { return TX + 64; }

}  // namespace lat
}  // namespace erlbm
