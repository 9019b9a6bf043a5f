// lattice.hip -- plain (one lattice per handle) E-RLBM collide-stream path.
//
// Device layout: populations indexed f[i * si + row * sr + x], row = y + ny z.
// Two layouts share this form: population planes (si = n_voxels padded,
// sr = nx) and row blocks (si = nx, sr = q nx padded: the q populations of
// one lattice row are contiguous, so a step streams through a few adjacent
// row blocks instead of 2q far-apart planes).
// Solid voxels hold no state.  A step is one fused kernel: pull-stream
// (gather f_i from x - c_i, or the opposite population of x itself when
// x - c_i is solid: halfway bounce-back), boundary columns in registers,
// collision, optional density test, one coalesced store per population.
//
// Phase.  The reference pushes after colliding, so between steps its inif
// holds ref_k = S(C(BC(ref_{k-1}))).  A pull kernel stores the
// post-collision populations h_k = C(BC(ref_k)) instead, and the gather of
// step k+1 reproduces ref_{k+1} = S(h_k) bit for bit in registers.  The first
// step after an upload gathers with the identity (its input is ref_0), and
// downloads/monitors of "the state" gather S(h) on the fly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <type_traits>

#include "collide.hpp"
#include "common.hpp"
#include "lattice_impl.hpp"
#include "walk.hpp"

using namespace erlbm;
using namespace erlbm::lat;

namespace erlbm {

static thread_local std::string g_last_error;

void set_last_error(const char *what, hipError_t e)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
}
void set_last_error_msg(const char *msg) { g_last_error = msg; }

}  // namespace erlbm

extern "C" const char *erlbm_last_error(void) { return g_last_error.c_str(); }

extern "C" int erlbm_device_count(int *count)
{
    if (!count) return ERLBM_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return ERLBM_OK;
}

namespace {

// Gather the pre-collision populations of voxel (x,y,z): identity (first step
// after an upload) or pull-stream with periodic wrap and bounce-back.
template <int Q, bool STREAM, bool SOLID>
__device__ __forceinline__ void gather(const double *__restrict__ src, const Geometry &g, uint32_t nb,
                                       int x, int y, int z, double (&f)[Q])
{
    using V = VSet<Q>;
    const long long row = y + (long long)g.ny * z;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        if constexpr (!STREAM || i == 0) {
            f[i] = src[at(g, i, row, x)];
        } else {
            if (SOLID && (nb & (1u << i))) {
                f[i] = src[at(g, V::opp[i], row, x)];
            } else {
                const int xs = V::cx[i] == 0 ? x : wrap(x - V::cx[i], g.nx);
                const int ys = V::cy[i] == 0 ? y : wrap(y - V::cy[i], g.ny);
                const int zs = V::cz[i] == 0 ? z : wrap(z - V::cz[i], g.nz);
                f[i] = src[at(g, i, ys + (long long)g.ny * zs, xs)];
            }
        }
    });
}

// Boundary columns in registers: inlet, then outlet (the drivers' order,
// main_stability-map_obstacle.cpp:92-94).  The outlet copies its neighbour
// column's gathered (and, if that column is the inlet, Zou-He'd) state.
template <int Q, bool STREAM>
__device__ __forceinline__ void apply_bc(const StepArgs &a, int x, int y, int z, double (&f)[Q])
{
    using V = VSet<Q>;
    if (x == a.bc.inlet_x) zouhe<Q>(f, a.bc.u_in, a.bc.inlet_west);
    if (x == a.bc.outlet_x) {
        const int xn = x + a.bc.outlet_dx;
        const long long vn = xn + (long long)a.g.nx * (y + (long long)a.g.ny * z);
        const uint32_t nbn = a.nbr[vn];
        if (nbn & FLUID_BIT) {
            double fa[Q];
            gather<Q, STREAM, true>(a.src, a.g, nbn, xn, y, z, fa);
            if (xn == a.bc.inlet_x) zouhe<Q>(fa, a.bc.u_in, a.bc.inlet_west);
            const double rho_adj = density_exact<Q>(fa);
            const double fator = a.bc.rho_out / rho_adj;
            static_for<Q>([&](auto I) { f[(int)I] = fa[(int)I] * fator; });
        } else {
            static_for<Q>([&](auto I) { f[(int)I] = V::w[(int)I] * a.bc.rho_out; });
        }
    }
}

// General kernel: any geometry, boundary columns, body force, 1-D voxel
// mapping.
template <int Q, int COL, bool STREAM, bool FORCE = false>
__global__ __launch_bounds__(BLOCK) void k_step_general(StepArgs a)
{
    if (stopped(a.status, a.step)) return;
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
            double f[Q];
            gather<Q, STREAM, true>(a.src, a.g, nb, x, y, z, f);
            apply_bc<Q, STREAM>(a, x, y, z, f);
            double vx, vy, vz;
            collide_exact<Q, COL, FORCE>(f, a.k, vx, vy, vz, a.bf);
            if (a.dens_limit > 0.0) {
                const double d = density_exact<Q>(f);
                bad = !(fabs(d) <= a.dens_limit);
            }
            static_for<Q>([&](auto I) {
                __builtin_nontemporal_store(f[(int)I], &a.dst[at(a.g, (int)I, yz, x)]);
            });
        }
    }
    flag_density(a.status, a.step, bad);
}

// Periodic all-fluid kernel: 2-D grid (x chunks of BS*NPT nodes, rows).
// NPT nodes per thread (stride BS, so every load stays coalesced) raise the
// loads in flight per wave.  NTL 1 makes every load nontemporal; NTL 2 only
// the populations with c_x = 0, whose cache lines no other workgroup reads,
// so that the lines an x-shifted population shares with the neighbouring
// x chunk stay cached for that chunk's workgroup.  The stop test
// reads the launch-time value of the stop word after the loads are issued,
// so it costs no extra memory round trip.
template <int Q, int COL, bool STREAM, int NPT, int NTL, int BS>
__global__ __launch_bounds__(BS) void k_step_periodic(StepArgs a, int remap)
{
    using V = VSet<Q>;
    const unsigned nbx = gridDim.x, nb = gridDim.x * gridDim.y;
    unsigned b = blockIdx.y * nbx + blockIdx.x;
    if (remap) b = xcd_runs(b, nb);
    const int bx = (int)(b % nbx);
    const int row = (int)(b / nbx);
    const int y = row % a.g.ny;
    const int z = row / a.g.ny;
    double f[NPT][Q];
    int xs_[NPT];
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        const int x = bx * BS * NPT + k * BS + (int)threadIdx.x;
        xs_[k] = x;
        if (x < a.g.nx) {
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                long long src;
                if constexpr (!STREAM || i == 0) {
                    src = at(a.g, i, row, x);
                } else {
                    const int xs = V::cx[i] == 0 ? x : wrap(x - V::cx[i], a.g.nx);
                    const int ys = V::cy[i] == 0 ? y : wrap(y - V::cy[i], a.g.ny);
                    const int zs = V::cz[i] == 0 ? z : wrap(z - V::cz[i], a.g.nz);
                    src = at(a.g, i, ys + (long long)a.g.ny * zs, xs);
                }
                const double *p = &a.src[src];
                // NTL 1: every load nontemporal; 2: only the populations
                // whose cache lines no x-neighbouring workgroup reads again
                if constexpr (NTL == 1 || (NTL == 2 && V::cx[i] == 0)) f[k][i] = __builtin_nontemporal_load(p);
                else f[k][i] = *p;
            });
        }
    }
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        const int x = xs_[k];
        if (x < a.g.nx) {
            double vx, vy, vz;
            collide_sel<Q, COL>(f[k], a.k, vx, vy, vz);
            if (a.dens_limit > 0.0) {
                const double d = density_exact<Q>(f[k]);
                bad = bad || !(fabs(d) <= a.dens_limit);
            }
            static_for<Q>([&](auto I) {
                __builtin_nontemporal_store(f[k][(int)I], &a.dst[at(a.g, (int)I, row, x)]);
            });
        }
    }
    flag_density(a.status, a.step, bad);
}

// Two steps per launch (temporal blocking) for periodic all-fluid D2Q9 FP64
// lattices without stop tests.  A workgroup owns a TX x TY tile: step k is
// computed on the tile plus a one-node halo (pull from HBM, collide) into
// LDS, step k+1 on the tile from LDS (pull, collide) to HBM.  HBM is read
// about (TX+2)(TY+2)/(TX TY) times and written once per two steps instead
// of once each per step; the arithmetic is the single-step kernel's, so the
// populations are bitwise the same.
template <int TX, int TY>
struct Fuse2 {
    static constexpr int EX = TX + 2, EY = TY + 2, NE = EX * EY;
    static constexpr int PA = (NE + BLOCK - 1) / BLOCK;       // halo positions per thread
    static constexpr int PB = (TX * TY + BLOCK - 1) / BLOCK;  // tile positions per thread
    static constexpr size_t LDS_BYTES = sizeof(double) * 9 * NE;
};

template <int COL, int TX, int TY>
__global__ __launch_bounds__(BLOCK) void k_step2_d2q9(StepArgs a, int remap)
{
    using V = VSet<9>;
    using F = Fuse2<TX, TY>;
    extern __shared__ double h1[];   // [9][NE]: post-collision of step k on the haloed tile
    // 1-D grid of tiles, x fastest; remap: XCD r (workgroups r, r+8, ...)
    // takes a contiguous run of tiles, so tiles sharing halo lines share an L2
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if (remap) b = xcd_runs(b, nb);
    const int ntx = a.g.nx / TX;
    const int x0 = (int)(b % (unsigned)ntx) * TX;
    const int y0 = (int)(b / (unsigned)ntx) * TY;
    const int tid = (int)threadIdx.x;
    const int nx = a.g.nx, ny = a.g.ny;
    // ---- step k on the haloed tile: loads of every position first
    double f[F::PA][9];
#pragma unroll
    for (int p = 0; p < F::PA; p++) {
        const int e = tid + p * BLOCK;
        if (e < F::NE) {
            const int x = wrap(x0 - 1 + e % F::EX, nx);
            const int y = wrap(y0 - 1 + e / F::EX, ny);
            static_for<9>([&](auto I) {
                constexpr int i = I;
                const int xs = V::cx[i] == 0 ? x : wrap(x - V::cx[i], nx);
                const int ys = V::cy[i] == 0 ? y : wrap(y - V::cy[i], ny);
                const double *q = &a.src[at(a.g, i, ys, xs)];
                if constexpr (V::cx[i] == 0) f[p][i] = __builtin_nontemporal_load(q);
                else f[p][i] = *q;
            });
        }
    }
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
#pragma unroll
    for (int p = 0; p < F::PA; p++) {
        const int e = tid + p * BLOCK;
        if (e < F::NE) {
            double vx, vy, vz;
            collide_sel<9, COL>(f[p], a.k, vx, vy, vz);
            static_for<9>([&](auto I) { h1[(int)I * F::NE + e] = f[p][(int)I]; });
        }
    }
    __syncthreads();
    // ---- step k+1 on the tile, pulled from LDS
#pragma unroll
    for (int p = 0; p < F::PB; p++) {
        const int o = tid + p * BLOCK;
        if (o < TX * TY) {
            const int ox = o % TX, oy = o / TX;
            const int e = (oy + 1) * F::EX + (ox + 1);
            double g[9];
            static_for<9>([&](auto I) {
                constexpr int i = I;
                g[i] = h1[i * F::NE + e - V::cx[i] - V::cy[i] * F::EX];
            });
            double vx, vy, vz;
            collide_sel<9, COL>(g, a.k, vx, vy, vz);
            static_for<9>([&](auto I) {
                __builtin_nontemporal_store(g[(int)I], &a.dst[at(a.g, (int)I, y0 + oy, x0 + ox)]);
            });
        }
    }
}

// Two steps per launch, walking a TX-wide x-strip down SEG output rows:
// only a ring of recent step-k rows (walk.hpp) lives in LDS, so the tile
// can be long in y (halo 2 rows per SEG) and wide in x (halo 2 columns per
// TX) without the LDS of a whole 2-D tile.  Thread t computes step-k position
// x0-1+t (t < TX+2) and step-(k+1) position x0+t (t < TX); the loads of the
// next PD step-k rows are in flight while the current output row is
// computed.  (PD = 2 and 4 -- 138 and 202 VGPRs -- measured 14-25% slower
// than PD = 1 at L = 2048, so PD = 1 is the one instantiated.)

template <int COL, int TX, int SEG, int PD>
__global__ __launch_bounds__(strip_threads<TX>()) void k_step2_strip(StepArgs a, int remap)
{
    static_assert(PD >= 1 && SEG % PD == 0, "rows in flight must divide the walk");
    using V = VSet<9>;
    constexpr int EX = TX + 2;
    extern __shared__ double ring[];   // [RING_ROWS][EX]
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
    // Memory instructions take no branches (threads without a position get
    // out-of-range offsets), so the wait for a row's loads at the top of the
    // walk never has to drain the stores issued after them.
    const __amdgpu_buffer_rsrc_t rs = pop_buffer<8>(a.src, a.g), rd = pop_buffer<8>(a.dst, a.g);
    const unsigned oob1 = act1 ? 0u : BUF_OOB;
    const unsigned vo[3] = {8u * wrap(x1 + 1, nx) | oob1, 8u * x1 | oob1, 8u * wrap(x1 - 1, nx) | oob1};
    const unsigned vst = t < TX ? 8u * (x0 + t) : 0u;   // column of the store, or-ed with
    const unsigned sto = t < TX ? 0u : BUF_OOB;          // the out-of-range bit
    // pulls of step-k row y (logical, -1 .. ny + PD); skip: no memory access
    auto load_row = [&](int y, double (&f)[9], bool skip) {
        const unsigned off = skip ? BUF_OOB : 0u;
        const int yw = wrap(y, ny);
        static_for<9>([&](auto I) {
            constexpr int i = I;
            const int ys = V::cy[i] == 0 ? yw : wrap(yw - V::cy[i], ny);
            f[i] = buf_load<V::cx[i] == 0 ? BUF_NT : 0>(rs, vo[V::cx[i] + 1] | off, (int)(8 * at(a.g, i, ys, 0)));
        });
    };
    // ring row of population i of step-k row y
    auto ring_row = [&](int i, int y) { return ring_base<1>(i) + (y - ya + 1) % ring_depth<1>(i); };
    auto collide_to_ring = [&](int y, double (&f)[9]) {
        if (!act1) return;
        double vx, vy, vz;
        collide_sel<9, COL>(f, a.k, vx, vy, vz);
        static_for<9>([&](auto I) { ring[ring_row((int)I, y) * EX + t] = f[(int)I]; });
    };
    auto output_row = [&](int y) {            // step k+1, row y, from the ring
        double g[9] = {};
        if (t < TX) {
            static_for<9>([&](auto I) {
                constexpr int i = I;
                g[i] = ring[ring_row(i, y - V::cy[i]) * EX + (t + 1 - V::cx[i])];
            });
            double vx, vy, vz;
            collide_sel<9, COL>(g, a.k, vx, vy, vz);
        }
        static_for<9>([&](auto I) { buf_store_nt(g[(int)I], rd, vst | sto, (int)(8 * at(a.g, (int)I, y, 0))); });
    };
    // PD rows of step-k loads in flight: register set J holds row y+1 when
    // output row y = ya + r, r = J (mod PD), is next
    double fq[PD][9], fb[9];
    load_row(ya - 1, fq[0], false);
    load_row(ya, fb, false);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    collide_to_ring(ya - 1, fq[0]);
    collide_to_ring(ya, fb);
    static_for<PD>([&](auto J) { load_row(ya + 1 + (int)J, fq[(int)J], false); });
    nine_dropped_stores<double>(rd);
    // Row y+1 overwrites rows that output row y-1 read in the previous
    // iteration, hence the second barrier.  (Four slots per population and
    // one barrier per row measured the same, with more LDS.)
    for (int r = 0; r < SEG; r += PD) {
        static_for<PD>([&](auto J) {
            const int y = ya + r + (int)J;
            collide_to_ring(y + 1, fq[(int)J]);                                   // step k, row y+1
            load_row(y + 1 + PD, fq[(int)J], y + 1 + PD > ya + SEG);            // in flight meanwhile
            lds_barrier();
            output_row(y);
            lds_barrier();
        });
    }
}

// Three steps per launch, the strip walk with two LDS rings (the 18-row
// layout, or with LEAN the 12-row one of walk.hpp, which writes the
// populations a pull reads from the older two rows one barrier late, from
// registers, so three workgroups share a CU instead of two): the step-k
// results of positions x0-2 .. x0+TX+1 (ring 1) and the step-(k+1) results
// of x0-1 .. x0+TX (ring 2), each keeping a population only while it is
// read (18 rows: 3 / 2 / 1 row slots for the populations whose source row
// is the oldest / middle / newest of the three a pull reads, which depends
// on the walking direction).  Step r of the walk collides step k of row r
// into ring 1 (phase 1), step k+1 of the row before it from ring 1 into
// ring 2 (phase 2) and step k+2 of the row two before from ring 2 to HBM
// (phase 3); the loads of the next row are in flight meanwhile.  Two
// barriers per row: phase 1 of the next row overwrites ring-1 rows that
// phase 2 read, and phase 2 ring-2 rows that phase 3 of the previous row
// read, each across a barrier.  Odd segments walk upwards, so that the two
// halo rows neighbouring segments share are read by both at the start (or
// both at the end) of their walks, while they are still in L2.  The strip
// is w <= TX positions wide (the rings are laid out for TX).

template <int COL, int TX, int DIR, bool LEAN>
__device__ __forceinline__ void strip3_walk(const StepArgs &a, int x0, int w, int ya, int rows, double *ring1,
                                            double *ring2)
{
    using V = VSet<9>;
    constexpr int E1 = TX + 4, E2 = TX + 2;
    const int t = (int)threadIdx.x;
    const int nx = a.g.nx, ny = a.g.ny;
    const bool act1 = t < w + 4, act2 = t < w + 2, act3 = t < w;
    const int x1 = act1 ? wrap(x0 - 2 + t, nx) : 0;
    const __amdgpu_buffer_rsrc_t rs = pop_buffer<8>(a.src, a.g), rd = pop_buffer<8>(a.dst, a.g);
    const unsigned oob1 = act1 ? 0u : BUF_OOB;
    const unsigned vo[3] = {8u * wrap(x1 + 1, nx) | oob1, 8u * x1 | oob1, 8u * wrap(x1 - 1, nx) | oob1};
    const unsigned vst = act3 ? 8u * (x0 + t) : 0u;
    auto load_row = [&](int y, double (&f)[9], bool skip) {   // pulls of step-k row y (-3 .. ny+2)
        const unsigned off = skip ? BUF_OOB : 0u;
        const int yw = wrap(y, ny);
        static_for<9>([&](auto I) {
            constexpr int i = I;
            const int ys = V::cy[i] == 0 ? yw : wrap(yw - V::cy[i], ny);
            f[i] = buf_load<V::cx[i] == 0 ? BUF_NT : 0>(rs, vo[V::cx[i] + 1] | off, (int)(8 * at(a.g, i, ys, 0)));
        });
    };
    auto rrow = [&](int i, int y) {
        if constexpr (LEAN) return lean_base<DIR>(i) + (y - ya + 2) % lean_depth<DIR>(i);
        else return ring_base<DIR>(i) + (y - ya + 2) % ring_depth<DIR>(i);
    };
    // lean rings: the populations read from the newest of the three rows a
    // pull reads (c_y = -DIR) are written at once, the others one barrier
    // later
    auto now = [](int i) { return !LEAN || V::cy[i] * DIR < 0; };
    auto collide = [&](double (&f)[9]) {
        double vx, vy, vz;
        collide_sel<9, COL>(f, a.k, vx, vy, vz);
    };
    // step-k rows ya-2 .. ya+rows+1 in walking order
    const int r0 = DIR > 0 ? ya - 2 : ya + rows + 1;
    double fq[9];
    double h2[9];   // lean: the previous walking step's step-(k+1) row, still to be written to ring 2
    load_row(r0, fq, false);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    nine_dropped_stores<double>(rd);
    for (int n = 0; n < rows + 4; n++) {
        const int r = r0 + DIR * n;
        double f[9];
        static_for<9>([&](auto I) { f[(int)I] = fq[(int)I]; });
        load_row(r + DIR, fq, n + 1 >= rows + 4);                    // in flight meanwhile
        if (act1) {                                                   // phase 1: step k, row r
            collide(f);
            static_for<9>([&](auto I) {
                if (now(I)) ring1[rrow((int)I, r) * E1 + t] = f[(int)I];
            });
        }
        lds_barrier();
        if constexpr (LEAN) {
            if (act2 && n >= 3) {                                     // the rest of row r-2 DIR
                static_for<9>([&](auto I) {
                    if (!now(I)) ring2[rrow((int)I, r - 2 * DIR) * E2 + t] = h2[(int)I];
                });
            }
        }
        if (act2 && n >= 2) {                                         // phase 2: step k+1, row r-DIR
            const int y = r - DIR;
            double g[9];
            static_for<9>([&](auto I) {
                constexpr int i = I;
                g[i] = ring1[rrow(i, y - V::cy[i]) * E1 + (t + 1 - V::cx[i])];
            });
            collide(g);
            static_for<9>([&](auto I) {
                if (now(I)) ring2[rrow((int)I, y) * E2 + t] = g[(int)I];
                if constexpr (LEAN) h2[(int)I] = g[(int)I];
            });
        }
        lds_barrier();
        if constexpr (LEAN) {
            if (act1) {                                               // the rest of row r
                static_for<9>([&](auto I) {
                    if (!now(I)) ring1[rrow((int)I, r) * E1 + t] = f[(int)I];
                });
            }
        }
        const bool out = n >= 4;                                      // phase 3: step k+2, row r-2 DIR
        const int y = out ? r - 2 * DIR : ya;
        double g[9] = {};
        if (act3 && out) {
            static_for<9>([&](auto I) {
                constexpr int i = I;
                g[i] = ring2[rrow(i, y - V::cy[i]) * E2 + (t + 1 - V::cx[i])];
            });
            collide(g);
        }
        const unsigned sto = act3 && out ? 0u : BUF_OOB;
        static_for<9>([&](auto I) { buf_store_nt(g[(int)I], rd, vst | sto, (int)(8 * at(a.g, (int)I, y, 0))); });
    }
}

// One launch: nstrip x nseg workgroups, the rows of a strip split evenly
// among its segments.
template <int COL, int TX, bool LEAN>
__global__ __launch_bounds__(strip_threads<TX>()) void k_step3_strip(StepArgs a, int remap, int nseg)
{
    constexpr int E1 = TX + 4;
    constexpr int ROWS = LEAN ? LEAN_ROWS : RING_ROWS;
    extern __shared__ double rings[];   // ring 1 [ROWS][TX+4], ring 2 [ROWS][TX+2]
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if (remap) b = xcd_runs(b, nb);
    const int nstrip = a.g.nx / TX;
    const int x0 = (int)(b % (unsigned)nstrip) * TX;
    const int seg = (int)(b / (unsigned)nstrip);
    const int ya = (int)((long long)seg * a.g.ny / nseg);
    const int rows = (int)((long long)(seg + 1) * a.g.ny / nseg) - ya;
    if (seg & 1) strip3_walk<COL, TX, -1, LEAN>(a, x0, TX, ya, rows, rings, rings + ROWS * E1);
    else strip3_walk<COL, TX, 1, LEAN>(a, x0, TX, ya, rows, rings, rings + ROWS * E1);
}

// The same walk on strips of any width: NW wavefronts per workgroup, nstrip
// strips of at most 64 NW - 4 positions, widths split evenly (a strip of
// that width fills every wavefront in all three phases, where a 256-wide
// strip leaves its fifth wavefront with 4 or 2 active lanes in phases 1
// and 2), lean rings.
template <int COL, int NW>
__global__ __launch_bounds__(64 * NW) void k_step3_gstrip(StepArgs a, int remap, int nstrip, int nseg)
{
    constexpr int TX = 64 * NW - 4;
    extern __shared__ double rings[];   // ring 1 [LEAN_ROWS][TX+4], ring 2 [LEAN_ROWS][TX+2]
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if (remap) b = xcd_runs(b, nb);
    const int s = (int)(b % (unsigned)nstrip);
    const int x0 = (int)((long long)s * a.g.nx / nstrip);
    const int w = (int)((long long)(s + 1) * a.g.nx / nstrip) - x0;
    const int seg = (int)(b / (unsigned)nstrip);
    const int ya = (int)((long long)seg * a.g.ny / nseg);
    const int rows = (int)((long long)(seg + 1) * a.g.ny / nseg) - ya;
    if (seg & 1) strip3_walk<COL, TX, -1, true>(a, x0, w, ya, rows, rings, rings + LEAN_ROWS * (TX + 4));
    else strip3_walk<COL, TX, 1, true>(a, x0, w, ya, rows, rings, rings + LEAN_ROWS * (TX + 4));
}

// D3Q19, two steps per launch: a workgroup owns a TX x TY tile of columns
// and walks it along z through a segment of planes.  Phase 1 pulls step k
// of plane r from HBM and collides it on the haloed (TX+2) x (TY+2) tile
// into an LDS ring of planes (lean layout: the populations a pull reads
// from the newest of its three planes at once, the others after the
// barrier that follows the step-(k+1) reads, so a population needs only
// 2 / 1 / 1 plane slots); phase 2 pulls step k+1 of plane r-DIR from the
// ring, collides it and stores it.  The loads of plane r+DIR are in flight
// meanwhile.  Each plane of a pull is addressed through a raw buffer based
// at that plane (the whole lattice may exceed 32-bit offsets).  Odd
// segments walk downwards so the halo planes neighbouring segments share
// are read while in L2.
template <int TX, int TY>
constexpr int zwalk_threads() { return ((TX + 2) * (TY + 2) + 63) / 64 * 64; }

// NTP: which step-k loads carry the nontemporal hint (0 none, the default:
// neighbouring tiles read a tile's halo lines again, and with the hint on
// the c_x = 0 populations (1) the reads were 1.59x the algorithmic bytes
// against 1.31x without, 1.200 against 1.100 ms per launch at 256^3;
// 2: c_x = c_y = 0 only, 1.43x, 1.134 ms).
template <int COL, int TX, int TY, int DIR, int NTP = 0>
__device__ __forceinline__ void zwalk(const StepArgs &a, int x0, int y0, int za, int rows, double *ring)
{
    using V = VSet<19>;
    constexpr int HX = TX + 2, NP = HX * (TY + 2);
    const int t = (int)threadIdx.x;
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const bool act1 = t < NP, act2 = t < TX * TY;
    const int hx = act1 ? t % HX : 0, hy = act1 ? t / HX : 0;
    const int x = wrap(x0 - 1 + hx, nx), y = wrap(y0 - 1 + hy, ny);
    // byte offsets within a plane: column and row of each c_x / c_y
    const unsigned oob1 = act1 ? 0u : BUF_OOB;
    const unsigned xo[3] = {8u * wrap(x + 1, nx), 8u * x, 8u * wrap(x - 1, nx)};
    const unsigned yo[3] = {8u * (unsigned)(wrap(y + 1, ny) * a.g.sr), 8u * (unsigned)(y * a.g.sr),
                            8u * (unsigned)(wrap(y - 1, ny) * a.g.sr)};
    const unsigned plane_bytes = (unsigned)(8 * (long long)ny * a.g.sr);
    auto plane_buf = [&](const double *p, int z) {
        return __builtin_amdgcn_make_buffer_rsrc((void *)(p + (long long)wrap(z, nz) * ny * a.g.sr), (short)0,
                                                 (int)plane_bytes, BUF_DWORD3);
    };
    auto load_plane = [&](int z, double (&f)[19], bool skip) {   // pulls of step-k plane z (-2 .. nz+1)
        const unsigned off = skip ? BUF_OOB : 0u;
        const __amdgpu_buffer_rsrc_t rp[3] = {plane_buf(a.src, z + 1), plane_buf(a.src, z), plane_buf(a.src, z - 1)};
        static_for<19>([&](auto I) {
            constexpr int i = I;
            constexpr bool nt = NTP == 1 ? V::cx[i] == 0 : (NTP == 2 ? V::cx[i] == 0 && V::cy[i] == 0 : false);
            f[i] = buf_load<nt ? BUF_NT : 0>(rp[V::cz[i] + 1], (yo[V::cy[i] + 1] + xo[V::cx[i] + 1]) | oob1 | off,
                                             (int)(8 * i * a.g.si));
        });
    };
    auto slot = [&](int i, int z) { return zring_base<DIR>(i) + (z - za + 2) % zring_depth<DIR>(i); };
    auto now = [](int i) { return V::cz[i] * DIR < 0; };
    auto collide = [&](double (&f)[19]) {
        double vx, vy, vz;
        collide_sel<19, COL>(f, a.k, vx, vy, vz);
    };
    const int ox = act2 ? t % TX : 0, oy = act2 ? t / TX : 0;
    const unsigned vst = 8u * (unsigned)((y0 + oy) * a.g.sr + x0 + ox);
    const unsigned sto = act2 ? 0u : BUF_OOB;
    // step-k planes za-1 .. za+rows in walking order
    const int r0 = DIR > 0 ? za - 1 : za + rows;
    double fq[19];
    load_plane(r0, fq, false);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    dropped_stores<double, 19>(plane_buf(a.dst, za));
    for (int n = 0; n < rows + 2; n++) {
        const int r = r0 + DIR * n;
        double f[19];
        static_for<19>([&](auto I) { f[(int)I] = fq[(int)I]; });
        load_plane(r + DIR, fq, n + 1 >= rows + 2);                  // in flight meanwhile
        if (act1) {                                                   // phase 1: step k, plane r
            collide(f);
            static_for<19>([&](auto I) {
                if (now(I)) ring[slot(I, r) * NP + t] = f[(int)I];
            });
        }
        lds_barrier();
        const bool out = n >= 2;                                      // phase 2: step k+1, plane r-DIR
        const int zo = out ? r - DIR : za;
        double g[19] = {};
        if (act2 && out) {
            static_for<19>([&](auto I) {
                constexpr int i = I;
                g[i] = ring[slot(i, zo - V::cz[i]) * NP + (oy + 1 - V::cy[i]) * HX + (ox + 1 - V::cx[i])];
            });
        }
        lds_barrier();
        if (act1) {                                                   // the rest of plane r
            static_for<19>([&](auto I) {
                if (!now(I)) ring[slot(I, r) * NP + t] = f[(int)I];
            });
        }
        if (act2 && out) collide(g);
        const __amdgpu_buffer_rsrc_t rd = plane_buf(a.dst, zo);
        const unsigned so = out ? sto : BUF_OOB;
        static_for<19>([&](auto I) {
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(buf_u2, g[(int)I]), rd, vst | so,
                                                  (int)(8 * (int)I * a.g.si), BUF_NT);
        });
    }
}

template <int COL, int TX, int TY, int NTP = 0>
__global__ __launch_bounds__((zwalk_threads<TX, TY>())) void k_step2_zwalk(StepArgs a, int remap, int nseg)
{
    extern __shared__ double zring[];   // [ZRING_PLANES][(TX+2) (TY+2)]
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if (remap) b = xcd_runs(b, nb);
    const int tx = a.g.nx / TX, ntile = tx * (a.g.ny / TY);
    const int tile = (int)(b % (unsigned)ntile), seg = (int)(b / (unsigned)ntile);
    const int za = (int)((long long)seg * a.g.nz / nseg);
    const int rows = (int)((long long)(seg + 1) * a.g.nz / nseg) - za;
    const int x0 = (tile % tx) * TX, y0 = (tile / tx) * TY;
    if (seg & 1) zwalk<COL, TX, TY, -1, NTP>(a, x0, y0, za, rows, zring);
    else zwalk<COL, TX, TY, 1, NTP>(a, x0, y0, za, rows, zring);
}

// Lane l of a wavefront receives the value of lane l-1 (WAVE_SHR1) or l+1
// (WAVE_SHL1): one DPP move per dword, no LDS.  Lanes with no source keep 0.
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;
template <int CTRL>
__device__ __forceinline__ double wave_shift(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Two steps per launch, one wavefront per strip, with neither LDS nor
// barriers: every wavefront walks its strip on its own.  Lane l computes
// step k at x = x0-1+l and keeps the step-k results of the last three rows
// in registers; the step-(k+1) pull takes a c_y = +-1 population from the
// row before / after in registers and a c_x = +-1 one from lane l-1 / l+1
// by a DPP move.  Lanes 1..62 write step k+1, so a strip is 62 wide (the
// last one ragged); a wavefront walks a segment of the strip's rows, and
// the loads of the next PD step-k rows are in flight while a row is output.
// Odd segments walk upwards: the two step-k rows that neighbouring segments
// both read are then read by both at the start or both at the end of their
// walks, while they are still in L2.
constexpr int WAVE_STRIP = 62;

template <int COL, int PD, int DIR>
__device__ __forceinline__ void wave_walk(const StepArgs &a, int x0, int ya, int rows)
{
    using V = VSet<9>;
    const int nx = a.g.nx, ny = a.g.ny;
    const int lane = (int)(threadIdx.x & 63);
    const int xk = wrap(x0 - 1 + lane, nx);     // step-k column of this lane
    const bool out = lane >= 1 && lane <= WAVE_STRIP && x0 + lane - 1 < nx;
    const int y0 = DIR > 0 ? ya : ya + rows - 1;   // first output row
    // Raw buffers over the population arrays: the per-lane column is the
    // VGPR offset, the wave-uniform plane + row the SGPR offset of a load
    // (the launcher checks that every byte offset is below 2^31).
    const __amdgpu_buffer_rsrc_t rs = pop_buffer<8>(a.src, a.g), rd = pop_buffer<8>(a.dst, a.g);
    const unsigned vo[3] = {8u * wrap(xk + 1, nx), 8u * xk, 8u * wrap(xk - 1, nx)};   // columns of c_x = -1, 0, +1
    const unsigned vst = out ? 8u * (x0 + lane - 1) : 0u;   // column of the store, or-ed with
    const unsigned sto = out ? 0u : BUF_OOB;                // the out-of-range bit
    // pulls of step-k row y (logical, -1 - PD .. ny + PD); skip: no memory access
    auto load_row = [&](int y, double (&f)[9], bool skip) {
        const unsigned off = skip ? BUF_OOB : 0u;
        const int yw = wrap(y, ny);
        static_for<9>([&](auto I) {
            constexpr int i = I;
            const int ys = V::cy[i] == 0 ? yw : wrap(yw - V::cy[i], ny);
            const int so = (int)(8 * at(a.g, i, ys, 0));
            f[i] = buf_load<V::cx[i] == 0 ? BUF_NT : 0>(rs, vo[V::cx[i] + 1] | off, so);
        });
    };
    auto collide = [&](double (&f)[9]) {
        double vx, vy, vz;
        collide_sel<9, COL>(f, a.k, vx, vy, vz);
    };
    // step-k results of output row y: hb row y-DIR, h0 row y, hn row y+DIR
    double hb[9], h0[9], hn[9], fq[PD][9];
    load_row(y0 - DIR, hb, false);
    load_row(y0, h0, false);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    static_for<PD>([&](auto J) { load_row(y0 + DIR * (1 + (int)J), fq[(int)J], (int)J >= rows); });
    collide(hb);
    collide(h0);
    // nine dropped stores: entering the loop, the prefetched rows are then
    // followed by nine stores as at its back edge, so the wait for them at
    // the top of the loop can leave the previous row's stores outstanding
    nine_dropped_stores<double>(rd);
    // output row number k of the walk from register set j
    auto row = [&](int k, auto J) {
        constexpr int j = J;
        const int y = y0 + DIR * k;
        static_for<9>([&](auto I) { hn[(int)I] = fq[j][(int)I]; });
        collide(hn);                                                       // step k, row y+DIR
        load_row(y + DIR * (1 + PD), fq[j], k + 1 + PD > rows);           // in flight meanwhile
        double g[9];
        static_for<9>([&](auto I) {
            constexpr int i = I;
            // population i of row y pulls from row y - c_y
            const double v = V::cy[i] == DIR ? hb[i] : (V::cy[i] == -DIR ? hn[i] : h0[i]);
            if constexpr (V::cx[i] > 0) g[i] = wave_shift<DPP_WAVE_SHR1>(v);
            else if constexpr (V::cx[i] < 0) g[i] = wave_shift<DPP_WAVE_SHL1>(v);
            else g[i] = v;
        });
        collide(g);                                                        // step k+1, row y
        static_for<9>([&](auto I) {
            buf_store_nt(g[(int)I], rd, vst | sto, (int)(8 * at(a.g, (int)I, y, 0)));
        });
        static_for<9>([&](auto I) {
            hb[(int)I] = h0[(int)I];
            h0[(int)I] = hn[(int)I];
        });
    };
    int k = 0;
    for (; k + PD <= rows; k += PD) static_for<PD>([&](auto J) { row(k + (int)J, J); });
    static_for<PD - 1>([&](auto J) {    // rows % PD left over
        if (k + (int)J < rows) row(k + (int)J, J);
    });
}

// One launch: nstrip x nseg wavefronts, the rows of a strip split evenly
// among its nseg segments (the launcher sizes nseg so that every wavefront
// is resident at once: a second, partial round of wavefronts would walk its
// rows at latency-bound speed after the first one is done).
template <int COL, int PD, int WPB, int WPE>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_step2_wave(StepArgs a, int nstrip, int nseg, int remap)
{
    const int ny = a.g.ny;
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if (remap) b = xcd_runs(b, nb);
    // the wavefront index is uniform across the wavefront; readfirstlane
    // lets the compiler keep everything derived from it in SGPRs
    const int wv = (int)b * WPB + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wv >= nstrip * nseg) return;
    const int seg = wv / nstrip;
    const int x0 = (wv % nstrip) * WAVE_STRIP;
    const int ya = (int)((long long)seg * ny / nseg);
    const int rows = (int)((long long)(seg + 1) * ny / nseg) - ya;
    if (seg & 1) wave_walk<COL, PD, -1>(a, x0, ya, rows);
    else wave_walk<COL, PD, 1>(a, x0, ya, rows);
}

// D2Q9 periodic kernel, two x-adjacent nodes per thread: every population
// plane is read with one 16-byte load per thread (8-byte aligned when the
// plane is x-shifted; multi-dword global loads need 4-byte alignment) and
// written with one aligned 16-byte store.  nx must be even.
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
typedef double d2a __attribute__((ext_vector_type(2)));

template <int COL, bool STREAM, bool NTL>
__global__ __launch_bounds__(BLOCK) void k_step_d2q9_pair(StepArgs a, int remap)
{
    using V = VSet<9>;
    if (stopped(a.status, a.step)) return;
    const unsigned nbx = gridDim.x, nb = gridDim.x * gridDim.y;
    unsigned b = blockIdx.y * nbx + blockIdx.x;
    if (remap) b = xcd_runs(b, nb);
    const int bx = (int)(b % nbx);
    const int y = (int)(b / nbx);
    const int x0 = 2 * (bx * BLOCK + (int)threadIdx.x);
    bool bad = false;
    if (x0 < a.g.nx) {
        const int nx = a.g.nx, ny = a.g.ny;
        double f[2][9];
        static_for<9>([&](auto I) {
            constexpr int i = I;
            int ys = y;
            if constexpr (STREAM && V::cy[i] != 0) ys = wrap(y - V::cy[i], ny);
            const double *row = a.src + at(a.g, i, ys, 0);
            const int xs = (STREAM ? x0 - V::cx[i] : x0);
            if (xs >= 0 && xs + 1 < nx) {
                const d2u *p = reinterpret_cast<const d2u *>(row + xs);
                d2u v;
                if constexpr (NTL) v = __builtin_nontemporal_load(p);
                else v = *p;
                f[0][i] = v.x;
                f[1][i] = v.y;
            } else {
                f[0][i] = row[wrap(xs, nx)];
                f[1][i] = row[wrap(xs + 1, nx)];
            }
        });
#pragma unroll
        for (int k = 0; k < 2; k++) {
            double vx, vy, vz;
            collide_sel<9, COL>(f[k], a.k, vx, vy, vz);
            if (a.dens_limit > 0.0) {
                const double d = density_exact<9>(f[k]);
                bad = bad || !(fabs(d) <= a.dens_limit);
            }
        }
        static_for<9>([&](auto I) {
            constexpr int i = I;
            d2a w = {f[0][i], f[1][i]};
            __builtin_nontemporal_store(w, reinterpret_cast<d2a *>(&a.dst[at(a.g, i, y, x0)]));
        });
    }
    flag_density(a.status, a.step, bad);
}

// ---- monitors -------------------------------------------------------------
struct MonArgs {
    const double *__restrict__ src;
    Geometry g;
    const uint32_t *__restrict__ nbr;   // null: all fluid
    const int *__restrict__ vox_of_fluid;  // null: identity
    int n_fluid;
    double *ke_node;      // per fluid site 0.5|u|^2 (ordered mode), or null
    double *partial_ke;   // per block
    double *partial_rho;  // per block max |rho|
    int *nonfinite;       // set if some rho non-finite
};

template <int Q, bool STREAM>
__global__ __launch_bounds__(BLOCK) void k_monitor(MonArgs a)
{
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    double e = 0.0, r_abs = 0.0;
    bool nf = false;
    if (p < a.n_fluid) {
        const long long v = a.vox_of_fluid ? a.vox_of_fluid[p] : p;
        const int x = (int)(v % a.g.nx);
        const long long yz = v / a.g.nx;
        const int y = (int)(yz % a.g.ny);
        const int z = (int)(yz / a.g.ny);
        double f[Q];
        if (a.nbr) gather<Q, STREAM, true>(a.src, a.g, a.nbr[v], x, y, z, f);
        else gather<Q, STREAM, false>(a.src, a.g, 0u, x, y, z, f);
        double rho, vx, vy, vz;
        moments_exact<Q>(f, rho, vx, vy, vz);
        e = ke_term<Q>(vx, vy, vz);
        const double d = density_exact<Q>(f);
        r_abs = fabs(d);
        nf = !isfinite(d);
        if (a.ke_node) a.ke_node[p] = e;
    }
    if (__any(nf) && (threadIdx.x & 63) == 0) atomicOr(a.nonfinite, 1);
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
        a.partial_ke[blockIdx.x] = te;
        a.partial_rho[blockIdx.x] = tr;
    }
}

// Final reduction, one workgroup.  ordered: the sum of ke_node in fluid
// order, ((0 + e_0) + e_1) + ... as the reference's loop forms it: waves 1-3
// stage the next chunk of ke_node into LDS (coalesced) while lane 0 of wave 0
// adds the current one in order, so the sum runs at the FP64 add latency, not
// the memory latency.  Else: a fixed-shape tree over the block partials
// (thread t sums partials t, t + FIN_THREADS, ... in order; then the wave
// tree; then the waves in order) -- reproducible, and not the reference's
// rounding.  If check_step != NO_STOP, applies the Taylor-Green map test
// (E > factor * E0 or non-finite -> stop before step check_step) with E0
// read from *e0.  out[0] = E, out[1] = max |rho|.
constexpr int FIN_THREADS = 256;
constexpr int FIN_CHUNK = 2048;   // doubles per LDS stage (two stages: 32 KB)

__global__ __launch_bounds__(FIN_THREADS) void k_finish(const double *ke_node, const double *partial_ke,
                                                        const double *partial_rho, int n_fluid, int n_blocks,
                                                        int ordered, double *out, const double *e0, double factor,
                                                        unsigned long long check_step, unsigned long long *status)
{
    __shared__ double stage[2][FIN_CHUNK];
    __shared__ double red_e[FIN_THREADS / 64], red_r[FIN_THREADS / 64];
    const int t = threadIdx.x;
    double e = 0.0;
    if (ordered) {
        const int nchunk = (n_fluid + FIN_CHUNK - 1) / FIN_CHUNK;
        for (int j = t; j < FIN_CHUNK && j < n_fluid; j += FIN_THREADS) stage[0][j] = ke_node[j];
        __syncthreads();
        for (int c = 0; c < nchunk; c++) {
            if (t >= 64) {
                const long long base = (long long)(c + 1) * FIN_CHUNK;
                for (int j = t - 64; j < FIN_CHUNK && base + j < n_fluid; j += FIN_THREADS - 64)
                    stage[(c + 1) & 1][j] = ke_node[base + j];
            } else if (t == 0) {
                const int m = min(FIN_CHUNK, n_fluid - c * FIN_CHUNK);
                const double *b = stage[c & 1];
                for (int j = 0; j < m; j++) e = e + b[j];
            }
            __syncthreads();
        }
    } else {
        for (int b = t; b < n_blocks; b += FIN_THREADS) e = e + partial_ke[b];
        e = wave_sum(e);
    }
    double r = 0.0;
    for (int b = t; b < n_blocks; b += FIN_THREADS) r = fmax(r, partial_rho[b]);
    r = wave_max(r);
    if ((t & 63) == 0) {
        red_e[t >> 6] = e;
        red_r[t >> 6] = r;
    }
    __syncthreads();
    if (t != 0) return;
    if (!ordered) {
        e = red_e[0];
        for (int w = 1; w < FIN_THREADS / 64; w++) e = e + red_e[w];
    }
    for (int w = 1; w < FIN_THREADS / 64; w++) r = fmax(r, red_r[w]);
    out[0] = e;
    out[1] = r;
    if (check_step != NO_STOP) {
        if (e > factor * (*e0) || !isfinite(e)) {
            atomicMin(&status[ST_STOP_STEP], check_step);
            atomicOr(&status[ST_REASON], (unsigned long long)ERLBM_STOP_ENERGY);
        }
    }
}

// AoS (reference fluid order) <-> SoA voxel layout, optionally streaming.
template <int Q, bool STREAM>
__global__ __launch_bounds__(BLOCK) void k_to_aos(const double *__restrict__ src, Geometry g,
                                                  const uint32_t *__restrict__ nbr,
                                                  const int *__restrict__ vox_of_fluid, int n_fluid,
                                                  double *__restrict__ aos)
{
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n_fluid) return;
    const long long v = vox_of_fluid ? vox_of_fluid[p] : p;
    const int x = (int)(v % g.nx);
    const long long yz = v / g.nx;
    const int y = (int)(yz % g.ny);
    const int z = (int)(yz / g.ny);
    double f[Q];
    if (nbr) gather<Q, STREAM, true>(src, g, nbr[v], x, y, z, f);
    else gather<Q, STREAM, false>(src, g, 0u, x, y, z, f);
    static_for<Q>([&](auto I) { aos[(long long)p * Q + (int)I] = f[(int)I]; });
}

template <int Q>
__global__ __launch_bounds__(BLOCK) void k_from_aos(const double *__restrict__ aos, Geometry g,
                                                    const int *__restrict__ vox_of_fluid, int n_fluid,
                                                    double *__restrict__ dst)
{
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n_fluid) return;
    const long long v = vox_of_fluid ? vox_of_fluid[p] : p;
    const int x = (int)(v % g.nx);
    const long long row = v / g.nx;
    static_for<Q>([&](auto I) { dst[at(g, (int)I, row, x)] = aos[(long long)p * Q + (int)I]; });
}

__global__ void k_reset_status(unsigned long long *status)
{
    if (threadIdx.x == 0) {
        status[ST_STOP_STEP] = NO_STOP;
        status[ST_REASON] = 0;
    }
}

}  // namespace

namespace {


// Steps actually run since the upload (launched, capped by a stop); syncs.
int steps_run(erlbm_lattice *h, long long *out, int *reason)
{
    unsigned long long st[ST_WORDS];
    ERLBM_HIP_TRY(hipMemcpyAsync(st, h->status, sizeof(st), hipMemcpyDeviceToHost, h->stream));
    ERLBM_HIP_TRY(hipStreamSynchronize(h->stream));
    long long r = h->launched;
    if (st[ST_STOP_STEP] != NO_STOP && (long long)st[ST_STOP_STEP] < r) r = (long long)st[ST_STOP_STEP];
    *out = r;
    if (reason) *reason = (int)st[ST_REASON];
    return ERLBM_OK;
}

int launch_finish(erlbm_lattice *h, bool ordered, double *out_dev, unsigned long long check_step, double factor)
{
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(FIN_THREADS), 0, h->stream, h->ke_node, h->partial_ke, h->partial_rho,
                       h->n_fluid, mon_blocks(h), ordered ? 1 : 0, out_dev, h->scalars + 2, factor, check_step,
                       h->status);
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

template <int Q>
int launch_monitor(erlbm_lattice *h, const double *src, bool stream, bool ordered)
{
    MonArgs m;
    m.src = src;
    m.g = geom(h);
    m.nbr = h->all_fluid ? nullptr : h->nbr;
    m.vox_of_fluid = h->vox_of_fluid;
    m.n_fluid = h->n_fluid;
    m.ke_node = ordered ? h->ke_node : nullptr;
    m.partial_ke = h->partial_ke;
    m.partial_rho = h->partial_rho;
    m.nonfinite = h->nonfinite;
    const int nb = mon_blocks(h);
    if (stream) hipLaunchKernelGGL((k_monitor<Q, true>), dim3(nb), dim3(BLOCK), 0, h->stream, m);
    else hipLaunchKernelGGL((k_monitor<Q, false>), dim3(nb), dim3(BLOCK), 0, h->stream, m);
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

// Energy / density reduction of the state in `src` (gathered through the
// stream when `stream`), result in out_dev[0..1]; with check_step set, the
// Taylor-Green map test stops the run before that step.
int monitor(erlbm_lattice *h, const void *src, bool stream, bool ordered, double *out_dev,
            unsigned long long check_step = NO_STOP, double factor = 2.0)
{
    ERLBM_HIP_TRY(hipMemsetAsync(h->nonfinite, 0, sizeof(int), h->stream));
    int rc;
    if (h->elem != 8) rc = lowp_monitor(h, src, stream, ordered);
    else if (h->q == 9) rc = launch_monitor<9>(h, static_cast<const double *>(src), stream, ordered);
    else rc = launch_monitor<19>(h, static_cast<const double *>(src), stream, ordered);
    if (rc) return rc;
    return launch_finish(h, ordered, out_dev, check_step, factor);
}

template <int Q, int COL, bool STREAM, int NPT, int NTL, int BS>
void launch_periodic_bs(erlbm_lattice *h, const StepArgs &a)
{
    const int gx = (h->nx + BS * NPT - 1) / (BS * NPT);
    const dim3 grid(gx, h->ny * h->nz);
    hipLaunchKernelGGL((k_step_periodic<Q, COL, STREAM, NPT, NTL, BS>), grid, dim3(BS), 0, h->stream, a,
                       h->tune_remap);
}

template <int Q, int COL, bool STREAM, int NPT, int NTL>
void launch_periodic(erlbm_lattice *h, const StepArgs &a)
{
    if constexpr (Q == 9 && NPT == 1) {
        if (h->tune_block == 128) return launch_periodic_bs<Q, COL, STREAM, NPT, NTL, 128>(h, a);
        if (h->tune_block == 512) return launch_periodic_bs<Q, COL, STREAM, NPT, NTL, 512>(h, a);
    }
    launch_periodic_bs<Q, COL, STREAM, NPT, NTL, BLOCK>(h, a);
}

template <int Q, int COL, bool STREAM>
void launch_periodic_tuned(erlbm_lattice *h, const StepArgs &a)
{
    if constexpr (Q == 9) {
        const bool ntl = h->tune_ntl != 0;
        if (h->tune_ntl == 2 && h->tune_npt == 1 && !h->tune_pair) {
            launch_periodic<Q, COL, STREAM, 1, 2>(h, a);
            return;
        }
        if (h->tune_pair && (h->nx % 2) == 0 && h->nz == 1) {
            const int gx = (h->nx / 2 + BLOCK - 1) / BLOCK;
            const dim3 grid(gx, h->ny);
            if (ntl) hipLaunchKernelGGL((k_step_d2q9_pair<COL, STREAM, true>), grid, dim3(BLOCK), 0, h->stream, a, h->tune_remap);
            else hipLaunchKernelGGL((k_step_d2q9_pair<COL, STREAM, false>), grid, dim3(BLOCK), 0, h->stream, a, h->tune_remap);
            return;
        }
        switch (h->tune_npt) {
        case 2: ntl ? launch_periodic<Q, COL, STREAM, 2, 1>(h, a) : launch_periodic<Q, COL, STREAM, 2, 0>(h, a); return;
        case 4: ntl ? launch_periodic<Q, COL, STREAM, 4, 1>(h, a) : launch_periodic<Q, COL, STREAM, 4, 0>(h, a); return;
        default: ntl ? launch_periodic<Q, COL, STREAM, 1, 1>(h, a) : launch_periodic<Q, COL, STREAM, 1, 0>(h, a); return;
        }
    } else {
        if (h->tune_ntl == 2) launch_periodic<Q, COL, STREAM, 1, 2>(h, a);
        else if (h->tune_ntl == 1) launch_periodic<Q, COL, STREAM, 1, 1>(h, a);
        else launch_periodic<Q, COL, STREAM, 1, 0>(h, a);
    }
}

template <int Q, int COL>
int launch_step_q(erlbm_lattice *h, const StepArgs &a, bool stream)
{
    if (h->forced) {
        const long long nb = (h->nvox + BLOCK - 1) / BLOCK;
        if (stream) hipLaunchKernelGGL((k_step_general<Q, COL, true, true>), dim3((unsigned)nb), dim3(BLOCK), 0, h->stream, a);
        else hipLaunchKernelGGL((k_step_general<Q, COL, false, true>), dim3((unsigned)nb), dim3(BLOCK), 0, h->stream, a);
    } else if (h->all_fluid && h->inlet_x < 0 && h->outlet_x < 0) {
        if (stream) launch_periodic_tuned<Q, COL, true>(h, a);
        else launch_periodic_tuned<Q, COL, false>(h, a);
    } else {
        const long long nb = (h->nvox + BLOCK - 1) / BLOCK;
        if (stream) hipLaunchKernelGGL((k_step_general<Q, COL, true>), dim3((unsigned)nb), dim3(BLOCK), 0, h->stream, a);
        else hipLaunchKernelGGL((k_step_general<Q, COL, false>), dim3((unsigned)nb), dim3(BLOCK), 0, h->stream, a);
    }
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

template <int Q>
int launch_step(erlbm_lattice *h, const StepArgs &a, bool stream)
{
    switch (h->col) {
    case ERLBM_COL_BGK: return launch_step_q<Q, COL_BGK>(h, a, stream);
    case ERLBM_COL_REG: return launch_step_q<Q, COL_REG>(h, a, stream);
    case ERLBM_COL_EXT_REG: return launch_step_q<Q, COL_EXT_REG>(h, a, stream);
    case ERLBM_COL_MRT:
        if constexpr (Q == 9) return launch_step_q<Q, COL_MRT>(h, a, stream);
        [[fallthrough]];
    default: set_last_error_msg("collision operator not supported on this path"); return ERLBM_ERR_UNSUPPORTED;
    }
}

// Size, storage layout and tuning knobs of a handle (environment overrides
// read here); no device call.
void set_layout(erlbm_lattice *h, int q, int nx, int ny, int nz, int precision)
{
    h->q = q;
    h->nx = nx;
    h->ny = ny;
    h->nz = nz;
    h->nvox = (long long)nx * ny * nz;
    auto env_ll = [](const char *name, long long dflt) {
        const char *v = getenv(name);
        return (v && *v) ? atoll(v) : dflt;
    };
    h->tune_pad = env_ll("ERLBM_PLANE_PAD", h->tune_pad);
    h->tune_npt = (int)env_ll("ERLBM_NPT", h->tune_npt);
    h->tune_ntl = (int)env_ll("ERLBM_NT_LOADS", h->tune_ntl);
    h->tune_remap = (int)env_ll("ERLBM_XCD_REMAP", h->tune_remap);
    h->tune_pair = (int)env_ll("ERLBM_PAIR", h->tune_pair);
    h->tune_block = (int)env_ll("ERLBM_BLOCK", h->tune_block);
    h->tune_vec = (int)env_ll("ERLBM_VEC", h->tune_vec);
    h->fuse_requested = getenv("ERLBM_FUSE") != nullptr;
    h->tune_fuse = (int)env_ll("ERLBM_FUSE", h->tune_fuse);
    h->tune_fuse_remap = (int)env_ll("ERLBM_FUSE_REMAP", h->tune_fuse_remap);
    h->tune_fuse_slots = (int)env_ll("ERLBM_FUSE_SLOTS", h->tune_fuse_slots);
    h->tune_fuse_prio = (int)env_ll("ERLBM_FUSE_PRIO", h->tune_fuse_prio);
    h->tune_zwalk_nt = (int)env_ll("ERLBM_ZWALK_NT", h->tune_zwalk_nt);
    h->tune_walk_nt = (int)env_ll("ERLBM_WALK_NT", h->tune_walk_nt);
    if (h->tune_pad < 0 || h->tune_pad > (1 << 24)) h->tune_pad = 0;
    h->layout = (int)env_ll("ERLBM_LAYOUT", h->layout);
    const long long pad = (h->tune_pad + 63) / 64 * 64;
    const long long nrows = (long long)ny * nz;
    if (h->layout == 0) {   // population planes
        h->si = (h->nvox + 63) / 64 * 64 + pad;
        h->sr = nx;
    } else {                // row blocks
        h->si = nx;
        h->sr = (long long)q * nx + pad;
    }
    h->extent = (q - 1) * h->si + (nrows - 1) * h->sr + nx;
    h->precision = precision;
    h->elem = precision == ERLBM_FP64 ? 8 : (precision == ERLBM_FP16_SHIFTED ? 2 : 4);
    h->f32_arith = precision == ERLBM_FP32_SHIFTED_F32;
}

}  // namespace

// ---- C ABI ------------------------------------------------------------------
extern "C" int erlbm_lattice_create(int q, int nx, int ny, int nz, const int *mask, int precision,
                                    void *stream, erlbm_lattice **out)
{
    if (!out || (q != 9 && q != 19) || nx < 1 || ny < 1 || nz < 1) {
        set_last_error_msg("erlbm_lattice_create: bad q or size");
        return ERLBM_ERR_ARG;
    }
    if (q == 9 && nz != 1) {
        set_last_error_msg("erlbm_lattice_create: D2Q9 needs nz == 1");
        return ERLBM_ERR_ARG;
    }
    if (precision != ERLBM_FP64 && precision != ERLBM_FP32_SHIFTED && precision != ERLBM_FP16_SHIFTED &&
        precision != ERLBM_FP32_SHIFTED_F32) {
        set_last_error_msg("erlbm_lattice_create: precision must be FP64, FP32_SHIFTED, FP16_SHIFTED or "
                           "FP32_SHIFTED_F32");
        return ERLBM_ERR_ARG;
    }
    const long long nvox = (long long)nx * ny * nz;
    if (nvox > (1ll << 31) - 1) {
        set_last_error_msg("erlbm_lattice_create: lattice too large");
        return ERLBM_ERR_ARG;
    }
    auto *h = new erlbm_lattice();
    set_layout(h, q, nx, ny, nz, precision);
    h->stream = (hipStream_t)stream;
    hipGetDevice(&h->device);
    if (hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess || h->n_cu < 1)
        h->n_cu = 256;

    // bounce mask: bit i set when the pull source x - c_i is solid
    std::vector<uint32_t> nbr(nvox, 0);
    std::vector<int> vof;
    int nf = 0;
    const int *cx = q == 9 ? VSet<9>::cx : VSet<19>::cx;
    const int *cy = q == 9 ? VSet<9>::cy : VSet<19>::cy;
    const int *cz = q == 9 ? VSet<9>::cz : VSet<19>::cz;
    for (long long v = 0; v < nvox; v++) {
        if (mask && !mask[v]) continue;
        nf++;
        const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long long)nx * ny));
        uint32_t b = FLUID_BIT;
        for (int i = 1; i < q; i++) {
            const int xs = ((x - cx[i]) % nx + nx) % nx;
            const int ys = ((y - cy[i]) % ny + ny) % ny;
            const int zs = ((z - cz[i]) % nz + nz) % nz;
            const long long vs = xs + (long long)nx * (ys + (long long)ny * zs);
            if (mask && !mask[vs]) b |= 1u << i;
        }
        nbr[v] = b;
    }
    h->n_fluid = nf;
    h->all_fluid = (nf == nvox);
    if (!h->all_fluid) {
        vof.reserve(nf);
        for (long long v = 0; v < nvox; v++)
            if (nbr[v] & FLUID_BIT) vof.push_back((int)v);
    }
    if (nf == 0) {
        delete h;
        set_last_error_msg("erlbm_lattice_create: no fluid site");
        return ERLBM_ERR_ARG;
    }
    auto fail = [&](hipError_t e, const char *what) {
        set_last_error(what, e);
        erlbm_lattice_destroy(h);
        return ERLBM_ERR_NOMEM;
    };
    hipError_t e;
    const size_t fbytes = (size_t)h->elem * (size_t)h->extent;
    if ((e = hipMalloc(&h->buf[0], fbytes)) != hipSuccess) return fail(e, "hipMalloc populations");
    if ((e = hipMalloc(&h->buf[1], fbytes)) != hipSuccess) return fail(e, "hipMalloc populations");
    if ((e = hipMalloc(&h->nbr, sizeof(uint32_t) * nvox)) != hipSuccess) return fail(e, "hipMalloc mask");
    if (!h->all_fluid && (e = hipMalloc(&h->vox_of_fluid, sizeof(int) * nf)) != hipSuccess)
        return fail(e, "hipMalloc map");
    const int nb = mon_blocks(h);
    if ((e = hipMalloc(&h->ke_node, sizeof(double) * nf)) != hipSuccess) return fail(e, "hipMalloc monitor");
    if ((e = hipMalloc(&h->partial_ke, sizeof(double) * nb)) != hipSuccess) return fail(e, "hipMalloc monitor");
    if ((e = hipMalloc(&h->partial_rho, sizeof(double) * nb)) != hipSuccess) return fail(e, "hipMalloc monitor");
    if ((e = hipMalloc(&h->scalars, sizeof(double) * 4)) != hipSuccess) return fail(e, "hipMalloc monitor");
    if ((e = hipMalloc(&h->nonfinite, sizeof(int))) != hipSuccess) return fail(e, "hipMalloc monitor");
    if ((e = hipMalloc(&h->status, sizeof(unsigned long long) * ST_WORDS)) != hipSuccess)
        return fail(e, "hipMalloc status");
    if ((e = hipMemcpyAsync(h->nbr, nbr.data(), sizeof(uint32_t) * nvox, hipMemcpyHostToDevice, h->stream)) !=
        hipSuccess)
        return fail(e, "upload mask");
    if (!h->all_fluid &&
        (e = hipMemcpyAsync(h->vox_of_fluid, vof.data(), sizeof(int) * nf, hipMemcpyHostToDevice, h->stream)) !=
            hipSuccess)
        return fail(e, "upload map");
    if ((e = hipMemsetAsync(h->buf[0], 0, fbytes, h->stream)) != hipSuccess) return fail(e, "clear");
    if ((e = hipMemsetAsync(h->buf[1], 0, fbytes, h->stream)) != hipSuccess) return fail(e, "clear");
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(64), 0, h->stream, h->status);
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return fail(e, "sync");
    *out = h;
    return ERLBM_OK;
}

extern "C" int erlbm_lattice_destroy(erlbm_lattice *h)
{
    if (!h) return ERLBM_OK;
    if (h->stream || h->buf[0]) hipStreamSynchronize(h->stream);
    hipFree(h->buf[0]);
    hipFree(h->buf[1]);
    hipFree(h->nbr);
    hipFree(h->vox_of_fluid);
    hipFree(h->ke_node);
    hipFree(h->partial_ke);
    hipFree(h->partial_rho);
    hipFree(h->scalars);
    hipFree(h->nonfinite);
    hipFree(h->status);
    delete h;
    return ERLBM_OK;
}

extern "C" int erlbm_lattice_fluid_count(const erlbm_lattice *h, int *n_fluid)
{
    if (!h || !n_fluid) return ERLBM_ERR_ARG;
    *n_fluid = h->n_fluid;
    return ERLBM_OK;
}

extern "C" int erlbm_set_params(erlbm_lattice *h, double tau, double tau_nh, double rho0, int collision)
{
    if (!h) return ERLBM_ERR_ARG;
    if (collision != ERLBM_COL_BGK && collision != ERLBM_COL_REG && collision != ERLBM_COL_EXT_REG &&
        collision != ERLBM_COL_MRT) {
        set_last_error_msg("erlbm_set_params: collision must be BGK, MRT, REG or EXT_REG");
        return ERLBM_ERR_ARG;
    }
    if (collision == ERLBM_COL_MRT && (h->q != 9 || h->f32_arith)) {
        // the reference defines coll_MRT for D2Q9 only (src/erlbm_common.h:75-77)
        set_last_error_msg("erlbm_set_params: MRT is defined for D2Q9 (FP64 arithmetic) only");
        return ERLBM_ERR_UNSUPPORTED;
    }
    if (!(tau > 0.0) || !(tau_nh > 0.0)) {
        set_last_error_msg("erlbm_set_params: tau and tau_nh must be positive");
        return ERLBM_ERR_ARG;
    }
    h->tau = tau;
    h->tau_nh = tau_nh;
    h->rho0 = rho0;
    h->col = collision;
    return ERLBM_OK;
}

extern "C" int erlbm_set_zouhe_inlet(erlbm_lattice *h, int pos_x, double vx)
{
    if (!h || pos_x >= h->nx) return ERLBM_ERR_ARG;
    h->inlet_x = pos_x < 0 ? -1 : pos_x;
    h->u_in = vx;
    return ERLBM_OK;
}

extern "C" int erlbm_set_outlet_rescale(erlbm_lattice *h, int pos_x, double rho)
{
    if (!h || pos_x >= h->nx) return ERLBM_ERR_ARG;
    if (pos_x >= 0) {
        const int dx = pos_x < h->nx / 2 ? 1 : -1;
        if (pos_x + dx < 0 || pos_x + dx >= h->nx) {
            set_last_error_msg("erlbm_set_outlet_rescale: neighbour column outside the lattice");
            return ERLBM_ERR_ARG;
        }
    }
    h->outlet_x = pos_x < 0 ? -1 : pos_x;
    h->rho_out = rho;
    return ERLBM_OK;
}

extern "C" int erlbm_set_body_force(erlbm_lattice *h, double ax, double ay, double az)
{
    if (!h) return ERLBM_ERR_ARG;
    if (!std::isfinite(ax) || !std::isfinite(ay) || !std::isfinite(az)) {
        set_last_error_msg("erlbm_set_body_force: accelerations must be finite");
        return ERLBM_ERR_ARG;
    }
    if (h->q == 9) az = 0.0;   // the 2-D operators read x and y only
    const bool forced = ax != 0.0 || ay != 0.0 || az != 0.0;
    if (forced && h->elem != 8) {
        set_last_error_msg("erlbm_set_body_force: a body force needs FP64 storage");
        return ERLBM_ERR_UNSUPPORTED;
    }
    h->acc[0] = ax;
    h->acc[1] = ay;
    h->acc[2] = az;
    h->forced = forced;
    return ERLBM_OK;
}

extern "C" int erlbm_set_arith(erlbm_lattice *h, int arith)
{
    if (!h) return ERLBM_ERR_ARG;
    if (arith != ERLBM_ARITH_EXACT && arith != ERLBM_ARITH_MOMENT) {
        set_last_error_msg("erlbm_set_arith: arith must be ERLBM_ARITH_EXACT or ERLBM_ARITH_MOMENT");
        return ERLBM_ERR_ARG;
    }
    if (arith == ERLBM_ARITH_MOMENT && (h->elem == 2 || h->f32_arith)) {
        set_last_error_msg("erlbm_set_arith: the moment-space arithmetic needs FP64 arithmetic on FP64 or FP32 storage");
        return ERLBM_ERR_UNSUPPORTED;
    }
    h->arith = arith;
    return ERLBM_OK;
}

extern "C" int erlbm_set_density_check(erlbm_lattice *h, double limit)
{
    if (!h) return ERLBM_ERR_ARG;
    h->dens_limit = limit > 0 ? limit : 0.0;
    return ERLBM_OK;
}

extern "C" int erlbm_upload_aos(erlbm_lattice *h, const double *inif)
{
    if (!h || !inif) return ERLBM_ERR_ARG;
    const size_t bytes = sizeof(double) * (size_t)h->n_fluid * h->q;
    const int nb = mon_blocks(h);
    if (h->elem == 8) {
        // stage in buffer 1, transpose into buffer 0
        double *stage = static_cast<double *>(h->buf[1]);
        double *dst = static_cast<double *>(h->buf[0]);
        ERLBM_HIP_TRY(hipMemcpyAsync(stage, inif, bytes, hipMemcpyHostToDevice, h->stream));
        if (h->q == 9)
            hipLaunchKernelGGL(k_from_aos<9>, dim3(nb), dim3(BLOCK), 0, h->stream, stage, geom(h), h->vox_of_fluid,
                               h->n_fluid, dst);
        else
            hipLaunchKernelGGL(k_from_aos<19>, dim3(nb), dim3(BLOCK), 0, h->stream, stage, geom(h), h->vox_of_fluid,
                               h->n_fluid, dst);
        ERLBM_HIP_TRY(hipGetLastError());
    } else {
        double *stage = nullptr;
        ERLBM_HIP_TRY(hipMalloc(&stage, bytes));
        h->shift = h->rho0;
        int rc = ERLBM_OK;
        if (hipMemcpyAsync(stage, inif, bytes, hipMemcpyHostToDevice, h->stream) != hipSuccess) rc = ERLBM_ERR_HIP;
        if (rc == ERLBM_OK) rc = lowp_from_aos(h, stage, h->buf[0]);
        if (hipStreamSynchronize(h->stream) != hipSuccess && rc == ERLBM_OK) rc = ERLBM_ERR_HIP;
        (void)hipFree(stage);
        if (rc) {
            set_last_error_msg("erlbm_upload_aos: shifted upload failed");
            return rc;
        }
    }
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(64), 0, h->stream, h->status);
    ERLBM_HIP_TRY(hipGetLastError());
    h->base = 0;
    h->launched = 0;
    h->checked = false;
    h->uploaded = true;
    // E0 for the energy test: the reference's energy_ini (ordered sum up to
    // 2^20 sites, the fixed-shape tree above that)
    int rc = monitor(h, h->buf[0], false, h->n_fluid <= (1 << 20), h->scalars + 2);
    if (rc) return rc;
    ERLBM_HIP_TRY(hipStreamSynchronize(h->stream));
    h->e0_valid = true;
    return ERLBM_OK;
}

namespace {

// Shapes of the two-step kernel (ERLBM_FUSE selects one; 0 = off).
enum FuseKind { FUSE_TILE, FUSE_STRIP, FUSE_WAVE, FUSE_STRIP3, FUSE_ZWALK, FUSE_GSTRIP3, FUSE_GSTRIP4, FUSE_GSTRIP5, FUSE_GSTRIP6 };
struct FuseShape {
    int tx, ty;       // tile; (strip) strip width and rows walked; (wave) rows walked
    FuseKind kind;
    int pd, wpb;      // (wave) rows of loads in flight, wavefronts per workgroup
};
constexpr FuseShape FUSE_SHAPES[] = {
    {0, 0, FUSE_TILE},     {64, 8, FUSE_TILE},    {32, 16, FUSE_TILE},   {32, 8, FUSE_TILE},
    {128, 4, FUSE_TILE},   {16, 16, FUSE_TILE},   {32, 32, FUSE_TILE},   {64, 4, FUSE_TILE},
    {128, 8, FUSE_TILE},   {256, 4, FUSE_TILE},   {64, 16, FUSE_TILE},   {256, 32, FUSE_STRIP},
    {256, 16, FUSE_STRIP}, {128, 32, FUSE_STRIP}, {256, 64, FUSE_STRIP}, {128, 64, FUSE_STRIP},
    // wavefront strips: ty = rounds of resident wavefronts per launch
    {WAVE_STRIP, 2, FUSE_WAVE, 1, 4}, {WAVE_STRIP, 1, FUSE_WAVE, 1, 4}, {WAVE_STRIP, 2, FUSE_WAVE, 2, 4},
    {256, 8, FUSE_STRIP},              {128, 16, FUSE_STRIP},            {128, 8, FUSE_STRIP},
    // three steps per launch
    {256, 32, FUSE_STRIP3},            {256, 16, FUSE_STRIP3},           {128, 32, FUSE_STRIP3},
    {256, 64, FUSE_STRIP3},
    // three steps per launch: lean rings, segments sized to the resident
    // workgroups (ty = 0: any height from 16 up) or of 32 rows; 29: 18-row
    // rings with sized segments
    {256, 0, FUSE_STRIP3},             {128, 0, FUSE_STRIP3},            {256, 32, FUSE_STRIP3},
    {256, 0, FUSE_STRIP3},
    // D3Q19, two steps per launch: x-y tiles walked along z
    {32, 8, FUSE_ZWALK},               {64, 4, FUSE_ZWALK},              {32, 4, FUSE_ZWALK},
    {64, 8, FUSE_ZWALK},               {128, 4, FUSE_ZWALK},             {32, 16, FUSE_ZWALK},
    // three steps per launch, lean rings, sized segments: strips of any
    // width up to tx = 64 NW - 4 for NW = 1 .. 4 wavefronts per workgroup
    {60, 0, FUSE_GSTRIP3},             {124, 0, FUSE_GSTRIP3},           {188, 0, FUSE_GSTRIP3},
    {252, 0, FUSE_GSTRIP3},
    // the same with the lanes aligned across the phases (c_x = 0 populations
    // in registers), 2, 3, 4, 8, 5 and 6 wavefronts per workgroup
    {124, 0, FUSE_GSTRIP3},            {188, 0, FUSE_GSTRIP3},           {252, 0, FUSE_GSTRIP3},
    {508, 0, FUSE_GSTRIP3},            {316, 0, FUSE_GSTRIP3},           {380, 0, FUSE_GSTRIP3},
    // four, five and six steps per launch: lane-aligned strips of any width
    // up to 64 NW - 2 (steps - 1), four wavefronts
    {250, 0, FUSE_GSTRIP4},            {248, 0, FUSE_GSTRIP5},           {246, 0, FUSE_GSTRIP6}};
constexpr int N_FUSE_SHAPES = sizeof(FUSE_SHAPES) / sizeof(FUSE_SHAPES[0]);

// Steps one launch of the shape advances.
int fuse_steps(int f)
{
    if (f <= 0 || f >= N_FUSE_SHAPES) return 2;
    const FuseKind k = FUSE_SHAPES[f].kind;
    if (k == FUSE_GSTRIP6) return 6;
    return k == FUSE_GSTRIP5 ? 5 : (k == FUSE_GSTRIP4 ? 4 : (k == FUSE_STRIP3 || k == FUSE_GSTRIP3 ? 3 : 2));
}

// Segments per strip for the wavefront-strip kernel: `rounds` times as many
// wavefronts as the GPU holds at once, at most one per row.
int wave_segments(const erlbm_lattice *h, int waves_per_cu, int rounds)
{
    const int nstrip = (h->nx + WAVE_STRIP - 1) / WAVE_STRIP;
    const long long cap = (long long)h->n_cu * waves_per_cu * rounds;
    return (int)std::max(1ll, std::min<long long>(h->ny, cap / nstrip));
}

// Whether the lattice splits into whole tiles / strips of the shape (the
// wavefront strips take ragged edges but need 64 distinct columns).
bool fuse_geometry_fits(const erlbm_lattice *h, const FuseShape &fs)
{
    if ((fs.kind == FUSE_ZWALK) != (h->q == 19)) return false;
    if (fs.kind == FUSE_ZWALK) {   // 31-bit offsets within a plane, segments of at least 16 planes
        const long long window = 8 * ((long long)(h->q - 1) * h->si + (long long)h->ny * h->sr);
        return h->elem == 8 && h->nz >= 16 && h->nx % fs.tx == 0 && h->ny % fs.ty == 0 && window < (1ll << 31);
    }
    if (h->nz != 1) return false;
    if (fs.kind == FUSE_WAVE) return h->nx >= 64 && h->extent * 8 < (1ll << 31);   // 31-bit buffer offsets
    if (fs.kind == FUSE_GSTRIP3 || fs.kind == FUSE_GSTRIP4 || fs.kind == FUSE_GSTRIP5 || fs.kind == FUSE_GSTRIP6)
        return h->extent * h->elem < (1ll << 31) && h->nx >= 16 && h->ny >= 16;
    if ((fs.kind == FUSE_STRIP || fs.kind == FUSE_STRIP3) && h->elem == 8 && h->extent * 8 >= (1ll << 31)) return false;
    if (fs.ty == 0) return h->nx % fs.tx == 0 && h->ny >= 16;
    return h->nx % fs.tx == 0 && h->ny % fs.ty == 0;
}

// Workgroups of one launch of the shape.
long long fuse_workgroups(const erlbm_lattice *h, const FuseShape &fs)
{
    if (fs.kind == FUSE_WAVE) {   // (as if four wavefronts per SIMD fit)
        const long long waves = (long long)((h->nx + WAVE_STRIP - 1) / WAVE_STRIP) * wave_segments(h, 16, fs.ty);
        return (waves + fs.wpb - 1) / fs.wpb;
    }
    if (fs.kind == FUSE_ZWALK) {   // segments sized to the resident workgroups (as if two fit a CU)
        const long long ntile = (long long)(h->nx / fs.tx) * (h->ny / fs.ty);
        return ntile * std::max(1ll, std::min<long long>(h->nz / 16, 2ll * h->n_cu / std::max(1ll, ntile)));
    }
    if (fs.ty == 0) {   // segments sized to the resident workgroups (as if two fit a CU)
        const long long nstrip = fs.kind == FUSE_GSTRIP3 ? (h->nx + fs.tx - 1) / fs.tx : h->nx / fs.tx;
        return nstrip * std::max(1ll, std::min<long long>(h->ny / 16, 2ll * h->n_cu / std::max(1ll, nstrip)));
    }
    return (long long)(h->nx / fs.tx) * (h->ny / fs.ty);
}

bool fuse_shape_fits(const erlbm_lattice *h, int f)
{
    if (f <= 0 || f >= N_FUSE_SHAPES) return false;
    const FuseShape fs = FUSE_SHAPES[f];
    // shifted storage: the strip walks only, FP64 arithmetic only (FP32
    // storage: two or three steps; FP16: two, by explicit request
    // (ERLBM_FUSE) only -- its 16-byte vector kernel is faster, 36.8 vs 53.2
    // us/step at L = 2048: the two collisions per launch make the walk
    // VALU-bound once the bytes are that few)
    if (h->elem != 8) {
        const bool two = f == FUSE2_STRIP_256x32 || f == FUSE2_STRIP_128x32;
        const bool three = f == FUSE3_STRIP_256x32 || f == FUSE3_STRIP_128x32 || f == FUSE3_ASTRIP_4W ||
                           f == FUSE4_ASTRIP_4W || f == FUSE5_ASTRIP_4W;
        if (h->f32_arith || !(two || (three && h->elem == 4))) return false;
        if (h->elem == 2 && !h->fuse_requested) return false;
        if (h->extent * h->elem >= (1ll << 31)) return false;   // 31-bit buffer offsets
    }
    return fuse_geometry_fits(h, fs);
}

// The multi-step shape a call of erlbm_step uses (0: one step per launch).
// Without an explicit ERLBM_FUSE, a shape must give every CU a workgroup
// (small lattices have too few strips or tiles to fill the chip: there the
// one-step kernel is faster).  FP64: the five-step lane-aligned walk, else
// the four-step one, else the three-step walk, lane-aligned on strips of any
// width or on 256-wide strips (the rule below), else on strips of any width
// (four wavefronts, lanes shifted), else the two-step walk, else 32x8 tiles
// where 64x8 ones do not fit or make fewer than two workgroups per CU, 64x4
// ones where they make 3 to 4.5, else the 64x8 tiles, else the wavefront
// strips (any width from 64 up); shifted FP32
// storage: the five-step lane-aligned walk (moment-space arithmetic only),
// else the four-step one, else one step per launch.
long long strip3_workgroups(const erlbm_lattice *h, int f);
long long astrip4_workgroups(const erlbm_lattice *h);
long long astrip5_workgroups(const erlbm_lattice *h);
int astrip_unit(const erlbm_lattice *h);

// Output columns the five-step walk's strips offer for this width (strips of
// at most 248 positions, edges on the lane-aligned walks' unit).
long long fw_cap(const erlbm_lattice *h)
{
    const int u = astrip_unit(h), units = h->nx / u;
    return (long long)((units + 248 / u - 1) / (248 / u)) * (248 / u) * u;
}

int fuse_shape(const erlbm_lattice *h, int energy_interval)
{
    if (!(h->all_fluid && !h->forced && h->inlet_x < 0 && h->outlet_x < 0 && h->dens_limit <= 0.0 &&
          energy_interval <= 0 && !h->checked))
        return 0;
    if (h->fuse_requested) return fuse_shape_fits(h, h->tune_fuse) ? h->tune_fuse : 0;
    if (h->q == 19) {
        for (int f : {FUSE2_ZWALK_64x8, FUSE2_ZWALK_64x4, FUSE2_ZWALK_32x8})
            if (fuse_shape_fits(h, f) && fuse_workgroups(h, FUSE_SHAPES[f]) >= h->n_cu) return f;
        return 0;
    }
    if (!(h->q == 9 && h->nz == 1)) return 0;
    const int first = h->tune_fuse;
    // FP64: the four-step lane-aligned walk where it makes at least 1.5
    // workgroups per CU (MLUPS, four-step / three-step rule below, exact and
    // moment-space arithmetic: L = 1280 81.0 / 74.3, 100.5 / 87.0; 1536 94.5 /
    // 84.4, 110.2 / 89.7; 2048 108.3 / 95.8, 129.6 / 102.8; 3000 114.5 /
    // 93.1, 123.0 / 96.4; 4096 120.6 / 105.1, 141.0 / 111.4; L = 1024, 320
    // workgroups: 52.9 / 60.1, 71.0 / 74.4)
    // FP64: the five-step walk where it makes at least 1.5 workgroups per CU
    // (MLUPS, five-step / four-step, exact then moment-space arithmetic:
    // L = 1280 80.2 / 79.9, 110.8 / 100.2; 2048 106.8 / 104.8, 147.5 / 129.2;
    // 4096 122.2 / 117.0, 165.7 / 142.3); FP32 storage in the moment-space
    // arithmetic likewise (L = 1280 116.7 / 115.1, 2048 173.5 / 170.0, 4096
    // 204.4 / 196.3), in the exact one not (80.2 / 77.8 four-step against
    // five-step at L = 1280, 114.1 / 113.7 at 2048)
    const bool five = (h->elem == 8 && first == FUSE3_STRIP_SIZED) || (h->elem == 4 && h->arith == ERLBM_ARITH_MOMENT);
    if (five && fuse_shape_fits(h, FUSE5_ASTRIP_4W)) {
        // FP64 below 1.5 workgroups per CU: from 0.75 per CU in the
        // moment-space arithmetic, from 0.85 in the exact one where 256-wide
        // strips do not divide the width (MLUPS, five-step / the shape
        // picked before, exact then moment-space: L = 768 (0.75 per CU)
        // 43.0 / 52.7, 58.2 / 52.9; 896 59.3 / 53.6, 76.6 / 53.7; 960 67.3 /
        // 56.3, 87.3 / 56.4; 1088 59.1 / 55.8, 83.0 / 69.2; 1152 66.0 /
        // 61.8, 91.5 / 73.5; at 1024, 256-wide strips, 52.6 / 60.1, 74.4 /
        // 74.0; at 704 (0.52 per CU) 36.6 / 48.7, 47.4 / 48.7) -- where its
        // strips are no more than the three-step walk's (L = 1000: five
        // strips of 200 against four of 250, 48.6 / 65.8, 69.4 / 80.7)
        const long long w = astrip5_workgroups(h), n = h->n_cu;
        const int u = astrip_unit(h), units = h->nx / u;
        const bool strips = (units + 248 / u - 1) / (248 / u) <= (units + 252 / u - 1) / (252 / u);
        const bool moment = h->arith == ERLBM_ARITH_MOMENT && h->col != ERLBM_COL_MRT;
        // between one and 1.5 workgroups per CU in the exact arithmetic some
        // CUs walk two five-step workgroups and others one: the two-wavefront
        // lane-aligned three-step walk balances them (MLUPS, it / five-step:
        // L = 1088 65.9 / 59.0, 1100 65.3 / 58.7, 1152 68.8 / 65.7, 1200
        // 72.8 / 70.9)
        if (h->elem == 8 && !moment && strips && h->nx % 256 != 0 && w > n && 2 * w < 3 * n &&
            fuse_shape_fits(h, FUSE3_ASTRIP_2W))
            return FUSE3_ASTRIP_2W;
        // in the exact arithmetic only where the strips use at least 80% of
        // the lanes (512 x 2048: three strips of about 170 on 256 lanes, 54.4
        // against 59.3 for the 256-wide three-step walk; the moment-space
        // walk, bound by HBM, stays the fastest there, 79.0 / 71.5)
        const bool lanes = moment || h->elem != 8 || 5 * h->nx >= 4 * fw_cap(h);
        if (lanes && (2 * w >= 3 * n || (strips && moment && 4 * w >= 3 * n) ||
                      (strips && h->elem == 8 && 20 * w >= 17 * n && h->nx % 256 != 0)))
            return FUSE5_ASTRIP_4W;
    }
    // FP32 storage likewise (L = 2048, MLUPS, four-step / three-step, exact
    // and moment-space arithmetic: 111.6 / 107.7, 162.6 / 155.2; L = 4096
    // 125.8 / 122.8, 193.7 / 180.4; L = 1280 78.7 / 79.4, 112.0 / 109.2)
    if (((h->elem == 8 && first == FUSE3_STRIP_SIZED && ((h->arith == ERLBM_ARITH_MOMENT && h->col != ERLBM_COL_MRT) || 5 * h->nx >= 4 * fw_cap(h))) ||
         h->elem == 4) &&
        fuse_shape_fits(h, FUSE4_ASTRIP_4W) && 2 * astrip4_workgroups(h) >= 3ll * h->n_cu)
        return FUSE4_ASTRIP_4W;
    // FP32 storage below that: one step per launch (MLUPS, one step / the
    // three-step walk, exact arithmetic: L = 1024 64.5 / 51.5, 1152 66.8 /
    // 65.1; moment-space, L = 512, 43.7 / 23.9; the five-step walk from
    // 0.75 workgroups per CU in the moment-space arithmetic, as above: 768
    // 58.3 / 56.7, 896 78.5 / 60.8, 1024 75.0 / 70.9, 1152 94.9 / 89.4)
    if (h->elem == 4) return 0;
    // FP64: the lane-aligned walk on strips of any width (four wavefronts,
    // four workgroups per CU) before the 256-wide walk where it puts no more
    // workgroups on a CU than that one, or makes three per CU or more
    // (MLUPS, aligned / 256-wide: L = 1024 49.6 / 57.3 (2 vs 1 per CU); 1280
    // 70.4 / 60.2; 1536 73.9 / 65.8; 2048 87.8 / 85.4; 4096 95.5 / 96.3)
    if (h->elem == 8 && first == FUSE3_STRIP_SIZED && fuse_shape_fits(h, FUSE3_ASTRIP_4W)) {
        const long long wa = strip3_workgroups(h, FUSE3_ASTRIP_4W);
        const bool wide = fuse_shape_fits(h, first) && fuse_workgroups(h, FUSE_SHAPES[first]) >= h->n_cu;
        const long long ww = wide ? strip3_workgroups(h, first) : 0;
        const long long n = h->n_cu;
        if (wide ? ((wa + n - 1) / n <= (ww + n - 1) / n || wa >= 3 * n) : wa >= (19 * n + 19) / 20)
            return FUSE3_ASTRIP_4W;
    }
    for (int f : {first, FUSE3_GSTRIP_4W, FUSE2_STRIP_256x32, 1, 16}) {
        // 32x8 tiles where 64x8 ones do not fit or make fewer than two
        // workgroups per CU (MLUPS, 32x8 / the shape before: L = 128 7.2 /
        // 4.9 (one step per launch), 256 26.7 / 17.2 (wavefront strips),
        // 384 41.0 / 34.7 (64x8), 448 49.1 / 46.8, 480 54.1 / 37.0; at 512
        // 52.4 against 57.9 for 64x8, at 576 45.8 / 49.6)
        if (f == 1 && h->elem == 8 && fuse_shape_fits(h, 3) &&
            (!fuse_shape_fits(h, 1) || fuse_workgroups(h, FUSE_SHAPES[1]) < 2ll * h->n_cu))
            return 3;
        // 64x4 tiles where 64x8 ones make 3 to 4.5 workgroups per CU (MLUPS,
        // 64x4 / 64x8, exact then moment-space: L = 640 45.8 / 42.6, 48.9 /
        // 45.3; 704 51.1 / 48.7, 55.4 / 53.0; at 576 (2.5 per CU) 45.4 /
        // 49.4, at 768 (4.5) 52.3 / 52.9)
        if (f == 1 && h->elem == 8 && fuse_shape_fits(h, 1) && fuse_shape_fits(h, 7)) {
            const long long w1 = fuse_workgroups(h, FUSE_SHAPES[1]);
            if (w1 >= 3ll * h->n_cu && 2 * w1 < 9ll * h->n_cu) return 7;
        }
        if (!fuse_shape_fits(h, f)) continue;
        // strips of any width: from 95% of a workgroup per CU (L = 1000:
        // 248 workgroups, 61.1k MLUPS against 46.4k for the wavefront
        // strips; L = 768: 192, 41.8k against 50.9k for the 64x8 tiles)
        const bool any = f == FUSE3_GSTRIP_4W || f == FUSE3_ASTRIP_4W;
        const long long need = any ? (19ll * h->n_cu + 19) / 20 : h->n_cu;
        if (fuse_workgroups(h, FUSE_SHAPES[f]) >= need) return f;
    }
    return 0;
}

template <int COL, int TX, int TY>
void launch_fused_t(erlbm_lattice *h, const StepArgs &a)
{
    constexpr size_t lds = Fuse2<TX, TY>::LDS_BYTES;
    if constexpr (lds > 65536)   // above the default dynamic-LDS limit (gfx950 has 160 KB)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_step2_d2q9<COL, TX, TY>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const unsigned nb = (unsigned)((h->nx / TX) * (h->ny / TY));
    hipLaunchKernelGGL((k_step2_d2q9<COL, TX, TY>), dim3(nb), dim3(BLOCK), lds, h->stream, a, h->tune_fuse_remap);
}

template <int COL, int TX, int SEG, int PD = 1>
void launch_strip_t(erlbm_lattice *h, const StepArgs &a)
{
    constexpr size_t lds = sizeof(double) * RING_ROWS * (TX + 2);
    if constexpr (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_step2_strip<COL, TX, SEG, PD>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const unsigned nb = (unsigned)((h->nx / TX) * (h->ny / SEG));
    hipLaunchKernelGGL((k_step2_strip<COL, TX, SEG, PD>), dim3(nb), dim3(strip_threads<TX>()), lds, h->stream, a,
                       h->tune_fuse_remap);
}

template <int COL, int PD, int WPB, int WPE>
void launch_wave_t(erlbm_lattice *h, const StepArgs &a, int rounds)
{
    static const int waves_per_cu = [] {
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, reinterpret_cast<const void *>(&k_step2_wave<COL, PD, WPB, WPE>),
                                                         64 * WPB, 0) != hipSuccess || nblk < 1)
            nblk = 1;
        return nblk * WPB;
    }();
    const int nstrip = (h->nx + WAVE_STRIP - 1) / WAVE_STRIP;
    const int nseg = wave_segments(h, waves_per_cu, rounds);
    const unsigned nb = (unsigned)(((long long)nstrip * nseg + WPB - 1) / WPB);
    hipLaunchKernelGGL((k_step2_wave<COL, PD, WPB, WPE>), dim3(nb), dim3(64 * WPB), 0, h->stream, a, nstrip, nseg,
                       h->tune_fuse_remap);
}

// SEG > 0: segments of SEG rows; SEG = 0: as many segments as make one
// workgroup per resident slot, rows split evenly, at least 16 per segment
// (a segment recomputes 4 + 2 halo rows of the first two steps).
template <int TX, bool LEAN>
constexpr size_t strip3_lds()
{
    return sizeof(double) * (LEAN ? LEAN_ROWS : RING_ROWS) * ((TX + 4) + (TX + 2));
}

// Workgroups of k_step3_strip<COL, TX, LEAN> that fit a CU (occupancy query,
// once; raises the kernel's dynamic-LDS limit where it needs more than 64 KB).
template <int COL, int TX, bool LEAN>
int strip3_per_cu()
{
    static const int per_cu = [] {
        constexpr size_t lds = strip3_lds<TX, LEAN>();
        const void *k = reinterpret_cast<const void *>(&k_step3_strip<COL, TX, LEAN>);
        if constexpr (lds > 65536) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, k, strip_threads<TX>(), lds) != hipSuccess || nblk < 1)
            nblk = 1;
        return nblk;
    }();
    return per_cu;
}

template <int COL, int TX, int SEG, bool LEAN = false>
void launch_strip3_t(erlbm_lattice *h, const StepArgs &a)
{
    constexpr size_t lds = strip3_lds<TX, LEAN>();
    const int per_cu = strip3_per_cu<COL, TX, LEAN>();
    const int nstrip = h->nx / TX;
    const int nseg = SEG > 0 ? h->ny / SEG
                             : (int)std::max(1ll, std::min<long long>(h->ny / 16, (long long)h->n_cu * per_cu / nstrip));
    hipLaunchKernelGGL((k_step3_strip<COL, TX, LEAN>), dim3((unsigned)(nstrip * nseg)), dim3(strip_threads<TX>()), lds,
                       h->stream, a, h->tune_fuse_remap, nseg);
}

// The three-step walk with the lanes aligned across the phases: lane t
// holds position x0 - 2 + t in all three (phase 1: lanes 0 .. w+3, phase 2:
// 1 .. w+2, phase 3 and the stores: 2 .. w+1), so a pull of a c_x = 0
// population reads the lane's own result of the step before, which stays in
// registers (the source row's age, in walking steps, is 1 + c_y DIR: 0 is
// the row the phase before has just computed).  The LDS rings keep only the
// six c_x != 0 populations (lean layout, 8 rows each): 32 KB for NW = 4, so
// four or five workgroups share a CU instead of three.  Barriers, lean late
// writes and walking directions as in strip3_walk.
// NTP 1: the c_x = 0 step-k loads carry the nontemporal hint; 0 (default):
// none -- neighbouring segments read their shared halo rows again from L2
// (L = 2048: reads 1.107x against 1.140x the algorithmic ones, 129.5 against
// 131.5 us per launch).
template <int COL, int NW, int DIR, int NTP = 0>
__device__ __forceinline__ void strip3_walk_aligned(const StepArgs &a, int x0, int w, int ya, int rows, double *ring1,
                                                    double *ring2, int prio)
{
    using V = VSet<9>;
    constexpr int E = 64 * NW;
    const int t = (int)threadIdx.x;
    const int nx = a.g.nx, ny = a.g.ny;
    const bool act1 = t < w + 4, act2 = t >= 1 && t < w + 3, act3 = t >= 2 && t < w + 2;
    const int x1 = act1 ? wrap(x0 - 2 + t, nx) : 0;
    const __amdgpu_buffer_rsrc_t rs = pop_buffer<8>(a.src, a.g), rd = pop_buffer<8>(a.dst, a.g);
    const unsigned oob1 = act1 ? 0u : BUF_OOB;
    const unsigned vo[3] = {8u * wrap(x1 + 1, nx) | oob1, 8u * x1 | oob1, 8u * wrap(x1 - 1, nx) | oob1};
    const unsigned vst = act3 ? 8u * (x0 - 2 + t) : 0u;
    auto load_row = [&](int y, double (&f)[9], bool skip) {   // pulls of step-k row y (-3 .. ny+2)
        const unsigned off = skip ? BUF_OOB : 0u;
        const int yw = wrap(y, ny);
        static_for<9>([&](auto I) {
            constexpr int i = I;
            const int ys = V::cy[i] == 0 ? yw : wrap(yw - V::cy[i], ny);
            f[i] = buf_load<NTP == 1 && V::cx[i] == 0 ? BUF_NT : 0>(rs, vo[V::cx[i] + 1] | off, (int)(8 * at(a.g, i, ys, 0)));
        });
    };
    auto rrow = [&](int i, int y) { return xring_base<DIR>(i) + (y - ya + 2) % xring_depth<DIR>(i); };
    auto inring = [](int i) { return V::cx[i] != 0; };
    auto now = [](int i) { return V::cy[i] * DIR < 0; };
    auto age = [](int i) { return 1 + V::cy[i] * DIR; };
    auto collide = [&](double (&f)[9]) {
        double vx, vy, vz;
        collide_sel<9, COL>(f, a.k, vx, vy, vz);
    };
    const int r0 = DIR > 0 ? ya - 2 : ya + rows + 1;
    const int nsteps = rows + 4;
    double fq[9];
    double h2[9];                  // the previous walking step's step-(k+1) row, still to be written to ring 2
    double k1a[9], k1b[9];         // c_x = 0: step-k results of the rows one and two walking steps back
    double k2a[9], k2b[9];         // c_x = 0: step-(k+1) results, likewise
    load_row(r0, fq, false);
    if (a.step >= a.stop_ro[ST_STOP_STEP]) return;
    nine_dropped_stores<double>(rd);
    for (int n = 0; n < nsteps; n++) {
        if (prio) progress_prio(n, nsteps, prio);
        const int r = r0 + DIR * n;
        double f[9];
        static_for<9>([&](auto I) { f[(int)I] = fq[(int)I]; });
        load_row(r + DIR, fq, n + 1 >= nsteps);                      // in flight meanwhile
        if (act1) {                                                   // phase 1: step k, row r
            collide(f);
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
        double g[9];
        if (act2 && n >= 2) {                                         // phase 2: step k+1, row r-DIR
            const int y = r - DIR;
            static_for<9>([&](auto I) {
                constexpr int i = I;
                if constexpr (V::cx[i] != 0) g[i] = ring1[rrow(i, y - V::cy[i]) * E + (t - V::cx[i])];
                else g[i] = age(i) == 0 ? f[i] : (age(i) == 1 ? k1a[i] : k1b[i]);
            });
            collide(g);
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
        double o[9] = {};
        if (act3 && out) {
            static_for<9>([&](auto I) {
                constexpr int i = I;
                if constexpr (V::cx[i] != 0) o[i] = ring2[rrow(i, y - V::cy[i]) * E + (t - V::cx[i])];
                else o[i] = age(i) == 0 ? h2[i] : (age(i) == 1 ? k2a[i] : k2b[i]);
            });
            collide(o);
        }
        static_for<9>([&](auto I) {
            if constexpr (V::cx[(int)I] == 0) {
                k2b[(int)I] = k2a[(int)I];
                k2a[(int)I] = h2[(int)I];
            }
        });
        const unsigned sto = act3 && out ? 0u : BUF_OOB;
        static_for<9>([&](auto I) { buf_store_nt(o[(int)I], rd, vst | sto, (int)(8 * at(a.g, (int)I, y, 0))); });
    }
}

// Strip edges on multiples of `unit` positions (16: 128-byte lines).
template <int COL, int NW, int NTP = 0>
__global__ __launch_bounds__(64 * NW) void k_step3_astrip(StepArgs a, int remap, int nstrip, int nseg, int unit,
                                                         int prio)
{
    constexpr int E = 64 * NW;
    extern __shared__ double rings[];   // ring 1 [XRING_ROWS][E], ring 2 [XRING_ROWS][E]
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if (remap) b = xcd_runs(b, nb);   // (L = 2048: 1,017 workgroups; 88.6k vs 87.9k MLUPS without)
    const int s = (int)(b % (unsigned)nstrip);
    const int units = a.g.nx / unit;
    const int x0 = (int)((long long)s * units / nstrip) * unit;
    const int w = (int)((long long)(s + 1) * units / nstrip) * unit - x0;
    const int seg = (int)(b / (unsigned)nstrip);
    const int ya = (int)((long long)seg * a.g.ny / nseg);
    const int rows = (int)((long long)(seg + 1) * a.g.ny / nseg) - ya;
    if (seg & 1) strip3_walk_aligned<COL, NW, -1, NTP>(a, x0, w, ya, rows, rings, rings + XRING_ROWS * E, prio);
    else strip3_walk_aligned<COL, NW, 1, NTP>(a, x0, w, ya, rows, rings, rings + XRING_ROWS * E, prio);
}

// Four and five steps per launch (stripP_walk_aligned): strip edges on
// multiples of `unit` positions, as k_step3_astrip; P-1 rings of
// XSLOTS x 64 NW values in dynamic LDS.
template <int COL, int NW>
__global__ __launch_bounds__(64 * NW) void k_step4_astrip(StepArgs a, int remap, int nstrip, int nseg, int unit,
                                                         int prio)
{
    extern __shared__ double rings[];
    auto collide = [&](double (&f)[9]) {
        double vx, vy, vz;
        collide_sel<9, COL>(f, a.k, vx, vy, vz);
    };
    astripP_workgroup<double, NW, 4>(a, remap, nstrip, nseg, unit, prio, rings, collide);
}

template <int COL, int NW>
__global__ __launch_bounds__(64 * NW) void k_step5_astrip(StepArgs a, int remap, int nstrip, int nseg, int unit,
                                                         int prio)
{
    extern __shared__ double rings[];
    auto collide = [&](double (&f)[9]) {
        double vx, vy, vz;
        collide_sel<9, COL>(f, a.k, vx, vy, vz);
    };
    astripP_workgroup<double, NW, 5>(a, remap, nstrip, nseg, unit, prio, rings, collide);
}

template <int COL, int NW>
__global__ __launch_bounds__(64 * NW) void k_step6_astrip(StepArgs a, int remap, int nstrip, int nseg, int unit,
                                                         int prio)
{
    extern __shared__ double rings[];
    auto collide = [&](double (&f)[9]) {
        double vx, vy, vz;
        collide_sel<9, COL>(f, a.k, vx, vy, vz);
    };
    astripP_workgroup<double, NW, 6>(a, remap, nstrip, nseg, unit, prio, rings, collide);
}

// Strips of at most 64 NW - 4 positions, as few as the width allows;
// segments as for launch_strip3_t with SEG = 0.
template <int NW, bool ALIGNED>
constexpr size_t gstrip3_lds()
{
    return ALIGNED ? sizeof(double) * XRING_ROWS * 2 * (64 * NW)
                   : sizeof(double) * LEAN_ROWS * ((64 * NW) + (64 * NW - 2));
}

// Workgroups of k_step3_gstrip / k_step3_astrip <COL, NW> that fit a CU.
template <int COL, int NW, bool ALIGNED>
int gstrip3_per_cu()
{
    static const int per_cu = [] {
        constexpr size_t lds = gstrip3_lds<NW, ALIGNED>();
        const void *k = ALIGNED ? reinterpret_cast<const void *>(&k_step3_astrip<COL, NW>)
                                : reinterpret_cast<const void *>(&k_step3_gstrip<COL, NW>);
        if constexpr (lds > 65536) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, k, 64 * NW, lds) != hipSuccess || nblk < 1) nblk = 1;
        return nblk;
    }();
    return per_cu;
}

// Strips and segments of one three-step launch on strips of at most tx
// positions (any width: as few strips as the width allows; else tx-wide
// strips) with per_cu workgroups fitting a CU: as many segments as make one
// workgroup per resident slot, at least 16 rows each.
// Whether and how the aligned walks set their issue priority by progress
// (progress_prio's scheme, 0 = not at all): where two or more workgroups
// share a CU (ERLBM_FUSE_PRIO: -1 so, 0 never, 1 equal quarters, 2 the
// 50 / 75 / 90% bands, the default; L = 2048 FP64 91.8-92.2k MLUPS against
// 90.8-91.0k with quarters, FP32 +0.9%).
int walk_prio(const erlbm_lattice *h, long long workgroups)
{
    return h->tune_fuse_prio >= 0 ? h->tune_fuse_prio : (workgroups >= 2ll * h->n_cu ? 2 : 0);
}

void strip3_grid(const erlbm_lattice *h, int tx, bool any_width, int per_cu, int &nstrip, int &nseg, int unit = 1)
{
    if (h->tune_fuse_slots > 0) per_cu = std::min(per_cu, h->tune_fuse_slots);
    nstrip = any_width ? (h->nx / unit + tx / unit - 1) / (tx / unit) : h->nx / tx;
    nseg = (int)std::max(1ll, std::min<long long>(h->ny / 16, (long long)h->n_cu * per_cu / nstrip));
}

// Strip edges of the lane-aligned walk on 128-byte lines where the width
// is a multiple of 16 positions (same-box MLUPS, edges anywhere / on lines:
// L = 2048 87.9k / 88.9k, 4096 96.5k / 99.4k, 1536 71.6k / 72.3k: no line
// at a strip edge is read twice or written in two parts).
int astrip_unit(const erlbm_lattice *h)
{
    return h->nx % 16 == 0 ? 16 : 1;
}

// Workgroups one launch of the 256-wide lean walk with sized segments (26)
// or of the lane-aligned walk with four wavefronts (42) makes here.
long long strip3_workgroups(const erlbm_lattice *h, int f)
{
    auto per_cu = [&](auto col) -> int {
        constexpr int C = decltype(col)::value;
        return f == FUSE3_ASTRIP_4W ? gstrip3_per_cu<C, 4, true>() : strip3_per_cu<C, 256, true>();
    };
    int pc = 1;
    switch (h->col) {
    case ERLBM_COL_BGK: pc = per_cu(std::integral_constant<int, COL_BGK>{}); break;
    case ERLBM_COL_MRT: pc = per_cu(std::integral_constant<int, COL_MRT>{}); break;
    case ERLBM_COL_REG: pc = per_cu(std::integral_constant<int, COL_REG>{}); break;
    default: pc = per_cu(std::integral_constant<int, COL_EXT_REG>{}); break;
    }
    int nstrip, nseg;
    strip3_grid(h, FUSE_SHAPES[f].tx, f == FUSE3_ASTRIP_4W, pc, nstrip, nseg, f == FUSE3_ASTRIP_4W ? astrip_unit(h) : 1);
    return (long long)nstrip * nseg;
}

template <int COL, int NW, bool ALIGNED = false>
void launch_gstrip3_t(erlbm_lattice *h, const StepArgs &a)
{
    constexpr size_t lds = gstrip3_lds<NW, ALIGNED>();
    const int unit = ALIGNED ? astrip_unit(h) : 1;
    int nstrip, nseg;
    // The moment-space build of the four-wavefront aligned walk fits five
    // workgroups per CU (92 VGPRs); its segments are sized for four all the
    // same (L = 2048, MLUPS: sized for five 83.7k, four 102.6k, three 101.2k):
    // shorter segments recompute more halo rows
    int per_cu = gstrip3_per_cu<COL, NW, ALIGNED>();
    if ((COL & COL_MOMENT) != 0 && ALIGNED && NW == 4) per_cu = std::min(per_cu, 4);
    strip3_grid(h, 64 * NW - 4, true, per_cu, nstrip, nseg, unit);
    if constexpr (ALIGNED) {
        if (NW == 4 && h->tune_walk_nt == 1)
            hipLaunchKernelGGL((k_step3_astrip<COL, NW, 1>), dim3((unsigned)(nstrip * nseg)), dim3(64 * NW), lds,
                               h->stream, a, h->tune_fuse_remap, nstrip, nseg, unit, walk_prio(h, nstrip * nseg));
        else
            hipLaunchKernelGGL((k_step3_astrip<COL, NW>), dim3((unsigned)(nstrip * nseg)), dim3(64 * NW), lds,
                               h->stream, a, h->tune_fuse_remap, nstrip, nseg, unit, walk_prio(h, nstrip * nseg));
    }
    else
        hipLaunchKernelGGL((k_step3_gstrip<COL, NW>), dim3((unsigned)(nstrip * nseg)), dim3(64 * NW), lds, h->stream,
                           a, h->tune_fuse_remap, nstrip, nseg);
}

// Four-step lane-aligned walk: three rings of one row per c_x != 0
// population, 64 NW positions each (36 KB for NW = 4).
template <int NW>
constexpr size_t astrip4_lds()
{
    return sizeof(double) * 3 * XSLOTS * (64 * NW);
}

template <int COL, int NW>
int astrip4_per_cu()
{
    static const int per_cu = [] {
        constexpr size_t lds = astrip4_lds<NW>();
        const void *k = reinterpret_cast<const void *>(&k_step4_astrip<COL, NW>);
        if constexpr (lds > 65536) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, k, 64 * NW, lds) != hipSuccess || nblk < 1) nblk = 1;
        return nblk;
    }();
    return per_cu;
}

// Workgroups one launch of the four-step walk makes here.
long long astrip4_workgroups(const erlbm_lattice *h)
{
    int pc = 1;
    switch (h->col) {
    case ERLBM_COL_BGK: pc = astrip4_per_cu<COL_BGK, 4>(); break;
    case ERLBM_COL_MRT: pc = astrip4_per_cu<COL_MRT, 4>(); break;
    case ERLBM_COL_REG: pc = astrip4_per_cu<COL_REG, 4>(); break;
    default: pc = astrip4_per_cu<COL_EXT_REG, 4>(); break;
    }
    int nstrip, nseg;
    strip3_grid(h, 250, true, pc, nstrip, nseg, astrip_unit(h));
    return (long long)nstrip * nseg;
}

// Five-step lane-aligned walk: four such rings (48 KB for NW = 4).
template <int NW>
constexpr size_t astrip5_lds()
{
    return sizeof(double) * 4 * XSLOTS * (64 * NW);
}

template <int COL, int NW>
int astrip5_per_cu()
{
    static const int per_cu = [] {
        constexpr size_t lds = astrip5_lds<NW>();
        const void *k = reinterpret_cast<const void *>(&k_step5_astrip<COL, NW>);
        if constexpr (lds >= 65536) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, k, 64 * NW, lds) != hipSuccess || nblk < 1) nblk = 1;
        return nblk;
    }();
    return per_cu;
}

// Workgroups one launch of the five-step walk makes here (counted for the
// moment-space E-RLBM build, the one the size rule gives this walk).
long long astrip5_workgroups(const erlbm_lattice *h)
{
    int nstrip, nseg;
    strip3_grid(h, 248, true, astrip5_per_cu<COL_EXT_REG | COL_MOMENT, 4>(), nstrip, nseg, astrip_unit(h));
    return (long long)nstrip * nseg;
}

template <int COL, int NW>
void launch_astrip5_t(erlbm_lattice *h, const StepArgs &a)
{
    constexpr size_t lds = astrip5_lds<NW>();
    const int unit = astrip_unit(h);
    int nstrip, nseg;
    strip3_grid(h, 64 * NW - 8, true, astrip5_per_cu<COL, NW>(), nstrip, nseg, unit);
    hipLaunchKernelGGL((k_step5_astrip<COL, NW>), dim3((unsigned)(nstrip * nseg)), dim3(64 * NW), lds, h->stream, a,
                       h->tune_fuse_remap, nstrip, nseg, unit, walk_prio(h, nstrip * nseg));
}

// Six-step lane-aligned walk: five such rings (60 KB for NW = 4).
template <int COL, int NW>
void launch_astrip6_t(erlbm_lattice *h, const StepArgs &a)
{
    constexpr size_t lds = sizeof(double) * 5 * XSLOTS * (64 * NW);
    static const int per_cu = [] {
        const void *k = reinterpret_cast<const void *>(&k_step6_astrip<COL, NW>);
        if constexpr (lds >= 65536) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, k, 64 * NW, lds) != hipSuccess || nblk < 1) nblk = 1;
        return nblk;
    }();
    const int unit = astrip_unit(h);
    int nstrip, nseg;
    strip3_grid(h, 64 * NW - 10, true, per_cu, nstrip, nseg, unit);
    hipLaunchKernelGGL((k_step6_astrip<COL, NW>), dim3((unsigned)(nstrip * nseg)), dim3(64 * NW), lds, h->stream, a,
                       h->tune_fuse_remap, nstrip, nseg, unit, walk_prio(h, nstrip * nseg));
}

template <int COL, int NW>
void launch_astrip4_t(erlbm_lattice *h, const StepArgs &a)
{
    constexpr size_t lds = astrip4_lds<NW>();
    const int unit = astrip_unit(h);
    int nstrip, nseg;
    strip3_grid(h, 64 * NW - 6, true, astrip4_per_cu<COL, NW>(), nstrip, nseg, unit);
    hipLaunchKernelGGL((k_step4_astrip<COL, NW>), dim3((unsigned)(nstrip * nseg)), dim3(64 * NW), lds, h->stream, a,
                       h->tune_fuse_remap, nstrip, nseg, unit, walk_prio(h, nstrip * nseg));
}

// Segments as many as make one workgroup per resident slot, at least 16
// planes each.
template <int COL, int TX, int TY, int NTP = 0>
void launch_zwalk_t(erlbm_lattice *h, const StepArgs &a)
{
    if constexpr (TX == 64 && TY == 8 && NTP == 0) {   // the default tile: load-hint policy knob
        if (h->tune_zwalk_nt == 1) return launch_zwalk_t<COL, TX, TY, 1>(h, a);
        if (h->tune_zwalk_nt == 2) return launch_zwalk_t<COL, TX, TY, 2>(h, a);
    }
    constexpr size_t lds = sizeof(double) * ZRING_PLANES * (TX + 2) * (TY + 2);
    const void *k = reinterpret_cast<const void *>(&k_step2_zwalk<COL, TX, TY, NTP>);
    static const int per_cu = [&] {
        if constexpr (lds > 65536) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        int nblk = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, k, zwalk_threads<TX, TY>(), lds) != hipSuccess || nblk < 1)
            nblk = 1;
        return nblk;
    }();
    const int ntile = (h->nx / TX) * (h->ny / TY);
    const int nseg = (int)std::max(1ll, std::min<long long>(h->nz / 16, (long long)h->n_cu * per_cu / ntile));
    hipLaunchKernelGGL((k_step2_zwalk<COL, TX, TY, NTP>), dim3((unsigned)(ntile * nseg)), dim3(zwalk_threads<TX, TY>()),
                       lds, h->stream, a, h->tune_fuse_remap, nseg);
}

template <int COL>
void launch_fused_c(erlbm_lattice *h, const StepArgs &a)
{
    if constexpr ((COL & COL_MOMENT) != 0) {
        // the moment-space collision on the shapes the size rule picks and
        // the two-step remainder; any other shape keeps the reference's order
        switch (h->fuse_active) {
        case FUSE2_ZWALK_64x8: launch_zwalk_t<COL, 64, 8>(h, a); return;
        case FUSE3_STRIP_SIZED: launch_strip3_t<COL, 256, 0, true>(h, a); return;
        case FUSE3_GSTRIP_4W: launch_gstrip3_t<COL, 4>(h, a); return;
        case FUSE3_ASTRIP_2W: launch_gstrip3_t<COL, 2, true>(h, a); return;
        case FUSE3_ASTRIP_3W: launch_gstrip3_t<COL, 3, true>(h, a); return;
        case FUSE3_ASTRIP_4W: launch_gstrip3_t<COL, 4, true>(h, a); return;
        case FUSE4_ASTRIP_4W: launch_astrip4_t<COL, 4>(h, a); return;
        case FUSE5_ASTRIP_4W: launch_astrip5_t<COL, 4>(h, a); return;
        case FUSE6_ASTRIP_4W: launch_astrip6_t<COL, 4>(h, a); return;
        case FUSE2_STRIP_256x32: launch_strip_t<COL, 256, 32>(h, a); return;
        case 1: launch_fused_t<COL, 64, 8>(h, a); return;   // the size rule's two-step tiles
        case 3: launch_fused_t<COL, 32, 8>(h, a); return;
        case 7: launch_fused_t<COL, 64, 4>(h, a); return;
        default: launch_fused_c<(COL & ~COL_MOMENT)>(h, a); return;
        }
    }
    if constexpr (COL != COL_MRT) {   // D3Q19 has no MRT
        switch (h->fuse_active) {
        case FUSE2_ZWALK_32x8: launch_zwalk_t<COL, 32, 8>(h, a); return;
        case FUSE2_ZWALK_64x4: launch_zwalk_t<COL, 64, 4>(h, a); return;
        case FUSE2_ZWALK_32x4: launch_zwalk_t<COL, 32, 4>(h, a); return;
        case FUSE2_ZWALK_64x8: launch_zwalk_t<COL, 64, 8>(h, a); return;
        case FUSE2_ZWALK_128x4: launch_zwalk_t<COL, 128, 4>(h, a); return;
        case FUSE2_ZWALK_32x16: launch_zwalk_t<COL, 32, 16>(h, a); return;
        default: break;
        }
    }
    switch (h->fuse_active) {
    case 22: launch_strip3_t<COL, 256, 32>(h, a); return;
    case 23: launch_strip3_t<COL, 256, 16>(h, a); return;
    case 24: launch_strip3_t<COL, 128, 32>(h, a); return;
    case 25: launch_strip3_t<COL, 256, 64>(h, a); return;
    case 26: launch_strip3_t<COL, 256, 0, true>(h, a); return;
    case 27: launch_strip3_t<COL, 128, 0, true>(h, a); return;
    case 28: launch_strip3_t<COL, 256, 32, true>(h, a); return;
    case 29: launch_strip3_t<COL, 256, 0, false>(h, a); return;
    case FUSE3_GSTRIP_1W: launch_gstrip3_t<COL, 1>(h, a); return;
    case FUSE3_GSTRIP_2W: launch_gstrip3_t<COL, 2>(h, a); return;
    case FUSE3_GSTRIP_3W: launch_gstrip3_t<COL, 3>(h, a); return;
    case FUSE3_GSTRIP_4W: launch_gstrip3_t<COL, 4>(h, a); return;
    case FUSE3_ASTRIP_2W: launch_gstrip3_t<COL, 2, true>(h, a); return;
    case FUSE3_ASTRIP_3W: launch_gstrip3_t<COL, 3, true>(h, a); return;
    case FUSE3_ASTRIP_4W: launch_gstrip3_t<COL, 4, true>(h, a); return;
    case FUSE3_ASTRIP_8W: launch_gstrip3_t<COL, 8, true>(h, a); return;
    case FUSE3_ASTRIP_5W: launch_gstrip3_t<COL, 5, true>(h, a); return;
    case FUSE3_ASTRIP_6W: launch_gstrip3_t<COL, 6, true>(h, a); return;
    case FUSE4_ASTRIP_4W: launch_astrip4_t<COL, 4>(h, a); return;
    case FUSE5_ASTRIP_4W: launch_astrip5_t<COL, 4>(h, a); return;
    case FUSE6_ASTRIP_4W: launch_astrip6_t<COL, 4>(h, a); return;
    case 16: launch_wave_t<COL, 1, 4, 1>(h, a, 2); return;
    case 17: launch_wave_t<COL, 1, 4, 1>(h, a, 1); return;
    case 18: launch_wave_t<COL, 2, 4, 1>(h, a, 2); return;
    case 11: launch_strip_t<COL, 256, 32>(h, a); return;
    case 12: launch_strip_t<COL, 256, 16>(h, a); return;
    case 13: launch_strip_t<COL, 128, 32>(h, a); return;
    case 14: launch_strip_t<COL, 256, 64>(h, a); return;
    case 15: launch_strip_t<COL, 128, 64>(h, a); return;
    case 19: launch_strip_t<COL, 256, 8>(h, a); return;
    case 20: launch_strip_t<COL, 128, 16>(h, a); return;
    case 21: launch_strip_t<COL, 128, 8>(h, a); return;
    default: break;
    }
    switch (h->fuse_active) {
    case 1: launch_fused_t<COL, 64, 8>(h, a); break;
    case 2: launch_fused_t<COL, 32, 16>(h, a); break;
    case 3: launch_fused_t<COL, 32, 8>(h, a); break;
    case 4: launch_fused_t<COL, 128, 4>(h, a); break;
    case 5: launch_fused_t<COL, 16, 16>(h, a); break;
    case 6: launch_fused_t<COL, 32, 32>(h, a); break;
    case 7: launch_fused_t<COL, 64, 4>(h, a); break;
    case 8: launch_fused_t<COL, 128, 8>(h, a); break;
    case 9: launch_fused_t<COL, 256, 4>(h, a); break;
    default: launch_fused_t<COL, 64, 16>(h, a); break;
    }
}

int launch_fused(erlbm_lattice *h, const StepArgs &a)
{
    if (h->elem != 8) return lowp_step2(h, a);
    if (h->arith == ERLBM_ARITH_MOMENT) {
        switch (h->col) {   // coll_MRT keeps the reference's order
        case ERLBM_COL_BGK: launch_fused_c<COL_BGK | COL_MOMENT>(h, a); break;
        case ERLBM_COL_MRT: launch_fused_c<COL_MRT>(h, a); break;
        case ERLBM_COL_REG: launch_fused_c<COL_REG | COL_MOMENT>(h, a); break;
        case ERLBM_COL_EXT_REG: launch_fused_c<COL_EXT_REG | COL_MOMENT>(h, a); break;
        default: return ERLBM_ERR_UNSUPPORTED;
        }
        ERLBM_HIP_TRY(hipGetLastError());
        return ERLBM_OK;
    }
    switch (h->col) {
    case ERLBM_COL_BGK: launch_fused_c<COL_BGK>(h, a); break;
    case ERLBM_COL_MRT: launch_fused_c<COL_MRT>(h, a); break;
    case ERLBM_COL_REG: launch_fused_c<COL_REG>(h, a); break;
    case ERLBM_COL_EXT_REG: launch_fused_c<COL_EXT_REG>(h, a); break;
    default: return ERLBM_ERR_UNSUPPORTED;
    }
    ERLBM_HIP_TRY(hipGetLastError());
    return ERLBM_OK;
}

}  // namespace

namespace {

StepArgs step_args(const erlbm_lattice *h)
{
    StepArgs a;
    a.g = geom(h);
    a.nbr = h->nbr;
    a.k.omega1 = 1.0 - 1.0 / h->tau;
    a.k.omega2 = 1.0 - 1.0 / h->tau_nh;
    a.k.oot = 1.0 / h->tau;
    a.k.m = moment_consts(h->q, h->col, h->tau, h->tau_nh);
    a.bf.ax = h->acc[0];
    a.bf.ay = h->acc[1];
    a.bf.az = h->acc[2];
    a.bf.factor = 1. - 1. / (2. * h->tau);   // source(), LBM_functions.cpp:5664
    a.bc.inlet_x = h->inlet_x;
    a.bc.inlet_west = h->inlet_x >= 0 && h->inlet_x < h->nx / 2;
    a.bc.u_in = h->u_in;
    a.bc.outlet_x = h->outlet_x;
    a.bc.outlet_dx = (h->outlet_x >= 0 && h->outlet_x < h->nx / 2) ? 1 : -1;
    a.bc.rho_out = h->rho_out;
    a.dens_limit = h->dens_limit;
    a.status = h->status;
    a.stop_ro = h->status;
    return a;
}

int dispatch_step(erlbm_lattice *h, const StepArgs &a, bool stream)
{
    if (vec_step_applies(h)) return vec_step(h, a, stream);
    if (h->elem != 8) return lowp_step(h, a, stream);
    return h->q == 9 ? launch_step<9>(h, a, stream) : launch_step<19>(h, a, stream);
}

}  // namespace

namespace {

// Five-step shape for a remainder of five steps after six-step launches:
// the five-step walk where the size rule would take it.  0: none.
long long astrip5_workgroups(const erlbm_lattice *h);
int remainder5_shape(const erlbm_lattice *h)
{
    if (fuse_steps(h->fuse_active) != 6) return 0;
    const int f = FUSE5_ASTRIP_4W;
    return fuse_shape_fits(h, f) && 2 * astrip5_workgroups(h) >= 3ll * h->n_cu ? f : 0;
}

// Four-step shape for a remainder of four steps after five- or six-step launches:
// the four-step walk where the size rule would take it.  0: none.
long long astrip4_workgroups(const erlbm_lattice *h);
int remainder4_shape(const erlbm_lattice *h)
{
    if (fuse_steps(h->fuse_active) < 5) return 0;
    const int f = FUSE4_ASTRIP_4W;
    return fuse_shape_fits(h, f) && 2 * astrip4_workgroups(h) >= 3ll * h->n_cu ? f : 0;
}

// Three-step shape for a remainder of three steps after four- or five-step launches:
// the lane-aligned three-step walk where it fills the chip.  0: none.
int remainder3_shape(const erlbm_lattice *h)
{
    if (fuse_steps(h->fuse_active) < 4) return 0;
    const int f = FUSE3_ASTRIP_4W;
    return fuse_shape_fits(h, f) && strip3_workgroups(h, f) >= (19ll * h->n_cu + 19) / 20 ? f : 0;
}

// Two-step shape for a remainder of two steps after three- or four-step launches (the
// two-step strip walk, else the 64x8 tiles, else the wavefront strips, where
// they fill the chip): one launch instead of two one-step launches (L = 2048:
// ~116 us against ~192 us), bitwise the same steps.  0: none.
int remainder2_shape(const erlbm_lattice *h)
{
    if (fuse_steps(h->fuse_active) < 3) return 0;
    for (int f : {FUSE2_STRIP_256x32, 1, 16})
        if (fuse_shape_fits(h, f) && fuse_workgroups(h, FUSE_SHAPES[f]) >= h->n_cu) return f;
    return 0;
}

}  // namespace

namespace {

// The multi-step kernels do not read the stop word, so they run only while no
// stop test can have stopped the run.  After calls with stop tests, a call
// without them reads the status once (the tests queued before it have ended
// by then): with no stop recorded, those kernels are allowed again.
int refresh_checked(erlbm_lattice *h, int energy_interval)
{
    if (!h->checked || energy_interval > 0 || h->dens_limit > 0.0) return ERLBM_OK;
    long long run = 0;
    int reason = 0;
    const int rc = steps_run(h, &run, &reason);
    if (rc) return rc;
    if (reason == 0) h->checked = false;
    return ERLBM_OK;
}

}  // namespace

extern "C" int erlbm_step(erlbm_lattice *h, int n_steps, int energy_interval, double energy_factor)
{
    if (!h || n_steps < 0) return ERLBM_ERR_ARG;
    if (!h->uploaded) {
        set_last_error_msg("erlbm_step: upload populations first");
        return ERLBM_ERR_STATE;
    }
    StepArgs a = step_args(h);
    if (int rc = refresh_checked(h, energy_interval)) return rc;
    if (energy_interval > 0 || h->dens_limit > 0.0) h->checked = true;
    h->fuse_active = fuse_shape(h, energy_interval);
    const bool fuse = h->fuse_active > 0;
    const int spl = fuse_steps(h->fuse_active);   // steps per fused launch
    const int rem2 = fuse ? remainder2_shape(h) : 0;
    const int rem3 = fuse ? remainder3_shape(h) : 0;
    const int rem4 = fuse ? remainder4_shape(h) : 0;
    const int rem5 = fuse ? remainder5_shape(h) : 0;
    for (int s = 0; s < n_steps;) {
        const long long k = h->launched;
        const int cur = (h->base + (int)(k & 1)) & 1;
        if (energy_interval > 0 && k > 0 && k % energy_interval == 0) {
            const bool ordered = h->n_fluid <= (1 << 20);
            int rc = monitor(h, h->buf[cur], true, ordered, h->scalars, (unsigned long long)k, energy_factor);
            if (rc) return rc;
        }
        a.src = static_cast<const double *>(h->buf[cur]);
        a.dst = static_cast<double *>(h->buf[cur ^ 1]);
        a.step = (unsigned long long)k;
        if (fuse && k > 0 && s + spl <= n_steps) {
            // steps k .. k+spl-1 in one launch, written to the other buffer:
            // after two steps that is the buffer a single step would have
            // written h_k to, so the parity of the buffers flips (base)
            int rc = launch_fused(h, a);
            if (rc) return rc;
            h->launched += spl;
            if (spl % 2 == 0) h->base ^= 1;
            s += spl;
            continue;
        }
        if (rem5 && k > 0 && n_steps - s == 5) {
            const int keep = h->fuse_active;
            h->fuse_active = rem5;
            int rc = launch_fused(h, a);
            h->fuse_active = keep;
            if (rc) return rc;
            h->launched += 5;
            s += 5;
            continue;
        }
        if (rem4 && k > 0 && n_steps - s == 4) {
            const int keep = h->fuse_active;
            h->fuse_active = rem4;
            int rc = launch_fused(h, a);
            h->fuse_active = keep;
            if (rc) return rc;
            h->launched += 4;
            h->base ^= 1;
            s += 4;
            continue;
        }
        if (rem3 && k > 0 && n_steps - s == 3) {
            const int keep = h->fuse_active;
            h->fuse_active = rem3;
            int rc = launch_fused(h, a);
            h->fuse_active = keep;
            if (rc) return rc;
            h->launched += 3;
            s += 3;
            continue;
        }
        if (rem2 && k > 0 && n_steps - s == 2) {
            const int keep = h->fuse_active;
            h->fuse_active = rem2;
            int rc = launch_fused(h, a);
            h->fuse_active = keep;
            if (rc) return rc;
            h->launched += 2;
            h->base ^= 1;
            s += 2;
            continue;
        }
        int rc = dispatch_step(h, a, k > 0);
        if (rc) return rc;
        h->launched++;
        s++;
    }
    return ERLBM_OK;
}

namespace {

// Launch shapes the autotuner tries for a lattice (all give identical
// populations; only speed differs).
struct LaunchShape {
    int ntl, block, npt, pair, vec;
    const char *name;
    int fuse = 0, fuse_remap = 0;   // two-step kernel tile (FUSE_SHAPES index), 0 = one step per launch
};

std::vector<LaunchShape> candidate_shapes(const erlbm_lattice *h)
{
    std::vector<LaunchShape> c;
    const bool periodic = h->all_fluid && !h->forced && h->inlet_x < 0 && h->outlet_x < 0;
    if (!periodic) return c;
    if (h->elem == 8 && h->q == 9) {
        struct Named {
            int f;
            const char *name[2];
        };
        static const Named fused[] = {
            {11, {"two steps, 256-wide strips x 32 rows", "two steps, 256-wide strips x 32 rows, XCD runs"}},
            {13, {"two steps, 128-wide strips x 32 rows", "two steps, 128-wide strips x 32 rows, XCD runs"}},
            {1, {"two steps, 64x8 tiles", "two steps, 64x8 tiles, XCD runs"}},
            {3, {"two steps, 32x8 tiles", "two steps, 32x8 tiles, XCD runs"}},
            {7, {"two steps, 64x4 tiles", "two steps, 64x4 tiles, XCD runs"}},
            {16, {"two steps, wavefront strips", "two steps, wavefront strips, XCD runs"}},
            {22, {"three steps, 256-wide strips x 32 rows", "three steps, 256-wide strips x 32 rows, XCD runs"}},
            {26, {"three steps, 256-wide strips, sized segments", "three steps, 256-wide strips, sized segments, XCD runs"}},
            {FUSE3_GSTRIP_4W, {"three steps, strips of any width (4 wavefronts)", "three steps, strips of any width (4 wavefronts), XCD runs"}},
            {FUSE3_GSTRIP_3W, {"three steps, strips of any width (3 wavefronts)", "three steps, strips of any width (3 wavefronts), XCD runs"}},
            {FUSE3_ASTRIP_4W, {"three steps, lane-aligned strips of any width (4 wavefronts)",
                               "three steps, lane-aligned strips of any width (4 wavefronts), XCD runs"}},
            {FUSE3_ASTRIP_3W, {"three steps, lane-aligned strips of any width (3 wavefronts)",
                               "three steps, lane-aligned strips of any width (3 wavefronts), XCD runs"}},
            {FUSE3_ASTRIP_2W, {"three steps, lane-aligned strips of any width (2 wavefronts)",
                               "three steps, lane-aligned strips of any width (2 wavefronts), XCD runs"}},
            {FUSE4_ASTRIP_4W, {"four steps, lane-aligned strips of any width (4 wavefronts)",
                               "four steps, lane-aligned strips of any width (4 wavefronts), XCD runs"}},
            {FUSE5_ASTRIP_4W, {"five steps, lane-aligned strips of any width (4 wavefronts)",
                               "five steps, lane-aligned strips of any width (4 wavefronts), XCD runs"}}};
        for (const Named &nm : fused) {
            if (!fuse_shape_fits(h, nm.f)) continue;
            for (int rm = 0; rm < 2; rm++) {
                LaunchShape ls{2, 256, 1, 0, 0, nm.name[rm], nm.f, rm};
                c.push_back(ls);
            }
        }
        c.push_back({2, 256, 1, 0, 0, "nt-loads=c_x0 block=256"});
        c.push_back({2, 512, 1, 0, 0, "nt-loads=c_x0 block=512"});
        c.push_back({2, 128, 1, 0, 0, "nt-loads=c_x0 block=128"});
        c.push_back({1, 256, 1, 0, 0, "nt-loads=all block=256"});
        c.push_back({0, 256, 1, 0, 0, "nt-loads=none block=256"});
        c.push_back({1, 256, 2, 0, 0, "nt-loads=all nodes/thread=2"});
        if (h->nx % 2 == 0) c.push_back({1, 256, 1, 1, 0, "nt-loads=all paired 16B"});
    } else if (h->elem == 8) {
        static const int zs[3] = {FUSE2_ZWALK_64x8, FUSE2_ZWALK_64x4, FUSE2_ZWALK_32x8};
        static const char *zn[3][2] = {{"two steps, 64x8 tiles walked along z", "two steps, 64x8 tiles walked along z, XCD runs"},
                                       {"two steps, 64x4 tiles walked along z", "two steps, 64x4 tiles walked along z, XCD runs"},
                                       {"two steps, 32x8 tiles walked along z", "two steps, 32x8 tiles walked along z, XCD runs"}};
        for (int k = 0; k < 3; k++) {
            if (!fuse_shape_fits(h, zs[k])) continue;
            for (int rm = 0; rm < 2; rm++) c.push_back({2, 256, 1, 0, 0, zn[k][rm], zs[k], rm});
        }
        c.push_back({2, 256, 1, 0, 0, "nt-loads=c_x0"});
        c.push_back({1, 256, 1, 0, 0, "nt-loads=all"});
        c.push_back({0, 256, 1, 0, 0, "nt-loads=none"});
    } else {
        if (h->q == 9 && !h->f32_arith && h->elem == 4) {
            static const char *names[4][2] = {{"two steps, 256-wide strips x 32 rows", "two steps, 256-wide strips x 32 rows, XCD runs"},
                                              {"two steps, 128-wide strips x 32 rows", "two steps, 128-wide strips x 32 rows, XCD runs"},
                                              {"three steps, 256-wide strips x 32 rows", "three steps, 256-wide strips x 32 rows, XCD runs"},
                                              {"three steps, 128-wide strips x 32 rows", "three steps, 128-wide strips x 32 rows, XCD runs"}};
            const int fs[4] = {FUSE2_STRIP_256x32, FUSE2_STRIP_128x32, FUSE3_STRIP_256x32, FUSE3_STRIP_128x32};
            // the same fit test the step path applies (whole strips, and
            // buffer offsets below 2^31 bytes for the raw-buffer walks)
            for (int k = 0; k < 4; k++) {
                if (!fuse_shape_fits(h, fs[k])) continue;
                for (int rm = 0; rm < 2; rm++) c.push_back({h->tune_ntl, h->tune_block, 1, 0, 0, names[k][rm], fs[k], rm});
            }
            static const char *aligned[3][2] = {
                {"three steps, lane-aligned strips of any width (4 wavefronts)",
                 "three steps, lane-aligned strips of any width (4 wavefronts), XCD runs"},
                {"four steps, lane-aligned strips of any width (4 wavefronts)",
                 "four steps, lane-aligned strips of any width (4 wavefronts), XCD runs"},
                {"five steps, lane-aligned strips of any width (4 wavefronts)",
                 "five steps, lane-aligned strips of any width (4 wavefronts), XCD runs"}};
            const int af[3] = {FUSE3_ASTRIP_4W, FUSE4_ASTRIP_4W, FUSE5_ASTRIP_4W};
            for (int k = 0; k < 3; k++)
                if (fuse_shape_fits(h, af[k]))
                    for (int rm = 0; rm < 2; rm++)
                        c.push_back({h->tune_ntl, h->tune_block, 1, 0, 0, aligned[k][rm], af[k], rm});
        }
        c.push_back({h->tune_ntl, h->tune_block, 1, 0, 0, "one node/thread"});
    }
    if (!h->f32_arith) {
        c.push_back({h->tune_ntl, h->tune_block, 1, 0, 1, "16B vectors, shuffled edges"});
        c.push_back({h->tune_ntl, h->tune_block, 1, 0, 2, "16B vectors, loaded edges"});
    }
    return c;
}

void apply_shape(erlbm_lattice *h, const LaunchShape &s)
{
    h->tune_ntl = s.ntl;
    h->tune_block = s.block;
    h->tune_npt = s.npt;
    h->tune_pair = s.pair;
    h->tune_vec = s.vec;
    h->tune_fuse = s.fuse;
    h->tune_fuse_remap = s.fuse_remap;
    h->fuse_requested = true;   // the measured choice stands, whatever the lattice size
}

}  // namespace

namespace {

// Description of a multi-step shape, in the autotuner's words.
std::string fuse_shape_desc(int f, int remap)
{
    if (f <= 0 || f >= N_FUSE_SHAPES) return "one step per launch";
    const FuseShape fs = FUSE_SHAPES[f];
    const std::string tx = std::to_string(fs.tx), ty = std::to_string(fs.ty);
    std::string d;
    switch (fs.kind) {
    case FUSE_TILE: d = "two steps, " + tx + "x" + ty + " tiles"; break;
    case FUSE_STRIP: d = "two steps, " + tx + "-wide strips x " + ty + " rows"; break;
    case FUSE_WAVE: d = "two steps, wavefront strips"; break;
    case FUSE_STRIP3:
        d = "three steps, " + tx + "-wide strips" + (fs.ty ? " x " + ty + " rows" : std::string(", sized segments"));
        break;
    case FUSE_GSTRIP3:
        d = std::string("three steps, ") + (f >= FUSE3_ASTRIP_2W ? "lane-aligned " : "") + "strips of any width (" +
            std::to_string((fs.tx + 4) / 64) + " wavefronts)";
        break;
    case FUSE_ZWALK: d = "two steps, " + tx + "x" + ty + " tiles walked along z"; break;
    case FUSE_GSTRIP4: d = "four steps, lane-aligned strips of any width (" + std::to_string((fs.tx + 6) / 64) + " wavefronts)"; break;
    case FUSE_GSTRIP5: d = "five steps, lane-aligned strips of any width (" + std::to_string((fs.tx + 8) / 64) + " wavefronts)"; break;
    case FUSE_GSTRIP6: d = "six steps, lane-aligned strips of any width (" + std::to_string((fs.tx + 10) / 64) + " wavefronts)"; break;
    }
    return remap ? d + ", XCD runs" : d;
}

}  // namespace

extern "C" int erlbm_step_shape(erlbm_lattice *h, int *shape, int *steps_per_launch, char *desc, int desc_len)
{
    if (!h) return ERLBM_ERR_ARG;
    if (h->uploaded)
        if (int rc = refresh_checked(h, 0)) return rc;
    const int f = fuse_shape(h, 0);
    if (shape) *shape = f;
    if (steps_per_launch) *steps_per_launch = f > 0 ? fuse_steps(f) : 1;
    if (desc && desc_len > 0) {
        const std::string d = fuse_shape_desc(f, h->tune_fuse_remap);
        std::snprintf(desc, (size_t)desc_len, "%s", d.c_str());
    }
    return ERLBM_OK;
}

extern "C" int erlbm_lattice_autotune(erlbm_lattice *h, int n_steps, char *choice, int choice_len, float *us_per_step)
{
    if (!h || n_steps < 1) return ERLBM_ERR_ARG;
    if (!h->uploaded) {
        set_last_error_msg("erlbm_lattice_autotune: upload populations first");
        return ERLBM_ERR_STATE;
    }
    const std::vector<LaunchShape> shapes = candidate_shapes(h);
    // the handle's launch choice as it was, restored as such if the tuning fails
    struct Kept {
        int ntl, block, npt, pair, vec, fuse, fuse_remap, fuse_active;
        bool fuse_requested;
    } const kept{h->tune_ntl, h->tune_block, h->tune_npt,  h->tune_pair,     h->tune_vec,
                 h->tune_fuse, h->tune_fuse_remap, h->fuse_active, h->fuse_requested};
    const int steps = (n_steps + 5) / 6 * 6;   // whole launches for the two- and three-step candidates
    // scratch copy of the current state, a private stop word: the handle's
    // run is not touched
    const size_t bytes = (size_t)h->extent * h->elem;
    void *tmp[2] = {nullptr, nullptr};
    unsigned long long *st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ERLBM_OK;
    auto hip = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == ERLBM_OK) {
            set_last_error(what, e);
            rc = e == hipErrorOutOfMemory ? ERLBM_ERR_NOMEM : ERLBM_ERR_HIP;
        }
        return rc == ERLBM_OK;
    };
    int best = -1;
    float best_ms = 0.f;
    if (!shapes.empty() && hip(hipMalloc(&tmp[0], bytes), "erlbm_lattice_autotune: hipMalloc") &&
        hip(hipMalloc(&tmp[1], bytes), "erlbm_lattice_autotune: hipMalloc") &&
        hip(hipMalloc(&st, 4 * sizeof(unsigned long long)), "erlbm_lattice_autotune: hipMalloc") &&
        hip(hipEventCreate(&e0), "hipEventCreate") && hip(hipEventCreate(&e1), "hipEventCreate")) {
        const int cur = (h->base + (int)(h->launched & 1)) & 1;
        hip(hipMemcpyAsync(tmp[0], h->buf[cur], bytes, hipMemcpyDeviceToDevice, h->stream), "hipMemcpyAsync");
        hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(64), 0, h->stream, st);
        StepArgs a = step_args(h);
        a.status = st;
        a.stop_ro = st;
        unsigned long long k = 0;
        for (size_t c = 0; c < shapes.size() && rc == ERLBM_OK; c++) {
            apply_shape(h, shapes[c]);
            const bool multi = shapes[c].fuse > 0;
            h->fuse_active = shapes[c].fuse;
            const int spl = multi ? fuse_steps(shapes[c].fuse) : 1;
            float ms = 0.f;
            // one untimed launch, then `steps` steps (a multiple of 6)
            for (int s = -spl; s < steps && rc == ERLBM_OK; s += spl) {
                if (s == 0) hip(hipEventRecord(e0, h->stream), "hipEventRecord");
                a.src = static_cast<const double *>(tmp[k & 1]);
                a.dst = static_cast<double *>(tmp[(k & 1) ^ 1]);
                a.step = 0;
                k++;   // a multi-step launch also writes the other buffer
                const int r = multi ? launch_fused(h, a) : dispatch_step(h, a, true);
                if (r && rc == ERLBM_OK) rc = r;
            }
            if (hip(hipEventRecord(e1, h->stream), "hipEventRecord") &&
                hip(hipEventSynchronize(e1), "hipEventSynchronize") &&
                hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) {
                if (best < 0 || ms < best_ms) {
                    best = (int)c;
                    best_ms = ms;
                }
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(tmp[0]);
    (void)hipFree(tmp[1]);
    (void)hipFree(st);
    if (rc != ERLBM_OK || best < 0) {
        h->tune_ntl = kept.ntl;
        h->tune_block = kept.block;
        h->tune_npt = kept.npt;
        h->tune_pair = kept.pair;
        h->tune_vec = kept.vec;
        h->tune_fuse = kept.fuse;
        h->tune_fuse_remap = kept.fuse_remap;
        h->fuse_active = kept.fuse_active;
        h->fuse_requested = kept.fuse_requested;
    } else {
        apply_shape(h, shapes[best]);
    }
    if (rc) return rc;
    if (choice && choice_len > 0) {
        const char *name = best >= 0 ? shapes[best].name : "fixed (no launch choice for this geometry)";
        std::strncpy(choice, name, (size_t)choice_len - 1);
        choice[choice_len - 1] = '\0';
    }
    if (us_per_step) *us_per_step = best >= 0 ? best_ms * 1000.f / steps : 0.f;
    return ERLBM_OK;
}

extern "C" int erlbm_autotune_candidates(int q, int nx, int ny, int nz, int precision, int n_cu, int *shapes,
                                         int cap, int *n)
{
    if (!n || (q != 9 && q != 19) || nx < 1 || ny < 1 || nz < 1 || (q == 9 && nz != 1) || n_cu < 1 ||
        (cap > 0 && !shapes) || (long long)nx * ny * nz > (1ll << 31) - 1 ||
        (precision != ERLBM_FP64 && precision != ERLBM_FP32_SHIFTED && precision != ERLBM_FP16_SHIFTED &&
         precision != ERLBM_FP32_SHIFTED_F32))
        return ERLBM_ERR_ARG;
    erlbm_lattice h;   // host-side description only: no device memory
    set_layout(&h, q, nx, ny, nz, precision);
    h.n_cu = n_cu;
    const std::vector<LaunchShape> c = candidate_shapes(&h);
    int k = 0;
    for (const LaunchShape &ls : c) {
        if (ls.fuse <= 0) continue;
        if (k < cap) shapes[k] = ls.fuse;
        k++;
    }
    *n = k;
    return ERLBM_OK;
}

extern "C" int erlbm_step_timed(erlbm_lattice *h, int n_steps, float *ms)
{
    if (!h || !ms || n_steps < 0) return ERLBM_ERR_ARG;
    hipEvent_t e0, e1;
    ERLBM_HIP_TRY(hipEventCreate(&e0));
    ERLBM_HIP_TRY(hipEventCreate(&e1));
    int rc = ERLBM_OK;
    if (hipEventRecord(e0, h->stream) != hipSuccess) rc = ERLBM_ERR_HIP;
    if (rc == ERLBM_OK) rc = erlbm_step(h, n_steps, 0, 0.0);
    if (rc == ERLBM_OK && hipEventRecord(e1, h->stream) != hipSuccess) rc = ERLBM_ERR_HIP;
    if (rc == ERLBM_OK && hipEventSynchronize(e1) != hipSuccess) rc = ERLBM_ERR_HIP;
    if (rc == ERLBM_OK && hipEventElapsedTime(ms, e0, e1) != hipSuccess) rc = ERLBM_ERR_HIP;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc == ERLBM_ERR_HIP) set_last_error_msg("erlbm_step_timed: HIP event failure");
    return rc;
}

extern "C" int erlbm_status(erlbm_lattice *h, long long *steps_done, int *stop_reason)
{
    if (!h) return ERLBM_ERR_ARG;
    long long r = 0;
    int reason = 0;
    int rc = steps_run(h, &r, &reason);
    if (rc) return rc;
    if (steps_done) *steps_done = r;
    if (stop_reason) *stop_reason = reason;
    return ERLBM_OK;
}

extern "C" int erlbm_synchronize(erlbm_lattice *h)
{
    if (!h) return ERLBM_ERR_ARG;
    ERLBM_HIP_TRY(hipStreamSynchronize(h->stream));
    return ERLBM_OK;
}

namespace {

// Buffer holding h_{r-1} (post-collision of the last step) or, if no step
// ran, the uploaded ref_0.
int current(erlbm_lattice *h, long long *run, int *cur)
{
    int rc = steps_run(h, run, nullptr);
    if (rc) return rc;
    *cur = (h->base + (int)(*run & 1)) & 1;
    return ERLBM_OK;
}

}  // namespace

namespace erlbm {
namespace lat {

// The populations `which` names, as device AoS in reference fluid order.
// *owned is set when *aos was allocated here (shifted storage); otherwise
// *aos is the handle's spare population buffer, valid until the next step.
int to_aos_device(erlbm_lattice *h, int which, double **aos, void **owned)
{
    *aos = nullptr;
    *owned = nullptr;
    if (!h->uploaded) return ERLBM_ERR_STATE;
    long long run;
    int cur;
    int rc = current(h, &run, &cur);
    if (rc) return rc;
    if (which == ERLBM_POST_COLLISION && run == 0) {
        set_last_error_msg("no step has run since the upload");
        return ERLBM_ERR_STATE;
    }
    const bool stream = (which == ERLBM_STATE) && run > 0;
    if (h->elem != 8) {
        double *stage = nullptr;
        ERLBM_HIP_TRY(hipMalloc(&stage, sizeof(double) * (size_t)h->n_fluid * h->q));
        rc = lowp_to_aos(h, h->buf[cur], stream, stage);
        if (rc) {
            (void)hipFree(stage);
            return rc;
        }
        *aos = stage;
        *owned = stage;
        return ERLBM_OK;
    }
    double *stage = static_cast<double *>(h->buf[cur ^ 1]);
    const double *src = static_cast<const double *>(h->buf[cur]);
    const int nb = mon_blocks(h);
    const uint32_t *nbr = h->all_fluid ? nullptr : h->nbr;
    if (h->q == 9) {
        if (stream)
            hipLaunchKernelGGL((k_to_aos<9, true>), dim3(nb), dim3(BLOCK), 0, h->stream, src, geom(h), nbr,
                               h->vox_of_fluid, h->n_fluid, stage);
        else
            hipLaunchKernelGGL((k_to_aos<9, false>), dim3(nb), dim3(BLOCK), 0, h->stream, src, geom(h), nbr,
                               h->vox_of_fluid, h->n_fluid, stage);
    } else {
        if (stream)
            hipLaunchKernelGGL((k_to_aos<19, true>), dim3(nb), dim3(BLOCK), 0, h->stream, src, geom(h), nbr,
                               h->vox_of_fluid, h->n_fluid, stage);
        else
            hipLaunchKernelGGL((k_to_aos<19, false>), dim3(nb), dim3(BLOCK), 0, h->stream, src, geom(h), nbr,
                               h->vox_of_fluid, h->n_fluid, stage);
    }
    ERLBM_HIP_TRY(hipGetLastError());
    *aos = stage;
    return ERLBM_OK;
}

}  // namespace lat
}  // namespace erlbm

extern "C" int erlbm_download_aos(erlbm_lattice *h, int which, double *inif)
{
    if (!h || !inif) return ERLBM_ERR_ARG;
    if (!h->uploaded) return ERLBM_ERR_STATE;
    double *aos = nullptr;
    void *owned = nullptr;
    int rc = to_aos_device(h, which, &aos, &owned);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * (size_t)h->n_fluid * h->q;
    if (hipMemcpyAsync(inif, aos, bytes, hipMemcpyDeviceToHost, h->stream) != hipSuccess) rc = ERLBM_ERR_HIP;
    if (hipStreamSynchronize(h->stream) != hipSuccess && rc == ERLBM_OK) rc = ERLBM_ERR_HIP;
    (void)hipFree(owned);
    if (rc) set_last_error_msg("erlbm_download_aos: copy failed");
    return rc;
}

extern "C" int erlbm_kinetic_energy(erlbm_lattice *h, int which, int ordered, double *energy)
{
    if (!h || !energy) return ERLBM_ERR_ARG;
    if (!h->uploaded) return ERLBM_ERR_STATE;
    long long run;
    int cur;
    int rc = current(h, &run, &cur);
    if (rc) return rc;
    if (which == ERLBM_POST_COLLISION && run == 0) return ERLBM_ERR_STATE;
    const bool stream = (which == ERLBM_STATE) && run > 0;
    rc = monitor(h, h->buf[cur], stream, ordered != 0, h->scalars);
    if (rc) return rc;
    ERLBM_HIP_TRY(hipMemcpyAsync(energy, h->scalars, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    ERLBM_HIP_TRY(hipStreamSynchronize(h->stream));
    return ERLBM_OK;
}

extern "C" int erlbm_density_monitor(erlbm_lattice *h, int which, double thresh, double *max_abs_rho,
                                     int *violated)
{
    if (!h) return ERLBM_ERR_ARG;
    if (!h->uploaded) return ERLBM_ERR_STATE;
    long long run;
    int cur;
    int rc = current(h, &run, &cur);
    if (rc) return rc;
    if (which == ERLBM_POST_COLLISION && run == 0) return ERLBM_ERR_STATE;
    const bool stream = (which == ERLBM_STATE) && run > 0;
    rc = monitor(h, h->buf[cur], stream, false, h->scalars);
    if (rc) return rc;
    double s[2];
    int nf = 0;
    ERLBM_HIP_TRY(hipMemcpyAsync(s, h->scalars, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    ERLBM_HIP_TRY(hipMemcpyAsync(&nf, h->nonfinite, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    ERLBM_HIP_TRY(hipStreamSynchronize(h->stream));
    if (max_abs_rho) *max_abs_rho = s[1];
    if (violated) *violated = (nf != 0) || (thresh > 0 && s[1] > thresh);
    return ERLBM_OK;
}
