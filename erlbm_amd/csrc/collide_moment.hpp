// collide_moment.hpp -- the BGK / regularized / extended-regularized
// collision re-associated in moment space (FP64; same operator, not bitwise).
//
// The operators of collide.hpp -- coll_Ext_Reg
// (SimBoltz_Functions/Collision_operators.cpp:1181-1307), coll_Reg (:859-967)
// and coll_BGK (:11-49), unforced -- are one formula,
//   f*_i = W f_i + K0 feq_i + K1 f1_i,
// with (W, K0, K1) = (w2, 1 - w2, w1 - w2) for E-RLBM
// (f* = feq + w1 f1 + w2 (f - feq - f1), w1 = 1 - 1/tau, w2 = 1 - 1/tau_nh),
// (0, 1, w1) for the regularized operator and (w1, 1/tau, 0) for BGK.  With
// the second-order equilibrium of dist_eq and the moment identities of D2Q9
// and D3Q19 (sum_i w_i c_i c_i = I/3, sum_i w_i c_ia c_ib c_ic c_id =
// (d_ab d_cd + d_ac d_bd + d_ad d_bc)/9), the equilibrium's stress is
// j u + rho I/3, so f1's stress Pi1 = sum_i c_i c_i (f_i - feq_i) equals
// Pi - j u - rho I/3 (Pi = sum_i c_i c_i f_i, j = sum_i c_i f_i, u = j/rho),
// and every term is a quadratic polynomial in c_i:
//   f*_i = W f_i + w_i (alpha + beta.c_i + c_i.Gamma.c_i)
//   alpha    = (K0 + D K1/2) rho - 1.5 T (j.u) - 1.5 K1 tr Pi
//   beta     = 3 K0 j
//   Gamma_ab = 4.5 T j_a u_b + 4.5 K1 Pi_ab - 1.5 K1 rho d_ab
// (T = K0 - K1 = 1/tau for all three operators, D the dimension).  Per node:
// ten moments from pair sums and differences of opposite populations, a
// handful of class values (rest, axis, and per coordinate plane the two
// diagonal classes), then one add and one fused multiply-add per population:
// 77 FP64 instructions per D2Q9 node against 175 for the reference's
// operation order.  The rounding differs from the reference's by a few units
// in the last place per step (the reference's own FMA build differs from
// its strict build likewise); the library's tests bound the difference at
// 1e-10 maximum relative population error after 1000 steps.
#pragma once

#include "collide.hpp"
#include "vset.hpp"

namespace erlbm {

namespace mom {

template <int Q>
struct Cls {
    using V = VSet<Q>;
    static constexpr int nz(int i) { return (V::cx[i] != 0) + (V::cy[i] != 0) + (V::cz[i] != 0); }
    // axis of an axis direction (0 x, 1 y, 2 z) and its sign
    static constexpr int axis(int i) { return V::cx[i] != 0 ? 0 : (V::cy[i] != 0 ? 1 : 2); }
    static constexpr int axis_sign(int i) { return V::cx[i] + V::cy[i] + V::cz[i]; }
    // plane of a diagonal: 0 xy, 1 xz, 2 yz; first and second axis; signs
    static constexpr int plane(int i) { return V::cz[i] == 0 ? 0 : (V::cy[i] == 0 ? 1 : 2); }
    static constexpr int first(int i) { return plane(i) == 2 ? V::cy[i] : V::cx[i]; }
    static constexpr int second(int i) { return plane(i) == 0 ? V::cy[i] : V::cz[i]; }
};

}  // namespace mom

// One collision in place (unforced BGK / REG / EXT_REG: COL as in
// collide.hpp, 1 / 3 / 4).  vx, vy, vz receive the pre-collision velocity.
template <int Q, int COL>
__device__ __forceinline__ void collide_moment(double (&g)[Q], const MomentConsts &m, double &vx, double &vy,
                                               double &vz)
{
    using V = VSet<Q>;
    using C = mom::Cls<Q>;
    constexpr int D = V::D;
    constexpr int NPL = D == 2 ? 1 : 3;
    constexpr double w0 = V::w[0];
    constexpr double wa = V::w[1];
    constexpr double wd = V::w[Q - 1];
    // pair sums / differences of opposite populations: the difference is
    // f(c) - f(-c) for the member whose first nonzero component is positive
    double as[3] = {0.0, 0.0, 0.0}, ad[3] = {0.0, 0.0, 0.0};   // axis pairs
    double ps[3][2], pd[3][2];   // diagonal pairs [plane][0: same signs, 1: opposite signs]
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * P + 1;
        const double sum = g[i] + g[i + 1];
        if constexpr (C::nz(i) == 1) {
            as[C::axis(i)] = sum;
            ad[C::axis(i)] = C::axis_sign(i) > 0 ? g[i] - g[i + 1] : g[i + 1] - g[i];
        } else {
            constexpr int p = C::plane(i), s = C::first(i) * C::second(i) > 0 ? 0 : 1;
            ps[p][s] = sum;
            pd[p][s] = C::first(i) > 0 ? g[i] - g[i + 1] : g[i + 1] - g[i];
        }
    });
    // per plane: sum of its four diagonals, Pi_ab, and the momentum of its
    // diagonals along the plane's first (ja) and second (jb) axis
    double dsum[3], pab[3], ja[3], jb[3];
    static_for<NPL>([&](auto P) {
        constexpr int p = P;
        dsum[p] = ps[p][0] + ps[p][1];
        pab[p] = ps[p][0] - ps[p][1];
        ja[p] = pd[p][0] + pd[p][1];
        jb[p] = pd[p][0] - pd[p][1];
    });
    double rho, j[3], pii[3];
    if constexpr (D == 2) {
        pii[0] = as[0] + dsum[0];
        pii[1] = as[1] + dsum[0];
        rho = (g[0] + pii[0]) + as[1];
        j[0] = ad[0] + ja[0];
        j[1] = ad[1] + jb[0];
        j[2] = 0.0;
        pii[2] = 0.0;
    } else {
        pii[0] = (as[0] + dsum[0]) + dsum[1];
        const double ty = as[1] + dsum[2];
        pii[1] = ty + dsum[0];
        pii[2] = (as[2] + dsum[1]) + dsum[2];
        rho = ((g[0] + pii[0]) + ty) + as[2];
        j[0] = (ad[0] + ja[0]) + ja[1];
        j[1] = (ad[1] + jb[0]) + ja[2];
        j[2] = (ad[2] + jb[1]) + jb[2];
    }
    double r;
    if (recip_range(rho)) r = recip_in_range(rho);
    else r = rho != 0.0 ? 1.0 / rho : 0.0;   // calcula's u = 0 at rho == 0
    double u[3];
    static_for<D>([&](auto A) { u[(int)A] = j[(int)A] * r; });
    vx = u[0];
    vy = u[1];
    vz = D == 3 ? u[2] : 0.0;
    double paa[3], sp, sP;
    static_for<D>([&](auto A) { paa[(int)A] = j[(int)A] * u[(int)A]; });
    if constexpr (D == 2) {
        sp = paa[0] + paa[1];
        sP = pii[0] + pii[1];
    } else {
        sp = (paa[0] + paa[1]) + paa[2];
        sP = (pii[0] + pii[1]) + pii[2];
    }
    const double alpha = __builtin_fma(m.cr, rho, -__builtin_fma(m.t15, sp, m.k15 * sP));
    const double kr = m.k15 * rho;
    double gaa[3], ag[3], e[3], b[3], bd[3];
    static_for<D>([&](auto A) {
        constexpr int a = A;
        gaa[a] = __builtin_fma(m.t45, paa[a], __builtin_fma(m.k45, pii[a], -kr));
        ag[a] = alpha + gaa[a];
        e[a] = wa * ag[a];
        b[a] = m.b_axis * j[a];
        bd[a] = m.b_diag * j[a];
    });
    // plane classes: PP / PM = w_d (alpha + G_aa + G_bb) +- 2 w_d G_ab,
    // S / Dd = w_d (beta_a +- beta_b)
    double pp[3], pm[3], sb[3], db[3];
    static_for<NPL>([&](auto P) {
        constexpr int p = P;
        constexpr int a = p == 2 ? 1 : 0, bb = p == 0 ? 1 : 2;
        const double pab_u = j[a] * u[bb];
        const double gab = __builtin_fma(m.t45, pab_u, m.k45 * pab[p]);
        const double pc = wd * (ag[a] + gaa[bb]);
        const double x = (2.0 * wd) * gab;
        pp[p] = pc + x;
        pm[p] = pc - x;
        sb[p] = bd[a] + bd[bb];
        db[p] = bd[a] - bd[bb];
    });
    const double a0 = w0 * alpha;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        double A;
        if constexpr (C::nz(i) == 0) {
            A = a0;
        } else if constexpr (C::nz(i) == 1) {
            constexpr int a = C::axis(i);
            A = C::axis_sign(i) > 0 ? e[a] + b[a] : e[a] - b[a];
        } else {
            constexpr int p = C::plane(i);
            if constexpr (C::first(i) * C::second(i) > 0) A = C::first(i) > 0 ? pp[p] + sb[p] : pp[p] - sb[p];
            else A = C::first(i) > 0 ? pm[p] + db[p] : pm[p] - db[p];
        }
        if constexpr (COL == 3) g[i] = A;   // regularized: W = 0
        else g[i] = __builtin_fma(m.W, g[i], A);
    });
}

// The kernels' collision: the reference's operation order (collide_exact),
// or the moment-space form where COL carries COL_MOMENT.
template <int Q, int COL>
__device__ __forceinline__ void collide_sel(double (&g)[Q], const CollideConsts &k, double &vx, double &vy, double &vz)
{
    if constexpr ((COL & COL_MOMENT) != 0) collide_moment<Q, (COL & ~COL_MOMENT)>(g, k.m, vx, vy, vz);
    else collide_exact<Q, COL>(g, k, vx, vy, vz);
}

}  // namespace erlbm
