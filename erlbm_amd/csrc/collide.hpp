// collide.hpp -- per-node FP64 collision, bitwise faithful to SimBoltz.
//
// Restates, in registers, the operation order of
//   calcula        SimBoltz_Functions/LBM_functions.cpp:3517-3589
//   dist_eq        SimBoltz_Functions/LBM_functions.cpp:2944-3000
//   coll_Ext_Reg   SimBoltz_Functions/Collision_operators.cpp:1181-1307
//   coll_Reg       SimBoltz_Functions/Collision_operators.cpp:859-967
//   coll_BGK       SimBoltz_Functions/Collision_operators.cpp:11-49
// so that, compiled with -ffp-contract=off and IEEE division, every finite
// population equals the reference's bit for bit.  Exact simplifications only:
//  * products by c in {-1, 0, +1} are not formed: x*(+-1) is exact, and a
//    +-0 term added to a running sum that starts at +0.0 never changes it;
//  * opposite directions share cv*3 and ((0.5 cv) cv 3) 3, which are odd and
//    even in cv under IEEE round-to-nearest;
//  * 1/rho is the division's own sequence without its no-op scaling steps
//    (recip_in_range), and the unforced momentum sums skip their leading 0.0 +
//    (only the sign of a zero velocity differs, which no population sees);
//  * the forcing term S (source(), LBM_functions.cpp:5645-5716) is +-0 at the
//    zero body force every driver passes, so the unforced operator (FORCE =
//    false) does not add it; with FORCE the equilibrium takes u + a/2, F = a rho
//    and S is added last, as coll_BGK / coll_Reg / coll_Ext_Reg do
//    (Collision_operators.cpp:24-49, 866-921, 1196-1254), and coll_MRT's own
//    forcing moments (:647-690) are formed.
// Non-finite states stay non-finite (possibly inf where the reference has
// NaN), which every stability test treats alike.
#pragma once

#include "vset.hpp"

namespace erlbm {

enum Collision { COL_BGK = 1, COL_MRT = 2, COL_REG = 3, COL_EXT_REG = 4 };
// Flag or-ed into a kernel's COL template argument: the operator in moment
// space (collide_moment.hpp) instead of the reference's operation order.
constexpr int COL_MOMENT = 16;

// Per-lattice constants of the moment-space collision (collide_moment.hpp),
// computed on the host from tau, tau_nh and the operator; b_axis / b_diag
// carry the velocity set's axis and diagonal weights.
struct MomentConsts {
    double W = 0.0;        // multiplier of f_i
    double cr = 0.0;       // K0 + D K1 / 2
    double t15 = 0.0;      // 1.5 T
    double k15 = 0.0;      // 1.5 K1
    double t45 = 0.0;      // 4.5 T
    double k45 = 0.0;      // 4.5 K1
    double b_axis = 0.0;   // 3 K0 w_axis
    double b_diag = 0.0;   // 3 K0 w_diag
};

struct CollideConsts {
    double omega1;   // 1 - 1/tau   (coll_Ext_Reg :1192; coll_Reg's (1 - 1/tau))
    double omega2;   // 1 - 1/tau_nh (coll_Ext_Reg :1194)
    double oot;      // 1/tau        (coll_BGK :24)
    MomentConsts m;  // the moment-space form of the same operator
};

// Body-force acceleration of colisao_sitio's acc_x / acc_y / acc_z arguments
// (src/erlbm_common.h:68-81) and source()'s factor 1 - 1/(2 tau)
// (LBM_functions.cpp:5664), computed on the host in that order.
struct BodyForce {
    double ax, ay, az;
    double factor;
};

// 1 / r, correctly rounded, for |r| in [2^-600, 2^600] (recip_range): the
// steps of the IEEE FP64 division sequence gfx950 code runs for 1.0 / r
// (reciprocal estimate, two Newton steps, remainder, final fused step)
// without its scaling and special-case steps (v_div_scale, v_div_fixup),
// which leave the quotient unchanged there: no operand or result comes near
// the overflow or subnormal range.
__device__ __forceinline__ bool recip_range(double r)
{
    const double ar = __builtin_fabs(r);
    return ar >= 0x1p-600 && ar <= 0x1p+600;
}

__device__ __forceinline__ double recip_in_range(double r)
{
    const double y0 = __builtin_amdgcn_rcp(r);
    const double y1 = __builtin_fma(y0, __builtin_fma(-r, y0, 1.0), y0);
    const double y2 = __builtin_fma(y1, __builtin_fma(-r, y1, 1.0), y1);
    return __builtin_fma(__builtin_fma(-r, y2, 1.0), y2, y2);
}

// rho, u as calcula() computes them (sums in velocity order from 0.0,
// u = m * (1/rho), u = 0 when rho == 0).  LEAD0 = false starts the momentum
// sums at their first term instead of 0.0 + that term: the two differ only
// in the sign of a zero sum, and a zero velocity component of either sign
// gives the same equilibrium (it enters as 1 +- 3 c.u and through u.u), so
// the unforced collision does not form the addition.  SHORT_RECIP: 1/rho by
// recip_in_range where it applies, else the division.
template <int Q, bool LEAD0 = true, bool SHORT_RECIP = true>
__device__ __forceinline__ void moments_exact(const double (&g)[Q], double &rho, double &vx,
                                              double &vy, double &vz)
{
    using V = VSet<Q>;
    double r = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    bool sx = LEAD0, sy = LEAD0, sz = LEAD0;   // sum started
    auto acc = [](double &m, bool &st, double v, int sign) {
        if (st) m = sign > 0 ? m + v : m - v;
        else m = sign > 0 ? v : -v;
        st = true;
    };
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        r = r + g[i];
        if constexpr (V::cx[i] != 0) acc(mx, sx, g[i], V::cx[i]);
        if constexpr (V::cy[i] != 0) acc(my, sy, g[i], V::cy[i]);
        if constexpr (V::cz[i] != 0) acc(mz, sz, g[i], V::cz[i]);
    });
    rho = r;
    auto scale = [&](double oor) {
        vx = mx * oor;
        vy = my * oor;
        vz = (V::D == 3) ? mz * oor : 0.0;
    };
    if (SHORT_RECIP && recip_range(r)) {
        scale(recip_in_range(r));
    } else if (r != 0.0) {
        scale(1.0 / r);
    } else {
        vx = 0.0;
        vy = 0.0;
        vz = 0.0;
    }
}

// c_i . u in dot_product's order (x, then y, then z), zero terms dropped.
template <int Q, int I>
__device__ __forceinline__ double cdotu(double vx, double vy, double vz)
{
    using V = VSet<Q>;
    double cv = 0.0;
    bool started = false;
    if constexpr (V::cx[I] != 0) {
        cv = V::cx[I] > 0 ? vx : -vx;
        started = true;
    }
    if constexpr (V::cy[I] != 0) {
        if (started) cv = V::cy[I] > 0 ? cv + vy : cv - vy;
        else cv = V::cy[I] > 0 ? vy : -vy;
        started = true;
    }
    if constexpr (V::cz[I] != 0) {
        if (started) cv = V::cz[I] > 0 ? cv + vz : cv - vz;
        else cv = V::cz[I] > 0 ? vz : -vz;
    }
    (void)started;
    return cv;
}

// f_eq = (w rho)(((1 + 3 cv) + ((0.5 cv) cv 3) 3) - (0.5 u.u) 3), dist_eq's order.
template <int Q>
__device__ __forceinline__ void feq_exact(double rho, double vx, double vy, double vz,
                                          double (&feq)[Q])
{
    using V = VSet<Q>;
    const double vquad = (V::D == 2) ? (vx * vx + vy * vy) : (vx * vx + vy * vy + vz * vz);
    const double s = (0.5 * vquad) * 3.0;
    feq[0] = (V::w[0] * rho) * (1.0 - s);
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * P + 1;  // i and i+1 are opposite
        const double cv = cdotu<Q, i>(vx, vy, vz);
        const double a = cv * 3.0;
        const double q = (((0.5 * cv) * cv) * 3.0) * 3.0;
        feq[i] = (V::w[i] * rho) * (((1.0 + a) + q) - s);
        feq[i + 1] = (V::w[i + 1] * rho) * (((1.0 - a) + q) - s);
    });
}

// source() (LBM_functions.cpp:5645-5716) at the velocity u the caller gives:
// F = a rho, S_i = ((PRE w_i) 3) ((cF + (3 cF) cv) - vF) with vF = F.u summed
// from 0.0 (dot_product, :2864-2876), cF = c_i.F from 0.0 without the zero-c
// terms (adding +-0 to a sum that starts at +0.0 leaves it unchanged) and
// cv = c_i.u as the equilibrium forms it.  PRE = 1 - 1/(2 tau) for the
// BGK / regularized operators; coll_MRT's own term (Collision_operators.cpp:
// 647-671) has no factor (MRT = true).
template <int Q, bool MRT>
__device__ __forceinline__ void source_exact(double rho, double ux, double uy, double uz, const BodyForce &bf,
                                             double (&S)[Q])
{
    using V = VSet<Q>;
    const double Fx = bf.ax * rho, Fy = bf.ay * rho, Fz = bf.az * rho;
    double vF = 0.0 + Fx * ux;
    vF = vF + Fy * uy;
    if constexpr (V::D == 3) vF = vF + Fz * uz;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        double cF = 0.0;
        if constexpr (V::cx[i] != 0) cF = V::cx[i] > 0 ? cF + Fx : cF - Fx;
        if constexpr (V::cy[i] != 0) cF = V::cy[i] > 0 ? cF + Fy : cF - Fy;
        if constexpr (V::D == 3 && V::cz[i] != 0) cF = V::cz[i] > 0 ? cF + Fz : cF - Fz;
        const double cv = cdotu<Q, i>(ux, uy, uz);
        const double pre = MRT ? V::w[i] * 3.0 : (bf.factor * V::w[i]) * 3.0;
        S[i] = pre * ((cF + (3.0 * cF) * cv) - vF);
    });
}

// coll_MRT, D2Q9 (Collision_operators.cpp:611-701): moments m = M f and
// m_eq = M f_eq, m <- m - s (m - m_eq) + (1 - s/2) F_m with F_m = +0 at zero
// force, f = M^-1 m; M and M^-1 of :449-471, rates of :614-622.  Entries
// of M are small integers, so M f is formed with exact +-f, 2f, 4f terms;
// zero entries of M and M^-1 are skipped (a +-0 term added to a running sum
// that starts at +0.0 never changes it).
struct MrtD2Q9 {
    static constexpr int M[9][9] = {
        {1, 1, 1, 1, 1, 1, 1, 1, 1},     {-4, -1, -1, -1, -1, 2, 2, 2, 2}, {4, -2, -2, -2, -2, 1, 1, 1, 1},
        {0, 1, -1, 0, 0, 1, -1, 1, -1},  {0, -2, 2, 0, 0, 1, -1, 1, -1},   {0, 0, 0, 1, -1, 1, -1, -1, 1},
        {0, 0, 0, -2, 2, 1, -1, -1, 1},  {0, 1, 1, -1, -1, 0, 0, 0, 0},    {0, 0, 0, 0, 0, 1, 1, -1, -1}};
    static constexpr double MI[9][9] = {
        {1. / 9., -1. / 9., 1. / 9., 0., 0., 0., 0., 0., 0.},
        {1. / 9., -1. / 36., -1. / 18., 1. / 6., -1. / 6., 0., 0., 1. / 4., 0.},
        {1. / 9., -1. / 36., -1. / 18., -1. / 6., 1. / 6., 0., 0., 1. / 4., 0.},
        {1. / 9., -1. / 36., -1. / 18., 0., 0., 1. / 6., -1. / 6., -1. / 4., 0.},
        {1. / 9., -1. / 36., -1. / 18., 0., 0., -1. / 6., 1. / 6., -1. / 4., 0.},
        {1. / 9., 1. / 18., 1. / 36., 1. / 6., 1. / 12., 1. / 6., 1. / 12., 0., 1. / 4.},
        {1. / 9., 1. / 18., 1. / 36., -1. / 6., -1. / 12., -1. / 6., -1. / 12., 0., 1. / 4.},
        {1. / 9., 1. / 18., 1. / 36., 1. / 6., 1. / 12., -1. / 6., -1. / 12., 0., -1. / 4.},
        {1. / 9., 1. / 18., 1. / 36., -1. / 6., -1. / 12., 1. / 6., 1. / 12., 0., -1. / 4.}};
    static constexpr double S_FIXED[9] = {0.0, 1.6, 1.8, 0.0, 1.9, 0.0, 1.9, -1.0, -1.0};   // -1: 1/tau
};

// FORCE: the forcing moments F_m = M S of coll_MRT's S (:680-690) instead of +0.
template <bool FORCE = false>
__device__ __forceinline__ void collide_mrt_d2q9(double (&g)[9], const double (&feq)[9], double oot,
                                                 const double (&S)[9])
{
    using T = MrtD2Q9;
    double m[9];
    static_for<9>([&](auto R) {
        constexpr int r = R;
        double a = 0.0, e = 0.0, fm = 0.0;
        static_for<9>([&](auto Cc) {
            constexpr int c = Cc;
            constexpr int k = T::M[r][c];
            if constexpr (k == 1) {
                a = a + g[c];
                e = e + feq[c];
                if constexpr (FORCE) fm = fm + S[c];
            } else if constexpr (k == -1) {
                a = a - g[c];
                e = e - feq[c];
                if constexpr (FORCE) fm = fm - S[c];
            } else if constexpr (k != 0) {
                a = a + (double)k * g[c];
                e = e + (double)k * feq[c];
                if constexpr (FORCE) fm = fm + (double)k * S[c];
            }
        });
        const double s = T::S_FIXED[r] < 0.0 ? oot : T::S_FIXED[r];
        m[r] = (a - s * (a - e)) + (1.0 - 0.5 * s) * fm;
    });
    static_for<9>([&](auto R) {
        constexpr int r = R;
        double a = 0.0;
        static_for<9>([&](auto Cc) {
            constexpr int c = Cc;
            if constexpr (T::MI[r][c] != 0.0) a = a + T::MI[r][c] * m[c];
        });
        g[r] = a;
    });
}

// One collision in place.  Returns nothing; g holds f* afterwards.
// vx, vy, vz receive the pre-collision velocity (a zero component of the
// unforced operator may carry either sign).  FORCE adds the body force bf
// (the operators' acc_x, acc_y, acc_z arguments).
// SHORT = false keeps the moments exactly as calcula() forms them (the
// division for 1/rho, sums from 0.0): the batched tiny-lattice kernel, where
// the short forms measured slower (obstacle map 8.47k against 8.68k points/s;
// the streaming walks: 94.3-94.7k against 93.9-94.1k MLUPS).
template <int Q, int COL, bool FORCE = false, bool SHORT = true>
__device__ __forceinline__ void collide_exact(double (&g)[Q], const CollideConsts &k, double &vx,
                                              double &vy, double &vz, const BodyForce &bf = BodyForce{})
{
    using V = VSet<Q>;
    double rho;
    moments_exact<Q, FORCE || !SHORT, SHORT>(g, rho, vx, vy, vz);
    // velocity of the equilibrium: u + a/2 under a body force (coll_Ext_Reg
    // :1198-1200; coll_MRT keeps u, :637-640)
    double ux = vx, uy = vy, uz = vz;
    double S[Q];
    if constexpr (FORCE) {
        if constexpr (COL != COL_MRT) {
            ux = vx + 0.5 * bf.ax;
            uy = vy + 0.5 * bf.ay;
            uz = vz + 0.5 * bf.az;
        }
        source_exact<Q, COL == COL_MRT>(rho, ux, uy, uz, bf, S);
    }
    double feq[Q];
    feq_exact<Q>(rho, ux, uy, uz, feq);

    if constexpr (COL == COL_BGK) {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            g[i] = g[i] + (feq[i] - g[i]) * k.oot;
            if constexpr (FORCE) g[i] = g[i] + S[i];
        });
        return;
    } else if constexpr (COL == COL_MRT) {
        static_assert(Q == 9, "coll_MRT exists for D2Q9 only");
        collide_mrt_d2q9<FORCE>(g, feq, k.oot, S);
        return;
    } else {
        double d[Q];
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            d[i] = g[i] - feq[i];
        });
        // Pi^(1)_ab = sum_i (f_i - f_eq_i) c_ia c_ib, index order, from 0.0
        double Pxx = 0.0, Pyy = 0.0, Pzz = 0.0, Pxy = 0.0, Pxz = 0.0, Pyz = 0.0;
        bool sxx = false, syy = false, szz = false, sxy = false, sxz = false, syz = false;
        auto acc = [](double &s, bool &st, double v, int sign) {
            if (!st) {
                s = sign > 0 ? v : -v;
                st = true;
            } else {
                s = sign > 0 ? s + v : s - v;
            }
        };
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            if constexpr (V::cx[i] != 0) acc(Pxx, sxx, d[i], 1);
            if constexpr (V::cy[i] != 0) acc(Pyy, syy, d[i], 1);
            if constexpr (V::D == 3 && V::cz[i] != 0) acc(Pzz, szz, d[i], 1);
            if constexpr (V::cx[i] * V::cy[i] != 0) acc(Pxy, sxy, d[i], V::cx[i] * V::cy[i]);
            if constexpr (V::D == 3 && V::cx[i] * V::cz[i] != 0) acc(Pxz, sxz, d[i], V::cx[i] * V::cz[i]);
            if constexpr (V::D == 3 && V::cy[i] * V::cz[i] != 0) acc(Pyz, syz, d[i], V::cy[i] * V::cz[i]);
        });
        // Q_aa = c_a c_a - 1/3: (1.0 - 1/3) or (0 - 1/3); 2 Q_ab = +-2 for a != b.
        constexpr double cs2 = 1. / 3.;
        constexpr double Q1 = 1.0 - cs2;
        constexpr double Q0 = 0.0 - cs2;
        const double axx1 = Q1 * Pxx, axx0 = Q0 * Pxx;
        const double ayy1 = Q1 * Pyy, ayy0 = Q0 * Pyy;
        const double azz1 = Q1 * Pzz, azz0 = Q0 * Pzz;
        const double bxy = 2.0 * Pxy, bxz = 2.0 * Pxz, byz = 2.0 * Pyz;
        double f1[Q];
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            double t = (V::cx[i] != 0 ? axx1 : axx0) + (V::cy[i] != 0 ? ayy1 : ayy0);
            if constexpr (V::D == 3) t = t + (V::cz[i] != 0 ? azz1 : azz0);
            if constexpr (V::cx[i] * V::cy[i] > 0) t = t + bxy;
            if constexpr (V::cx[i] * V::cy[i] < 0) t = t - bxy;
            if constexpr (V::D == 3 && V::cx[i] * V::cz[i] > 0) t = t + bxz;
            if constexpr (V::D == 3 && V::cx[i] * V::cz[i] < 0) t = t - bxz;
            if constexpr (V::D == 3 && V::cy[i] * V::cz[i] > 0) t = t + byz;
            if constexpr (V::D == 3 && V::cy[i] * V::cz[i] < 0) t = t - byz;
            constexpr double k45 = (0.5 * 3. * 3.) * V::w[i];
            f1[i] = k45 * t;
        });
        if constexpr (COL == COL_REG) {
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                g[i] = feq[i] + k.omega1 * f1[i];
                if constexpr (FORCE) g[i] = g[i] + S[i];
            });
        } else {
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                const double fgt1 = d[i] - f1[i];
                g[i] = (feq[i] + k.omega1 * f1[i]) + k.omega2 * fgt1;
                if constexpr (FORCE) g[i] = g[i] + S[i];
            });
        }
    }
}

// density(): sum in velocity order from 0.0 (LBM_functions.cpp:5543-5552).
template <int Q>
__device__ __forceinline__ double density_exact(const double (&g)[Q])
{
    double r = 0.0;
    static_for<Q>([&](auto I) { r = r + g[(int)I]; });
    return r;
}

// kinectic_energy()'s per-site term 0.5 (vx^2 + vy^2 + vz^2)
// (Other_functions.cpp:100-115); in 2D calcula sets vz = 0.
template <int Q>
__device__ __forceinline__ double ke_term(double vx, double vy, double vz)
{
    if constexpr (VSet<Q>::D == 2) return 0.5 * (vx * vx + vy * vy);
    else return 0.5 * (vx * vx + vy * vy + vz * vz);
}

}  // namespace erlbm
