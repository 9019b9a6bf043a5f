// host_init.cpp -- host-side initial conditions, map axes and geometries of
// the ERLBM drivers, in C++ with the same libm calls and operation order as
// the reference, so that the populations the GPU starts from are bitwise the
// reference's.  Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "../../include/erlbm.h"
#include "vset.hpp"

using namespace erlbm;

namespace {

// dist_eq (SimBoltz_Functions/LBM_functions.cpp:2944-3000), literal order.
template <int Q>
void dist_eq_host(double *feq, double vx, double vy, double vz, double rho)
{
    using V = VSet<Q>;
    const double ocs2 = 3.;
    const double vquad = (V::D == 2) ? (vx * vx + vy * vy) : (vx * vx + vy * vy + vz * vz);
    for (int i = 0; i < Q; i++) {
        double cv = 0.0;
        cv = cv + V::cx[i] * vx;
        cv = cv + V::cy[i] * vy;
        if (V::D == 3) cv = cv + V::cz[i] * vz;
        feq[i] = V::w[i] * rho * (1.0 + cv * ocs2 + 0.5 * cv * cv * ocs2 * ocs2 - 0.5 * vquad * ocs2);
    }
}

}  // namespace

extern "C" int erlbm_dist_eq(int q, double vx, double vy, double vz, double rho, double *feq)
{
    if (!feq) return ERLBM_ERR_ARG;
    if (q == 9) dist_eq_host<9>(feq, vx, vy, vz, rho);
    else if (q == 19) dist_eq_host<19>(feq, vx, vy, vz, rho);
    else return ERLBM_ERR_ARG;
    return ERLBM_OK;
}

// inicializa_taylor_green, src/main_taylor-green.cpp:57-100 (q = 9 or 19,
// periodic L x L x 1, fluid index x + y L).
extern "C" int erlbm_init_taylor_green(int q, int Li, double u_0, double tau, int poisson, int neq, double *inif)
{
    if (!inif || Li < 1 || (q != 9 && q != 19)) return ERLBM_ERR_ARG;
    const double L = (double)Li;
    const double k = 2. * M_PI / L;
    const double cs2 = 1. / 3.;
    const int *cx = q == 9 ? VSet<9>::cx : VSet<19>::cx;
    const int *cy = q == 9 ? VSet<9>::cy : VSet<19>::cy;
    const double *w = q == 9 ? VSet<9>::w : VSet<19>::w;
    for (int y = 0; y < Li; y++) {
        for (int x = 0; x < Li; x++) {
            double rho = 1.0;
            if (poisson) rho = 1.0 + 0.75 * u_0 * u_0 * (cos(4. * M_PI * x / L) + cos(4. * M_PI * y / L));
            const double u_x = u_0 * sin(2. * M_PI * x / L) * cos(2. * M_PI * y / L);
            const double u_y = -u_0 * cos(2. * M_PI * x / L) * sin(2. * M_PI * y / L);
            double *f = inif + (size_t)(x + y * Li) * q;
            erlbm_dist_eq(q, u_x, u_y, 0.0, rho, f);
            if (!neq) continue;
            const double Sxx = 2. * u_0 * k * cos(k * x) * cos(k * y);
            const double Pxx = -rho * cs2 * tau * Sxx;
            const double Pyy = -Pxx;
            for (int i = 0; i < q; i++) {
                const double ccx = cx[i];
                const double ccy = cy[i];
                f[i] += w[i] / (2. * cs2 * cs2) * ((ccx * ccx - cs2) * Pxx + (ccy * ccy - cs2) * Pyy);
            }
        }
    }
    return ERLBM_OK;
}

// initial_double_shear_layer, src/main_double-shear-layer.cpp:48-66.
extern "C" int erlbm_init_double_shear(int Li, double rho_ini, double *inif)
{
    if (!inif || Li < 1) return ERLBM_ERR_ARG;
    const double U_0 = 1. / 32., LAMBDA = 80., EPSILON = 0.05;
    const double L = (double)Li;
    for (int y = 0; y < Li; y++) {
        for (int x = 0; x < Li; x++) {
            double u_x = U_0 * tanh(LAMBDA * ((double)y / L - 1. / 4.));
            if (y > Li / 2) u_x = U_0 * tanh(LAMBDA * (3. / 4. - (double)y / L));
            const double u_y = EPSILON * U_0 * sin(2. * M_PI * ((double)x / L + 1. / 4.));
            dist_eq_host<9>(inif + (size_t)(x + y * Li) * 9, u_x, u_y, 0.0, rho_ini);
        }
    }
    return ERLBM_OK;
}

// Taylor-Green map-point populations, src/main_stability-map_taylor-green.cpp:49-61.
extern "C" int erlbm_init_tg_map(int Li, double u_0, double rho_ini, double *inif)
{
    if (!inif || Li < 1) return ERLBM_ERR_ARG;
    const double L = (double)Li;
    for (int y = 0; y < Li; y++) {
        for (int x = 0; x < Li; x++) {
            const double u_x = u_0 * sin(2. * M_PI * (double)x / L) * cos(2. * M_PI * (double)y / L);
            const double u_y = -u_0 * cos(2. * M_PI * (double)x / L) * sin(2. * M_PI * (double)y / L);
            dist_eq_host<19>(inif + (size_t)(x + y * Li) * 19, u_x, u_y, 0.0, rho_ini);
        }
    }
    return ERLBM_OK;
}

// The trigonometric tables of that field: sx[x] = sin(2 pi x / L),
// cx[x] = cos(2 pi x / L) for x < L, so that u_x = (u_0 sx[x]) cx[y] and
// u_y = ((-u_0) cx[x]) sx[y] reproduce the driver's expressions exactly.
extern "C" int erlbm_tg_map_tables(int Li, double *sx, double *cx)
{
    if (!sx || !cx || Li < 1) return ERLBM_ERR_ARG;
    const double L = (double)Li;
    for (int x = 0; x < Li; x++) {
        sx[x] = sin(2. * M_PI * (double)x / L);
        cx[x] = cos(2. * M_PI * (double)x / L);
    }
    return ERLBM_OK;
}

// geometria_obstaculo, src/main_stability-map_obstacle.cpp:42-70:
// 22 x 12 x 1, solid columns x = 0, 21 and the block x in {7,8}, y in 4..7.
extern "C" int erlbm_obstacle_mask(int *mask)
{
    if (!mask) return ERLBM_ERR_ARG;
    const int nx = 22, ny = 12;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            int s = 1;
            if (x == 0 || x == nx - 1) s = 0;
            if ((x == 7 || x == 8) && y >= 4 && y <= 7) s = 0;
            mask[x + y * nx] = s;
        }
    return ERLBM_OK;
}

// Map axes: Taylor-Green map (src/main_stability-map_taylor-green.cpp:134-164)
// and obstacle map (src/main_stability-map_obstacle.cpp:147-172).
extern "C" int erlbm_tg_map_point_params(int i, int j, int n_ma, int n_nu, int Li, int collision,
                                         double tau_nh_in, double *u_0, double *tau, double *tau_nh, int *n_steps)
{
    if (!u_0 || !tau || !tau_nh || n_ma < 1 || n_nu < 1 || Li < 1) return ERLBM_ERR_ARG;
    const double c_s = 1. / sqrt(3.);
    const double u_min = 0.05 * c_s;
    const double log_nu_min = -5.;
    const double log_nu_max = -2.;
    *u_0 = u_min + (i + 1) * (c_s - u_min) / n_ma;
    const double log_nu = log_nu_min + j * (log_nu_max - log_nu_min) / n_nu;
    *tau = 3. * pow(10., log_nu) + 0.5;
    *tau_nh = (collision == ERLBM_COL_REG) ? 1.0 : tau_nh_in;
    if (n_steps) *n_steps = (int)((double)(100 * Li) / *u_0);
    return ERLBM_OK;
}

extern "C" int erlbm_obstacle_map_point_params(int i, int j, int n_tau, int n_nu, double Ma, double *u_imp,
                                               double *tau, double *tau_nh, int *n_steps)
{
    if (!u_imp || !tau || !tau_nh || n_tau < 1 || n_nu < 1) return ERLBM_ERR_ARG;
    *u_imp = Ma / sqrt(3.);
    const double tau_nh_min = 0.5, tau_nh_max = 1.1;
    const double log_nu_min = -5., log_nu_max = -0.7;
    const double log_nu = log_nu_min + j * (log_nu_max - log_nu_min) / n_nu;
    *tau = 3. * pow(10., log_nu) + 0.5;
    *tau_nh = tau_nh_min + (double)i * (tau_nh_max - tau_nh_min) / n_tau;
    if (n_steps) *n_steps = (int)((double)(200 * 12) / *u_imp);
    return ERLBM_OK;
}

// Analytic-decay error of a Taylor-Green run, src/main_taylor-green.cpp:178-213:
// moments of the post-collision populations at step `step` against
// u_ex = u_0 e^{-2 nu k^2 t} (sin kx cos ky, -cos kx sin ky).
// out[0] = E_vector, out[1] = E_magnitude, out[2] = (nu_eff - nu) / nu.
extern "C" int erlbm_taylor_green_error(int q, int Li, double u_0, double visc, long long step, const double *inif,
                                        double *out)
{
    if (!inif || !out || Li < 1 || (q != 9 && q != 19)) return ERLBM_ERR_ARG;
    const double L = Li;
    const double k = 2. * M_PI / L;
    double s_mag = 0, s_vec = 0, s_r = 0, s_proj = 0, s_phi = 0;
    const double e = exp(-2. * visc * k * k * step);
    const int *cx = q == 9 ? VSet<9>::cx : VSet<19>::cx;
    const int *cy = q == 9 ? VSet<9>::cy : VSet<19>::cy;
    const int *cz = q == 9 ? VSet<9>::cz : VSet<19>::cz;
    for (int y = 0; y < Li; y++) {
        for (int x = 0; x < Li; x++) {
            // calcula (LBM_functions.cpp:3517-3589)
            const double *f = inif + (size_t)(x + y * Li) * q;
            double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
            for (int i = 0; i < q; i++) {
                rho = rho + f[i];
                mx = mx + cx[i] * f[i];
                my = my + cy[i] * f[i];
                mz = mz + cz[i] * f[i];
            }
            double vx = 0., vy = 0., vz = 0.;
            if (rho) {
                const double oor = 1.0 / (rho);
                vx = mx * oor;
                vy = my * oor;
                vz = q == 19 ? mz * oor : 0.0;
            }
            const double rx = u_0 * e * sin(k * x) * cos(k * y);
            const double ry = -u_0 * e * cos(k * x) * sin(k * y);
            const double us = sqrt(vx * vx + vy * vy + vz * vz), ur = sqrt(rx * rx + ry * ry);
            s_mag += (us - ur) * (us - ur);
            s_vec += (vx - rx) * (vx - rx) + (vy - ry) * (vy - ry) + vz * vz;
            s_r += ur * ur;
            const double px = sin(k * x) * cos(k * y), py = -cos(k * x) * sin(k * y);
            s_proj += vx * px + vy * py;
            s_phi += px * px + py * py;
        }
    }
    const double A = s_proj / s_phi / u_0;
    const double nu_ef = -log(A) / (2. * k * k * step);
    out[0] = sqrt(s_vec / s_r);
    out[1] = sqrt(s_mag / s_r);
    out[2] = nu_ef / visc - 1.;
    return ERLBM_OK;
}
