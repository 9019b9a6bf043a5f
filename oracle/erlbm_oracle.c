/*
 * erlbm_oracle.c -- CPU restatement of the SimBoltz E-RLBM collide-stream
 * path (loesantos/ERLBM).  TEST INFRASTRUCTURE: the checker for the HIP
 * path, and the CPU baseline timed beside it.  Never linked into the product.
 *
 * Built strict (-ffp-contract=off) so every result is bitwise that of the
 * reference compiled without FMA contraction; each function cites the
 * reference path:line whose operation order it follows.
 *
 * Pinned by: tests/test_oracle_pins.py (against the reference library built
 * from source into oracle/_ref, against tests/golden fixtures written from
 * it, against the published data/ tables and maps, and against closed forms:
 * mass/momentum conservation, rest-state invariance, tau_nh=1 == RLBM,
 * tau_nh=tau == BGK, streaming as a permutation).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORC_COL_BGK = 1, ORC_COL_MRT = 2, ORC_COL_REG = 3, ORC_COL_EXT_REG = 4 };

#define NV 9
#define DIM 2
#define SFX(name) name##_d2q9
#include "oracle_impl.h"
#undef NV
#undef DIM
#undef SFX

#define NV 19
#define DIM 3
#define SFX(name) name##_d3q19
#include "oracle_impl.h"
#undef NV
#undef DIM
#undef SFX

/* ---- ZouHeD3Q19_vx_x, Boundary_conditions.cpp:350-442 --------------------- */
void orc_zouhe_d3q19_vx_x(int pos_x, double vx, int nx, int ny, int nz,
                          const int *ini, double *inif)
{
    int x = pos_x;
    int infx = 0;
    if (pos_x < nx / 2) infx = 1;
    for (int y = 0; y < ny; y++)
        for (int z = 0; z < nz; z++) {
            int me = ini[x + y * nx + z * ny * nx];
            if (me && infx) {
                double *f = inif + (size_t)(me - 1) * 19;
                double sum_f_xminus = f[1] + f[7] + f[9] + f[11] + f[13];
                double sum_f_xzero = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] +
                                     f[16] + f[17] + f[18];
                double rho = (2.0 * sum_f_xminus + sum_f_xzero) / (1.0 - vx);
                double vy = 0.0, vz = 0.0;
                f[1] = f[1] + rho * vx / 3.0;
                f[7] = f[7] + rho * vx / 6.0 + 0.5 * rho * vy -
                       0.5 * (f[18] + f[3] + f[16] - f[15] - f[4] - f[17]);
                f[9] = f[9] + rho * vx / 6.0 - 0.5 * rho * vy +
                       0.5 * (f[18] + f[3] + f[16] - f[15] - f[4] - f[17]);
                f[11] = f[11] + rho * vx / 6.0 + 0.5 * rho * vz -
                        0.5 * (f[17] + f[5] + f[16] - f[15] - f[6] - f[18]);
                f[13] = f[13] + rho * vx / 6.0 - 0.5 * rho * vz +
                        0.5 * (f[17] + f[5] + f[16] - f[15] - f[6] - f[18]);
            } else if (me) {
                double *f = inif + (size_t)(me - 1) * 19;
                double sum_f_xplus = f[2] + f[8] + f[10] + f[12] + f[14];
                double sum_f_xzero = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] +
                                     f[16] + f[17] + f[18];
                double rho = (2.0 * sum_f_xplus + sum_f_xzero) / (1.0 - vx);
                double vy = 0.0, vz = 0.0;
                f[2] = f[2] - rho * vx / 3.0;
                f[10] = f[10] - rho * vx / 6.0 + 0.5 * rho * vy -
                        0.5 * (f[18] + f[3] + f[16] - f[15] - f[4] - f[17]);
                f[8] = f[8] - rho * vx / 6.0 - 0.5 * rho * vy +
                       0.5 * (f[18] + f[3] + f[16] - f[15] - f[4] - f[17]);
                f[14] = f[14] - rho * vx / 6.0 + 0.5 * rho * vz -
                        0.5 * (f[17] + f[5] + f[16] - f[15] - f[6] - f[18]);
                f[12] = f[12] - rho * vx / 6.0 - 0.5 * rho * vz +
                        0.5 * (f[17] + f[5] + f[16] - f[15] - f[6] - f[18]);
            }
        }
}

/* ---- ZouHeD2Q9_vx_x, Boundary_conditions.cpp:710-735 ---------------------- */
void orc_zouhe_d2q9_vx_x(int pos_x, double vx, int nx, int ny,
                         const int *ini, double *inif)
{
    int x = pos_x;
    for (int y = 0; y < ny; y++) {
        int me = ini[x + y * nx];
        if (!me) continue;
        double *f = inif + (size_t)(me - 1) * 9;
        double rho = (f[0] + f[3] + f[4] + 2 * (f[2] + f[8] + f[6])) / (1. - vx);
        f[1] = (rho / 9.0) * (1.0 + 3.0 * vx + 3.0 * vx * vx);
        f[5] = (rho / 36.0) * (1.0 + 3.0 * vx + 3.0 * vx * vx);
        f[7] = (rho / 36.0) * (1.0 + 3.0 * vx + 3.0 * vx * vx);
    }
}

/* ---- driver initial conditions --------------------------------------------- */

/* inicializa_taylor_green, src/main_taylor-green.cpp:57-100 (D2Q9 build,
 * periodic L x L, fluid index = x + y L).  Poisson density and the
 * f_eq + f^(1) initialisation are both switchable as in the driver.         */
void orc_init_taylor_green_d2q9(int Li, double u_0, double tau, int poisson,
                                int neq, double *inif)
{
    setup_d2q9();
    const double L = (double)Li;
    const double k = 2. * M_PI / L;
    const double cs2 = 1. / 3.;
    for (int y = 0; y < Li; y++)
        for (int x = 0; x < Li; x++) {
            double rho = 1.0;
            if (poisson)
                rho = 1.0 + 0.75 * u_0 * u_0 *
                                (cos(4. * M_PI * x / L) + cos(4. * M_PI * y / L));
            const double u_x = u_0 * sin(2. * M_PI * x / L) * cos(2. * M_PI * y / L);
            const double u_y = -u_0 * cos(2. * M_PI * x / L) * sin(2. * M_PI * y / L);
            double *f = inif + (size_t)(x + y * Li) * 9;
            dist_eq_d2q9(f, u_x, u_y, 0.0, rho);
            if (!neq) continue;
            const double Sxx = 2. * u_0 * k * cos(k * x) * cos(k * y);
            const double Pxx = -rho * cs2 * tau * Sxx;
            const double Pyy = -Pxx;
            for (int i = 0; i < 9; i++) {
                const double cx = C_d2q9[i * 2];
                const double cy = C_d2q9[i * 2 + 1];
                f[i] += W_d2q9[i] / (2. * cs2 * cs2) *
                        ((cx * cx - cs2) * Pxx + (cy * cy - cs2) * Pyy);
            }
        }
}

/* initial_double_shear_layer, src/main_double-shear-layer.cpp:48-66
 * (U_0 = 1/32, lambda = 80, epsilon = 0.05, :38-40).                        */
void orc_init_double_shear_d2q9(int Li, double rho_ini, double *inif)
{
    setup_d2q9();
    const double U_0 = 1. / 32., LAMBDA = 80., EPSILON = 0.05;
    const double L = (double)Li;
    for (int y = 0; y < Li; y++)
        for (int x = 0; x < Li; x++) {
            double u_x = U_0 * tanh(LAMBDA * ((double)y / L - 1. / 4.));
            if (y > Li / 2) u_x = U_0 * tanh(LAMBDA * (3. / 4. - (double)y / L));
            const double u_y = EPSILON * U_0 * sin(2. * M_PI * ((double)x / L + 1. / 4.));
            dist_eq_d2q9(inif + (size_t)(x + y * Li) * 9, u_x, u_y, 0.0, rho_ini);
        }
}

/* Taylor-Green map-point field, src/main_stability-map_taylor-green.cpp:49-61
 * (D3Q19, periodic L x L x 1, equilibrium at rho_ini).                      */
void orc_init_tg_map_d3q19(int Li, double u_0, double rho_ini, double *inif)
{
    setup_d3q19();
    const double L = (double)Li;
    for (int y = 0; y < Li; y++)
        for (int x = 0; x < Li; x++) {
            const double u_x = u_0 * sin(2. * M_PI * (double)x / L) *
                               cos(2. * M_PI * (double)y / L);
            const double u_y = -u_0 * cos(2. * M_PI * (double)x / L) *
                               sin(2. * M_PI * (double)y / L);
            dist_eq_d3q19(inif + (size_t)(x + y * Li) * 19, u_x, u_y, 0.0, rho_ini);
        }
}

/* Rest equilibrium at rho (dist_eq(f, u=0, rho)) for every fluid site:
 * src/main_stability-map_obstacle.cpp:81-82.                                 */
void orc_init_rest_d3q19(int n_fluid, double rho, double *inif)
{
    setup_d3q19();
    for (int p = 0; p < n_fluid; p++)
        dist_eq_d3q19(inif + (size_t)p * 19, 0., 0., 0., rho);
}

/* geometria_obstaculo, src/main_stability-map_obstacle.cpp:42-70: 22 x 12 x 1
 * voxels, solid columns x = 0 and x = 21, block x in {7,8}, y in 4..7.      */
void orc_obstacle_mask(int *mask)
{
    const int nx = 22, ny = 12;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            int s = 1;
            if (x == 0 || x == nx - 1) s = 0;
            if ((x == 7 || x == 8) && y >= 4 && y <= 7) s = 0;
            mask[x + y * nx] = s;
        }
}

/* ---- map axes ---------------------------------------------------------------
 * Taylor-Green map, src/main_stability-map_taylor-green.cpp:134-164.        */
void orc_tg_map_params(int i, int j, int n_ma, int n_nu, int col_is_reg,
                       double tau_nh_in, double *u_0, double *tau,
                       double *tau_nh)
{
    const double c_s = 1. / sqrt(3.);
    const double u_min = 0.05 * c_s;
    const double log_nu_min = -5.;
    const double log_nu_max = -2.;
    *u_0 = u_min + (i + 1) * (c_s - u_min) / n_ma;
    const double log_nu = log_nu_min + j * (log_nu_max - log_nu_min) / n_nu;
    *tau = 3. * pow(10., log_nu) + 0.5;
    *tau_nh = col_is_reg ? 1.0 : tau_nh_in;
}

/* Obstacle map, src/main_stability-map_obstacle.cpp:147-172.                */
void orc_obstacle_map_params(int i, int j, int n_tau, int n_nu, double Ma,
                             double *u_imp, double *tau, double *tau_nh)
{
    *u_imp = Ma / sqrt(3.);
    const double tau_nh_min = 0.5, tau_nh_max = 1.1;
    const double log_nu_min = -5., log_nu_max = -0.7;
    const double log_nu = log_nu_min + j * (log_nu_max - log_nu_min) / n_nu;
    *tau = 3. * pow(10., log_nu) + 0.5;
    *tau_nh = tau_nh_min + (double)i * (tau_nh_max - tau_nh_min) / n_tau;
}

/* ---- one Taylor-Green map point, roda_ponto,
 * src/main_stability-map_taylor-green.cpp:41-104.  Returns 1 if unstable;
 * *stop_step is the step at which the run ended (n_steps when stable).      */
int orc_run_tg_map_point(int Li, int col, double u_0, double tau, double tau_nh,
                         int *stop_step, double *e_last)
{
    setup_d3q19();
    const int nf = Li * Li;
    double *a = (double *)malloc(sizeof(double) * nf * 19);
    double *b = (double *)malloc(sizeof(double) * nf * 19);
    int *mask = (int *)malloc(sizeof(int) * nf);
    int *ini = (int *)malloc(sizeof(int) * nf);
    int *st = (int *)malloc(sizeof(int) * nf * 19);
    for (int v = 0; v < nf; v++) mask[v] = 1;
    orc_number_fluid_d3q19(mask, nf, ini);
    orc_stream_map_d3q19(Li, Li, 1, ini, st);
    orc_init_tg_map_d3q19(Li, u_0, 1.0, a);
    const int n_steps = (int)((double)(100 * Li) / u_0);
    const double e0 = orc_kinetic_energy_d3q19(a, nf);
    double e = e0;
    int unstable = 0, step;
    for (step = 0; step < n_steps; step++) {
        if (step % 100 == 0 && step > 0) {
            e = orc_kinetic_energy_d3q19(a, nf);
            if (e > 2. * e0 || !isfinite(e)) {
                unstable = 1;
                break;
            }
        }
        orc_collide_stream_d3q19(a, b, st, nf, col, tau, tau_nh, 0.0, 1);
        double *t = a; a = b; b = t;
    }
    if (stop_step) *stop_step = step;
    if (e_last) *e_last = e;
    free(a); free(b); free(mask); free(ini); free(st);
    return unstable;
}

/* ---- one obstacle map point, roda_ponto,
 * src/main_stability-map_obstacle.cpp:75-123.  *stop_step = steps run.      */
int orc_run_obstacle_map_point(double u_imp, double tau, double tau_nh,
                               int *stop_step)
{
    setup_d3q19();
    const int nx = 22, ny = 12, nz = 1, nv = nx * ny;
    int mask[22 * 12], ini[22 * 12];
    orc_obstacle_mask(mask);
    const int nf = orc_number_fluid_d3q19(mask, nv, ini);
    double *a = (double *)malloc(sizeof(double) * nf * 19);
    double *b = (double *)malloc(sizeof(double) * nf * 19);
    int *st = (int *)malloc(sizeof(int) * nf * 19);
    orc_stream_map_d3q19(nx, ny, nz, ini, st);
    const double rho_ini = 1.0;
    orc_init_rest_d3q19(nf, rho_ini, a);
    const int n_steps = (int)((double)(200 * ny) / u_imp);
    int unstable = 0, step;
    for (step = 0; step < n_steps && !unstable; step++) {
        orc_zouhe_d3q19_vx_x(1, u_imp, nx, ny, nz, ini, a);
        orc_devnull_rho_x_d3q19(nx - 2, rho_ini, nx, ny, nz, ini, a);
        unstable = orc_collide_stream_d3q19(a, b, st, nf, ORC_COL_EXT_REG, tau,
                                            tau_nh, 10. * rho_ini, 1);
        double *t = a; a = b; b = t;
    }
    if (stop_step) *stop_step = step;
    free(a); free(b); free(st);
    return unstable;
}

/* Run a whole map (or a subsample of it) with OpenMP over points, as the
 * reference map loops do (schedule(dynamic), one private lattice per point).
 * kind 0: Taylor-Green (col, tau_nh used); kind 1: obstacle (Ma).
 * Points are (ii[k], jj[k]), k < n_pts; out[k] = 0/1, steps[k] = steps run. */
void orc_run_map(int kind, int Li, int col, double tau_nh_in, double Ma,
                 int n_i, int n_j, const int *ii, const int *jj, int n_pts,
                 int *out, int *steps, int nthreads)
{
    setup_d3q19();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
    for (int k = 0; k < n_pts; k++) {
        double a, tau, tnh;
        if (kind == 0) {
            orc_tg_map_params(ii[k], jj[k], n_i, n_j, col == ORC_COL_REG,
                              tau_nh_in, &a, &tau, &tnh);
            out[k] = orc_run_tg_map_point(Li, col, a, tau, tnh, &steps[k], 0);
        } else {
            orc_obstacle_map_params(ii[k], jj[k], n_i, n_j, Ma, &a, &tau, &tnh);
            out[k] = orc_run_obstacle_map_point(a, tau, tnh, &steps[k]);
        }
    }
}

/* ---- Taylor-Green analytic error, src/main_taylor-green.cpp:178-213 --------
 * out[0] = E_vector, out[1] = E_magnitude, out[2] = (nu_eff - nu)/nu.       */
void orc_tg_error_d2q9(const double *inif, int Li, double u_0, double visc,
                       int step, double *out)
{
    setup_d2q9();
    const double L = Li;
    const double k = 2. * M_PI / L;
    double s_mag = 0, s_vec = 0, s_r = 0, s_proj = 0, s_phi = 0;
    const double e = exp(-2. * visc * k * k * step);
    for (int y = 0; y < Li; y++)
        for (int x = 0; x < Li; x++) {
            double vx, vy, vz, rho;
            calcula_d2q9(inif + (size_t)(x + y * Li) * 9, &vx, &vy, &vz, &rho);
            const double rx = u_0 * e * sin(k * x) * cos(k * y);
            const double ry = -u_0 * e * cos(k * x) * sin(k * y);
            const double us = sqrt(vx * vx + vy * vy + vz * vz),
                         ur = sqrt(rx * rx + ry * ry);
            s_mag += (us - ur) * (us - ur);
            s_vec += (vx - rx) * (vx - rx) + (vy - ry) * (vy - ry) + vz * vz;
            s_r += ur * ur;
            const double px = sin(k * x) * cos(k * y), py = -cos(k * x) * sin(k * y);
            s_proj += vx * px + vy * py;
            s_phi += px * px + py * py;
        }
    const double A = s_proj / s_phi / u_0;
    const double nu_ef = -log(A) / (2. * k * k * step);
    out[0] = sqrt(s_vec / s_r);
    out[1] = sqrt(s_mag / s_r);
    out[2] = nu_ef / visc - 1.;
}

/* ---- whole Taylor-Green run, src/main_taylor-green.cpp:104-217 (D2Q9) ------
 * Fig. 3 mode (Re given) when tau_in <= 0, kappa_4 mode (--tau) otherwise.
 * out[0..2] as orc_tg_error_d2q9; out[3] = tau; out[4] = steps nT.          */
void orc_run_taylor_green_d2q9(int Li, int col, double tau_nh_in, double Re,
                               double u_0, double tau_in, int neq, int poisson,
                               int nthreads, double *out)
{
    setup_d2q9();
    double tau, visc;
    const int modo_tau = tau_in > 0;
    if (modo_tau) {
        tau = tau_in;
        visc = (tau - 0.5) / 3.;
    } else {
        visc = u_0 * Li / Re;
        tau = 0.5 + 3. * visc;
    }
    const double tau_nh = (col == ORC_COL_REG) ? 1.0 : (col == ORC_COL_BGK ? tau : tau_nh_in);
    const int nf = Li * Li;
    int *mask = (int *)malloc(sizeof(int) * nf);
    int *ini = (int *)malloc(sizeof(int) * nf);
    int *st = (int *)malloc(sizeof(int) * nf * 9);
    double *a = (double *)malloc(sizeof(double) * nf * 9);
    double *b = (double *)malloc(sizeof(double) * nf * 9);
    for (int v = 0; v < nf; v++) mask[v] = 1;
    orc_number_fluid_d2q9(mask, nf, ini);
    orc_stream_map_d2q9(Li, Li, 1, ini, st);
    orc_init_taylor_green_d2q9(Li, u_0, tau, poisson, neq, a);
    const double L = Li;
    const double k = 2. * M_PI / L;
    const int nT = modo_tau ? (int)round(1. / (2. * visc * k * k)) : (int)round(L / u_0);
    for (int step = 0; step <= nT; step++) {
        /* collide in place, measure at step nT before the push (the measured
         * moments are those of the post-collision populations, :160-178) */
        if (step == nT) {
            for (int p = 0; p < nf; p++)
                collide_d2q9(a + (size_t)p * 9, 0., 0., 0., col, tau, tau_nh);
            orc_tg_error_d2q9(a, Li, u_0, visc, step, out);
            break;
        }
        orc_collide_stream_d2q9(a, b, st, nf, col, tau, tau_nh, 0.0, nthreads);
        double *t = a; a = b; b = t;
    }
    out[3] = tau;
    out[4] = nT;
    free(mask); free(ini); free(st); free(a); free(b);
}

/* Run n_steps collide-stream sweeps of a whole lattice (no BCs), ping-pong,
 * under the body force (ax, ay, az). */
void orc_run_steps_d2q9(double *inif, int n_fluid, const int *ini_stream,
                        int n_steps, int col, double tau, double tau_nh,
                        double ax, double ay, double az, int nthreads)
{
    double *b = (double *)malloc(sizeof(double) * (size_t)n_fluid * 9);
    double *a = inif;
    for (int s = 0; s < n_steps; s++) {
        orc_collide_stream_acc_d2q9(a, b, ini_stream, n_fluid, col, tau, tau_nh, ax, ay, az, 0.0, nthreads);
        double *t = a; a = b; b = t;
    }
    if (a != inif) {
        memcpy(inif, a, sizeof(double) * (size_t)n_fluid * 9);
        free(a);
    } else {
        free(b);
    }
}

void orc_run_steps_d3q19(double *inif, int n_fluid, const int *ini_stream,
                         int n_steps, int col, double tau, double tau_nh,
                         double ax, double ay, double az, int nthreads)
{
    double *b = (double *)malloc(sizeof(double) * (size_t)n_fluid * 19);
    double *a = inif;
    for (int s = 0; s < n_steps; s++) {
        orc_collide_stream_acc_d3q19(a, b, ini_stream, n_fluid, col, tau, tau_nh, ax, ay, az, 0.0, nthreads);
        double *t = a; a = b; b = t;
    }
    if (a != inif) {
        memcpy(inif, a, sizeof(double) * (size_t)n_fluid * 19);
        free(a);
    } else {
        free(b);
    }
}

/* General D3Q19 step loop with the obstacle driver's boundary conditions:
 * per step ZouHe at inlet_x (if >= 0), devnull at outlet_x (if >= 0), then
 * collide (+density check when dens_limit > 0; body force ax, ay, az) and
 * push.  Returns the number of steps run; stops after the first step whose
 * check fails.                                                             */
int orc_run_bc_steps_d3q19(double *inif, int nx, int ny, int nz, const int *ini,
                           const int *ini_stream, int n_fluid, int n_steps,
                           int col, double tau, double tau_nh, int inlet_x,
                           double u_in, int outlet_x, double rho_out,
                           double dens_limit, double ax, double ay, double az,
                           int *violated)
{
    double *b = (double *)malloc(sizeof(double) * (size_t)n_fluid * 19);
    double *a = inif;
    int bad = 0, s;
    for (s = 0; s < n_steps && !bad; s++) {
        if (inlet_x >= 0) orc_zouhe_d3q19_vx_x(inlet_x, u_in, nx, ny, nz, ini, a);
        if (outlet_x >= 0) orc_devnull_rho_x_d3q19(outlet_x, rho_out, nx, ny, nz, ini, a);
        bad = orc_collide_stream_acc_d3q19(a, b, ini_stream, n_fluid, col, tau, tau_nh, ax, ay, az, dens_limit, 1);
        double *t = a; a = b; b = t;
    }
    if (a != inif) {
        memcpy(inif, a, sizeof(double) * (size_t)n_fluid * 19);
        free(a);
    } else {
        free(b);
    }
    if (violated) *violated = bad;
    return s;
}

/* D2Q9 analogue with ZouHeD2Q9_vx_x and devnull_rho_x. */
int orc_run_bc_steps_d2q9(double *inif, int nx, int ny, const int *ini,
                          const int *ini_stream, int n_fluid, int n_steps,
                          int col, double tau, double tau_nh, int inlet_x,
                          double u_in, int outlet_x, double rho_out,
                          double dens_limit, double ax, double ay, double az,
                          int *violated)
{
    double *b = (double *)malloc(sizeof(double) * (size_t)n_fluid * 9);
    double *a = inif;
    int bad = 0, s;
    for (s = 0; s < n_steps && !bad; s++) {
        if (inlet_x >= 0) orc_zouhe_d2q9_vx_x(inlet_x, u_in, nx, ny, ini, a);
        if (outlet_x >= 0) orc_devnull_rho_x_d2q9(outlet_x, rho_out, nx, ny, 1, ini, a);
        bad = orc_collide_stream_acc_d2q9(a, b, ini_stream, n_fluid, col, tau, tau_nh, ax, ay, az, dens_limit, 1);
        double *t = a; a = b; b = t;
    }
    if (a != inif) {
        memcpy(inif, a, sizeof(double) * (size_t)n_fluid * 9);
        free(a);
    } else {
        free(b);
    }
    if (violated) *violated = bad;
    return s;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
