// ref_harness.cpp -- TEST INFRASTRUCTURE.
//
// Compiles the reference's own SimBoltz library and ERLBM drivers, as they
// lie under the reference checkout, into a shared library with a C ABI, so
// that the oracle restatement can be checked against the real thing.  The
// reference sources are #included from their own directory (the -I paths of
// oracle/Makefile); nothing of them is copied.  The drivers' main() functions
// are renamed so the per-point runners and initialisers can be called.
//
// Built twice per velocity set: "strict" (-ffp-contract=off) and "fma"
// (the reference Makefile's default contraction with FMA hardware).

#ifndef LATT
#define LATT 9
#endif

#if LATT == 9
#define nvel 9
#define dim 2
#define main ref_main_taylor_green
#include "main_taylor-green.cpp"
#undef main
#define main ref_main_double_shear
#include "main_double-shear-layer.cpp"
#undef main
#define main ref_main_error_calculator
#include "error_calculator.cpp"
#undef main
#else
#define nvel 19
#define dim 3
#define main ref_main_tg_map
#include "main_stability-map_taylor-green.cpp"
#undef main
#define main ref_main_obstacle_map
#include "main_stability-map_obstacle.cpp"
#undef main
#endif

#if LATT == 9
#define RSFX(n) n##_d2q9
#else
#define RSFX(n) n##_d3q19
#endif

namespace {

// A lattice description built with the reference's own helpers
// (src/erlbm_common.h:89-148).
struct RefLattice {
    GEOMETRY g;
    LATTICE l;
};

void build(RefLattice &r, int nx, int ny, int nz, const int *mask)
{
    r.g.file = "";
    r.g.ftesc = 1.0;
    r.g.nx = nx;
    r.g.ny = ny;
    r.g.nz = nz;
    r.g.ini = new int[nx * ny * nz];
    for (int i = 0; i < nx * ny * nz; i++) r.g.ini[i] = mask ? (mask[i] ? 1 : 0) : 1;
    numera_fluido(r.g);
    define_rede(r.g, r.l);
}

void release(RefLattice &r)
{
    delete[] r.g.ini;
    delete[] r.l.c_i;
    delete[] r.l.Q_i;
    delete[] r.l.ini_stream;
}

PARAMETERS params(double tau, double tau_nh)
{
    PARAMETERS p;
    p.tau = tau;
    p.visc = (tau - 0.5) / 3.;
    p.tau_nh = tau_nh;
    p.rho_ini = 1.0;
    return p;
}

}  // namespace

extern "C" {

// Velocity set, weights and Q tensor as define_rede() builds them.
void RSFX(ref_lattice)(double *c, double *w, double *q)
{
    RefLattice r;
    build(r, 2, 2, dim == 3 ? 2 : 1, nullptr);
    for (int i = 0; i < nvel * dim; i++) c[i] = r.l.c_i[i];
    for (int i = 0; i < nvel; i++) w[i] = r.l.w[i];
    for (int i = 0; i < nvel * dim * dim; i++) q[i] = r.l.Q_i[i];
    release(r);
}

// numera_fluido + def_dir_prop on a mask; returns the fluid count.
int RSFX(ref_stream_map)(int nx, int ny, int nz, const int *mask, int *ini_out,
                         int *stream_out)
{
    RefLattice r;
    build(r, nx, ny, nz, mask);
    for (int i = 0; i < nx * ny * nz; i++) ini_out[i] = r.g.ini[i];
    for (int i = 0; i < r.g.fluid * nvel; i++) stream_out[i] = r.l.ini_stream[i];
    int n = r.g.fluid;
    release(r);
    return n;
}

// colisao_sitio on n_sites AoS sites.
void RSFX(ref_collide)(double *f, int n_sites, double ax, double ay, double az,
                       int col, double tau, double tau_nh)
{
    RefLattice r;
    build(r, 2, 2, dim == 3 ? 2 : 1, nullptr);
    PARAMETERS p = params(tau, tau_nh);
    for (int s = 0; s < n_sites; s++) colisao_sitio(f + s * nvel, ax, ay, az, r.l, p, col);
    release(r);
}

void RSFX(ref_calcula)(double *f, double *out4)
{
    RefLattice r;
    build(r, 2, 2, dim == 3 ? 2 : 1, nullptr);
    calcula(f, out4[0], out4[1], out4[2], out4[3], r.l);
    release(r);
}

void RSFX(ref_dist_eq)(double *feq, double vx, double vy, double vz, double rho)
{
    RefLattice r;
    build(r, 2, 2, dim == 3 ? 2 : 1, nullptr);
    dist_eq(feq, vx, vy, vz, rho, r.l);
    release(r);
}

double RSFX(ref_density)(double *f) { return density(f); }

double RSFX(ref_kinetic_energy)(int nx, int ny, int nz, const int *mask, double *inif)
{
    RefLattice r;
    build(r, nx, ny, nz, mask);
    r.l.inif = inif;
    double e = kinectic_energy(r.g, r.l);
    release(r);
    return e;
}

// A boundary routine applied once to inif (fluid-ordered AoS).
// which: 0 devnull_rho_x, 1 ZouHeD3Q19_vx_x, 2 ZouHeD2Q9_vx_x.
void RSFX(ref_boundary)(int which, int nx, int ny, int nz, const int *mask,
                        double *inif, int pos_x, double value)
{
    RefLattice r;
    build(r, nx, ny, nz, mask);
    r.l.inif = inif;
    if (which == 0) devnull_rho_x(pos_x, value, r.g, r.l);
#if nvel == 19
    if (which == 1) ZouHeD3Q19_vx_x(pos_x, value, r.g, r.l);
#else
    if (which == 2) ZouHeD2Q9_vx_x(pos_x, value, r.g, r.l);
#endif
    r.l.inif = nullptr;
    release(r);
}

// The drivers' step loop: [ZouHe inlet] -> [devnull outlet] -> per site
// {collide; [density check]; propag_site} -> swap, n_steps times, stopping
// after a step whose check fails (the obstacle driver's loop condition).
// ax, ay, az: the acc_x, acc_y, acc_z passed to colisao_sitio (the drivers
// pass 0).
int RSFX(ref_run_steps)(int nx, int ny, int nz, const int *mask, double *inif,
                        int n_steps, int col, double tau, double tau_nh,
                        int inlet_x, double u_in, int outlet_x, double rho_out,
                        double dens_limit, double ax, double ay, double az,
                        int *violated)
{
    RefLattice r;
    build(r, nx, ny, nz, mask);
    PARAMETERS p = params(tau, tau_nh);
    const int nf = r.g.fluid;
    double *a = new double[nf * nvel];
    double *b = new double[nf * nvel];
    for (int i = 0; i < nf * nvel; i++) a[i] = inif[i];
    r.l.inif = a;
    r.l.inif_new = b;
    int bad = 0, s;
    for (s = 0; s < n_steps && !bad; s++) {
#if nvel == 19
        if (inlet_x >= 0) ZouHeD3Q19_vx_x(inlet_x, u_in, r.g, r.l);
#else
        if (inlet_x >= 0) ZouHeD2Q9_vx_x(inlet_x, u_in, r.g, r.l);
#endif
        if (outlet_x >= 0) devnull_rho_x(outlet_x, rho_out, r.g, r.l);
        for (int pto = 0; pto < nf; pto++) {
            double *f = r.l.inif + pto * nvel;
            colisao_sitio(f, ax, ay, az, r.l, p, col);
            if (dens_limit > 0) {
                const double d = density(f);
                if (fabs(d) > dens_limit || !std::isfinite(d)) bad = 1;
            }
            double mx, my, mz;
            propag_site(r.l, pto, mx, my, mz);
        }
        double *t = r.l.inif;
        r.l.inif = r.l.inif_new;
        r.l.inif_new = t;
    }
    for (int i = 0; i < nf * nvel; i++) inif[i] = r.l.inif[i];
    delete[] a;
    delete[] b;
    r.l.inif = r.l.inif_new = nullptr;
    release(r);
    if (violated) *violated = bad;
    return s;
}

// The double-shear driver's OpenMP step loop (src/main_double-shear-layer.cpp:
// 162-180, 210-214) on a periodic nx x ny x nz lattice: n_steps sweeps of
// `omp parallel for` {colisao_sitio; propag_site}, then the pointer swap.
// Used as the timed CPU baseline.  Returns the wall time in seconds.
double RSFX(ref_run_steps_omp)(int nx, int ny, int nz, double *inif, int n_steps, int col, double tau,
                               double tau_nh, int nthreads)
{
    RefLattice r;
    build(r, nx, ny, nz, nullptr);
    PARAMETERS p = params(tau, tau_nh);
    const int n_fluid = r.g.fluid;
    double *b = new double[(size_t)n_fluid * nvel];
    r.l.inif = inif;
    r.l.inif_new = b;
    LATTICE lattice = r.l;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const double t0 = omp_get_wtime();
    for (int step = 0; step < n_steps; step++) {
#pragma omp parallel for
        for (int pto = 0; pto < n_fluid; pto++) {
            double *f = lattice.inif + pto * nvel;
            colisao_sitio(f, 0., 0., 0., lattice, p, col);
            double mx, my, mz;
            propag_site(lattice, pto, mx, my, mz);
        }
        double *temp = lattice.inif;
        lattice.inif = lattice.inif_new;
        lattice.inif_new = temp;
    }
    const double t1 = omp_get_wtime();
    if (lattice.inif != inif) {
        for (size_t i = 0; i < (size_t)n_fluid * nvel; i++) inif[i] = lattice.inif[i];
    }
    delete[] b;
    r.l.inif = r.l.inif_new = nullptr;
    release(r);
    return t1 - t0;
}

#if LATT == 9
// inicializa_taylor_green (src/main_taylor-green.cpp:57-100).
void ref_init_taylor_green_d2q9(int Li, double u_0, double tau, int poisson, int neq, double *inif)
{
    RefLattice r;
    build(r, Li, Li, 1, nullptr);
    r.l.inif = inif;
    inicializa_taylor_green(r.g, r.l, u_0, tau, poisson != 0, neq != 0);
    r.l.inif = nullptr;
    release(r);
}

// initial_double_shear_layer (src/main_double-shear-layer.cpp:48-66).
void ref_init_double_shear_d2q9(int Li, double *inif)
{
    RefLattice r;
    build(r, Li, Li, 1, nullptr);
    r.l.inif = inif;
    initial_double_shear_layer(r.g, r.l, (double)Li, 1.0);
    r.l.inif = nullptr;
    release(r);
}

// The whole Taylor-Green driver (prints its one result line to stdout).
int ref_taylor_green_main(int argc, char **argv) { return ref_main_taylor_green(argc, argv); }
#else
// roda_ponto of the Taylor-Green map (src/main_stability-map_taylor-green.cpp:41-104)
// over a list of map points, with the driver's axis expressions (:134-164).
void ref_tg_map_points(int col, double tau_nh_in, int Li, int n_ma, int n_nu,
                       const int *ii, const int *jj, int n_pts, int *out)
{
    GEOMETRY geometry;
    geometria_periodica(geometry, Li, Li, 1);
    LATTICE lattice;
    define_rede(geometry, lattice);
    const double c_s = 1. / sqrt(3.);
    const double u_min = 0.05 * c_s;
    const double log_nu_min = -5.;
    const double log_nu_max = -2.;
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < n_pts; k++) {
        const int i = ii[k], j = jj[k];
        const double u_0 = u_min + (i + 1) * (c_s - u_min) / n_ma;
        const double log_nu = log_nu_min + j * (log_nu_max - log_nu_min) / n_nu;
        PARAMETERS parameters;
        parameters.tau = 3. * pow(10., log_nu) + 0.5;
        parameters.visc = pow(10., log_nu);
        parameters.tau_nh = (col == COL_REG) ? 1.0 : tau_nh_in;
        parameters.rho_ini = 1.0;
        out[k] = roda_ponto(geometry, lattice, parameters, col, u_0);
    }
}

// roda_ponto of the obstacle map (src/main_stability-map_obstacle.cpp:75-123)
// with the driver's axis expressions (:147-172).
void ref_obstacle_map_points(double Ma, int n_tau, int n_nu, const int *ii,
                             const int *jj, int n_pts, int *out)
{
    GEOMETRY geometry;
    geometria_obstaculo(geometry);
    LATTICE lattice;
    define_rede(geometry, lattice);
    const double u_imp = Ma / sqrt(3.);
    const double tau_nh_min = 0.5, tau_nh_max = 1.1;
    const double log_nu_min = -5., log_nu_max = -0.7;
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < n_pts; k++) {
        const int i = ii[k], j = jj[k];
        const double log_nu = log_nu_min + j * (log_nu_max - log_nu_min) / n_nu;
        PARAMETERS parameters;
        parameters.tau = 3. * pow(10., log_nu) + 0.5;
        parameters.visc = pow(10., log_nu);
        parameters.tau_nh = tau_nh_min + (double)i * (tau_nh_max - tau_nh_min) / n_tau;
        parameters.rho_ini = 1.0;
        out[k] = roda_ponto(geometry, lattice, parameters, u_imp);
    }
}

// The obstacle geometry as geometria_obstaculo builds it (numbered ini).
// (declared below, D3Q19 only)
#endif

// rec_velocity / rec_vorticity of inif, written into the current directory
// under the names the reference gives them.
void RSFX(ref_rec_velocity)(int nx, int ny, int nz, const int *mask, double *inif, double tempo)
{
    RefLattice r;
    build(r, nx, ny, nz, mask);
    r.l.inif = inif;
    rec_velocity(r.g, r.l, tempo);
    release(r);
}

void RSFX(ref_rec_vorticity)(int nx, int ny, int nz, const int *mask, double *inif, unsigned int passo)
{
    RefLattice r;
    build(r, nx, ny, nz, mask);
    r.l.inif = inif;
    rec_vorticity(r.g, r.l, passo);
    release(r);
}

double RSFX(ref_enstrophy)(int nx, int ny, int nz, const int *mask, double *inif)
{
    RefLattice r;
    build(r, nx, ny, nz, mask);
    r.l.inif = inif;
    const double e = calc_enstrophy(r.g, r.l);
    release(r);
    return e;
}

// A driver's main() with the given arguments (argv[0] included).
// D2Q9 library: 0 main_taylor-green, 1 main_double-shear-layer,
// 2 tools/error_calculator; D3Q19: 0 main_stability-map_taylor-green,
// 1 main_stability-map_obstacle.
int RSFX(ref_driver_main)(int which, int argc, char **argv)
{
#if LATT == 9
    if (which == 0) return ref_main_taylor_green(argc, argv);
    if (which == 1) return ref_main_double_shear(argc, argv);
    if (which == 2) return ref_main_error_calculator(argc, argv);
#else
    if (which == 0) return ref_main_tg_map(argc, argv);
    if (which == 1) return ref_main_obstacle_map(argc, argv);
#endif
    return -1;
}

#if LATT != 9
// The obstacle geometry as geometria_obstaculo builds it (numbered ini).
int ref_obstacle_geometry(int *ini_out)
{
    GEOMETRY geometry;
    geometria_obstaculo(geometry);
    for (int i = 0; i < geometry.nx * geometry.ny * geometry.nz; i++) ini_out[i] = geometry.ini[i];
    int n = geometry.fluid;
    delete[] geometry.ini;
    return n;
}
#endif

}  // extern "C"
