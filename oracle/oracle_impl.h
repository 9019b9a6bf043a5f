/*
 * CPU restatement of the SimBoltz E-RLBM hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This header is included twice by erlbm_oracle.c, once per velocity set, with
 *   NV   number of velocities (9 or 19)
 *   DIM  spatial dimension (2 or 3)
 *   SFX(name)  suffixing macro (name##_d2q9 / name##_d3q19)
 * defined.  Every function restates one reference function in the same
 * floating-point operation order (the file is compiled with
 * -ffp-contract=off), so that its results are bitwise those of the reference
 * built without FMA contraction.  Citations are path:line in loesantos/ERLBM.
 *
 * Only tests/, the smoke check and bench.py's CPU-baseline leg use this code,
 * as the checker.  The GPU product path never calls it.
 */

/* ---- velocity set, weights, Q tensor -------------------------------------
 * SimBoltz_Functions/LBM_functions.cpp:577-708 (def_lattice_d3q19),
 * :1437-1501 (def_lattice_d2q9), :5815-5829 (calc_Q).                      */

static double SFX(C)[NV * DIM];
static double SFX(W)[NV];
static double SFX(QT)[NV * DIM * DIM];
static int SFX(OPP)[NV];
static int SFX(ready) = 0;

static void SFX(setup)(void)
{
    if (SFX(ready)) return;
#if NV == 9
    static const int cc[9][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1},
                                 {1, 1}, {-1, -1}, {1, -1}, {-1, 1}};
    for (int i = 0; i < 9; i++) {
        SFX(C)[i * 2 + 0] = cc[i][0];
        SFX(C)[i * 2 + 1] = cc[i][1];
    }
    SFX(W)[0] = 4. / 9.;
    for (int i = 1; i < 5; i++) SFX(W)[i] = 1. / 9.;
    for (int i = 5; i < 9; i++) SFX(W)[i] = 1. / 36.;
#else
    static const int cc[19][3] = {
        {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
        {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
        {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
        {0, -1, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}};
    for (int i = 0; i < 19; i++)
        for (int a = 0; a < 3; a++) SFX(C)[i * 3 + a] = cc[i][a];
    SFX(W)[0] = 1. / 3.;
    for (int i = 1; i < 7; i++) SFX(W)[i] = 1. / 18.;
    for (int i = 7; i < 19; i++) SFX(W)[i] = 1. / 36.;
#endif
    const double c_s2 = 1. / 3.;
    for (int i = 0; i < NV; i++)
        for (int a = 0; a < DIM; a++)
            for (int b = 0; b < DIM; b++)
                SFX(QT)[a + b * DIM + i * DIM * DIM] =
                    SFX(C)[i * DIM + a] * SFX(C)[i * DIM + b] - c_s2 * (a == b);
    /* opposite directions, LBM_functions.cpp:2035-2055 / :2196-2219 */
    for (int i = 0; i < NV; i++)
        for (int j = 0; j < NV; j++) {
            int same = 1;
            for (int a = 0; a < DIM; a++)
                if (SFX(C)[j * DIM + a] != -SFX(C)[i * DIM + a]) same = 0;
            if (same) SFX(OPP)[i] = j;
        }
    SFX(ready) = 1;
}

void SFX(orc_lattice)(double *c, double *w, double *q, int *opp)
{
    SFX(setup)();
    memcpy(c, SFX(C), sizeof(SFX(C)));
    memcpy(w, SFX(W), sizeof(SFX(W)));
    memcpy(q, SFX(QT), sizeof(SFX(QT)));
    memcpy(opp, SFX(OPP), sizeof(SFX(OPP)));
}

/* ---- dot_product, LBM_functions.cpp:2864-2876 ---------------------------- */
static inline double SFX(dot)(const double *a, const double *b)
{
    double r = 0.0;
    for (int k = 0; k < DIM; k++) r = r + a[k] * b[k];
    return r;
}

/* ---- calcula, LBM_functions.cpp:3517-3589 -------------------------------- */
static inline void SFX(calcula)(const double *f, double *vx, double *vy,
                                double *vz, double *rho)
{
    const double *c = SFX(C);
    double mx = 0.0, my = 0.0, mz = 0.0, r = 0.0;
    for (int i = 0; i < NV; i++) {
        r = r + f[i];
        mx = mx + c[i * DIM + 0] * f[i];
        my = my + c[i * DIM + 1] * f[i];
#if DIM == 3
        mz = mz + c[i * DIM + 2] * f[i];
#endif
    }
    *rho = r;
    if (r) {
        double oor = 1.0 / (r);
        *vx = mx * oor;
        *vy = my * oor;
#if DIM == 3
        *vz = mz * oor;
#else
        *vz = 0.0;
#endif
    } else {
        *vx = 0.;
        *vy = 0.;
        *vz = 0.;
    }
    (void)mz;
}

void SFX(orc_calcula)(const double *f, double *out4)
{
    SFX(setup)();
    SFX(calcula)(f, &out4[0], &out4[1], &out4[2], &out4[3]);
}

/* ---- dist_eq, LBM_functions.cpp:2944-3000 -------------------------------- */
static inline void SFX(dist_eq)(double *feq, double vx, double vy, double vz,
                                double rho)
{
    const double ocs2 = 3.;
    double v[3] = {vx, vy, vz};
#if DIM == 2
    double vquad = (vx * vx + vy * vy);
#else
    double vquad = (vx * vx + vy * vy + vz * vz);
#endif
    for (int i = 0; i < NV; i++) {
        double cv = SFX(dot)(&SFX(C)[i * DIM], v);
        feq[i] = SFX(W)[i] * rho *
                 (1.0 + cv * ocs2 + 0.5 * cv * cv * ocs2 * ocs2 -
                  0.5 * vquad * ocs2);
    }
}

void SFX(orc_dist_eq)(double *feq, double vx, double vy, double vz, double rho)
{
    SFX(setup)();
    SFX(dist_eq)(feq, vx, vy, vz, rho);
}

/* ---- dist_eq(feq, T), LBM_functions.cpp:3015-3024 (rest equilibrium) ----- */
static inline void SFX(dist_eq_rest)(double *feq, double t)
{
    for (int i = 0; i < NV; i++) feq[i] = SFX(W)[i] * t;
}

/* ---- source, LBM_functions.cpp:5645-5716 --------------------------------- */
static inline void SFX(source)(double Fx, double Fy, double Fz, double vx,
                               double vy, double vz, double rho, double tau,
                               double *S)
{
    const double ocs2 = 3.;
    double v[3] = {vx, vy, vz};
    double F[3] = {Fx, Fy, Fz};
    double factor = (1. - 1. / (2. * tau));
    double vF = SFX(dot)(F, v);
    for (int i = 0; i < NV; i++) {
        double cF = SFX(dot)(&SFX(C)[i * DIM], F);
        double cv = SFX(dot)(&SFX(C)[i * DIM], v);
        S[i] = factor * SFX(W)[i] * ocs2 * (cF + ocs2 * cF * cv - vF);
    }
    (void)rho;
}

/* ---- density, LBM_functions.cpp:5543-5552 -------------------------------- */
static inline double SFX(density)(const double *f)
{
    double r = 0.0;
    for (int i = 0; i < NV; i++) r = r + f[i];
    return r;
}

double SFX(orc_density)(const double *f) { return SFX(density)(f); }

/* Shared front half of coll_Reg / coll_Ext_Reg: moments, source, f_eq and the
 * non-equilibrium flux tensor Pi^(1), then f^(1).
 * Collision_operators.cpp:1190-1244 (2D) and :1257-1294 (3D);
 * coll_Reg :866-912 / :925-960 has the same operation order.               */
static inline void SFX(reg_front)(const double *f, double ax, double ay,
                                  double az, double tau, double *S,
                                  double *f_eq, double *f_1)
{
    double vx, vy, vz, rho;
    SFX(calcula)(f, &vx, &vy, &vz, &rho);
    double vx_alt = vx + 0.5 * ax;
    double vy_alt = vy + 0.5 * ay;
    double vz_alt = vz + 0.5 * az;
    double Fx = ax * rho;
    double Fy = ay * rho;
    double Fz = az * rho;
    SFX(source)(Fx, Fy, Fz, vx_alt, vy_alt, vz_alt, rho, tau, S);
    SFX(dist_eq)(f_eq, vx_alt, vy_alt, vz_alt, rho);
    const double *c = SFX(C);
    const double *Q = SFX(QT);
    const double one_over_2cs4 = 0.5 * 3. * 3.;
#if DIM == 2
    double Pxx = 0.0, Pyy = 0.0, Pxy = 0.0;
    for (int i = 0; i < NV; i++) {
        Pxx = Pxx + (f[i] - f_eq[i]) * c[i * 2] * c[i * 2];
        Pyy = Pyy + (f[i] - f_eq[i]) * c[i * 2 + 1] * c[i * 2 + 1];
        Pxy = Pxy + (f[i] - f_eq[i]) * c[i * 2] * c[i * 2 + 1];
    }
    for (int i = 0; i < NV; i++) {
        int ind = i * 4;
        f_1[i] = one_over_2cs4 * SFX(W)[i] *
                 (Q[ind] * Pxx + Q[ind + 3] * Pyy + 2 * Q[ind + 1] * Pxy);
    }
#else
    double Pxx = 0.0, Pyy = 0.0, Pzz = 0.0, Pxy = 0.0, Pxz = 0.0, Pyz = 0.0;
    for (int i = 0; i < NV; i++) {
        Pxx = Pxx + (f[i] - f_eq[i]) * c[i * 3] * c[i * 3];
        Pyy = Pyy + (f[i] - f_eq[i]) * c[i * 3 + 1] * c[i * 3 + 1];
        Pzz = Pzz + (f[i] - f_eq[i]) * c[i * 3 + 2] * c[i * 3 + 2];
        Pxy = Pxy + (f[i] - f_eq[i]) * c[i * 3] * c[i * 3 + 1];
        Pxz = Pxz + (f[i] - f_eq[i]) * c[i * 3] * c[i * 3 + 2];
        Pyz = Pyz + (f[i] - f_eq[i]) * c[i * 3 + 1] * c[i * 3 + 2];
    }
    for (int i = 0; i < NV; i++) {
        int ind = i * 9;
        f_1[i] = one_over_2cs4 * SFX(W)[i] *
                 (Q[ind] * Pxx + Q[ind + 4] * Pyy + Q[ind + 8] * Pzz +
                  2 * Q[ind + 1] * Pxy + 2 * Q[ind + 2] * Pxz +
                  2 * Q[ind + 5] * Pyz);
    }
#endif
}

/* ---- coll_Ext_Reg, Collision_operators.cpp:1181-1307 ----------------------
 * f* = f_eq + (1-1/tau) f1 + (1-1/tau_nh) f_gt1 + S,  f_gt1 = f - f_eq - f1 */
static inline void SFX(coll_ext_reg)(double *f, double ax, double ay,
                                     double az, double tau, double tau_nh)
{
    double S[NV], f_eq[NV], f_1[NV], f_gt1[NV];
    double omega_1 = 1.0 - 1.0 / tau;
    double omega_gt1 = 1.0 - 1.0 / tau_nh;
    SFX(reg_front)(f, ax, ay, az, tau, S, f_eq, f_1);
    for (int i = 0; i < NV; i++) f_gt1[i] = f[i] - f_eq[i] - f_1[i];
    for (int i = 0; i < NV; i++)
        f[i] = f_eq[i] + omega_1 * f_1[i] + omega_gt1 * f_gt1[i] + S[i];
}

/* ---- coll_Reg, Collision_operators.cpp:859-967 ---------------------------- */
static inline void SFX(coll_reg)(double *f, double ax, double ay, double az,
                                 double tau)
{
    double S[NV], f_eq[NV], f_1[NV];
    double one_over_tau = 1.0 / tau;
    SFX(reg_front)(f, ax, ay, az, tau, S, f_eq, f_1);
    for (int i = 0; i < NV; i++)
        f[i] = f_eq[i] + (1.0 - one_over_tau) * f_1[i] + S[i];
}

/* ---- coll_BGK, Collision_operators.cpp:11-49 ------------------------------ */
static inline void SFX(coll_bgk)(double *f, double ax, double ay, double az,
                                 double tau)
{
    double S[NV], f_eq[NV];
    double vx, vy, vz, rho;
    double one_over_tau = 1.0 / tau;
    SFX(calcula)(f, &vx, &vy, &vz, &rho);
    double vx_alt = vx + 0.5 * ax;
    double vy_alt = vy + 0.5 * ay;
    double vz_alt = vz + 0.5 * az;
    double Fx = ax * rho, Fy = ay * rho, Fz = az * rho;
    SFX(source)(Fx, Fy, Fz, vx_alt, vy_alt, vz_alt, rho, tau, S);
    SFX(dist_eq)(f_eq, vx_alt, vy_alt, vz_alt, rho);
    for (int i = 0; i < NV; i++) {
        double op = (f_eq[i] - f[i]) * one_over_tau;
        f[i] = f[i] + op + S[i];
    }
}

#if NV == 9
/* ---- coll_MRT, D2Q9 branch: Collision_operators.cpp:611-701, with the
 * moment matrix and its inverse of :449-471 (moments rho, e, epsilon, jx,
 * qx, jy, qy, pxx, pxy over the velocity order of def_lattice_d2q9) and the
 * relaxation rates of :614-622.  The reference dispatches MRT for D2Q9
 * only (src/erlbm_common.h:75-77).                                          */
static const double SFX(MRT_M)[9][9] = {
    {1., 1., 1., 1., 1., 1., 1., 1., 1.},
    {-4., -1., -1., -1., -1., 2., 2., 2., 2.},
    {4., -2., -2., -2., -2., 1., 1., 1., 1.},
    {0., 1., -1., 0., 0., 1., -1., 1., -1.},
    {0., -2., 2., 0., 0., 1., -1., 1., -1.},
    {0., 0., 0., 1., -1., 1., -1., -1., 1.},
    {0., 0., 0., -2., 2., 1., -1., -1., 1.},
    {0., 1., 1., -1., -1., 0., 0., 0., 0.},
    {0., 0., 0., 0., 0., 1., 1., -1., -1.}};
static const double SFX(MRT_MI)[9][9] = {
    {1. / 9., -1. / 9., 1. / 9., 0., 0., 0., 0., 0., 0.},
    {1. / 9., -1. / 36., -1. / 18., 1. / 6., -1. / 6., 0., 0., 1. / 4., 0.},
    {1. / 9., -1. / 36., -1. / 18., -1. / 6., 1. / 6., 0., 0., 1. / 4., 0.},
    {1. / 9., -1. / 36., -1. / 18., 0., 0., 1. / 6., -1. / 6., -1. / 4., 0.},
    {1. / 9., -1. / 36., -1. / 18., 0., 0., -1. / 6., 1. / 6., -1. / 4., 0.},
    {1. / 9., 1. / 18., 1. / 36., 1. / 6., 1. / 12., 1. / 6., 1. / 12., 0., 1. / 4.},
    {1. / 9., 1. / 18., 1. / 36., -1. / 6., -1. / 12., -1. / 6., -1. / 12., 0., 1. / 4.},
    {1. / 9., 1. / 18., 1. / 36., 1. / 6., 1. / 12., -1. / 6., -1. / 12., 0., -1. / 4.},
    {1. / 9., 1. / 18., 1. / 36., -1. / 6., -1. / 12., 1. / 6., 1. / 12., 0., -1. / 4.}};

static inline void SFX(coll_mrt)(double *f, double ax, double ay, double az, double tau)
{
    const double s[9] = {0.0, 1.6, 1.8, 0.0, 1.9, 0.0, 1.9, 1. / tau, 1. / tau};
    double f_eq[9], vx, vy, vz, rho;
    SFX(calcula)(f, &vx, &vy, &vz, &rho);
    SFX(dist_eq)(f_eq, vx, vy, vz, rho);
    double S[9] = {0.0};
    if (ax != 0.0 || ay != 0.0 || az != 0.0) {
        const double ocs2 = 3.;
        double v[2] = {vx, vy};
        double F[2] = {ax * rho, ay * rho};
        const double vF = SFX(dot)(F, v);
        for (int i = 0; i < 9; i++) {
            const double cF = SFX(dot)(&SFX(C)[i * 2], F);
            const double cv = SFX(dot)(&SFX(C)[i * 2], v);
            S[i] = SFX(W)[i] * ocs2 * (cF + ocs2 * cF * cv - vF);
        }
    }
    double m[9], meq[9], fm[9];
    for (int r = 0; r < 9; r++) {
        m[r] = 0.0;
        meq[r] = 0.0;
        fm[r] = 0.0;
        for (int c = 0; c < 9; c++) {
            m[r] = m[r] + SFX(MRT_M)[r][c] * f[c];
            meq[r] = meq[r] + SFX(MRT_M)[r][c] * f_eq[c];
            fm[r] = fm[r] + SFX(MRT_M)[r][c] * S[c];
        }
    }
    for (int r = 0; r < 9; r++) m[r] = m[r] - s[r] * (m[r] - meq[r]) + (1.0 - 0.5 * s[r]) * fm[r];
    for (int r = 0; r < 9; r++) {
        f[r] = 0.0;
        for (int c = 0; c < 9; c++) f[r] = f[r] + SFX(MRT_MI)[r][c] * m[c];
    }
}
#endif

/* ---- colisao_sitio dispatch, src/erlbm_common.h:68-81 --------------------- */
static inline void SFX(collide)(double *f, double ax, double ay, double az,
                                int col, double tau, double tau_nh)
{
    switch (col) {
    case ORC_COL_BGK: SFX(coll_bgk)(f, ax, ay, az, tau); break;
    case ORC_COL_REG: SFX(coll_reg)(f, ax, ay, az, tau); break;
    case ORC_COL_EXT_REG: SFX(coll_ext_reg)(f, ax, ay, az, tau, tau_nh); break;
#if NV == 9
    case ORC_COL_MRT: SFX(coll_mrt)(f, ax, ay, az, tau); break;
#endif
    default: break;
    }
}

void SFX(orc_collide)(double *f, int n_sites, double ax, double ay, double az,
                      int col, double tau, double tau_nh)
{
    SFX(setup)();
    for (int s = 0; s < n_sites; s++)
        SFX(collide)(f + (size_t)s * NV, ax, ay, az, col, tau, tau_nh);
}

/* ---- numera_fluido, src/erlbm_common.h:89-99 ------------------------------
 * mask (1 fluid / 0 solid) -> ini (0 solid, k = fluid index k-1), x fastest. */
int SFX(orc_number_fluid)(const int *mask, int n_vox, int *ini)
{
    int n = 0;
    for (int i = 0; i < n_vox; i++) ini[i] = mask[i] ? ++n : 0;
    return n;
}

/* ---- def_dir_prop, LBM_functions.cpp:1967-2311 -----------------------------
 * Push target of population i at fluid site p: ini_stream[p*NV+i].  Periodic
 * wrap ((round(x+c)+n)%n, :2094-2095) and halfway bounce-back (reverse the
 * step, land on the same site with the opposite index, :2099-2128).         */
static int SFX(round_number)(double num) /* Other_functions.cpp:5-13 */
{
    int n;
    if (num < 0) n = (int)(num - 0.5);
    else n = (int)(num + 0.5);
    return n;
}

void SFX(orc_stream_map)(int nx, int ny, int nz, const int *ini,
                         int *ini_stream)
{
    SFX(setup)();
    const double *c = SFX(C);
    double step_x[NV], step_y[NV], step_z[NV];
    int steps[NV];
    for (int i = 1; i < NV; i++) {
        int cxi = SFX(round_number)(c[i * DIM + 0]);
        int cyi = SFX(round_number)(c[i * DIM + 1]);
        int czi = DIM == 3 ? SFX(round_number)(c[i * DIM + DIM - 1]) : 0;
        steps[i] = abs(cxi);
        if (abs(cyi) > steps[i]) steps[i] = abs(cyi);
        if (abs(czi) > steps[i]) steps[i] = abs(czi);
        step_x[i] = c[i * DIM + 0] / steps[i];
        step_y[i] = c[i * DIM + 1] / steps[i];
        step_z[i] = DIM == 3 ? c[i * DIM + DIM - 1] / steps[i] : 0.0;
    }
    const int *opp = SFX(OPP);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                int me = ini[x + y * nx + z * nx * ny];
                if (!me) continue;
                int *dir = ini_stream + (size_t)(me - 1) * NV;
                dir[0] = (me - 1) * NV;
                for (int i = 1; i < NV; i++) {
                    int inv = 0;
                    double sx = step_x[i], sy = step_y[i], sz = step_z[i];
                    double xf = (double)x, yf = (double)y, zf = (double)z;
                    int xp = x, yp = y, zp = z;
                    int dest = me;
                    for (int stp = 1; stp < steps[i] + 1; stp++) {
                        xf = xf + sx;
                        yf = yf + sy;
                        zf = zf + sz;
                        xp = (SFX(round_number)(xf) + nx) % nx;
                        yp = (SFX(round_number)(yf) + ny) % ny;
                        zp = (SFX(round_number)(zf) + nz) % nz;
                        dest = ini[xp + yp * nx + zp * nx * ny];
                        if (dest == 0) {
                            sx = -sx;
                            sy = -sy;
                            sz = -sz;
                            xf = xf + sx;
                            yf = yf + sy;
                            zf = zf + sz;
                            xp = (SFX(round_number)(xf) + nx) % nx;
                            yp = (SFX(round_number)(yf) + ny) % ny;
                            zp = (SFX(round_number)(zf) + nz) % nz;
                            dest = ini[xp + yp * nx + zp * nx * ny];
                            inv++;
                        }
                    }
                    if (inv % 2 == 0) dir[i] = i + (dest - 1) * NV;
                    else dir[i] = opp[i] + (dest - 1) * NV;
                }
            }
}

/* ---- propag_site, LBM_functions.cpp:3885-3955 (monophasic branch) -------- */
static inline void SFX(propag_site)(const double *inif, double *inif_new,
                                    const int *ini_stream, int pto)
{
    const double *f = inif + (size_t)pto * NV;
    const int *dir = ini_stream + (size_t)pto * NV;
    for (int i = 0; i < NV; i++) inif_new[dir[i]] = f[i];
}

void SFX(orc_propagate)(const double *inif, double *inif_new,
                        const int *ini_stream, int n_fluid)
{
    for (int p = 0; p < n_fluid; p++)
        SFX(propag_site)(inif, inif_new, ini_stream, p);
}

/* ---- kinectic_energy, Other_functions.cpp:100-115 (sequential sum) ------- */
double SFX(orc_kinetic_energy)(const double *inif, int n_fluid)
{
    SFX(setup)();
    double sum = 0;
    for (int p = 0; p < n_fluid; p++) {
        double vx, vy, vz, rho;
        SFX(calcula)(inif + (size_t)p * NV, &vx, &vy, &vz, &rho);
        sum = sum + 0.5 * (vx * vx + vy * vy + vz * vz);
    }
    return sum;
}

/* ---- devnull_rho_x, Boundary_conditions.cpp:68-141 ------------------------
 * Column pos_x copies the neighbour column (x+1 west half, x-1 east half)
 * rescaled to density rho; a solid neighbour gives the rest equilibrium.    */
void SFX(orc_devnull_rho_x)(int pos_x, double rho, int nx, int ny, int nz,
                            const int *ini, double *inif)
{
    SFX(setup)();
    int infx = -1;
    if (pos_x < nx / 2) infx = 1;
    for (int y = 0; y < ny; y++)
        for (int z = 0; z < nz; z++) {
            int me = ini[pos_x + y * nx + z * ny * nx];
            if (!me) continue;
            double *fp = inif + (size_t)(me - 1) * NV;
            int adj = ini[(pos_x + infx) + y * nx + z * ny * nx];
            if (adj) {
                const double *fa = inif + (size_t)(adj - 1) * NV;
                double rho_adj = SFX(density)(fa);
                double fator = rho / rho_adj;
                for (int i = 0; i < NV; i++) fp[i] = fa[i] * fator;
            } else {
                SFX(dist_eq_rest)(fp, rho);
            }
        }
}

/* ---- one driver step: [BCs] -> per site {collide; [density check]; push} ---
 * Mirrors the step loops of src/main_double-shear-layer.cpp:167-180,
 * src/main_stability-map_taylor-green.cpp:83-92 and
 * src/main_stability-map_obstacle.cpp:98-111, with the body-force
 * accelerations ax, ay, az passed to every collision (colisao_sitio's
 * acc_x, acc_y, acc_z, src/erlbm_common.h:68-81).  Returns 1 when some site's
 * post-collision density is non-finite or |rho| > dens_limit (checked only
 * when dens_limit > 0).  The caller swaps the buffers.                       */
int SFX(orc_collide_stream_acc)(double *inif, double *inif_new,
                                const int *ini_stream, int n_fluid, int col,
                                double tau, double tau_nh, double ax,
                                double ay, double az, double dens_limit,
                                int nthreads)
{
    SFX(setup)();
    int bad = 0;
#ifdef _OPENMP
    if (nthreads > 1) {
#pragma omp parallel for num_threads(nthreads) reduction(| : bad)
        for (int p = 0; p < n_fluid; p++) {
            double *f = inif + (size_t)p * NV;
            SFX(collide)(f, ax, ay, az, col, tau, tau_nh);
            if (dens_limit > 0) {
                double d = SFX(density)(f);
                if (fabs(d) > dens_limit || !isfinite(d)) bad |= 1;
            }
            SFX(propag_site)(inif, inif_new, ini_stream, p);
        }
        return bad;
    }
#endif
    (void)nthreads;
    for (int p = 0; p < n_fluid; p++) {
        double *f = inif + (size_t)p * NV;
        SFX(collide)(f, ax, ay, az, col, tau, tau_nh);
        if (dens_limit > 0) {
            double d = SFX(density)(f);
            if (fabs(d) > dens_limit || !isfinite(d)) bad = 1;
        }
        SFX(propag_site)(inif, inif_new, ini_stream, p);
    }
    return bad;
}

/* The same sweep with the drivers' zero body force. */
int SFX(orc_collide_stream)(double *inif, double *inif_new,
                            const int *ini_stream, int n_fluid, int col,
                            double tau, double tau_nh, double dens_limit,
                            int nthreads)
{
    return SFX(orc_collide_stream_acc)(inif, inif_new, ini_stream, n_fluid, col,
                                       tau, tau_nh, 0., 0., 0., dens_limit,
                                       nthreads);
}
