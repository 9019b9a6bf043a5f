/*
 * erlbm.h -- C ABI of the MI355X-native E-RLBM collide-stream library.
 *
 * The reference (loesantos/ERLBM, SimBoltz) exposes per-site C++ functions
 * that its drivers call inside their own step loops.  A per-site call has no
 * meaning on a GPU, so this boundary sits at step-loop granularity: one
 * lattice handle owns its populations in HBM and a call runs the drivers'
 * whole step  [Zou-He inlet] -> [outlet rescale] -> collide (+check) ->
 * stream -> swap.  Each entry point names the reference interface it
 * replaces (path:line in the reference checkout).
 *
 * Conventions
 *   - every function returns ERLBM_OK (0) or an ERLBM_ERR_* code; the text of
 *     the last error is available from erlbm_last_error();
 *   - host arrays are borrowed for the duration of the call only;
 *   - a handle is bound to the HIP stream given at creation (NULL = default
 *     stream); distinct handles may be driven from distinct host threads;
 *   - populations cross the boundary in the reference's own layout: AoS,
 *     f[pto * q + i], pto = fluid index in x-fastest voxel order
 *     (numera_fluido, src/erlbm_common.h:89-99), velocity order of
 *     def_lattice_d2q9 / def_lattice_d3q19.
 */
#ifndef ERLBM_H
#define ERLBM_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ERLBM_OK = 0,
    ERLBM_ERR_ARG = 1,     /* invalid argument */
    ERLBM_ERR_HIP = 2,     /* HIP runtime error */
    ERLBM_ERR_STATE = 3,   /* call not valid in the handle's current state */
    ERLBM_ERR_NOMEM = 4,   /* device memory exhausted / over budget */
    ERLBM_ERR_UNSUPPORTED = 5
};

/* Collision operators: the COLISAO enum of src/erlbm_common.h:42. */
enum { ERLBM_COL_BGK = 1, ERLBM_COL_MRT = 2, ERLBM_COL_REG = 3, ERLBM_COL_EXT_REG = 4 };

/* Population storage precision.  The shifted formats store f_i - w_i rho0
 * and compute in FP64 (non-packed FP32 issues at the FP64 rate on MI355X,
 * and FP32 arithmetic biases the equilibrium); ERLBM_FP32_SHIFTED_F32
 * computes in FP32 with compensated moment sums.                          */
enum { ERLBM_FP64 = 0, ERLBM_FP32_SHIFTED = 1, ERLBM_FP16_SHIFTED = 2, ERLBM_FP32_SHIFTED_F32 = 4 };

/* Which populations a monitor reads. */
enum {
    ERLBM_STATE = 0,          /* the populations the reference's inif holds between steps */
    ERLBM_POST_COLLISION = 1  /* post-collision populations of the last step, before the
                                 push (what main_double-shear-layer.cpp:187-196 and
                                 main_taylor-green.cpp:176-213 measure) */
};

/* Stop reasons reported by erlbm_status. */
enum { ERLBM_STOP_DENSITY = 1, ERLBM_STOP_ENERGY = 2 };

typedef struct erlbm_lattice erlbm_lattice;

/* Text of the last error on this thread ("" if none). */
const char *erlbm_last_error(void);

/* Number of HIP devices visible (0 when no GPU). */
int erlbm_device_count(int *count);

/* ---- lattice lifecycle ---------------------------------------------------
 * Replaces geometria_periodica / numera_fluido / define_rede /
 * aloca_populacoes (src/erlbm_common.h:89-148) and def_dir_prop
 * (SimBoltz_Functions/LBM_functions.cpp:1967-2311): the stream map becomes a
 * per-voxel bounce mask (periodic wrap, halfway bounce-back off solids).
 * q in {9 (D2Q9, nz = 1), 19 (D3Q19)}; mask[x + y nx + z nx ny] = 1 fluid,
 * 0 solid; mask = NULL means all fluid.  stream: a hipStream_t or NULL.   */
int erlbm_lattice_create(int q, int nx, int ny, int nz, const int *mask, int precision,
                         void *stream, erlbm_lattice **out);
/* libera_populacoes, src/erlbm_common.h:150-157. */
int erlbm_lattice_destroy(erlbm_lattice *h);
/* geometry.fluid (Definitions.cpp:44). */
int erlbm_lattice_fluid_count(const erlbm_lattice *h, int *n_fluid);

/* ---- parameters and boundary conditions ----------------------------------
 * PARAMETERS.tau / tau_nh / rho_ini (Definitions.cpp:119-239) and the
 * collision choice of colisao_sitio (src/erlbm_common.h:68-81).            */
int erlbm_set_params(erlbm_lattice *h, double tau, double tau_nh, double rho0, int collision);
/* ZouHeD3Q19_vx_x (Boundary_conditions.cpp:350-442) / ZouHeD2Q9_vx_x
 * (:710-735) applied at column pos_x before every collision; pos_x < 0
 * disables it.                                                              */
int erlbm_set_zouhe_inlet(erlbm_lattice *h, int pos_x, double vx);
/* devnull_rho_x (Boundary_conditions.cpp:68-141) at column pos_x after the
 * inlet and before every collision; pos_x < 0 disables it.                */
int erlbm_set_outlet_rescale(erlbm_lattice *h, int pos_x, double rho);
/* Body force: the accelerations acc_x, acc_y, acc_z that colisao_sitio
 * (src/erlbm_common.h:68-81) passes to every collision operator
 * (coll_Ext_Reg, Collision_operators.cpp:1182; coll_Reg :859; coll_BGK :11;
 * coll_MRT :487).  F = a rho, the equilibrium at u + a/2 and the source term
 * of source() (SimBoltz_Functions/LBM_functions.cpp:5645-5716) added last
 * (Collision_operators.cpp:1196-1254); coll_MRT forms its own forcing
 * moments (:647-690).  D2Q9 ignores az, as the 2-D operators do.  All zero
 * (the default, what every reference driver passes) is the unforced
 * operator.  A forced lattice needs FP64 storage (ERLBM_ERR_UNSUPPORTED
 * otherwise) and advances one step per launch.                             */
int erlbm_set_body_force(erlbm_lattice *h, double ax, double ay, double az);
/* Collision arithmetic of the multi-step kernels (the reference chooses its
 * rounding at build time: its Makefile's FMA contraction or not).
 * ERLBM_ARITH_EXACT (default): the reference's operation order, every finite
 * population bitwise its -ffp-contract=off build.  ERLBM_ARITH_MOMENT: the
 * same BGK / RLBM / E-RLBM operator (coll_BGK, coll_Reg, coll_Ext_Reg,
 * Collision_operators.cpp:11-49, 859-967, 1181-1307) re-associated in moment
 * space, f*_i = W f_i + w_i (alpha + beta.c_i + c_i.Gamma.c_i), with less
 * than half the FP64 instructions; within 1e-10 maximum relative population
 * error of the reference after 1000 steps.  Applies to the multi-step
 * kernels of periodic lattices in FP64 storage and in FP32 shifted storage
 * (FP64 arithmetic); one-step launches, coll_MRT and lattices with
 * boundary columns, solids, stop tests or a body force keep the
 * reference's order.  FP16 storage and FP32 arithmetic:
 * ERLBM_ERR_UNSUPPORTED.                                                   */
enum { ERLBM_ARITH_EXACT = 0, ERLBM_ARITH_MOMENT = 1 };
int erlbm_set_arith(erlbm_lattice *h, int arith);
/* Per-step post-collision density test of main_stability-map_obstacle.cpp:
 * 104-106: a step whose some site has |rho| > limit or non-finite rho stops
 * the run after that step completes.  limit <= 0 disables it.             */
int erlbm_set_density_check(erlbm_lattice *h, double limit);

/* ---- transfer --------------------------------------------------------------
 * Host AoS in reference fluid order, n_fluid * q doubles.  Upload resets the
 * run (step counter, stop state).                                          */
int erlbm_upload_aos(erlbm_lattice *h, const double *inif);
/* which = ERLBM_STATE or ERLBM_POST_COLLISION (the latter needs >= 1 step). */
int erlbm_download_aos(erlbm_lattice *h, int which, double *inif);

/* ---- stepping ----------------------------------------------------------------
 * n_steps repetitions of the drivers' step (e.g.
 * main_double-shear-layer.cpp:162-229, main_stability-map_obstacle.cpp:
 * 88-118): [inlet] -> [outlet] -> collide (+density check) -> stream ->
 * swap.  Asynchronous on the handle's stream.  If energy_interval > 0, the
 * Taylor-Green map check of main_stability-map_taylor-green.cpp:71-81 runs
 * before every step whose global index is a positive multiple of
 * energy_interval: the run stops there if E > energy_factor * E0 or E is not
 * finite, E0 being the energy of the uploaded state.                       */
int erlbm_step(erlbm_lattice *h, int n_steps, int energy_interval, double energy_factor);
/* erlbm_step(h, n_steps, 0, 0) bracketed by HIP events recorded on the
 * handle's stream; *ms = elapsed milliseconds (synchronises).            */
int erlbm_step_timed(erlbm_lattice *h, int n_steps, float *ms);
/* Autotuner (no counterpart in the reference): times n_steps steps of every
 * launch shape this lattice's kernels offer (nontemporal-load policy, block
 * size, nodes per thread, 16-byte vector kernels) on a scratch copy of the
 * current state, keeps the fastest for the handle and reports it
 * (`choice`, may be NULL) with its time per step.  The run itself is not
 * advanced.  Geometries with solids or boundary columns have one kernel
 * and nothing to choose.  Needs uploaded populations; synchronises.      */
int erlbm_lattice_autotune(erlbm_lattice *h, int n_steps, char *choice, int choice_len, float *us_per_step);
/* Diagnostics (no counterpart in the reference): the multi-step shapes
 * (ERLBM_FUSE values, each once per XCD-run setting) that
 * erlbm_lattice_autotune would time for an all-fluid periodic lattice of
 * this size and precision on a GPU of n_cu compute units -- the same fit
 * tests as erlbm_step (whole strips or tiles, buffer offsets below 2^31
 * bytes) -- computed on the host without any device call.  *n = the count;
 * at most cap are written to shapes.                                      */
int erlbm_autotune_candidates(int q, int nx, int ny, int nz, int precision, int n_cu, int *shapes, int cap,
                              int *n);
/* Diagnostics (no counterpart in the reference): the multi-step kernel
 * shape (the ERLBM_FUSE value; 0 = one step per launch), the steps one of
 * its launches advances and its description (`desc`, may be NULL) that
 * erlbm_step(h, n, 0, 0) would use now for all but a first step after an
 * upload and the remainder.                                              */
int erlbm_step_shape(erlbm_lattice *h, int *shape, int *steps_per_launch, char *desc, int desc_len);
/* Global steps run so far and the stop reason (0 if none); synchronises. */
int erlbm_status(erlbm_lattice *h, long long *steps_done, int *stop_reason);
int erlbm_synchronize(erlbm_lattice *h);

/* ---- monitors ------------------------------------------------------------------
 * kinectic_energy (Other_functions.cpp:100-115): sum over fluid sites of
 * 0.5 |u|^2, FP64.  ordered != 0 sums in fluid order exactly as the
 * reference does (bitwise equal); ordered == 0 uses a fixed-shape
 * wavefront/block tree (reproducible, differs in the last bits).         */
int erlbm_kinetic_energy(erlbm_lattice *h, int which, int ordered, double *energy);
/* density (LBM_functions.cpp:5543-5552) over all fluid sites of `which`:
 * max |rho| and whether some rho is non-finite or above thresh.          */
int erlbm_density_monitor(erlbm_lattice *h, int which, double thresh, double *max_abs_rho,
                          int *violated);

/* ---- output fields and files --------------------------------------------------
 * The drivers' output, computed on the GPU from the populations `which`
 * names (ERLBM_STATE / ERLBM_POST_COLLISION) and copied to the host once.
 * Vector fields are 3 doubles per voxel, x fastest, as the VTK files hold
 * them.                                                                     */
/* calcula at every fluid site (LBM_functions.cpp:3517-3589) as rec_velocity
 * (:4218-4288) writes it: u at fluid voxels, 0 at solid ones.              */
int erlbm_velocity_field(erlbm_lattice *h, int which, double *vel);
/* vorticity (LBM_functions.cpp:3761-3872) as rec_vorticity (:4302-4346)
 * writes it: central differences with periodic wrap, a solid neighbour
 * counting as u = 0, 0 at solid voxels.                                    */
int erlbm_vorticity_field(erlbm_lattice *h, int which, double *vort);
/* calc_enstrophy (Other_functions.cpp:119-142): sum of 0.5 |w|^2 over every
 * voxel in the reference's order (x outermost), bitwise.                   */
int erlbm_enstrophy(erlbm_lattice *h, int which, double *enstrophy);
/* The files rec_velocity / rec_vorticity write (legacy ASCII VTK, six
 * significant digits), byte for byte.  Names: "vel_" + std::to_string(t)
 * + ".vtk" and "vor_" + step + ".vtk" in the reference.                    */
int erlbm_write_velocity_vtk(const char *path, int nx, int ny, int nz, const double *vel);
int erlbm_write_vorticity_vtk(const char *path, int nx, int ny, int nz, const double *vort);
/* read_velocity of tools/error_calculator.cpp:125-168: dimensions, and with
 * vel != NULL (cap >= 3 nx ny nz doubles) the vectors.                     */
int erlbm_read_velocity_vtk(const char *path, double *vel, long long cap, int *nx, int *ny, int *nz);
/* Each value replaced by what a write-then-read of the VTK text gives. */
int erlbm_vtk_round(double *v, long long n);
/* tools/error_calculator.cpp:35-114: relative L2 errors of a run's velocity
 * (nx x ny) against a reference run on an integer refinement of its grid,
 * sampled at coinciding nodes: E_vector (the paper's) and E_magnitude.     */
int erlbm_velocity_l2_error(const double *sim, int nx, int ny, const double *ref, int nx_ref, int ny_ref,
                            double *e_vector, double *e_magnitude);
/* The map files of src/main_stability-map_taylor-green.cpp:181-200 and
 * src/main_stability-map_obstacle.cpp:188-207 (map[i + j n_i], 1 unstable);
 * tau_nh is the driver's --tau-nh value, as it prints it.                  */
int erlbm_write_tg_map_vtk(const char *path, int n_ma, int n_nu, int collision, double tau_nh, const int *map);
int erlbm_write_obstacle_map_vtk(const char *path, int n_tau, int n_nu, double Ma, const int *map);

/* ---- host-side driver helpers ---------------------------------------------
 * Initial conditions, geometries and map axes of the reference drivers,
 * computed on the host with the reference's libm calls and operation order
 * (the GPU then starts from bitwise the reference's populations).          */
/* dist_eq (SimBoltz_Functions/LBM_functions.cpp:2944-3000): q doubles. */
int erlbm_dist_eq(int q, double vx, double vy, double vz, double rho, double *feq);
/* inicializa_taylor_green (src/main_taylor-green.cpp:57-100), L*L*q doubles. */
int erlbm_init_taylor_green(int q, int L, double u_0, double tau, int poisson, int neq, double *inif);
/* initial_double_shear_layer (src/main_double-shear-layer.cpp:48-66), D2Q9. */
int erlbm_init_double_shear(int L, double rho_ini, double *inif);
/* Taylor-Green map point (src/main_stability-map_taylor-green.cpp:49-61), D3Q19. */
int erlbm_init_tg_map(int L, double u_0, double rho_ini, double *inif);
/* sin / cos tables of that field (L entries each). */
int erlbm_tg_map_tables(int L, double *sin_tab, double *cos_tab);
/* geometria_obstaculo (src/main_stability-map_obstacle.cpp:42-70): 22*12 ints. */
int erlbm_obstacle_mask(int *mask);
/* Map axes (src/main_stability-map_taylor-green.cpp:134-164, :63;
 * src/main_stability-map_obstacle.cpp:147-172, :84).                       */
int erlbm_tg_map_point_params(int i, int j, int n_ma, int n_nu, int L, int collision, double tau_nh,
                              double *u_0, double *tau, double *tau_nh_out, int *n_steps);
int erlbm_obstacle_map_point_params(int i, int j, int n_tau, int n_nu, double Ma, double *u_imp, double *tau,
                                    double *tau_nh, int *n_steps);
/* Taylor-Green analytic-decay error (src/main_taylor-green.cpp:178-213) of
 * post-collision populations at `step`: out = {E_vector, E_magnitude,
 * (nu_eff - nu) / nu}.                                                     */
int erlbm_taylor_green_error(int q, int L, double u_0, double visc, long long step, const double *inif,
                             double *out);

/* ---- batched tiny lattices (stability maps, small requests) ------------------
 * Replaces the map drivers' `omp parallel for` over map points, each calling
 * roda_ponto (src/main_stability-map_taylor-green.cpp:41-104, 150-177;
 * src/main_stability-map_obstacle.cpp:75-123, 161-186): thousands of
 * independent D3Q19 lattices of at most 256 fluid sites (nz = 1) sharing one
 * geometry run in one persistent kernel, one lattice per workgroup resident
 * in LDS, jobs taken from a device-side queue.  Per job: initial state,
 * tau / tau_nh, collision, step budget and the map driver's stop test.    */
enum { ERLBM_INIT_REST = 0,          /* dist_eq(u = 0, rho0) everywhere (obstacle map :81-82) */
       ERLBM_INIT_TAYLOR_GREEN = 1   /* equilibrium Taylor-Green field of amplitude u0 (:49-61) */ };
/* precision codes for jobs: ERLBM_FP64, ERLBM_FP32_SHIFTED, or ERLBM_FP_AUTO =
 * FP32-shifted unless tau - 1/2 < 1e-3 (near-singular viscosity: FP64).   */
enum { ERLBM_FP_AUTO = 3 };

typedef struct {
    double tau, tau_nh;
    double u0;             /* Taylor-Green amplitude, and the Zou-He inlet velocity */
    double rho0;           /* initial / outlet density */
    double energy_factor;  /* energy test: stop if E > energy_factor * E0 or E non-finite */
    double dens_limit;     /* > 0: stop after a step with |rho| > dens_limit or non-finite rho */
    long long n_steps;     /* step budget */
    int check_every;       /* energy test before steps k = check_every, 2 check_every, ... (0: none) */
    int init;              /* ERLBM_INIT_* */
    int collision;         /* ERLBM_COL_BGK / REG / EXT_REG */
    int precision;         /* ERLBM_FP64 / ERLBM_FP32_SHIFTED / ERLBM_FP32_SHIFTED_F32 / ERLBM_FP_AUTO */
} erlbm_job;

typedef struct {
    int unstable;          /* 1 if a stop test fired (the map's 1 = instavel) */
    int steps;             /* steps run (the step index the run stopped at) */
    int done;              /* 1 when retired (stopped or budget used) */
    int precision;         /* precision the job ran in */
    double e0, energy;     /* initial and last measured kinetic energy (energy test only) */
    double margin;         /* closeness to the stop test: max E/E0 over the energy tests, or
                              max |rho| / dens_limit over all steps and sites */
} erlbm_job_result;

typedef struct erlbm_batch erlbm_batch;

/* Geometry shared by the batch: nx * ny voxels (nz = 1), mask as in
 * erlbm_lattice_create (NULL = all fluid, periodic), at most 256 fluid
 * sites.  inlet_x >= 0: ZouHeD3Q19_vx_x at that column with vx = job.u0;
 * outlet_x >= 0: devnull_rho_x at that column with rho = job.rho0.       */
int erlbm_batch_create(int nx, int ny, const int *mask, int inlet_x, int outlet_x, void *stream,
                       erlbm_batch **out);
int erlbm_batch_destroy(erlbm_batch *b);
/* Runs jobs[0..n_jobs) to retirement; results[k] for jobs[k].  Synchronous. */
int erlbm_batch_run(erlbm_batch *b, const erlbm_job *jobs, int n_jobs, erlbm_job_result *results);
/* No counterpart in the reference: erlbm_batch_run for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same jobs, same kernels, identical
 * results; the call itself allocates nothing: the device-side job records,
 * results, queue order and queue counters live in a static device array of
 * the library, about 1.9 MB, that each GPU the library runs on holds until
 * the process ends.  At most ERLBM_NOALLOC_MAX_JOBS jobs per call
 * (ERLBM_ERR_ARG beyond).  Runs in the one-launch mode whatever
 * erlbm_batch_set_chunk says (the results are the same).  Calls that overlap
 * on one GPU share that array and race: serialise them.  Synchronous.     */
enum { ERLBM_NOALLOC_MAX_JOBS = 16384 };
int erlbm_batch_run_noalloc(erlbm_batch *b, const erlbm_job *jobs, int n_jobs, erlbm_job_result *results);
/* Launch shape of the persistent kernels: workgroups per CU for the FP64 and
 * FP32 kernels (0 = occupancy maximum).  Chosen by the autotuner.       */
int erlbm_batch_set_launch(erlbm_batch *b, int wg_per_cu_f64, int wg_per_cu_f32);
/* chunk_steps = 0 (default): each (precision, collision) group of a run is
 * one persistent launch that takes jobs until every job has retired.
 * chunk_steps > 0: each group runs as a captured HIP graph [queue reset,
 * persistent kernel advancing every live job by <= chunk_steps steps with
 * its state saved in HBM, compaction of the live-job list], replayed until
 * no job is live.  Results are identical either way.                     */
int erlbm_batch_set_chunk(erlbm_batch *b, int chunk_steps);
/* Autotuner (no counterpart in the reference): runs the sample with every
 * job's budget cut to max_steps (so all jobs do equal work and the time
 * measures per-step throughput, not the longest job) under every
 * workgroups-per-CU value up to the occupancy limit, per precision class,
 * and keeps the fastest on the batch.  The chunking set by
 * erlbm_batch_set_chunk is left as it is.  Out pointers may be NULL;
 * *points_per_s is the sample's rate at the choice.                      */
int erlbm_batch_autotune(erlbm_batch *b, const erlbm_job *sample, int n_sample, long long max_steps,
                         int *wg_per_cu_f64, int *wg_per_cu_f32, double *points_per_s);

/* Job lists of the two published maps, with the drivers' axis expressions:
 * Taylor-Green map (src/main_stability-map_taylor-green.cpp:134-164; L = 16,
 * 100 T, energy test every 100 steps) and obstacle map
 * (src/main_stability-map_obstacle.cpp:147-172; 200 ny / U0 steps, density
 * test |rho| > 10 rho0).  Points (i, j) for i in [i0, i1), j in [j0, j1),
 * i fastest; jobs must hold (i1-i0)*(j1-j0) entries.                      */
int erlbm_tg_map_jobs(int n_ma, int n_nu, int collision, double tau_nh, int precision, int i0, int i1, int j0,
                      int j1, erlbm_job *jobs);
int erlbm_obstacle_map_jobs(int n_tau, int n_nu, double Ma, int precision, int i0, int i1, int j0, int j1,
                            erlbm_job *jobs);

/* ---- serving: continuous batching of uneven solves ----------------------------
 * No single counterpart in the reference, whose drivers run one solve per
 * process (scripts/run_*.sh): the four drivers' solves as requests to one
 * device-resident loop.  Map points (src/main_stability-map_*.cpp:41-123)
 * share persistent LDS-resident kernels in slot pools, one graph-replayed
 * chunk of `chunk_steps` steps per round; D2Q9 runs
 * (src/main_taylor-green.cpp, src/main_double-shear-layer.cpp) are plain
 * lattices on their own streams, one time slice per round.  Admission by
 * HBM budget (bytes of resident request state), first fit in arrival order;
 * each request stays resident until its stop test or budget retires it.    */
enum {
    ERLBM_REQ_TG_MAP_POINT = 1,        /* roda_ponto of the Taylor-Green map at (i, j) of n_i x n_j */
    ERLBM_REQ_OBSTACLE_MAP_POINT = 2,  /* roda_ponto of the obstacle map at (i, j) of n_i x n_j (Ma 0.1) */
    ERLBM_REQ_TAYLOR_GREEN = 3,        /* D2Q9 Taylor-Green, Fig. 3 mode (Re 50, U0 = 2/L) at L, to t = T */
    ERLBM_REQ_DOUBLE_SHEAR = 4         /* D2Q9 double shear layer (Re 80000) at L, to t* = t_final */
};

typedef struct {
    int kind;              /* ERLBM_REQ_* */
    int i, j, n_i, n_j;    /* map point and grid (map requests) */
    int L;                 /* lattice size (D2Q9 runs) */
    int collision;         /* ERLBM_COL_* (the obstacle map is E-RLBM) */
    int precision;         /* ERLBM_FP64 / ERLBM_FP32_SHIFTED / ERLBM_FP_AUTO / ... */
    double tau_nh;         /* E-RLBM tau_nh (Taylor-Green map, D2Q9 runs) */
    double t_final;        /* double shear layer: t* = step U0 / L at the end */
} erlbm_request;

typedef struct {
    long long id;
    int kind, status;      /* status: ERLBM_OK, or the error that rejected the request */
    int unstable;          /* map points: 1 = unstable; D2Q9 runs: 1 = stopped by a stop test */
    int precision;         /* storage precision it ran in */
    long long steps;       /* steps run */
    double value[3];       /* map: {margin, E0, last E}; Taylor-Green: {E_vector, E_magnitude,
                              (nu_eff - nu)/nu}; double shear: {kinetic energy, 0, 0} */
    double t_submit, t_admit, t_done;   /* seconds on the server clock */
} erlbm_response;

typedef struct erlbm_server erlbm_server;

/* stream: the pools' HIP stream (NULL: the server makes one); tiny_slots:
 * slots per pool (a pool reserves them from the budget on first use).    */
int erlbm_server_create(long long hbm_budget_bytes, int chunk_steps, int tiny_slots, void *stream,
                        erlbm_server **out);
int erlbm_server_destroy(erlbm_server *s);
/* Queue a request (never blocks); *id identifies its response. */
int erlbm_server_submit(erlbm_server *s, const erlbm_request *req, long long *id);
/* One scheduling round: admit what fits, run one chunk / slice of every
 * resident request, retire finished ones.  Synchronous.                  */
int erlbm_server_pump(erlbm_server *s, int *in_flight, int *queued);
/* Move up to max_out finished responses to out (completion order). */
int erlbm_server_poll(erlbm_server *s, erlbm_response *out, int max_out, int *n_out);
int erlbm_server_clock(erlbm_server *s, double *t);
int erlbm_server_memory(erlbm_server *s, long long *used, long long *budget);

/* ---- measurement -------------------------------------------------------------
 * No counterpart in the reference: a plain device-to-device copy of `bytes`
 * (nontemporal double2 loads/stores), timed with HIP events over `iters`
 * launches.  Calibrates the HBM traffic counters and gives the achievable
 * bandwidth beside the collide-stream kernel.                              */
int erlbm_hbm_copy_bench(long long bytes, int iters, float *ms_per_iter);

#ifdef __cplusplus
}
#endif

#endif /* ERLBM_H */
