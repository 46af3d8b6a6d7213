/*
 * mcl.h — C-ABI of libmcl_hip.so, the MI355X (gfx950) implementation of the
 * Mantle_Convection_Lab per-time-step hot path.
 *
 * The reference calls its hot path as K&R `void f()` functions that read and
 * write the global struct G (function_def.h:73,89,91,102,177).  A reference
 * build that uses this library keeps those signatures and forwards each call to
 * the mcl_* entry point with the same name:
 *
 *   buoyancy_force()               Buoyancy.c:3        -> mcl_buoyancy_force
 *   viscosity()                    Viscosity.c:3       -> mcl_viscosity
 *   velocity_pressure_KGF_matrix() Build_matrix.c:4    -> mcl_velocity_pressure_KGF_matrix
 *   solve_velocity_pressure()      Solver.c:3          -> mcl_solve_velocity_pressure
 *   update_temperature()           Update_temperature.c:3 -> mcl_update_temperature
 *
 * Mesh arrays are passed once per mesh (first step, or every step when AMR=1,
 * Main.c:33-52) through mcl_mesh_upload as borrowed host pointers in the
 * reference's own layouts; the library copies them and keeps no host pointer.
 * Fields stay resident in device memory between calls; mcl_set_field /
 * mcl_get_field move them between the caller's G arrays and the device.
 *
 * Every entry point returns 0 on success or a positive MCL_E_* code; the
 * message of the last failure is available from mcl_last_error().  Calls on one
 * context are not re-entrant; distinct contexts are independent.
 */
#ifndef MCL_H
#define MCL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCL_ABI_VERSION 1

enum mcl_error {
    MCL_OK = 0,
    MCL_E_ARG = 1,        /* invalid argument or inconsistent mesh */
    MCL_E_DEVICE = 2,     /* HIP runtime error */
    MCL_E_STATE = 3,      /* call out of order (e.g. solve before mesh upload) */
    MCL_E_NUMERIC = 4,    /* NaN/Inf detected in a solve */
    MCL_E_UNSUPPORTED = 5 /* option not available */
};

/* Inner solvers (the reference's `solver` key, Solver.c:82-89). */
enum mcl_solver { MCL_SOLVER_JC = 0, MCL_SOLVER_GS = 1, MCL_SOLVER_CG = 2, MCL_SOLVER_MG = 3 };

/* Schur-complement diagonal use in the Uzawa iteration. */
enum mcl_schur {
    MCL_SCHUR_REFERENCE = 0, /* z = BAB * r, as Solver.c:133-134 */
    MCL_SCHUR_DIVIDE = 1     /* z = r / BAB */
};

/* Arithmetic of the Stokes solve. */
enum mcl_precision {
    MCL_PREC_FP64 = 0,       /* FP64 throughout, reference operation order */
    MCL_PREC_MIXED = 1       /* FP64 inner CG preconditioned by an FP32 multigrid V-cycle, on the
                                augmented velocity block K + 3 G W G^T while Uzawa counts are high
                                (the returned pressure is the plain system's) */
};

/* Fields that can be moved between host and device (per quadtree level). */
enum mcl_field {
    MCL_FIELD_T = 0,          /* nno   temperature            G.T[lev] */
    MCL_FIELD_VIS = 1,        /* nno   viscosity              G.vis[lev] */
    MCL_FIELD_BUOYANCY = 2,   /* nno   buoyancy               G.buoyancy[lev] */
    MCL_FIELD_V = 3,          /* 2*nno nodal velocity [d][node] G.V[lev] */
    MCL_FIELD_D = 4,          /* neq   velocity solution      G.d[lev] */
    MCL_FIELD_P = 5,          /* nel   pressure               G.p[lev] */
    MCL_FIELD_F = 6,          /* neq   force vector           G.F[lev] */
    MCL_FIELD_KDIAG = 7,      /* neq   diagonal of K          G.Kdiag[lev] */
    MCL_FIELD_STRESS = 8,     /* nno   nodal stress           G.stress[lev] */
    MCL_FIELD_HEAT_D = 9,     /* neqT  heat solution          G.heat[lev].d */
    MCL_FIELD_HEAT_DDOT = 10, /* neqT  heat rate              G.heat[lev].ddot */
    MCL_FIELD_HEAT_F = 11,    /* neqT  heat force             G.heat[lev].F */
    MCL_FIELD_TAU = 12        /* 2*nel SUPG tau [d][e]        G.heat[lev].tau */
};

/* Assembled matrices that can be read back (row-slot form, Mesh.c:455-511). */
enum mcl_matrix {
    MCL_MATRIX_K = 0,         /* velocity stiffness, neq rows   G.Krow_* */
    MCL_MATRIX_G = 1,         /* gradient, neq rows, element columns  G.Grow_* */
    MCL_MATRIX_HEAT_K = 2,    /* heat advection-diffusion, neqT rows */
    MCL_MATRIX_HEAT_C = 3     /* heat (SUPG) capacity, neqT rows */
};

/* One quadtree level of the reference mesh, in the reference's layouts
 * (global_variables.h:315-422), flattened row-major:
 *   Mcode[e]                 (morton << 4) | element level      G.Mcode[lev]
 *   ien[a*nel+e]             global node numbers, a = 0..3       G.ien[lev][a]
 *   node[i]                  sorted global node ids              G.node[lev]
 *   id[d*nno+i]              velocity equation, -1 = fixed       G.id[lev][d]
 *   lm[(d*4+a)*nel+e]        id at element corner                G.lm[lev][d][a]
 *   heat_id[i], heat_lm[a*nel+e]                                 G.heat[lev].id/lm
 *   v_bc[d*nno+i], v_bc_val  'V' fixed, 'S' free slip, 'N'       G.bc.v_bc[lev][d]
 *   T_bc[i], T_bc_val        'T' fixed, 'F' flux, 'N'            G.bc.T_bc[lev]
 *   X[d*nno+i]               node coordinates                    G.X[lev][d]
 *   dx[l], dz[l]             element sizes of quadtree level l   G.dx, G.dz
 *   h_ni[k], h_ci[c*nh+k]    hanging node k and its two parents  G.hangle[lev].ni/ci
 * MG transfer maps (only for MCL_SOLVER_MG, else NULL), from Connection.c:708-1500:
 *   interp_n[u], interp_eq[k*neq+u]  eq u of this level <- mean of interp_n[u]
 *                            eqs of the level below (-1 = fixed)  G.interp[lev]
 *   proj_l[k], proj_u[k]     eq proj_l[k] of the level below <- eq proj_u[k] of
 *                            this level (injection)               G.project[lev-1]
 */
typedef struct mcl_mesh_view {
    int L;            /* G.max_level */
    int lev;          /* level index of this view */
    int nel, nno, neq, neqT;
    int gnox, gnoz;
    const int *Mcode;
    const int *ien;
    const int *node;
    const int *id;
    const int *lm;
    const int *heat_id;
    const int *heat_lm;
    const signed char *v_bc;
    const double *v_bc_val;
    const signed char *T_bc;
    const double *T_bc_val;
    const double *X;
    const double *dx;
    const double *dz;
    int nh;
    const int *h_ni;
    const int *h_ci;
    const int *interp_n;
    const int *interp_eq;
    int nproj;
    const int *proj_l;
    const int *proj_u;
} mcl_mesh_view;

/* The reference's run-time controls used on the hot path (Input.c:3-98), plus
 * the options this library adds (schur, precision, warm_start). */
typedef struct mcl_control {
    double Ra;                 /* Rayleigh */
    double A, refT, lm_jump;   /* viscosity law, Viscosity.c:13-16 */
    double accuracy;           /* inner-solver tolerance (and SUPG small-alpha threshold) */
    double citcom_accu;        /* Uzawa dv/dp tolerance, citcom_solver_accu */
    double theta;              /* heat time weight */
    double Q;                  /* internal heating */
    int max_citcom_cycle;
    int max_cg_cycle;
    int cg_precondition;
    int check_v_only;
    int solver;                /* enum mcl_solver */
    int auto_accuracy;
    int zero_P;
    int AMR;
    int mg_cycle, mg_finest_iteration, mg_coarse_iteration;
    int schur;                 /* enum mcl_schur */
    int precision;             /* enum mcl_precision */
    int warm_start;            /* 1: keep d and p between steps (overrides zero_P); MCL_PREC_MIXED
                                  extrapolates them in model time from the last two solutions */
    double inner_rtol;         /* MCL_PREC_MIXED: relative tolerance of each inner solve (0: 3e-7) */
    int exact_dots;            /* FP64 path: 1 = dot products summed in the reference's sequential
                                  order (bit-for-bit with the reference; slower), 0 = compensated
                                  parallel sums */
} mcl_control;

typedef struct mcl_stokes_stats {
    int uzawa_cycles;
    int converged;             /* dv, dp below citcom_accu before the cycle cap */
    double dv, dp;
    long long inner_iterations;
    long long spmv;
    double accuracy_used;
    double res_momentum;       /* |F - K d - G p| / |F| after the solve (the mixed path on meshes
                                  with hanging nodes: |C^T (F - K d - G p)| / |C^T F| in the
                                  constrained space it solves in) */
    double res_continuity;     /* |G^T d| / |F| after the solve */
    double device_ms;          /* device time of the persistent Uzawa kernel (HIP events on its stream) */
} mcl_stokes_stats;

typedef struct mcl_heat_stats {
    int cg_cycles;
    int init_cg_cycles;
    double dt;
    double model_time;
    double device_ms;
} mcl_heat_stats;

typedef struct mcl_ctx mcl_ctx;

/* Library / context lifetime. */
int mcl_abi_version(void);
int mcl_create(int device, mcl_ctx **out);
void mcl_destroy(mcl_ctx *ctx);
const char *mcl_last_error(const mcl_ctx *ctx);

/* Mesh (Main.c:33-52 setup products).  levels[0..nlev-1] = G.min_level..G.max_level-1. */
int mcl_mesh_upload(mcl_ctx *ctx, const mcl_mesh_view *levels, int nlev);
int mcl_set_control(mcl_ctx *ctx, const mcl_control *control);

/* Host <-> device transfer of fields (count elements of double). */
int mcl_set_field(mcl_ctx *ctx, int lev, int field, const double *src, size_t count);
int mcl_get_field(mcl_ctx *ctx, int lev, int field, double *dst, size_t count);
/* Assembled matrix in row-slot form: ncol[row], col/val[row*width+slot]. */
int mcl_get_matrix(mcl_ctx *ctx, int lev, int matrix, int width, int *ncol, int *col, double *val);
int mcl_set_model_time(mcl_ctx *ctx, double model_time);
int mcl_get_model_time(mcl_ctx *ctx, double *model_time);

/* The per-step hot path, in the reference's call order (Main.c:58-64). */
int mcl_buoyancy_force(mcl_ctx *ctx);                 /* Buoyancy.c:3-22 */
int mcl_viscosity(mcl_ctx *ctx);                      /* Viscosity.c:3-24 */
int mcl_velocity_pressure_KGF_matrix(mcl_ctx *ctx);   /* Build_matrix.c:4-40 */
int mcl_solve_velocity_pressure(mcl_ctx *ctx, mcl_stokes_stats *stats); /* Solver.c:3-12 */
int mcl_update_temperature(mcl_ctx *ctx, mcl_heat_stats *stats);        /* Update_temperature.c:3-133 */

/* Synchronise the context's stream. */
int mcl_synchronize(mcl_ctx *ctx);

/* Most workgroups any persistent kernel of this context may launch (0: as
 * many as fit on the device, the default).  For contexts that share a device
 * with other work at the same time; results depend on the workgroup count
 * only through the order of the parallel sums. */
int mcl_set_launch_cap(mcl_ctx *ctx, int max_workgroups);

/* ---------------------------------------------------------------------------
 * Batches (no counterpart in the reference, which advances one instance per
 * job): many independent instances, one context each on the same device,
 * advanced together.  mcl_batch_create captures each instance's time step
 * (buoyancy, viscosity, assembly, the Stokes solve with nodal velocity and
 * stress, the heat step) into a HIP graph; mcl_batch_step replays every
 * instance's graph on one of `lanes` concurrent streams and waits for all.
 * Every instance must have taken its first step through the ordinary entry
 * points; auto_accuracy is not available.  While the batch exists, each of its
 * contexts launches at most (device capacity / lanes) workgroups per
 * persistent kernel (or its own smaller launch cap), so the persistent kernels
 * of all lanes fit on the device at once. */
typedef struct mcl_batch mcl_batch;
int mcl_batch_create(mcl_ctx *const *ctxs, int n, int lanes, mcl_batch **out);
void mcl_batch_destroy(mcl_batch *batch);
const char *mcl_batch_last_error(const mcl_batch *batch);
/* one time step of every instance with skip[i] == 0 (skip may be NULL); per
 * instance the Stokes and heat statistics and the step's device time (ms,
 * events around its graph on its lane); any output array may be NULL */
int mcl_batch_step(mcl_batch *batch, const int *skip, mcl_stokes_stats *stokes, mcl_heat_stats *heat,
                   double *step_ms);

/* mcl_solve_velocity_pressure without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * FP64 Uzawa solve (precision MCL_PREC_FP64; MCL_E_UNSUPPORTED otherwise) with
 * the same kernels and results as mcl_solve_velocity_pressure, but its work
 * space -- grid-barrier words, reduction partials, statistics, residual norms
 * and the 6 n_v + 9 n_p work vectors -- lives in a static device array of
 * 48 MiB per GPU (held until the process ends), located at mcl_create; no
 * hipMalloc, hipFree or event is created, buffers are cleared by a kernel and
 * stats->device_ms comes from the kernel's own clock.  Meshes whose work space
 * exceeds the array (above about 512 x 512 elements) are rejected with
 * MCL_E_ARG.  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU race. */
int mcl_solve_velocity_pressure_noalloc(mcl_ctx *ctx, mcl_stokes_stats *stats);

/* ---------------------------------------------------------------------------
 * Learned initial guess (no counterpart in the reference: new in this library).
 * A U-Net maps the temperature anomaly (T minus its row mean, scaled to unit
 * RMS s) and log10 viscosity at the element centres of a uniform H x W element
 * mesh to (u_x, u_z, p - mean p) / (Ra s).  It is trained on the device on
 * snapshots of the solver's own fields; mcl_stokes_guess proposes (d, p) for
 * the next solve and keeps whichever of (network guess, current fields) has
 * the smaller initial Schur residual |G^T K^-1 (F - G p)|.
 */
typedef struct mcl_net mcl_net;

int mcl_net_create(int H, int W, int batch, unsigned long long seed, mcl_net **out);
void mcl_net_destroy(mcl_net *net);
const char *mcl_net_last_error(const mcl_net *net);
long long mcl_net_num_params(const mcl_net *net);
/* flat FP32 parameters: per convolution w[Cout][3][3][Cin], b[Cout]; head w[3][32], b[3] */
int mcl_net_get_params(mcl_net *net, float *host);
int mcl_net_set_params(mcl_net *net, const float *host);
int mcl_net_get_grads(mcl_net *net, float *host);
/* output scale of channel 0..2 (RMS of that training target, set by the first mcl_net_train) */
double mcl_net_get_scale(const mcl_net *net, int channel);
int mcl_net_set_scale(mcl_net *net, int channel, double scale);
/* host arrays: x [B][H][W][2], y and target [B][H][W][3] */
int mcl_net_forward(mcl_net *net, const float *x_host, float *y_host);
/* one Adam step on the mean squared error (lr = 0: gradients only); loss is the batch MSE */
int mcl_net_train_step(mcl_net *net, const float *x_host, const float *t_host, float lr, float *loss);
/* capture the context's current T, viscosity, velocity and pressure as a snapshot */
int mcl_net_capture(mcl_ctx *ctx, mcl_net *net);
long long mcl_net_dataset_size(const mcl_net *net);
/* epochs of shuffled mini-batches over the captured snapshots, each sample mirrored in x with
 * probability 1/2; losses[epoch] = mean batch loss */
int mcl_net_train(mcl_net *net, int epochs, float lr, unsigned long long seed, float *losses);
/* propose the network's (d, p) for the next solve of ctx (precision=1, warm_start=1, after
 * mcl_velocity_pressure_KGF_matrix): for each candidate one inner solve gives
 * d = K^-1 (F - G p) and r = |G^T d|; the candidate with the smaller r is left in the
 * context's d, p (*accepted = 1: the network's); both r are returned if non-NULL */
int mcl_stokes_guess(mcl_ctx *ctx, mcl_net *net, int *accepted, double *res_net, double *res_prev);

#ifdef __cplusplus
}
#endif
#endif
