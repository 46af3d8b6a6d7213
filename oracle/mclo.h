/*
 * mclo — CPU FP64 restatement of the Mantle_Convection_Lab per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP library
 * (libmcl_hip.so): only tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it.  No product code path calls it.
 *
 * Every function follows the reference's arithmetic in order (same loop nests,
 * same summation order, no FMA contraction) so that, built with the reference's
 * own flags, results can be compared bit for bit with the reference's golden
 * dumps.  Each function cites the reference file:line it restates.
 *
 * Mesh arrays use the reference's layouts for one quadtree level `lev`
 * (global_variables.h:315-422), flattened:
 *   ien[a*nel+e]           global node numbers of element e, a = 0..3
 *   node[i]                sorted global node ids (local index = position)
 *   id[d*nno+i]            velocity equation of dof d at local node i, -1 = fixed
 *   lm[(d*4+a)*nel+e]      = id[d][local(ien[a][e])]
 *   heat_id[i], heat_lm[a*nel+e]
 *   v_bc[d*nno+i], v_bc_val[d*nno+i], T_bc[i], T_bc_val[i]  ('V','S','N' / 'T','F','N')
 *   X[d*nno+i]             node coordinates
 *   dx[l], dz[l]           element sizes per quadtree level l = 0..L-1
 *   Mcode[e]               (morton location << 4) | element level
 *   h_ni[k], h_ci[c*nh+k]  hanging node k: local index and two parent local indices
 * Row-slot matrices (Mesh.c:455-511): ncol[row], col[row*W+slot], val[row*W+slot]
 * with W = MCLO_VW (36) for velocity rows, MCLO_TW (18) for heat rows.
 */
#ifndef MCLO_H
#define MCLO_H

#ifdef __cplusplus
extern "C" {
#endif

#define MCLO_VW 36
#define MCLO_TW 18

enum { MCLO_SOLVER_JC = 0, MCLO_SOLVER_GS = 1, MCLO_SOLVER_CG = 2, MCLO_SOLVER_MG = 3 };

typedef struct mclo_mesh {
    int L;          /* max_level: quadtree levels 0..L-1 */
    int lev;        /* the level these arrays describe (L-1 = finest) */
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
} mclo_mesh;

/* Geometric multigrid transfer maps between level l (coarse) and l+1 (fine),
 * the local parts of Connection.c:708-1500 at one rank:
 *   interp_n[u], interp_eq[k*neq_fine+u]   fine eq u <- mean of interp_n[u] coarse eqs (-1 = fixed)
 *   proj_l[k], proj_u[k]                   coarse eq proj_l[k] <- fine eq proj_u[k] (injection),
 *                                          held by the coarse level */
typedef struct mclo_mg_level {
    const mclo_mesh *mesh;
    int *Kncol; int *Kcol; double *Kval; double *Kdiag;    /* assembled per level */
    const int *interp_n; const int *interp_eq;             /* maps to this level from the one below */
    int nproj; const int *proj_l; const int *proj_u;       /* injection into this level from the one above */
} mclo_mg_level;

typedef struct mclo_ctrl {
    double Ra;
    double A, refT, lm_jump;      /* viscosity law, Viscosity.c:13-16 */
    double accuracy;              /* inner tolerance (and SUPG small-alpha threshold) */
    double citcom_accu;           /* Uzawa dv/dp tolerance */
    double theta;                 /* time-integration weight */
    double Q;                     /* internal heating */
    int max_citcom_cycle;
    int max_cg_cycle;
    int cg_precondition;          /* Jacobi-PCG (1) or plain CG (0), velocity and heat */
    int check_v_only;
    int solver;                   /* MCLO_SOLVER_* */
    int auto_accuracy;
    int schur_divide;             /* 0: z = BAB*r as the reference; 1: z = r/BAB ("plain-fixed") */
    int mg_cycle, mg_finest_iteration, mg_coarse_iteration;
} mclo_ctrl;

typedef struct mclo_stokes_stats {
    int uzawa_cycles;
    double dv, dp;
    long long inner_iterations;   /* total inner-solver iterations */
    long long spmv;               /* K products */
    double accuracy_used;
} mclo_stokes_stats;

typedef struct mclo_heat_stats {
    int cg_cycles;
    int init_cg_cycles;           /* first call only (initial_T_dot) */
    double dt;
} mclo_heat_stats;

/* Buoyancy.c:3-22 + Parallel.c:354-454 */
int mclo_buoyancy(const mclo_mesh *m, const double *T, double Ra, double *buoy);
/* Viscosity.c:3-24 (grain-damage branch excluded) */
int mclo_viscosity(const mclo_mesh *m, const double *T, double A, double refT, double lm_jump, double *vis);
/* Build_matrix.c:4-381, 720-744: K (36-slot rows, sorted), G, F, Kdiag */
int mclo_assemble_stokes(const mclo_mesh *m, const double *vis, const double *buoy,
                         int *Kncol, int *Kcol, double *Kval, double *Kdiag,
                         int *Gncol, int *Gcol, double *Gval, double *F);
/* Solver.c:14-249 (citcom_solver) with the inner solver chosen by ctrl->solver.
 * d and p hold the initial guess (the reference zeroes both, Solver.c:251-278).
 * For MCLO_SOLVER_MG, mg/nlev describe levels min..max (mg[nlev-1] = finest). */
int mclo_stokes_solve(const mclo_mesh *m, const mclo_ctrl *c,
                      const int *Kncol, const int *Kcol, const double *Kval, const double *Kdiag,
                      const int *Gncol, const int *Gcol, const double *Gval, const double *F,
                      double *d, double *p, mclo_mg_level *mg, int nlev, mclo_stokes_stats *st);
/* Solver.c:1319-1343 */
void mclo_v_to_nodes(const mclo_mesh *m, const double *d, double *V);
/* Grain_size_evolution.c:154-213 + Solver.c:1345-1366 */
void mclo_compute_stress(const mclo_mesh *m, const double *vis, const double *V, double *stress);
/* Update_temperature.c:135-177 (without the 0.75 factor applied at :37) */
double mclo_advection_dt(const mclo_mesh *m, const double *V);
/* Update_temperature.c:241-291 */
void mclo_artificial_diffusivity(const mclo_mesh *m, const double *V, double accuracy, double *tau);
/* Update_temperature.c:293-594, 788-824: heat K, C (18-slot sorted rows) and F */
int mclo_assemble_heat(const mclo_mesh *m, const mclo_ctrl *c, const double *V, const double *tau,
                       int *Kncol, int *Kcol, double *Kval, int *Cncol, int *Ccol, double *Cval, double *F);
/* Update_temperature.c:3-133 with the first-call initial_T_dot (:179-239).
 * heat_d/heat_ddot are the persistent heat vectors (neqT); on first_call they are
 * initialised from T.  T (nno) is overwritten with the new temperature.
 * If the optional outputs are non-NULL the step's tau (2*nel), heat K/C rows and F
 * are returned as well. */
int mclo_update_temperature(const mclo_mesh *m, const mclo_ctrl *c, const double *V,
                            double *T, double *heat_d, double *heat_ddot, int first_call,
                            double *model_time, mclo_heat_stats *st,
                            double *tau_out, int *Kncol, int *Kcol, double *Kval,
                            int *Cncol, int *Ccol, double *Cval, double *F_out);

/* Row-slot sparse product y = A x (Solver.c:360-386 / Update_temperature.c:935-962). */
void mclo_rows_prod(int n, int W, const int *ncol, const int *col, const double *val,
                    const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
