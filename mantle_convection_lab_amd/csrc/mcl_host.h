/*
 * mcl_host.h — the C host layer around libmcl_hip.so: the reference's input
 * format, mesh set-up and driver support, restated for a whole mesh held by one
 * device (no MPI decomposition).
 *
 *   mclh_read_input        Parsing.c (getpar: name=value, quotes, '#', last
 *                          occurrence wins) + the key table of Input.c:3-98,
 *                          plus the keys this build adds (see mcl_host.c)
 *   mclh_build_mesh        Tree.c (quadtree recipes of develop_mode, Mcode
 *                          levels, projection to coarser levels), Mesh.c
 *                          (sizes, node positions), Construct_ien_id_lm.c,
 *                          Boundary_condition.c, Connection.c (hanging nodes,
 *                          multigrid maps)
 *   mclh_initial_temperature  Initial_condition.c (+ a seeded perturbation)
 *   mclh_write_output      Output.c file layout
 *
 * The virtual ranks of an nprox x nproz input are laid side by side in one
 * mesh: each virtual rank applies its recipe to its own block (as each MPI rank
 * does), and the union of the blocks is the mesh.
 */
#ifndef MCL_HOST_H
#define MCL_HOST_H

#include "mcl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mclh_input {
    char data_dir[256];
    char solver[8];
    int max_level, min_level, npx, npz;
    double length, height;
    char top_v_bc[2], bot_v_bc[2], lef_v_bc[2], rig_v_bc[2];
    double top_v_bc_val[2], bot_v_bc_val[2], lef_v_bc_val[2], rig_v_bc_val[2];
    char top_T_bc, bot_T_bc, lef_T_bc, rig_T_bc;
    double top_T_bc_val, bot_T_bc_val, lef_T_bc_val, rig_T_bc_val;
    int T_ini;
    double T_mantle, T_mantle_perturb;
    int T_wave_number;
    double Ra;
    int use_hash;
    double A, refT, lm_jump, highest_vis, lowest_vis;
    int use_dense_matrix, mg_coarse_iteration, mg_finest_iteration;
    int max_timestep, save_spacing;
    double Q, theta;
    int AMR, mpi_KF_method, cg_precondition, Tcg_precondition;
    double accuracy;
    int auto_accuracy;
    double citcom_accu;
    int dev_mode, check_v_only, max_citcom_cycle, mg_cycle, max_cg_cycle;
    char old_T_file[1000];
    int old_T_level, old_T_timestep;
    int lg_mul, grain_damage, show_convg, zero_P;
    /* keys added by this build */
    int schur;                   /* schur_mode: 0 reference (multiply), 1 divide */
    int precision;               /* 0 fp64, 1 mixed */
    int warm_start;
    double inner_rtol;
    int exact_dots;              /* FP64 path: sequential dot products (bit-for-bit) */
    unsigned long long T_seed;   /* T_ini = 4: seeded perturbation */
    int device;
    int write_output;            /* 0: skip Output.c files */
} mclh_input;

typedef struct mclh_level {
    int L, lev, nel, nno, neq, neqT, gnox, gnoz;
    int *Mcode;      /* [nel] */
    int *ien;        /* [4][nel] */
    int *node;       /* [nno] */
    int *id;         /* [2][nno] */
    int *lm;         /* [2][4][nel] */
    int *heat_id;    /* [nno] */
    int *heat_lm;    /* [4][nel] */
    int *veq_node;   /* [neq] global node */
    int *veq_dof;    /* [neq] */
    int *heat_eq_node; /* [neqT] global node */
    signed char *v_bc; double *v_bc_val;   /* [2][nno] */
    signed char *T_bc; double *T_bc_val;   /* [nno] */
    double *X;       /* [2][nno] */
    double *area;    /* [nel] */
    int *p2n_nel;    /* [nno] */
    int *p2n_el;     /* [4][nno] */
    int nh;
    int *h_node, *h_ni, *h_cnode, *h_ci;   /* [nh], [nh], [2][nh], [2][nh] */
    /* multigrid maps to this level from the one below (solver = "mg"):
     * interp: eq u <- mean of interp_n[u] coarse eqs; proj: coarse eq
     * proj_l[k] <- eq proj_u[k] of this level */
    int *interp_n;   /* [neq] or NULL */
    int *interp_eq;  /* [4][neq] */
    int nproj;
    int *proj_l, *proj_u;
} mclh_level;

typedef struct mclh_mesh {
    int L, min_level, nlev;      /* levels min_level..L-1 stored as lev[0..nlev-1] */
    int gnox, gnoz;
    double dx[32], dz[32];
    mclh_level *lev;
} mclh_mesh;

/* 0 on success; positive on error with a message in err (if non-NULL). */
int mclh_read_input(const char *path, mclh_input *in, char *err, int errlen);
void mclh_default_input(mclh_input *in);
int mclh_set_key(mclh_input *in, const char *name, const char *value);
int mclh_build_mesh(const mclh_input *in, mclh_mesh **out, char *err, int errlen);
void mclh_free_mesh(mclh_mesh *m);
/* uniform mesh of leaves at `depth` (0 .. max_level-1) with the input's levels and ranks */
int mclh_build_mesh_depth(const mclh_input *in, int depth, mclh_mesh **out, char *err, int errlen);
/* Quadtree adaptation of the finest level of `old` (no counterpart in the reference,
 * whose adaptive_mesh_refine, Tree.c:95-99, is empty): leaves with |grad T| h >
 * refine_thr are split (up to max_depth), complete sibling families with all
 * |grad T| h < coarsen_thr are merged (down to min_depth), then leaves next to one
 * two or more levels deeper are split until the tree is 2:1 balanced.  T_old is the
 * finest level's nodal temperature; *T_new (malloc'd, free with mclh_free_array) gets
 * the new finest level's: old nodes keep their value, new nodes the bilinear
 * interpolant of the old leaf containing them. */
int mclh_adapt_mesh(const mclh_input *in, const mclh_mesh *old, const double *T_old, int min_depth, int max_depth,
                    double refine_thr, double coarsen_thr, mclh_mesh **out, double **T_new, int *nrefined,
                    int *ncoarsened, char *err, int errlen);
void mclh_free_array(void *p);
/* sizeof of mclh_input, mcl_control, mcl_mesh_view, mcl_stokes_stats, mcl_heat_stats,
 * mclh_level, mclh_mesh (which = 0..6), for checking binding layouts */
long mclh_sizeof(int which);
int mclh_initial_temperature(const mclh_input *in, const mclh_mesh *m, int k, double *T, char *err, int errlen);
void mclh_fill_view(const mclh_mesh *m, int k, mcl_mesh_view *v);
void mclh_fill_control(const mclh_input *in, mcl_control *c);
/* Seeded perturbation: u = top 53 bits of SplitMix64(seed + node) * 2^-53. */
double mclh_seeded_unit(unsigned long long seed, long long node);
/* Output.c:4-156 files for one step (coord, temperature, stress, viscosity, velocity). */
int mclh_write_output(const mclh_input *in, const mclh_level *l, int timestep, const double *T,
                      const double *stress, const double *vis, const double *V, char *err, int errlen);
/* Diagnostics this build adds (the reference has none): Vrms = sqrt(mean |u|^2)
 * over the box (bilinear quadrature), top Nusselt = -mean dT/dz at z = H
 * divided by the conductive flux (T_bot - T_top)/H. */
int mclh_diagnostics(const mclh_input *in, const mclh_level *l, const double *T, const double *V,
                     double *vrms, double *nu_top);

#ifdef __cplusplus
}
#endif
#endif
