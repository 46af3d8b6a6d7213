// Device-side data layout of one quadtree level and the parameter blocks of
// the persistent kernels.  Shared by the kernel translation units and the
// host C-ABI implementation (mcl_api.cpp).
#pragma once
#include "mcl_device.h"

// One quadtree level resident in HBM.  Arrays indexed [k][i] are stored
// slot-major (k * n + i) so that consecutive threads (consecutive i) read
// consecutive addresses.
struct LevelDev {
    int L, lev;
    int nel, nno, neq, neqT, gnox, gnoz;
    // level tables: element sizes per quadtree level and the element gradient
    // vector g[level][p] of get_element_G (Build_matrix.c:346-381)
    const double *lvl_dx, *lvl_dz;   // [MCL_MAX_LEVELS]
    const double *gtab;              // [MCL_MAX_LEVELS][8]
    // elements
    const int *ienl;                 // [4][nel] local node index of corner a
    const int *elev;                 // [nel] element quadtree level
    const int *lm;                   // [8][nel] velocity eq of (dof d, corner a) at (d*4+a)
    const int *gt_row;               // [8][nel] free velocity rows of the element, ascending, -1 padded
    const signed char *gt_p;         // [8][nel] element dof index 2a+d of gt_row
    // nodes
    const int *node;                 // [nno] global node id
    const int *p2n_cnt;              // [nno] number of elements at the node
    const int *p2n_el;               // [4][nno] those elements, ascending
    const int *p2n_a;                // [4][nno] corner index of the node in each
    const int *id;                   // [2][nno]
    const int *heat_id;              // [nno]
    const signed char *v_bc;         // [2][nno]
    const double *v_bc_val;          // [2][nno]
    const signed char *T_bc;         // [nno]
    const double *T_bc_val;          // [nno]
    const double *X;                 // [2][nno]
    // velocity rows (ELL, sorted columns)
    const int *veq_node;             // [neq] local node of the equation
    const int *veq_dof;              // [neq]
    int WK;
    const int *Kcol;                 // [WK][neq]
    const int *Kncol;                // [neq]
    const signed char *Kslot;        // [32][neq]: slot of (adjacent element k, element dof q)
    double *Kval;                    // [WK][neq]
    double *Kdiag, *F;               // [neq]
    // heat rows
    const int *heq_node;             // [neqT] local node of the heat equation
    int WT;
    const int *Tcol;                 // [WT][neqT]
    const int *Tncol;                // [neqT]
    const signed char *Tslot;        // [16][neqT]
    const int *Tdiag_slot;           // [neqT]
    // hanging nodes
    int nh, h_serial;
    const int *h_ni, *h_ci;          // [nh], [2][nh]
    // Galerkin constraint u = C u_free of the mixed path: hanging equation h
    // has hp_n[h] (1..4) free parents hp_i[k][h] with weights hp_w[k][h]
    // (hp_n = 0 for free equations; chains of hanging parents are expanded,
    // Dirichlet parents dropped); ch_* is the transpose (children of each
    // free equation, CSR)
    const int *hp_n, *hp_i;          // [neq], [4][neq]
    const double *hp_w;              // [4][neq]
    const int *ch_ptr, *ch_i;        // [neq + 1], [...]
    const double *ch_w;
};

#define MCL_PHASES 16
struct DevStats {
    int cycles, init_cycles, aborted, nan_flag;
    double dv, dp, dt;
    long long inner, spmv;
    long long phase[MCL_PHASES];   // 100 MHz ticks per solver phase (builds with MCL_PHASE_TIMERS)
    double probe_r1;               // |G^T d|^2 after the first inner solve (probe launches)
    long long t_begin, t_end;      // device wall clock (wall_clock64) at the FP64 Uzawa kernel's start and end
};

// One quadtree level for the reference's stationary inner solvers
// (Gauss-Seidel and the stand-alone multigrid, Solver.c:668-1316,1866-1928):
// the level's FP64 K in ELL form, the Gauss-Seidel wavefront schedule (rows
// grouped so that every row comes after each lower-numbered row it couples to
// and before each higher-numbered one: updating a group's rows in parallel
// reproduces the sequential sweep exactly), the transfer maps and the level's
// vectors.
struct GsLevelDev {
    int n, WK;
    const int *col, *ncol;   // [WK][n], [n]
    const double *val;       // [WK][n]
    int nsched;
    const int *sptr;         // [nsched + 1]
    const int *srows;        // [n] rows of group g: srows[sptr[g] .. sptr[g+1])
    // interp from the level below: row u <- mean of pn[u] entries pcol[a][u] (-1 skipped, still counted)
    const int *pn, *pcol;    // [n], [4][n]
    // injection into the level below: its row proj_l[k] <- this level's row proj_u[k]
    int nproj;
    const int *proj_l, *proj_u;
    double *err, *rhs, *resu;   // [n]: correction (G.err[lev]), right-hand side, residual
};

#define MCL_INNER_JC 0
#define MCL_INNER_GS 1
#define MCL_INNER_CG 2
#define MCL_INNER_MG 3

struct UzawaParams {
    LevelDev L;
    GridSync gs;
    RedBuf rb;
    DevStats *stats;
    double *d, *p, *u1, *Ahv, *Ahp, *BAB;
    double *r0, *r1, *r2, *z0, *z1, *s1, *s2;
    double *cg_r, *cg_s, *cg_d, *cg_q;
    double accuracy, citcom_accu;
    int max_citcom_cycle, max_cg_cycle, cg_precondition, check_v_only, schur;
    int exact_dots;          // dot products in the reference's sequential order
    // inner solver (MCL_INNER_*): Jacobi runs on the whole grid; Gauss-Seidel
    // and multigrid need a launch of one workgroup
    int solver, mg_cycle, mg_finest_iteration, mg_coarse_iteration;
    double *xt, *xt2;        // [n_v] previous iterates (Jacobi ping-pong, sweep and V-cycle differences)
    const GsLevelDev *gsl;   // [ngsl] levels, coarsest first (device memory)
    int ngsl;
};

struct HeatParams {
    LevelDev L;
    GridSync gs;
    RedBuf rb;
    DevStats *stats;
    const double *V;
    double *T, *tau, *hK, *hC, *hA, *hF, *hd, *hdd, *d0, *B;
    double *cg_r, *cg_s, *cg_d, *cg_q;
    double *model_time;
    double accuracy, theta, Q, Llast0, Llast1, diff_dt;
    int first_call, cg_precondition;
    int exact_dots;
};

// One multigrid level of the FP32 inner solver (mixed precision).  Level 0 is
// the coarsest; the finest level shares its rows with the Stokes velocity
// equations.  Levels with block != 0 are small enough to be processed by one
// workgroup (LDS-speed barriers); the others are processed by the whole grid.
// Rows (summed over levels) of the coarse multigrid levels that every
// workgroup of the mixed solve keeps in LDS (three float vectors per row).
#ifndef MCL_MG_LDS_ROWS
#define MCL_MG_LDS_ROWS 1536
#endif
// Matrix entries (ELL slots summed over levels) of those levels, also kept in
// LDS (FP32 values and 16-bit columns, 6 bytes an entry).
#ifndef MCL_MG_LDS_NNZ
#define MCL_MG_LDS_NNZ 16384
#endif

struct MgLevelDev {
    int n, WK;
    int block;               // coarse level run per workgroup from LDS
    int lds_off;             // offset of its vectors in the LDS block (floats)
    int moff;                // offset of its matrix in the LDS matrix block (entries)
    const unsigned short *col16;  // [WK][n] col as 16-bit indices (LDS levels)
    const int *col;          // [WK][n] (the level's FP64 ELL pattern)
    const int *ncol;         // [n]
    float *val;              // [WK][n] FP32 copy of the level's K
    float *dinv;             // [n] 1 / diag(K)
    const double *val64;     // [WK][n] the FP64 K of the level
    const double *diag64;    // [n]
    // prolongation from level-1: fine row u <- mean of pn[u] coarse rows pcol[k][u] (-1 = fixed)
    const int *pn;           // [n]
    const int *pcol;         // [4][n]
    // restriction to level-1 (= P^T), stored on this level's coarse partner:
    // coarse row c <- sum_k rval[k][c] * r_fine[rcol[k][c]]
    int rw;
    const int *rcol;         // [rw][n_coarse]
    const float *rval;       // [rw][n_coarse]
    float *x, *b, *r, *t;    // level vectors
    // fused V-cycle phases of a grid level, with per-workgroup row lists for
    // the workgroup count of the mixed solve (a workgroup owns rows O of
    // own_rows; W = [O | H1] adds the K columns of O, V = [W | H2] those of W;
    // F = the fine rows its coarse rows restrict from, T = [O | Th] adds the
    // K columns of F, U = [T | Uh] those of T):
    //  fuse_up:   x' = x + P x_c on W, then sweep pre + 1 on O from x' in LDS;
    //  fuse_down: the residual on F into LDS, then the coarse right-hand side;
    //  deep_up (pre = 2): x' on V, sweep 3 on W, sweep 4 on O;
    //  deep_down (pre = 2): sweep 2 on T from x_1 on U, the residual on F,
    //      the coarse right-hand side (above the LDS levels: T = W, U = V and
    //      the residual on O).
    int fuse_up, fuse_down, deep_up, deep_down;
    const int *hu_ptr, *hu_row;        // [nb + 1], H1
    const unsigned short *kl;          // [WK][n] K columns of O as local indices (into W, V)
    const int *h2_ptr, *h2_row;        // [nb + 1], H2
    const unsigned short *hm;          // H1 rows' K columns -> V, block hu_ptr[g] WK, slot-major
    const int *hd_ptr, *hd_row;        // [nb + 1], F
    const short *rl;                   // [rw][n] (coarse level) restriction columns as indices into F
    const int *th_ptr, *th_row;        // [nb + 1], Th
    const int *uh_ptr, *uh_row;        // [nb + 1], Uh
    const unsigned short *mt;          // T rows' K columns -> U, block (g R + th_ptr[g]) WK, slot-major
    const unsigned short *mf;          // F rows' K columns -> T, block hd_ptr[g] WK, slot-major
};

// floats of the LDS buffer of the fused V-cycle phases (per workgroup)
#ifndef MCL_HALO_LDS
#define MCL_HALO_LDS 4096
#endif

// Multigrid levels the mixed solve can hold (every workgroup keeps a copy of
// their descriptors in LDS).
#define MCL_MG_MAX_LEVELS 14

struct MixedParams {
    UzawaParams U;
    const MgLevelDev *mg;    // [nmg] (device memory)
    const float *lam;        // [nmg] Gershgorin bounds of D^-1 K per level
    int nmg;
    const double *cinv;      // dense inverse of the coarsest K [nc][nc]
    int nc;
    double *rr;              // FP64 residual of the inner CG
    double inner_rtol;       // relative residual tolerance of each inner solve
    double relax, relax_max; // relax > 0: Uzawa-loop inner solves to min(relax_max, max(inner_rtol, relax / dv)),
                             // then one correction solve of d for the final p
    int max_inner;
    int probe;               // 1: stop after the first inner solve and report |G^T d|^2
    // point-to-point waits in the V-cycle: neighbour workgroups of each
    // workgroup (CSR over workgroups, at most MCL_BLOCK each); use_nbr = 0
    // makes every wait a grid barrier
    const int *nbr_ptr, *nbr_idx;
    int use_nbr;
    int pre;                 // Jacobi sweeps per level on each leg of the V-cycle (>= 1)
    float omega;             // damped-Jacobi weight of the smoother
    // augmented Lagrangian (al_state != null): U.L's K is K + g G W G^T with
    // W = diag(al_w) and g = *al_state (0 or al_gamma), and the solve returns
    // p + g W G^T d; it then sets *al_state for the next solve: al_gamma if
    // its Uzawa count reached al_keep (g > 0) or al_start (g = 0), else 0
    double *al_state;
    const double *al_w;      // [nel] 1 / BAB of the plain K
    double al_gamma;
    int al_keep, al_start;
    // warm-start history (null: none): ws[0] solves since the history was
    // reset, ws[1], ws[2] the model time of the last two; the solve appends
    // itself after its last phase (model_time: the device's model time)
    double *ws;
    const double *model_time;
};

// Reduction slots in RedBuf (each slot holds MCL_MAX_WG partials).
#define MCL_RED_SLOTS 16

extern "C" hipError_t mcl_upload_shape(const ShapeTables *s);

__global__ void k_viscosity(int nno, const double *T, const double *X, double A, double refT, double lm_jump,
                            double *vis);
__global__ void k_buoy_terms(int nent, int nno, const int *ent_e, const signed char *ent_xi,
                             const signed char *ent_eta, const int *ienl, int nel, const double *T, double Ra,
                             double *terms);
__global__ void k_buoy_rows(int gnoz, int gnox, const int *ptr, const double *terms, double *avg);
__global__ void k_buoy_apply(int nno, int gnoz, const int *node, const double *T, double Ra, const double *avg,
                             double *buoy);
// One level of the all-levels assembly launch: its rows occupy the
// workgroups [blk0, blk0 + ceil(neq / MCL_BLOCK)) of the grid.
struct AsmLevel {
    LevelDev L;
    const double *vis, *buoy;
    int blk0;
};
__global__ void k_assemble_K(const AsmLevel *lv, int nlev);
__global__ void k_uzawa_fp64(UzawaParams P);
__global__ void k_v_to_nodes(LevelDev L, const double *d, double *V);
__global__ void k_stress_elem(LevelDev L, const double *vis, const double *V, double *es, double *area);
__global__ void k_stress_node(LevelDev L, const double *es, const double *area, double *stress);
__global__ void k_residuals(LevelDev L, RedBuf rb, const double *d, const double *p, int galerkin);
__global__ void k_collect3(RedBuf rb, int slot0, int nblocks, double *out);
__global__ void k_heat_fp64(HeatParams H);
__global__ void k_uzawa_mixed(MixedParams P);
__global__ void k_uzawa_probe(MixedParams P);
// One multigrid level of the all-levels FP32 preparation launch: its rows
// occupy the workgroups [blk0, blk0 + ceil(n / MCL_BLOCK)).
struct PrepLevel {
    int n, W;
    const double *val, *diag;
    float *val32, *dinv;
    unsigned *bound;
    int blk0;
};
__global__ void k_prepare_fp32(const PrepLevel *lv, int nlev);
// One multigrid level of the augmented-Lagrangian launches: K_al = K +
// gamma G W G^T (W = diag(w), w_e = 1 / sum_i G_ie^2 / K_ii, the BAB of the
// plain K), same ELL pattern, gamma read from the device.  Elements occupy workgroups [eblk0, ...) of
// k_al_weights, rows [rblk0, ...) of k_al_rows.
struct AlLevel {
    LevelDev L;
    double *val, *diag, *w;
    const double *gamma;     // device scalar (0: K_al = K)
    int eblk0, rblk0;
};
__global__ void k_al_weights(const AlLevel *lv, int nlev);
// warm start from the last two solutions (MixedParams::ws): d, p extrapolated
// linearly in model time from themselves and the previous solution h ([nv +
// np]: d then p) once `after` solves (at least 2) have followed a reset of
// the history, step ratio at most rmax; then h = d, p
__global__ void k_extrapolate(int nv, int np, double *d, double *p, double *h, const double *ws,
                              const double *model_time, int after, double rmax);
__global__ void k_al_rows(const AlLevel *lv, int nlev);
__global__ void k_coarse_inverse(int n, int W, const int *col, const int *ncol, const double *val, double *inv);
__global__ void k_inject(int nc, const int *c2f, const double *fine, double *coarse);
__global__ void k_zero_words(unsigned *p, long long n);
