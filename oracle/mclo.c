/*
 * mclo — CPU FP64 restatement of the Mantle_Convection_Lab per-step hot path.
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP library); see mclo.h.
 *
 * Parity status of each function (how it is pinned) is listed next to it:
 *   "golden"  compared with dumps of the reference built from its own sources;
 *   "closed"  compared with closed forms / invariants in the tests.
 */
#include "mclo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NEN 4
#define NSD 2
#define NEE 8

/* ---------------------------------------------------------------- shape tables
 * Shape_function.c:17-121: bilinear N on the reference square with node order
 * (-1,-1), (-1,1), (1,1), (1,-1); Gauss points = nodes / sqrt(3), same order. */
typedef struct {
    double N[NEN][NEN];       /* N[a][g] at Gauss point g */
    double dxi[NEN][NEN];     /* dN_a/dxi at Gauss g */
    double deta[NEN][NEN];    /* dN_a/deta at Gauss g */
    double Nc[NEN];           /* centre values */
    double dxic[NEN], detac[NEN];
} shape_t;

static void shape_init(shape_t *s)
{
    const double pts[NEN][2] = {{-1.0, -1.0}, {-1.0, 1.0}, {1.0, 1.0}, {1.0, -1.0}};
    double beta = 1.0 / sqrt(3.0);
    int g;
    for (g = 0; g < NEN; g++) {
        double xi = pts[g][0] * beta, eta = pts[g][1] * beta;
        s->N[0][g] = 0.25 * (1 - xi) * (1 - eta);
        s->N[1][g] = 0.25 * (1 - xi) * (1 + eta);
        s->N[2][g] = 0.25 * (1 + xi) * (1 + eta);
        s->N[3][g] = 0.25 * (1 + xi) * (1 - eta);
        s->dxi[0][g] = 0.25 * (-1.0) * (1 - eta);
        s->dxi[1][g] = 0.25 * (-1.0) * (1 + eta);
        s->dxi[2][g] = 0.25 * (1.0) * (1 + eta);
        s->dxi[3][g] = 0.25 * (1.0) * (1 - eta);
        s->deta[0][g] = 0.25 * (1.0 - xi) * (-1.0);
        s->deta[1][g] = 0.25 * (1.0 - xi) * (1.0);
        s->deta[2][g] = 0.25 * (1.0 + xi) * (1.0);
        s->deta[3][g] = 0.25 * (1.0 + xi) * (-1.0);
    }
    for (g = 0; g < NEN; g++) s->Nc[g] = 0.25;
    s->dxic[0] = 0.25 * (-1.0) * (1 - 0.0);
    s->dxic[1] = 0.25 * (-1.0) * (1 + 0.0);
    s->dxic[2] = 0.25 * (1.0) * (1 + 0.0);
    s->dxic[3] = 0.25 * (1.0) * (1 - 0.0);
    s->detac[0] = 0.25 * (1.0 - 0.0) * (-1.0);
    s->detac[1] = 0.25 * (1.0 - 0.0) * (1.0);
    s->detac[2] = 0.25 * (1.0 + 0.0) * (1.0);
    s->detac[3] = 0.25 * (1.0 + 0.0) * (-1.0);
}

/* Shape_function.c:180-197 (Na) */
static double shape_at(int a, double xi, double eta)
{
    switch (a) {
    case 0: return 0.25 * (1.0 - xi) * (1.0 - eta);
    case 1: return 0.25 * (1.0 - xi) * (1.0 + eta);
    case 2: return 0.25 * (1.0 + xi) * (1.0 + eta);
    default: return 0.25 * (1.0 + xi) * (1.0 - eta);
    }
}

static int elem_level(int mcode) { return mcode & 15; }

/* Array_work.c:106-129 (node_index, bsearch branch) */
static int local_index(const mclo_mesh *m, int gnode)
{
    int lo = 0, hi = m->nno - 1;
    if (gnode == -1) return -1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (m->node[mid] == gnode) return mid;
        if (m->node[mid] < gnode) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

static int *local_ien(const mclo_mesh *m)
{
    int *li = malloc((size_t)NEN * m->nel * sizeof(int));
    int a, e;
    for (a = 0; a < NEN; a++)
        for (e = 0; e < m->nel; e++)
            li[a * m->nel + e] = local_index(m, m->ien[a * m->nel + e]);
    return li;
}

/* Element Jacobian terms, Shape_function.c:123-167: dx_dxi = dx[level]/2,
 * dz_deta = dz[level]/2, dx_deta = dz_dxi = 0, J = dx_dxi*dz_deta - 0*0. */
static void elem_geom(const mclo_mesh *m, int e, double *dxdxi, double *dzdeta, double *jac)
{
    int level = elem_level(m->Mcode[e]);
    *dxdxi = 0.5 * m->dx[level];
    *dzdeta = 0.5 * m->dz[level];
    *jac = (*dxdxi) * (*dzdeta) - 0.0 * 0.0;
}

/* ------------------------------------------------------------ row-slot helpers */
static int slot_find(const int *col, int n, int c)
{
    int j;
    for (j = 0; j < n; j++)
        if (col[j] == c) return j;
    return -1;
}

/* Insert or accumulate value v at column c in a row (Build_matrix.c:75-88). */
static void slot_add(int *ncol, int *col, double *val, int c, double v)
{
    int k = slot_find(col, *ncol, c);
    if (k >= 0) {
        val[k] += v;
    } else {
        col[*ncol] = c;
        val[*ncol] += v;
        (*ncol)++;
    }
}

/* Bubble sort of one row by column (Build_matrix.c:720-738): stable, so the
 * result is the row ordered by ascending column. */
static void slot_sort(int n, int *col, double *val)
{
    int swapped = 1, j;
    while (swapped) {
        swapped = 0;
        for (j = 0; j < n - 1; j++) {
            if (col[j] > col[j + 1]) {
                int tc = col[j]; double tv = val[j];
                col[j] = col[j + 1]; val[j] = val[j + 1];
                col[j + 1] = tc; val[j + 1] = tv;
                swapped = 1;
            }
        }
    }
}

void mclo_rows_prod(int n, int W, const int *ncol, const int *col, const double *val,
                    const double *x, double *y)
{
    int i;
    /* rows are independent and each keeps its sequential sum: the threads
     * change nothing in the arithmetic */
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) {
        int j;
        double s = 0.0;
        for (j = 0; j < ncol[i]; j++) s += val[(size_t)i * W + j] * x[col[(size_t)i * W + j]];
        y[i] = s;
    }
}

/* ------------------------------------------------------------------- buoyancy
 * Buoyancy.c:3-22, get_node_horiz_avg Parallel.c:354-437 (integer xi, eta: the
 * quotients are int divisions, exact only for unit element steps),
 * remove_node_horiz_avg Parallel.c:439-454.  [golden] */
int mclo_buoyancy(const mclo_mesh *m, const double *T, double Ra, double *buoy)
{
    int i, e, a, k;
    double *avg = calloc((size_t)m->gnoz, sizeof(double));
    int *li = local_ien(m);
    if (!avg || !li) return 1;
    for (i = 0; i < m->nno; i++) buoy[i] = Ra * T[i];
    for (e = 0; e < m->nel; e++) {
        int n0 = m->ien[0 * m->nel + e], n2 = m->ien[2 * m->nel + e];
        int ix0 = n0 / m->gnoz, iz0 = n0 - ix0 * m->gnoz;
        int ix1 = n2 / m->gnoz, iz1 = n2 - ix1 * m->gnoz;
        int idx[NEN];
        double xi, eta;
        for (a = 0; a < NEN; a++) idx[a] = li[a * m->nel + e];
        for (i = ix0; i < ix1; i++) {
            xi = (2 * i - ix0 - ix1) / (ix1 - ix0);
            for (k = iz0; k < iz1; k++) {
                eta = (2 * k - iz0 - iz1) / (iz1 - iz0);
                for (a = 0; a < NEN; a++) avg[k] += shape_at(a, xi, eta) * buoy[idx[a]];
            }
        }
        if (iz1 == m->gnoz - 1) {
            k = iz1;
            for (i = ix0; i < ix1; i++) {
                xi = (2 * i - ix0 - ix1) / (ix1 - ix0);
                eta = 1.0;
                for (a = 0; a < NEN; a++) avg[k] += shape_at(a, xi, eta) * buoy[idx[a]];
            }
        }
        if (ix1 == m->gnox - 1) {
            xi = 1.0;
            for (k = iz0; k < iz1; k++) {
                eta = (2 * k - iz0 - iz1) / (iz1 - iz0);
                for (a = 0; a < NEN; a++) avg[k] += shape_at(a, xi, eta) * buoy[idx[a]];
            }
        }
        if (n2 == m->gnox * m->gnoz - 1) avg[m->gnoz - 1] += buoy[m->nno - 1];
    }
    for (i = 0; i < m->gnoz; i++) avg[i] = avg[i] / m->gnox;
    for (i = 0; i < m->nno; i++) {
        int node = m->node[i];
        int ix = node / m->gnoz, iz = node - ix * m->gnoz;
        buoy[i] -= avg[iz];
    }
    free(avg);
    free(li);
    return 0;
}

/* Viscosity.c:3-24  [golden] */
int mclo_viscosity(const mclo_mesh *m, const double *T, double A, double refT, double lm_jump, double *vis)
{
    int i;
    for (i = 0; i < m->nno; i++) {
        vis[i] = exp(A * (refT - T[i]));
        if (m->X[1 * m->nno + i] < 0.77) vis[i] *= lm_jump;
    }
    return 0;
}

/* ------------------------------------------------------------ Stokes assembly */

/* get_element_K, Build_matrix.c:192-271: 8x8 element stiffness, upper node
 * pairs (b >= a) integrated at 2x2 Gauss points with eta linearly interpolated
 * from nodal viscosity, divided by J, then mirrored.  [golden] */
static void element_K(const mclo_mesh *m, const shape_t *s, const int *li, const double *vis,
                      int e, double K[NEE][NEE])
{
    int a, b, i, j, g, d, p, q;
    double dxdxi, dzdeta, jac;
    const double dzdxi = 0.0, dxdeta = 0.0;
    double D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    int nd[NEN];
    elem_geom(m, e, &dxdxi, &dzdeta, &jac);
    for (d = 0; d < NEN; d++) nd[d] = li[d * m->nel + e];
    for (a = 0; a < NEN; a++)
        for (i = 0; i < NSD; i++) {
            p = NSD * a + i;
            for (b = a; b < NEN; b++)
                for (j = 0; j < NSD; j++) {
                    q = NSD * b + j;
                    K[p][q] = 0.0;
                    for (g = 0; g < NEN; g++) {
                        double eta = 0.0, ax, az, bx, bz;
                        for (d = 0; d < NEN; d++) eta += s->N[d][g] * vis[nd[d]];
                        D[0][0] = D[1][1] = 2.0 * eta;
                        D[2][2] = eta;
                        ax = s->dxi[a][g] * dzdeta - s->deta[a][g] * dzdxi;
                        az = s->deta[a][g] * dxdxi - s->dxi[a][g] * dxdeta;
                        bx = s->dxi[b][g] * dzdeta - s->deta[b][g] * dzdxi;
                        bz = s->deta[b][g] * dxdxi - s->dxi[b][g] * dxdeta;
                        if (i == 0 && j == 0)
                            K[p][q] += ((ax * D[0][0] + az * D[2][0]) * bx + (ax * D[0][2] + az * D[2][2]) * bz);
                        else if (i == 1 && j == 0)
                            K[p][q] += ((az * D[1][0] + ax * D[2][0]) * bx + (az * D[1][2] + ax * D[2][2]) * bz);
                        else if (i == 0 && j == 1)
                            K[p][q] += ((ax * D[0][1] + az * D[2][1]) * bz + (ax * D[0][2] + az * D[2][2]) * bx);
                        else
                            K[p][q] += ((az * D[1][1] + ax * D[2][1]) * bz + (az * D[1][2] + ax * D[2][2]) * bx);
                    }
                    K[p][q] /= jac;
                }
        }
    for (p = 0; p < NEE; p++)
        for (q = 0; q < NEE; q++) K[q][p] = K[p][q];
}

/* get_element_F, Build_matrix.c:273-344: vertical buoyancy load plus the
 * free-slip natural-BC term (zero for zero boundary values).  [golden] */
static void element_F(const mclo_mesh *m, const shape_t *s, const int *li, const double *buoy,
                      int e, double F[NEE])
{
    int a, d, i, p, nat[NSD], idx[NEN];
    double fg[NEN], jac, dxdxi, dzdeta;
    elem_geom(m, e, &dxdxi, &dzdeta, &jac);
    for (a = 0; a < NEN; a++) idx[a] = li[a * m->nel + e];
    for (d = 0; d < NEN; d++) {
        fg[d] = 0.0;
        for (a = 0; a < NEN; a++) fg[d] += s->N[a][d] * buoy[idx[a]];
    }
    nat[0] = nat[1] = 0;
    for (i = 0; i < NSD; i++) {
        for (a = 0; a < NEN; a++)
            if (m->v_bc[i * m->nno + idx[a]] == 'S') nat[i] += a;
        if (nat[i] == 3) nat[i] = 1;
        else if (nat[i] == 1 || nat[i] == 5) nat[i] = 2;
        else if (nat[i] == 0 || nat[i] == 1 || nat[i] == 2 || nat[i] == 3) nat[i] = 0;
    }
    for (i = 0; i < NSD; i++)
        for (a = 0; a < NEN; a++) {
            p = NSD * a + i;
            F[p] = 0.0;
            if (i == 1) {
                for (d = 0; d < NEN; d++) F[p] += s->N[a][d] * fg[d];
                F[p] *= jac;
            }
            if (nat[i] != 0) {
                double L = m->X[i * m->nno + idx[2]] - m->X[i * m->nno + idx[0]];
                F[p] += 0.5 * L * m->v_bc_val[i * m->nno + idx[a]];
            }
        }
}

/* get_element_G, Build_matrix.c:346-381: g = -sum_gauss J*dN/dx_i.  [golden] */
static void element_G(const mclo_mesh *m, const shape_t *s, int e, double g[NEE])
{
    int a, gs;
    double dxdxi, dzdeta, jac;
    const double dzdxi = 0.0, dxdeta = 0.0;
    elem_geom(m, e, &dxdxi, &dzdeta, &jac);
    for (a = 0; a < NEN; a++) {
        g[NSD * a] = 0.0;
        for (gs = 0; gs < NEN; gs++) g[NSD * a] -= s->dxi[a][gs] * dzdeta - s->deta[a][gs] * dzdxi;
    }
    for (a = 0; a < NEN; a++) {
        g[NSD * a + 1] = 0.0;
        for (gs = 0; gs < NEN; gs++) g[NSD * a + 1] -= s->deta[a][gs] * dxdxi - s->dxi[a][gs] * dxdeta;
    }
}

/* velocity_pressure_KGF_matrix at one level: K_matrix (Build_matrix.c:42-115),
 * F_matrix (:117-140), G_matrix (:142-190), then the column sort and Kdiag of
 * remote_KF_contribution (:720-744).  [golden] */
int mclo_assemble_stokes(const mclo_mesh *m, const double *vis, const double *buoy,
                         int *Kncol, int *Kcol, double *Kval, double *Kdiag,
                         int *Gncol, int *Gcol, double *Gval, double *F)
{
    shape_t s;
    int *li, e, i, a, j, b, P, Q, p, q;
    const int W = MCLO_VW;
    shape_init(&s);
    li = local_ien(m);
    if (!li) return 1;
    for (i = 0; i < m->neq; i++) {
        F[i] = 0.0;
        Kncol[i] = 0;
        Gncol[i] = 0;
        for (j = 0; j < W; j++) { Kval[(size_t)i * W + j] = 0.0; Gval[(size_t)i * W + j] = 0.0; }
    }
    for (e = 0; e < m->nel; e++) {
        double K[NEE][NEE];
        element_K(m, &s, li, vis, e, K);
        for (i = 0; i < NSD; i++)
            for (a = 0; a < NEN; a++) {
                P = m->lm[(i * NEN + a) * m->nel + e];
                p = NSD * a + i;
                for (j = 0; j < NSD; j++)
                    for (b = 0; b < NEN; b++) {
                        Q = m->lm[(j * NEN + b) * m->nel + e];
                        q = NSD * b + j;
                        if (P >= 0 && Q >= 0)
                            slot_add(&Kncol[P], &Kcol[(size_t)P * W], &Kval[(size_t)P * W], Q, K[p][q]);
                        if (P >= 0) {
                            int nb = li[b * m->nel + e];
                            if (m->v_bc[j * m->nno + nb] == 'V')
                                F[P] -= K[p][q] * m->v_bc_val[j * m->nno + nb];
                        }
                    }
            }
    }
    for (e = 0; e < m->nel; e++) {
        double Fe[NEE];
        element_F(m, &s, li, buoy, e, Fe);
        for (i = 0; i < NSD; i++)
            for (a = 0; a < NEN; a++) {
                P = m->lm[(i * NEN + a) * m->nel + e];
                if (P >= 0) F[P] += Fe[NSD * a + i];
            }
    }
    for (e = 0; e < m->nel; e++) {
        double g[NEE];
        element_G(m, &s, e, g);
        for (i = 0; i < NSD; i++)
            for (a = 0; a < NEN; a++) {
                P = m->lm[(i * NEN + a) * m->nel + e];
                if (P == -1) continue;
                slot_add(&Gncol[P], &Gcol[(size_t)P * W], &Gval[(size_t)P * W], e, g[NSD * a + i]);
            }
    }
    for (i = 0; i < m->neq; i++) {
        int k;
        slot_sort(Kncol[i], &Kcol[(size_t)i * W], &Kval[(size_t)i * W]);
        k = slot_find(&Kcol[(size_t)i * W], Kncol[i], i);
        Kdiag[i] = k >= 0 ? Kval[(size_t)i * W + k] : 0.0;
    }
    free(li);
    return 0;
}

/* ---------------------------------------------------------------- Stokes solve */

static double vdot(int n, const double *x, const double *y)
{
    /* global_v_prod / global_p_prod at one rank, Array_work.c:287-341 (veq_nsh = 1) */
    double s = 0.0;
    int i;
    for (i = 0; i < n; i++) s += x[i] * y[i] / 1;
    return s;
}

static double pdot(int n, const double *x, const double *y)
{
    double s = 0.0;
    int i;
    for (i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

/* enforce_constrain_veq, Solver.c:1368-1437 (one rank): hanging dof = mean of
 * its two parents (boundary value where a parent dof is fixed).  A hanging dof
 * that is itself fixed is left alone.  [golden, meshes with hanging nodes] */
static void constrain_v(const mclo_mesh *m, double *x)
{
    int k, d, c;
    for (k = 0; k < m->nh; k++) {
        double t[NSD];
        int inode = m->h_ni[k];
        for (d = 0; d < NSD; d++) {
            t[d] = 0.0;
            for (c = 0; c < 2; c++) {
                int kn = m->h_ci[c * m->nh + k];
                int eqc = m->id[d * m->nno + kn];
                if (eqc != -1) t[d] += x[eqc];
                else t[d] += m->v_bc_val[d * m->nno + kn];
            }
            t[d] *= 0.5;
        }
        if (inode >= 0)
            for (d = 0; d < NSD; d++) {
                int eq = m->id[d * m->nno + inode];
                if (eq >= 0) x[eq] = t[d];
            }
    }
}

typedef struct {
    const int *ncol; const int *col; const double *val; const double *diag; int n;
} vrows;

static void kprod(const vrows *A, const double *x, double *y, long long *spmv)
{
    mclo_rows_prod(A->n, MCLO_VW, A->ncol, A->col, A->val, x, y);
    if (spmv) (*spmv)++;
}

/* Conj_Grad, Solver.c:1589-1797: Jacobi-PCG (cg_precondition) or plain CG, stop
 * on delta <= accu*delta0, the cycle cap, or max|alpha r_i| < accu (absolute,
 * after the update); the true residual is recomputed when cycle % 50 == 0.
 * Returns the iteration count.  [golden: per-solve Uzawa counts and fields] */
static int conj_grad(const mclo_mesh *m, const vrows *A, const double *b, double *x,
                     double accu, int max_cycle, int precond, long long *spmv)
{
    int n = A->n, i, cycle = 0;
    double *r = malloc(n * sizeof(double)), *dd = malloc(n * sizeof(double));
    double *q = malloc(n * sizeof(double)), *s = malloc(n * sizeof(double));
    double delta_new, delta_old, delta0, alpha, beta, t1;
    kprod(A, x, r, spmv);
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) {
        r[i] = b[i] - r[i];
        dd[i] = precond ? r[i] / A->diag[i] : r[i];
    }
    delta_new = vdot(n, r, dd);
    delta0 = delta_new;
    while (cycle < max_cycle && delta_new > accu * delta0) {
        double res = 0.0;
        kprod(A, dd, q, spmv);
        t1 = vdot(n, dd, q);
        alpha = delta_new / t1;
#pragma omp parallel for schedule(static)
        for (i = 0; i < n; i++) x[i] += alpha * dd[i];
        constrain_v(m, x);
        if (cycle % 50 == 0) {
            kprod(A, x, r, spmv);
#pragma omp parallel for schedule(static)
            for (i = 0; i < n; i++) r[i] = b[i] - r[i];
        } else {
#pragma omp parallel for schedule(static)
            for (i = 0; i < n; i++) r[i] -= alpha * q[i];
        }
        if (precond) {
#pragma omp parallel for schedule(static)
            for (i = 0; i < n; i++) s[i] = r[i] / A->diag[i];
            delta_old = delta_new;
            delta_new = vdot(n, r, s);
            beta = delta_new / delta_old;
#pragma omp parallel for schedule(static)
            for (i = 0; i < n; i++) dd[i] = s[i] + beta * dd[i];
        } else {
            delta_old = delta_new;
            delta_new = vdot(n, r, r);
            beta = delta_new / delta_old;
#pragma omp parallel for schedule(static)
            for (i = 0; i < n; i++) dd[i] = r[i] + beta * dd[i];
        }
        cycle++;
        /* a maximum does not depend on the order it is taken in */
#pragma omp parallel for schedule(static) reduction(max : res)
        for (i = 0; i < n; i++)
            if (fabs(alpha * r[i]) > res) res = fabs(alpha * r[i]);
        if (res < accu) break;
    }
    free(r); free(dd); free(q); free(s);
    return cycle;
}

/* jacobi, Solver.c:1800-1864 (one rank): x_i = (b_i - sum_{j!=i} A_ij x_j^old)/A_ii,
 * stop when max|x - x_old| < accu.  [golden: jc_16] */
static int jacobi_solve(const vrows *A, const double *b, double *x, double accu, int max_cycle,
                        long long *spmv)
{
    int n = A->n, i, j, cycle = 0;
    double *xo = malloc(n * sizeof(double)), *pr = malloc(n * sizeof(double));
    while (cycle < max_cycle) {
        double diff = 0.0;
        for (i = 0; i < n; i++) { xo[i] = x[i]; pr[i] = 0.0; }
        for (i = 0; i < n; i++) {
            double diag = 1.0;
            for (j = 0; j < A->ncol[i]; j++) {
                int c = A->col[(size_t)i * MCLO_VW + j];
                if (c != i) pr[i] += A->val[(size_t)i * MCLO_VW + j] * xo[c];
                else diag = A->val[(size_t)i * MCLO_VW + j];
            }
            x[i] = (b[i] - pr[i]) / diag;
        }
        if (spmv) (*spmv)++;
        cycle++;
        for (i = 0; i < n; i++) {
            double t = fabs(x[i] - xo[i]);
            if (t > diff) diff = t;
        }
        if (diff < accu) break;
    }
    free(xo); free(pr);
    return cycle;
}

/* gauss_seidel, Solver.c:1866-1928: as jacobi but using updated x in place.
 * [golden: gs_16, and as the smoother of mg_16 / mg_32] */
static int gs_solve(const vrows *A, const double *b, double *x, double accu, int max_cycle,
                    long long *spmv)
{
    int n = A->n, i, j, cycle = 0;
    double *xo = malloc(n * sizeof(double)), *pr = malloc(n * sizeof(double));
    while (cycle < max_cycle) {
        double diff = 0.0;
        for (i = 0; i < n; i++) { xo[i] = x[i]; pr[i] = 0.0; }
        for (i = 0; i < n; i++) {
            double diag = 1.0;
            for (j = 0; j < A->ncol[i]; j++) {
                int c = A->col[(size_t)i * MCLO_VW + j];
                if (c != i) pr[i] += A->val[(size_t)i * MCLO_VW + j] * x[c];
                else diag = A->val[(size_t)i * MCLO_VW + j];
            }
            x[i] = (b[i] - pr[i]) / diag;
        }
        if (spmv) (*spmv)++;
        cycle++;
        for (i = 0; i < n; i++) {
            double t = fabs(x[i] - xo[i]);
            if (t > diff) diff = t;
        }
        if (diff < accu) break;
    }
    free(xo); free(pr);
    return cycle;
}

/* Multigrid V-cycle as a stand-alone inner solver, Solver.c:668-769 at one
 * rank: Gauss-Seidel pre/post smoothing, injection restriction (project_res,
 * :770-832), averaging prolongation (interp_error, :1009-1143), coarse solve by
 * Gauss-Seidel (1000 sweeps); cycles until max|x - x_old| < accu.  Below the
 * finest level x aliases err[lev], so the prolongation overwrites the smoothed
 * coarse iterate before it is added (x = 2 P e there), as in the reference's
 * MG(), where x is G.err[lev] too.  [golden: mg_16, mg_32] */
static void mg_vcycle(const mclo_ctrl *c, mclo_mg_level *mg, int lev, const double *rhs, double *x,
                      double **err, long long *spmv)
{
    const mclo_mesh *mu = mg[lev].mesh, *ml = mg[lev - 1].mesh;
    vrows Au = {mg[lev].Kncol, mg[lev].Kcol, mg[lev].Kval, mg[lev].Kdiag, mu->neq};
    vrows Al = {mg[lev - 1].Kncol, mg[lev - 1].Kcol, mg[lev - 1].Kval, mg[lev - 1].Kdiag, ml->neq};
    double *res_u = malloc(mu->neq * sizeof(double)), *res_l = malloc(ml->neq * sizeof(double));
    int i, k;
    gs_solve(&Au, rhs, x, c->accuracy, c->mg_finest_iteration, spmv);
    /* get_residual, Solver.c:1291-1316 */
    for (i = 0; i < mu->neq; i++) {
        res_u[i] = rhs[i];
        for (k = 0; k < Au.ncol[i]; k++)
            res_u[i] -= Au.val[(size_t)i * MCLO_VW + k] * x[Au.col[(size_t)i * MCLO_VW + k]];
    }
    for (k = 0; k < mg[lev - 1].nproj; k++) res_l[mg[lev - 1].proj_l[k]] = res_u[mg[lev - 1].proj_u[k]];
    if (lev - 1 == 0)
        gs_solve(&Al, res_l, err[lev - 1], c->accuracy, 1000, spmv);
    else
        mg_vcycle(c, mg, lev - 1, res_l, err[lev - 1], err, spmv);
    /* interp_error local part */
    for (i = 0; i < mu->neq; i++) {
        int n = mg[lev].interp_n[i];
        if (n > 0) {
            double v = 0.0;
            for (k = 0; k < n; k++) {
                int l = mg[lev].interp_eq[k * mu->neq + i];
                if (l != -1) v += err[lev - 1][l];
            }
            err[lev][i] = v / n;
        }
    }
    for (i = 0; i < mu->neq; i++) x[i] += err[lev][i];
    gs_solve(&Au, rhs, x, c->accuracy, c->mg_coarse_iteration, spmv);
    free(res_u); free(res_l);
}

static int mg_solve(const mclo_ctrl *c, mclo_mg_level *mg, int nlev, const double *rhs, double *x,
                    long long *spmv)
{
    int top = nlev - 1, cycle = 0, l, i, conv = 0;
    int n = mg[top].mesh->neq;
    double **err = malloc(nlev * sizeof(double *));
    double *xt = malloc(n * sizeof(double));
    for (l = 0; l < nlev; l++) err[l] = calloc((size_t)mg[l].mesh->neq, sizeof(double));
    while (!conv && cycle < c->mg_cycle) {
        double diff = 0.0;
        for (i = 0; i < n; i++) xt[i] = x[i];
        mg_vcycle(c, mg, top, rhs, x, err, spmv);
        for (i = 0; i < n; i++) {
            double t = fabs(x[i] - xt[i]);
            if (t > diff) diff = t;
        }
        conv = diff < c->accuracy;
        cycle++;
    }
    for (l = 0; l < nlev; l++) free(err[l]);
    free(err); free(xt);
    return cycle;
}

static int inner_solve(const mclo_mesh *m, const mclo_ctrl *c, const vrows *A, const double *b,
                       double *x, mclo_mg_level *mg, int nlev, long long *spmv)
{
    switch (c->solver) {
    case MCLO_SOLVER_JC: return jacobi_solve(A, b, x, c->accuracy, 1000, spmv);
    case MCLO_SOLVER_GS: return gs_solve(A, b, x, c->accuracy, 1000, spmv);
    case MCLO_SOLVER_MG: return mg_solve(c, mg, nlev, b, x, spmv);
    default: return conj_grad(m, A, b, x, c->accuracy, c->max_cg_cycle, c->cg_precondition, spmv);
    }
}

/* citcom_solver, Solver.c:14-249: Uzawa pressure-CG on the Schur complement
 * G^T K^-1 G with the diagonal BAB = sum_i G_ie^2 / Kdiag_i, applied as
 * z = BAB * r (reference) or z = r / BAB (schur_divide); stop when the update
 * ratios dv, dp fall below citcom_accu or at the cycle cap.  [golden; auto_accuracy:
 * golden auto_16; schur_divide (not in the reference): closed, its converged solve
 * equals a direct sparse solve of the saddle-point system] */
int mclo_stokes_solve(const mclo_mesh *m, const mclo_ctrl *c,
                      const int *Kncol, const int *Kcol, const double *Kval, const double *Kdiag,
                      const int *Gncol, const int *Gcol, const double *Gval, const double *F,
                      double *d, double *p, mclo_mg_level *mg, int nlev, mclo_stokes_stats *st)
{
    int nel = m->nel, neq = m->neq, i, j, cycle = 0;
    const int W = MCLO_VW;
    double *r0 = malloc(nel * sizeof(double)), *r1 = malloc(nel * sizeof(double));
    double *r2 = malloc(nel * sizeof(double)), *z0 = malloc(nel * sizeof(double));
    double *z1 = malloc(nel * sizeof(double)), *s1 = malloc(nel * sizeof(double));
    double *s2 = malloc(nel * sizeof(double)), *BAB = malloc(nel * sizeof(double));
    double *Ah = malloc(neq * sizeof(double)), *u1 = malloc(neq * sizeof(double));
    double dv = 1.0, dp = 1.0, r1z1, r0z0, delta, s2Ah, alpha, t1, t2, t3, t4, *sh;
    mclo_ctrl cc = *c;
    vrows A = {Kncol, Kcol, Kval, Kdiag, neq};
    long long inner = 0, spmv = 0;

    if (c->auto_accuracy) {
        /* Solver.c:41-51 */
        double v = sqrt(vdot(neq, F, F) / neq);
        cc.accuracy = c->citcom_accu * v;
    }
    for (i = 0; i < nel; i++) BAB[i] = 0.0;
    for (i = 0; i < neq; i++)
        for (j = 0; j < Gncol[i]; j++) {
            int cc_ = Gcol[(size_t)i * W + j];
            BAB[cc_] += Gval[(size_t)i * W + j] * Gval[(size_t)i * W + j] / Kdiag[i];
        }
    kprod(&A, d, u1, &spmv);
    /* G_prod_p, Solver.c:569-591 */
    for (i = 0; i < neq; i++) {
        Ah[i] = 0.0;
        for (j = 0; j < Gncol[i]; j++) Ah[i] += Gval[(size_t)i * W + j] * p[Gcol[(size_t)i * W + j]];
    }
    for (i = 0; i < neq; i++) Ah[i] = F[i] - Ah[i] - u1[i];
    inner += inner_solve(m, &cc, &A, Ah, u1, mg, nlev, &spmv);
    for (i = 0; i < neq; i++) d[i] += u1[i];
    /* GT_prod_d, Solver.c:645-666 */
    for (i = 0; i < nel; i++) r1[i] = 0.0;
    for (i = 0; i < neq; i++)
        for (j = 0; j < Gncol[i]; j++) r1[Gcol[(size_t)i * W + j]] += Gval[(size_t)i * W + j] * d[i];

    while ((dv > c->citcom_accu || (dp > c->citcom_accu && c->check_v_only == 0)) &&
           cycle < c->max_citcom_cycle) {
        for (i = 0; i < nel; i++) z1[i] = c->schur_divide ? r1[i] / BAB[i] : BAB[i] * r1[i];
        r1z1 = pdot(nel, r1, z1);
        if (cycle == 0) {
            for (i = 0; i < nel; i++) s2[i] = z1[i];
        } else {
            r0z0 = pdot(nel, r0, z0);
            delta = r1z1 / r0z0;
            for (i = 0; i < nel; i++) s2[i] = z1[i] + delta * s1[i];
        }
        for (i = 0; i < neq; i++) {
            Ah[i] = 0.0;
            for (j = 0; j < Gncol[i]; j++) Ah[i] += Gval[(size_t)i * W + j] * s2[Gcol[(size_t)i * W + j]];
        }
        for (i = 0; i < neq; i++) u1[i] = 0.0;
        inner += inner_solve(m, &cc, &A, Ah, u1, mg, nlev, &spmv);
        for (i = 0; i < nel; i++) Ah[i] = 0.0;
        for (i = 0; i < neq; i++)
            for (j = 0; j < Gncol[i]; j++) Ah[Gcol[(size_t)i * W + j]] += Gval[(size_t)i * W + j] * u1[i];
        s2Ah = pdot(nel, s2, Ah);
        alpha = r1z1 / s2Ah;
        for (i = 0; i < nel; i++) {
            r2[i] = r1[i] - alpha * Ah[i];
            p[i] += alpha * s2[i];
        }
        for (i = 0; i < neq; i++) d[i] -= alpha * u1[i];
        t1 = pdot(nel, s2, s2);
        t2 = pdot(nel, p, p);
        t3 = vdot(neq, u1, u1);
        t4 = vdot(neq, d, d);
        dp = alpha * sqrt(t1 / (1.0e-32 + t2));
        dv = alpha * sqrt(t3 / (1.0e-32 + t4));
        cycle++;
        sh = s1; s1 = s2; s2 = sh;
        sh = r0; r0 = r1; r1 = r2; r2 = sh;
        sh = z0; z0 = z1; z1 = sh;
    }
    if (st) {
        st->uzawa_cycles = cycle;
        st->dv = dv;
        st->dp = dp;
        st->inner_iterations = inner;
        st->spmv = spmv;
        st->accuracy_used = cc.accuracy;
    }
    free(r0); free(r1); free(r2); free(z0); free(z1); free(s1); free(s2); free(BAB);
    free(Ah); free(u1);
    return 0;
}

/* v_to_nodes, Solver.c:1319-1343  [golden] */
void mclo_v_to_nodes(const mclo_mesh *m, const double *d, double *V)
{
    int i, k;
    for (i = 0; i < m->nno; i++)
        for (k = 0; k < NSD; k++) {
            int j = m->id[k * m->nno + i];
            V[k * m->nno + i] = j >= 0 ? d[j] : m->v_bc_val[k * m->nno + i];
        }
}

/* compute_stress, Grain_size_evolution.c:154-213, and p_to_nodes,
 * Solver.c:1345-1366: element 2 eta sqrt(I2) (with u2_1 = u1_2), averaged to
 * nodes with element-area weights.  [golden] */
void mclo_compute_stress(const mclo_mesh *m, const double *vis, const double *V, double *stress)
{
    shape_t s;
    int e, a, i;
    int *li = local_ien(m);
    double *es = malloc(m->nel * sizeof(double)), *area = malloc(m->nel * sizeof(double));
    double *tot = calloc((size_t)m->nno, sizeof(double));
    shape_init(&s);
    for (e = 0; e < m->nel; e++) {
        double eta = 0.0, u11 = 0.0, u22 = 0.0, u12 = 0.0, u21, sr, dxdxi, dzdeta, jac;
        int n0 = li[0 * m->nel + e], n2 = li[2 * m->nel + e];
        elem_geom(m, e, &dxdxi, &dzdeta, &jac);
        for (a = 0; a < NEN; a++) eta += s.Nc[a] * vis[li[a * m->nel + e]];
        for (a = 0; a < NEN; a++) {
            int k = li[a * m->nel + e];
            double ax = s.dxic[a] * dzdeta - s.detac[a] * 0.0;
            double az = s.detac[a] * dxdxi - s.dxic[a] * 0.0;
            u11 += ax * V[0 * m->nno + k];
            u22 += az * V[1 * m->nno + k];
            u12 += az * V[0 * m->nno + k];
        }
        u11 /= jac; u22 /= jac; u12 /= jac;
        u21 = u12;
        sr = u11 * u11 + u22 * u22 + u12 * u21 * 2.0;
        sr = sqrt(0.5 * sr);
        es[e] = 2.0 * eta * sr;
        /* element_area, Construct_ien_id_lm.c:198-218 */
        area[e] = (m->X[0 * m->nno + n2] - m->X[0 * m->nno + n0]) * (m->X[1 * m->nno + n2] - m->X[1 * m->nno + n0]);
    }
    /* p_to_nodes walks each node's elements in ascending element order
     * (what_elements_node_in, Construct_ien_id_lm.c:220-244). */
    for (i = 0; i < m->nno; i++) stress[i] = 0.0;
    for (e = 0; e < m->nel; e++)
        for (a = 0; a < NEN; a++) {
            i = li[a * m->nel + e];
            stress[i] += area[e] * es[e];
            tot[i] += area[e];
        }
    for (i = 0; i < m->nno; i++) stress[i] /= tot[i];
    free(li); free(es); free(area); free(tot);
}

/* ------------------------------------------------------------------ heat step */

/* advection_dt, Update_temperature.c:135-177: the advective loop reuses L of
 * the last element visited by the diffusive loop.  [golden] */
double mclo_advection_dt(const mclo_mesh *m, const double *V)
{
    double diff = 1e10, dt = 1e10, L[NSD] = {0.0, 0.0};
    int e, a, d;
    int *li = local_ien(m);
    for (e = 0; e < m->nel; e++) {
        int level = elem_level(m->Mcode[e]);
        L[0] = m->dx[level];
        L[1] = m->dz[level];
        for (d = 0; d < NSD; d++) {
            double ts = L[d] * L[d];
            if (diff > ts) diff = ts;
        }
    }
    for (e = 0; e < m->nel; e++) {
        double u1 = 0.0, u2 = 0.0, u, step;
        for (a = 0; a < NEN; a++) {
            int k = li[a * m->nel + e];
            u1 += 0.25 * V[0 * m->nno + k];
            u2 += 0.25 * V[1 * m->nno + k];
        }
        u = fabs(u1) / L[0] + fabs(u2) / L[1];
        step = 0.5 / u;
        dt = step < dt ? step : dt;
    }
    dt = dt < diff ? dt : diff;
    free(li);
    return dt;
}

/* calculate_artificial_diffusivity, Update_temperature.c:241-291:
 * alpha_d = u_d L_d / 2, xi_d = coth(alpha_d) - 1/alpha_d (0 when |alpha_d| <
 * accuracy), k = (xi_0 u_0 L_0 + xi_1 u_1 L_1)/2, tau_d = k u_d / |u|^2.
 * [golden; closed forms for the xi limits] */
void mclo_artificial_diffusivity(const mclo_mesh *m, const double *V, double accuracy, double *tau)
{
    int e, a, d;
    int *li = local_ien(m);
    for (e = 0; e < m->nel; e++) {
        int level = elem_level(m->Mcode[e]);
        double L[NSD], u[NSD], al[NSD], xi[NSD], us, kh;
        L[0] = m->dx[level];
        L[1] = m->dz[level];
        for (d = 0; d < NSD; d++) {
            u[d] = 0.0;
            for (a = 0; a < NEN; a++) u[d] += 0.25 * V[d * m->nno + li[a * m->nel + e]];
            al[d] = 0.5 * u[d] * L[d] / 1.0;
            if (fabs(al[d]) < accuracy) xi[d] = 0.0;
            else xi[d] = cosh(al[d]) / sinh(al[d]) - 1.0 / al[d];
        }
        us = (u[0] * u[0] + u[1] * u[1]);
        kh = (xi[0] * u[0] * L[0] + xi[1] * u[1] * L[1]) * 0.5;
        for (d = 0; d < NSD; d++) tau[d * m->nel + e] = kh * u[d] / us;
    }
    free(li);
}

/* heat K, F, C assembly, Update_temperature.c:293-594, and the column sort of
 * other_KFC_heat_contribution (:788-824).  [golden] */
int mclo_assemble_heat(const mclo_mesh *m, const mclo_ctrl *c, const double *V, const double *tau,
                       int *Kncol, int *Kcol, double *Kval, int *Cncol, int *Ccol, double *Cval, double *F)
{
    shape_t s;
    const int W = MCLO_TW;
    int *li, e, i, j, a, b, g, P, Q;
    shape_init(&s);
    li = local_ien(m);
    for (i = 0; i < m->neqT; i++) {
        Kncol[i] = 0; Cncol[i] = 0; F[i] = 0.0;
        for (j = 0; j < W; j++) { Kval[(size_t)i * W + j] = 0.0; Cval[(size_t)i * W + j] = 0.0; }
    }
    /* K_matrix_heat + get_element_K_heat (:293-423) */
    for (e = 0; e < m->nel; e++) {
        double k[NEN][NEN], u1 = 0.0, u2 = 0.0, dxdxi, dzdeta, jac;
        elem_geom(m, e, &dxdxi, &dzdeta, &jac);
        for (a = 0; a < NEN; a++) {
            int kn = li[a * m->nel + e];
            u1 += V[0 * m->nno + kn] * 0.25;
            u2 += V[1 * m->nno + kn] * 0.25;
        }
        for (a = 0; a < NEN; a++)
            for (b = 0; b < NEN; b++) {
                k[a][b] = 0.0;
                for (g = 0; g < NEN; g++) {
                    double ax = s.dxi[a][g] * dzdeta - s.deta[a][g] * 0.0;
                    double az = s.deta[a][g] * dxdxi - s.dxi[a][g] * 0.0;
                    double bx = s.dxi[b][g] * dzdeta - s.deta[b][g] * 0.0;
                    double bz = s.deta[b][g] * dxdxi - s.dxi[b][g] * 0.0;
                    double nh = tau[0 * m->nel + e] * ax + tau[1 * m->nel + e] * az;
                    double t1 = u1 * s.N[a][g] * bx + u2 * s.N[a][g] * bz;
                    double t2 = u1 * nh * bx + u2 * nh * bz;
                    double t3 = ax * 1.0 * bx + ax * 0.0 * bz + az * 0.0 * bx + az * 1.0 * bz;
                    k[a][b] += t1 + (t2 + t3) / jac;
                }
            }
        for (a = 0; a < NEN; a++) {
            P = m->heat_lm[a * m->nel + e];
            for (b = 0; b < NEN; b++) {
                Q = m->heat_lm[b * m->nel + e];
                if (P >= 0 && Q >= 0)
                    slot_add(&Kncol[P], &Kcol[(size_t)P * W], &Kval[(size_t)P * W], Q, k[a][b]);
                if (P >= 0) {
                    int kn = li[b * m->nel + e];
                    if (m->T_bc[kn] == 'T') F[P] -= k[a][b] * m->T_bc_val[kn];
                }
            }
        }
    }
    /* F_matrix_heat + get_element_F_heat (:425-508) */
    for (e = 0; e < m->nel; e++) {
        double f[NEN], dxdxi, dzdeta, jac;
        int nat = 0, level = elem_level(m->Mcode[e]);
        elem_geom(m, e, &dxdxi, &dzdeta, &jac);
        for (a = 0; a < NEN; a++)
            if (m->T_bc[li[a * m->nel + e]] == 'F') nat += a;
        if (nat == 3) nat = 1;
        else if (nat == 1 || nat == 5) nat = 2;
        else if (nat == 0 || nat == 1 || nat == 2 || nat == 3) nat = 0;
        for (a = 0; a < NEN; a++) {
            f[a] = 0.0;
            for (g = 0; g < NEN; g++) {
                double ax = s.dxi[a][g] * dzdeta - s.deta[a][g] * 0.0;
                double az = s.deta[a][g] * dxdxi - s.dxi[a][g] * 0.0;
                double nh = tau[0 * m->nel + e] * ax + tau[1 * m->nel + e] * az;
                f[a] += (s.N[a][g] + nh) * c->Q;
            }
            f[a] *= jac;
            if (nat != 0) f[a] -= 0.5 * m->dz[level] * m->T_bc_val[li[a * m->nel + e]];
        }
        for (a = 0; a < NEN; a++) {
            P = m->heat_lm[a * m->nel + e];
            if (P >= 0) F[P] += f[a];
        }
    }
    /* C_matrix_heat + get_element_C_heat (:510-594) */
    for (e = 0; e < m->nel; e++) {
        double cm[NEN][NEN], dxdxi, dzdeta, jac;
        elem_geom(m, e, &dxdxi, &dzdeta, &jac);
        for (a = 0; a < NEN; a++)
            for (b = 0; b < NEN; b++) {
                cm[a][b] = 0.0;
                for (g = 0; g < NEN; g++) {
                    double ax = s.dxi[a][g] * dzdeta - s.deta[a][g] * 0.0;
                    double az = s.deta[a][g] * dxdxi - s.dxi[a][g] * 0.0;
                    double nh = tau[0 * m->nel + e] * ax + tau[1 * m->nel + e] * az;
                    cm[a][b] += (s.N[a][g] + nh) * s.N[b][g];
                }
                cm[a][b] *= jac;
            }
        for (a = 0; a < NEN; a++) {
            P = m->heat_lm[a * m->nel + e];
            for (b = 0; b < NEN; b++) {
                Q = m->heat_lm[b * m->nel + e];
                if (P >= 0 && Q >= 0)
                    slot_add(&Cncol[P], &Ccol[(size_t)P * W], &Cval[(size_t)P * W], Q, cm[a][b]);
            }
        }
    }
    for (i = 0; i < m->neqT; i++) {
        slot_sort(Kncol[i], &Kcol[(size_t)i * W], &Kval[(size_t)i * W]);
        slot_sort(Cncol[i], &Ccol[(size_t)i * W], &Cval[(size_t)i * W]);
    }
    free(li);
    return 0;
}

/* enforce_constrain_Teq, Update_temperature.c:829-893 (one rank) */
static void constrain_T(const mclo_mesh *m, double *x)
{
    int k, c;
    for (k = 0; k < m->nh; k++) {
        double t = 0.0;
        int inode = m->h_ni[k];
        for (c = 0; c < 2; c++) {
            int kn = m->h_ci[c * m->nh + k];
            int eqc = m->heat_id[kn];
            if (eqc != -1) t += x[eqc];
            else t += m->T_bc_val[kn];
        }
        t *= 0.5;
        if (inode >= 0) {
            int eq = m->heat_id[inode];
            if (eq >= 0) x[eq] = t;
        }
    }
}

/* Conj_Grad_heat, Update_temperature.c:1140-1354: as Conj_Grad but on heat rows,
 * diagonal looked up per row, and the plain-CG branch breaks on the absolute
 * test only after more than 10 cycles.  [golden] */
static int conj_grad_heat(const mclo_mesh *m, int n, const int *ncol, const int *col, const double *val,
                          const double *b, double *x, double accu, int max_cycle, int precond)
{
    const int W = MCLO_TW;
    int i, cycle = 0;
    double *r = malloc(n * sizeof(double)), *dd = malloc(n * sizeof(double));
    double *q = malloc(n * sizeof(double)), *s = malloc(n * sizeof(double));
    double *diag = malloc(n * sizeof(double));
    double dn, dold, d0, alpha, beta, t1;
    for (i = 0; i < n; i++) {
        int k = slot_find(&col[(size_t)i * W], ncol[i], i);
        diag[i] = val[(size_t)i * W + k];
    }
    mclo_rows_prod(n, W, ncol, col, val, x, r);
    for (i = 0; i < n; i++) {
        r[i] = b[i] - r[i];
        dd[i] = precond ? r[i] / diag[i] : r[i];
    }
    dn = vdot(n, r, dd);
    d0 = dn;
    while (cycle < max_cycle && dn > accu * d0) {
        double res = 0.0;
        mclo_rows_prod(n, W, ncol, col, val, dd, q);
        t1 = vdot(n, dd, q);
        alpha = dn / t1;
        for (i = 0; i < n; i++) x[i] += alpha * dd[i];
        constrain_T(m, x);
        if (cycle % 50 == 0) {
            mclo_rows_prod(n, W, ncol, col, val, x, r);
            for (i = 0; i < n; i++) r[i] = b[i] - r[i];
        } else {
            for (i = 0; i < n; i++) r[i] -= alpha * q[i];
        }
        if (precond) {
            for (i = 0; i < n; i++) s[i] = r[i] / diag[i];
            dold = dn;
            dn = vdot(n, r, s);
            beta = dn / dold;
            for (i = 0; i < n; i++) dd[i] = s[i] + beta * dd[i];
        } else {
            dold = dn;
            dn = vdot(n, r, r);
            beta = dn / dold;
            for (i = 0; i < n; i++) dd[i] = r[i] + beta * dd[i];
        }
        cycle++;
        for (i = 0; i < n; i++)
            if (fabs(alpha * r[i]) > res) res = fabs(alpha * r[i]);
        if (precond ? res < accu : (res < accu && cycle > 10)) break;
    }
    free(r); free(dd); free(q); free(s); free(diag);
    return cycle;
}

/* update_temperature, Update_temperature.c:3-133: the CTdot term of B is
 * dropped exactly as in the reference (the statement ends before it, :64-65),
 * so B = F + C d /(theta dt); A = C/(theta dt) + K on K's column pattern; one
 * Jacobi-PCG solve with accu 1e-3 and cap 1000.  [golden] */
int mclo_update_temperature(const mclo_mesh *m, const mclo_ctrl *c, const double *V,
                            double *T, double *hd, double *hdd, int first_call,
                            double *model_time, mclo_heat_stats *st,
                            double *tau_out, int *Kncol_o, int *Kcol_o, double *Kval_o,
                            int *Cncol_o, int *Ccol_o, double *Cval_o, double *F_out)
{
    const int W = MCLO_TW;
    int n = m->neqT, i, j;
    size_t slots = (size_t)n * W;
    double *tau = malloc(2 * (size_t)m->nel * sizeof(double));
    int *Kncol = malloc(n * sizeof(int)), *Kcol = malloc(slots * sizeof(int));
    int *Cncol = malloc(n * sizeof(int)), *Ccol = malloc(slots * sizeof(int));
    double *Kval = malloc(slots * sizeof(double)), *Cval = malloc(slots * sizeof(double));
    double *F = malloc(n * sizeof(double)), *B = malloc(n * sizeof(double));
    double *d0 = malloc(n * sizeof(double)), *CT = malloc(n * sizeof(double));
    double *CTd = malloc(n * sizeof(double)), *Aval = malloc(slots * sizeof(double));
    double dt, td;
    int init_cycles = 0, cycles;

    mclo_artificial_diffusivity(m, V, c->accuracy, tau);
    mclo_assemble_heat(m, c, V, tau, Kncol, Kcol, Kval, Cncol, Ccol, Cval, F);
    if (first_call) {
        /* initial_T_dot, :179-239 */
        j = 0;
        for (i = 0; i < m->nno; i++)
            if (m->heat_id[i] >= 0) hd[j++] = T[i];
        mclo_rows_prod(n, W, Kncol, Kcol, Kval, hd, B);
        for (i = 0; i < n; i++) B[i] = F[i] - B[i];
        for (i = 0; i < n; i++) hdd[i] = 0.0;
        init_cycles = conj_grad_heat(m, n, Cncol, Ccol, Cval, B, hdd, 1e-3, 400, c->cg_precondition);
    }
    dt = mclo_advection_dt(m, V) * 0.75;
    *model_time += dt;
    for (i = 0; i < n; i++) d0[i] = hd[i];
    mclo_rows_prod(n, W, Cncol, Ccol, Cval, hd, CT);
    mclo_rows_prod(n, W, Cncol, Ccol, Cval, hdd, CTd);
    td = c->theta * dt;
    for (i = 0; i < n; i++) B[i] = F[i] + CT[i] / td;
    for (i = 0; i < n; i++)
        for (j = 0; j < Kncol[i]; j++)
            Aval[(size_t)i * W + j] = Cval[(size_t)i * W + j] / td + Kval[(size_t)i * W + j];
    cycles = conj_grad_heat(m, n, Kncol, Kcol, Aval, B, hd, 1e-3, 1000, c->cg_precondition);
    for (i = 0; i < m->nno; i++) {
        int k = m->heat_id[i];
        T[i] = k >= 0 ? hd[k] : m->T_bc_val[i];
    }
    for (i = 0; i < n; i++)
        hdd[i] = 1.0 / (c->theta * dt) * (hd[i] - d0[i]) - (1.0 - c->theta) / c->theta * hdd[i];
    if (st) { st->cg_cycles = cycles; st->init_cg_cycles = init_cycles; st->dt = dt; }
    if (tau_out) memcpy(tau_out, tau, 2 * (size_t)m->nel * sizeof(double));
    if (Kncol_o) {
        memcpy(Kncol_o, Kncol, n * sizeof(int)); memcpy(Kcol_o, Kcol, slots * sizeof(int));
        memcpy(Kval_o, Kval, slots * sizeof(double));
    }
    if (Cncol_o) {
        memcpy(Cncol_o, Cncol, n * sizeof(int)); memcpy(Ccol_o, Ccol, slots * sizeof(int));
        memcpy(Cval_o, Cval, slots * sizeof(double));
    }
    if (F_out) memcpy(F_out, F, n * sizeof(double));
    free(tau); free(Kncol); free(Kcol); free(Cncol); free(Ccol); free(Kval); free(Cval);
    free(F); free(B); free(d0); free(CT); free(CTd); free(Aval);
    return 0;
}
