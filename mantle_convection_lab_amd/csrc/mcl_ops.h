// Device operators shared by the FP64 and mixed-precision kernels: ELL row
// products, the gradient G and divergence G^T applied from element tables, the
// hanging-node constraints.  Sums run in the reference's order (sorted ELL
// columns; a row's elements ascending; an element's rows ascending).
#pragma once
#include "mcl_kernels.h"

// ------------------------------------------------------------- sparse products

__device__ __forceinline__ double ell_row(const int *__restrict__ col, const double *__restrict__ val, int ncol,
                                          int n, int row, const double *__restrict__ x)
{
    // loads issued nine entries at a time (one memory round trip per chunk
    // instead of per entry); the sum keeps the ascending-column order
    double s = 0.0;
    int j = 0;
    for (; j + 9 <= ncol; j += 9) {
        int c[9];
        double v[9];
#pragma unroll
        for (int u = 0; u < 9; u++) {
            c[u] = col[(j + u) * n + row];
            v[u] = val[(j + u) * n + row];
        }
#pragma unroll
        for (int u = 0; u < 9; u++) s += v[u] * x[c[u]];
    }
    for (; j < ncol; j++) s += val[j * n + row] * x[col[j * n + row]];
    return s;
}

// (G p)[P], G_prod_p (Solver.c:569-591): columns are the row's elements, ascending.
// (G p) for velocity row P with the element values read through load(e)
template <typename LD>
__device__ __forceinline__ double g_row_ld(const LevelDev &L, int P, LD load)
{
    int n = L.veq_node[P], ip = L.veq_dof[P];
    double s = 0.0;
    for (int k = 0; k < L.p2n_cnt[n]; k++) {
        int e = L.p2n_el[k * L.nno + n], a = L.p2n_a[k * L.nno + n];
        s += L.gtab[L.elev[e] * 8 + 2 * a + ip] * load(e);
    }
    return s;
}

__device__ __forceinline__ double g_row(const LevelDev &L, int P, const double *__restrict__ p)
{
    return g_row_ld(L, P, [&](int e) { return p[e]; });
}

// (G^T x)[e], GT_prod_d (Solver.c:645-666): rows of the element ascending,
// with the velocity values read through load(r).
template <typename LD>
__device__ __forceinline__ double gt_elem_ld(const LevelDev &L, int e, LD load)
{
    double s = 0.0;
    int lev = L.elev[e];
    for (int t = 0; t < 8; t++) {
        int r = L.gt_row[t * L.nel + e];
        if (r < 0) break;
        s += L.gtab[lev * 8 + L.gt_p[t * L.nel + e]] * load(r);
    }
    return s;
}

__device__ __forceinline__ double gt_elem(const LevelDev &L, int e, const double *__restrict__ x)
{
    return gt_elem_ld(L, e, [&](int r) { return x[r]; });
}

// ------------------------------------------------------------ persistent Uzawa

struct CgVecs {
    double *r, *s, *d, *q;
};

// Hanging-dof constraint, enforce_constrain_veq (Solver.c:1368-1437).
static __device__ __forceinline__ void constrain_v(const LevelDev &L, double *__restrict__ x, int gtid, int gsize)
{
    if (L.nh == 0) return;
    if (L.h_serial) {
        if (gtid != 0) return;
        for (int k = 0; k < L.nh; k++) {
            for (int d = 0; d < 2; d++) {
                double t = 0.0;
                for (int c = 0; c < 2; c++) {
                    int kn = L.h_ci[c * L.nh + k];
                    int eqc = L.id[d * L.nno + kn];
                    t += eqc != -1 ? x[eqc] : L.v_bc_val[d * L.nno + kn];
                }
                t *= 0.5;
                int eq = L.id[d * L.nno + L.h_ni[k]];
                if (eq >= 0) x[eq] = t;
            }
        }
        return;
    }
    for (int k = gtid; k < L.nh; k += gsize) {
        for (int d = 0; d < 2; d++) {
            double t = 0.0;
            for (int c = 0; c < 2; c++) {
                int kn = L.h_ci[c * L.nh + k];
                int eqc = L.id[d * L.nno + kn];
                t += eqc != -1 ? x[eqc] : L.v_bc_val[d * L.nno + kn];
            }
            t *= 0.5;
            int eq = L.id[d * L.nno + L.h_ni[k]];
            if (eq >= 0) x[eq] = t;
        }
    }
}

static __device__ __forceinline__ void constrain_T(const LevelDev &L, double *__restrict__ x, int gtid, int gsize)
{
    if (L.nh == 0) return;
    int start = L.h_serial ? (gtid == 0 ? 0 : L.nh) : gtid;
    int step = L.h_serial ? 1 : gsize;
    for (int k = start; k < L.nh; k += step) {
        double t = 0.0;
        for (int c = 0; c < 2; c++) {
            int kn = L.h_ci[c * L.nh + k];
            int eqc = L.heat_id[kn];
            t += eqc != -1 ? x[eqc] : L.T_bc_val[kn];
        }
        t *= 0.5;
        int eq = L.heat_id[L.h_ni[k]];
        if (eq >= 0) x[eq] = t;
    }
}

// An ELL matrix with its diagonal (explicit array or slot index per row).
struct EllMat {
    const int *col;
    const int *ncol;
    const double *val;
    const double *diag;      // diagonal values, or NULL to use diag_slot
    const int *diag_slot;
    int n;
};

__device__ __forceinline__ double ell_diag(const EllMat &A, int i)
{
    return A.diag ? A.diag[i] : A.val[A.diag_slot[i] * A.n + i];
}

