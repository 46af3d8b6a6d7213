// Mixed-precision Stokes solve of libmcl_hip.so (gfx950).
//
// The Uzawa outer iteration (Solver.c:14-249) runs in FP64 exactly as in the
// FP64 path; only the inner velocity solves K u = b change:
//   FP64 iterative refinement      r = b - K u (FP64 ELL),  u += |r| e
//   around an FP32 CG              K32 e = r / |r| to a relative tolerance,
//   preconditioned by a V-cycle    geometric multigrid over the quadtree levels
//                                  (damped-Jacobi smoothing, restriction P^T,
//                                  bilinear prolongation of the reference's
//                                  interp maps, dense inverse on the coarsest
//                                  grid); Jacobi when there is a single level.
// Levels small enough for one workgroup are processed by workgroup 0 alone with
// workgroup barriers; larger levels use the whole grid and grid barriers.
#include "mcl_kernels.h"
#include "mcl_ops.h"

// ------------------------------------------------------------ set-up kernels

// FP32 copy of an FP64 ELL matrix, 1/diag, and the Gershgorin bound of D^-1 K
// (max over rows of sum_j |K_ij| / K_ii), written as float bits with atomicMax
// (positive floats order like their bit patterns).
__global__ void k_to_fp32(int n, int W, const double *__restrict__ val, const double *__restrict__ diag,
                          float *__restrict__ val32, float *__restrict__ dinv)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int s = 0; s < W; s++) val32[s * n + i] = (float)val[s * n + i];
        dinv[i] = (float)(1.0 / diag[i]);
    }
}

__global__ void k_gershgorin(int n, int W, const double *__restrict__ val, const double *__restrict__ diag,
                             unsigned *__restrict__ bound_bits)
{
    float m = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < W; k++) s += fabs(val[k * n + i]);
        m = fmaxf(m, (float)(s / diag[i]));
    }
    for (int o = MCL_WAVE / 2; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, MCL_WAVE));
    if ((threadIdx.x & (MCL_WAVE - 1)) == 0) atomicMax(bound_bits, __float_as_uint(m));
}

// Dense inverse of the coarsest K by in-place Gauss-Jordan elimination (K is
// symmetric positive definite: no pivoting).  One workgroup, n <= 64.
__global__ void k_coarse_inverse(int n, int W, const int *__restrict__ col, const int *__restrict__ ncol,
                                 const double *__restrict__ val, double *__restrict__ inv)
{
    __shared__ double A[64 * 64];
    __shared__ double rowk[64], colk[64];
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) A[t] = 0.0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        for (int s = 0; s < ncol[i]; s++) A[i * n + col[s * n + i]] = val[s * n + i];
    __syncthreads();
    for (int k = 0; k < n; k++) {
        double piv = A[k * n + k];
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            rowk[j] = (j == k) ? 1.0 / piv : A[k * n + j] / piv;
            colk[j] = A[j * n + k];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
            int i = t / n, j = t - i * n;
            if (i == k) {
                A[t] = rowk[j];
            } else {
                double f = colk[i];
                A[t] = (j == k) ? -f * rowk[k] : A[t] - f * rowk[j];
            }
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) inv[t] = A[t];
}

// coarse[c] = fine[c2f[c]]: nodal values injected from the finest level.
__global__ void k_inject(int nc, const int *__restrict__ c2f, const double *__restrict__ fine,
                         double *__restrict__ coarse)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) coarse[c] = fine[c2f[c]];
}

// --------------------------------------------------------------- level ops

struct Ex {
    int t, nt;
};

__device__ __forceinline__ float ell32(const MgLevelDev &A, int i, const float *__restrict__ x)
{
    float s = 0.f;
    for (int k = 0; k < A.WK; k++) s += A.val[k * A.n + i] * x[A.col[k * A.n + i]];
    return s;
}

// out = (first ? 0 : in) + w D^-1 (b - K in)
__device__ __forceinline__ void jacobi_sweep(const MgLevelDev &A, const float *__restrict__ in, float *__restrict__ out,
                                             float w, bool first, Ex ex)
{
    for (int i = ex.t; i < A.n; i += ex.nt) {
        float bi = A.b[i];
        out[i] = first ? w * A.dinv[i] * bi : in[i] + w * A.dinv[i] * (bi - ell32(A, i, in));
    }
}

__device__ __forceinline__ void residual(const MgLevelDev &A, const float *__restrict__ x, Ex ex)
{
    for (int i = ex.t; i < A.n; i += ex.nt) A.r[i] = A.b[i] - ell32(A, i, x);
}

// b_c = R r_f with R = P^T (restriction arrays live on the coarse level)
__device__ __forceinline__ void restrict_to(const MgLevelDev &C, const MgLevelDev &F, Ex ex)
{
    for (int c = ex.t; c < C.n; c += ex.nt) {
        float s = 0.f;
        for (int k = 0; k < C.rw; k++) {
            int u = C.rcol[k * C.n + c];
            if (u >= 0) s += C.rval[k * C.n + c] * F.r[u];
        }
        C.b[c] = s;
    }
}

// x_f += P x_c
__device__ __forceinline__ void prolong_add(const MgLevelDev &F, const MgLevelDev &C, Ex ex)
{
    for (int u = ex.t; u < F.n; u += ex.nt) {
        int m = F.pn[u];
        if (m == 0) continue;
        float s = 0.f;
        for (int k = 0; k < m; k++) {
            int l = F.pcol[k * F.n + u];
            if (l >= 0) s += C.x[l];
        }
        F.x[u] += s / (float)m;
    }
}

__device__ __forceinline__ bool lsync(bool grid, const GridSync &gs, unsigned &gen, bool ok)
{
    if (grid) return ok && grid_sync(gs, gen);
    __syncthreads();
    return ok;
}

// Pre-smoothing from x = 0 with an even number of sweeps, result in A.x.
__device__ __forceinline__ void smooth_pre(const MgLevelDev &A, float w, int sweeps, bool grid, Ex ex, const GridSync &gs,
                           unsigned &gen, bool &ok)
{
    float *cur = A.t, *nxt = A.x;
    if (sweeps % 2 == 1) { cur = A.x; nxt = A.t; }
    jacobi_sweep(A, nullptr, cur, w, true, ex);
    ok = lsync(grid, gs, gen, ok);
    for (int s = 1; s < sweeps; s++) {
        jacobi_sweep(A, cur, nxt, w, false, ex);
        ok = lsync(grid, gs, gen, ok);
        float *tmp = cur; cur = nxt; nxt = tmp;
    }
}

// Post-smoothing from the current A.x with an even number of sweeps.
__device__ __forceinline__ void smooth_post(const MgLevelDev &A, float w, int sweeps, bool grid, Ex ex, const GridSync &gs,
                            unsigned &gen, bool &ok)
{
    float *cur = A.x, *nxt = A.t;
    for (int s = 0; s < sweeps; s++) {
        jacobi_sweep(A, cur, nxt, w, false, ex);
        ok = lsync(grid, gs, gen, ok);
        float *tmp = cur; cur = nxt; nxt = tmp;
    }
}

// V-cycle on levels lo..top (all with the same execution mode), solving
// A_top x = b_top; level 0 is solved with the dense inverse.
__device__ __forceinline__ void vcycle_range(const MixedParams &P, const float *lam, int top, int lo, bool grid, Ex ex,
                             const GridSync &gs, unsigned &gen, bool &ok)
{
    const MgLevelDev *mg = P.mg;
    for (int l = top; l > lo; l--) {
        const MgLevelDev &A = mg[l];
        float w = P.omega / lam[l];
        smooth_pre(A, w, P.pre, grid, ex, gs, gen, ok);
        residual(A, A.x, ex);
        ok = lsync(grid, gs, gen, ok);
        restrict_to(mg[l - 1], A, ex);
        ok = lsync(grid, gs, gen, ok);
    }
    if (lo == 0) {
        const MgLevelDev &C = mg[0];
        for (int i = ex.t; i < C.n; i += ex.nt) {
            double s = 0.0;
            for (int j = 0; j < C.n; j++) s += P.cinv[i * C.n + j] * (double)C.b[j];
            C.x[i] = (float)s;
        }
    } else {
        const MgLevelDev &A = mg[lo];
        smooth_pre(A, P.omega / lam[lo], P.pre + P.post, grid, ex, gs, gen, ok);
    }
    ok = lsync(grid, gs, gen, ok);
    for (int l = lo + 1; l <= top; l++) {
        const MgLevelDev &A = mg[l];
        prolong_add(A, mg[l - 1], ex);
        ok = lsync(grid, gs, gen, ok);
        smooth_post(A, P.omega / lam[l], P.post, grid, ex, gs, gen, ok);
    }
}

// z = M r on the finest level: mg[top].b holds r, the result is in mg[top].x.
__device__ __forceinline__ void vcycle(const MixedParams &P, const float *lam, const GridSync &gs, unsigned &gen, bool &ok)
{
    const int top = P.nmg - 1;
    int lb = 0;  // level 0 is always processed by one workgroup
    for (int l = 0; l <= top; l++)
        if (P.mg[l].block) lb = l;
    Ex eg{(int)(blockIdx.x * blockDim.x + threadIdx.x), (int)(gridDim.x * blockDim.x)};
    Ex eb{(int)threadIdx.x, (int)blockDim.x};
    if (lb < top) {
        // grid-wide levels above the block section; a V-cycle over them whose
        // bottom (lb+1's coarse partner) is handled by workgroup 0
        const MgLevelDev *mg = P.mg;
        for (int l = top; l > lb; l--) {
            const MgLevelDev &A = mg[l];
            smooth_pre(A, P.omega / lam[l], P.pre, true, eg, gs, gen, ok);
            residual(A, A.x, eg);
            ok = lsync(true, gs, gen, ok);
            restrict_to(mg[l - 1], A, eg);
            ok = lsync(true, gs, gen, ok);
        }
        if (blockIdx.x == 0) vcycle_range(P, lam, lb, 0, false, eb, gs, gen, ok);
        ok = lsync(true, gs, gen, ok);
        for (int l = lb + 1; l <= top; l++) {
            const MgLevelDev &A = mg[l];
            prolong_add(A, mg[l - 1], eg);
            ok = lsync(true, gs, gen, ok);
            smooth_post(A, P.omega / lam[l], P.post, true, eg, gs, gen, ok);
        }
    } else {
        if (blockIdx.x == 0) vcycle_range(P, lam, top, 0, false, eb, gs, gen, ok);
        ok = lsync(true, gs, gen, ok);
    }
}

// ---------------------------------------------------------- FP32 inner CG

// Solve K32 e = r (r in P.cg_r, overwritten), e in P.cg_e, to |r| <= rtol |r0|.
// Returns the iteration count.
__device__ __forceinline__ int cg32(const MixedParams &P, const float *lam, const GridSync &gs, unsigned &gen, bool &ok,
                    dd (*lds)[MCL_BLOCK / MCL_WAVE])
{
    const MgLevelDev &A = P.mg[P.nmg - 1];
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    const int n = A.n;
    const bool use_mg = P.nmg > 1;
    float *r = P.cg_r, *e = P.cg_e, *d = P.cg_d, *q = P.cg_q;
    float *z = use_mg ? A.x : P.cg_z;
    // z = M r
    if (use_mg) {
        for (int i = gtid; i < n; i += gsize) A.b[i] = r[i];
        vcycle(P, lam, gs, gen, ok);
    } else {
        for (int i = gtid; i < n; i += gsize) z[i] = A.dinv[i] * r[i];
    }
    {
        dd acc[2] = {dd_zero(), dd_zero()};
        for (int i = gtid; i < n; i += gsize) {
            float zi = z[i], ri = r[i];
            d[i] = zi;
            e[i] = 0.f;
            acc[0] = dd_add_prod(acc[0], ri, zi);
            acc[1] = dd_add_prod(acc[1], ri, ri);
        }
        reduce_publish<2>(P.U.rb, 2, acc, lds);
    }
    ok = ok && grid_sync(gs, gen);
    double delta, rr0;
    {
        double o[2];
        reduce_collect<2>(P.U.rb, 2, gridDim.x, o, lds);
        delta = o[0];
        rr0 = o[1];
    }
    const double stop2 = (double)P.cg_rtol * (double)P.cg_rtol * rr0;
    int it = 0;
    while (ok && it < P.max_cg32 && rr0 > 0.0) {
        {
            dd acc[1] = {dd_zero()};
            for (int i = gtid; i < n; i += gsize) {
                float qi = ell32(A, i, d);
                q[i] = qi;
                acc[0] = dd_add_prod(acc[0], d[i], qi);
            }
            reduce_publish<1>(P.U.rb, 12, acc, lds);
        }
        ok = ok && grid_sync(gs, gen);
        double dq;
        {
            double o[1];
            reduce_collect<1>(P.U.rb, 12, gridDim.x, o, lds);
            dq = o[0];
        }
        const float alpha = (float)(delta / dq);
        double rr, rz;
        {
            dd acc[2] = {dd_zero(), dd_zero()};
            for (int i = gtid; i < n; i += gsize) {
                e[i] += alpha * d[i];
                float ri = r[i] - alpha * q[i];
                r[i] = ri;
                acc[0] = dd_add_prod(acc[0], ri, ri);
                if (!use_mg) {
                    float zi = A.dinv[i] * ri;
                    z[i] = zi;
                    acc[1] = dd_add_prod(acc[1], ri, zi);
                } else {
                    A.b[i] = ri;
                }
            }
            reduce_publish<2>(P.U.rb, 13, acc, lds);
        }
        ok = ok && grid_sync(gs, gen);
        {
            double o[2];
            reduce_collect<2>(P.U.rb, 13, gridDim.x, o, lds);
            rr = o[0];
            rz = o[1];
        }
        it++;
        if (rr <= stop2) break;
        if (use_mg) {
            vcycle(P, lam, gs, gen, ok);
            dd acc[1] = {dd_zero()};
            for (int i = gtid; i < n; i += gsize) acc[0] = dd_add_prod(acc[0], r[i], z[i]);
            reduce_publish<1>(P.U.rb, 15, acc, lds);
            ok = ok && grid_sync(gs, gen);
            double o[1];
            reduce_collect<1>(P.U.rb, 15, gridDim.x, o, lds);
            rz = o[0];
        }
        const float beta = (float)(rz / delta);
        delta = rz;
        for (int i = gtid; i < n; i += gsize) d[i] = z[i] + beta * d[i];
        ok = ok && grid_sync(gs, gen);
    }
    return it;
}

// Inner solve K x = b in FP64 by iterative refinement over cg32.
__device__ __forceinline__ int inner_mixed(const MixedParams &P, const float *lam, const double *__restrict__ b,
                           double *__restrict__ x, bool x_zero, const GridSync &gs, unsigned &gen, bool &ok,
                           dd (*lds)[MCL_BLOCK / MCL_WAVE], long long &spmv)
{
    const LevelDev &L = P.U.L;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    const int n = L.neq;
    double bn = 0.0;
    int total = 0;
    for (int ref = 0; ok && ref <= P.max_refine; ref++) {
        const bool zero = x_zero && ref == 0;
        {
            dd acc[2] = {dd_zero(), dd_zero()};
            for (int i = gtid; i < n; i += gsize) {
                double bi = b[i];
                double ri = zero ? bi : bi - ell_row(L.Kcol, L.Kval, L.Kncol[i], n, i, x);
                P.rr[i] = ri;
                acc[0] = dd_add_prod(acc[0], ri, ri);
                if (ref == 0) acc[1] = dd_add_prod(acc[1], bi, bi);
            }
            if (!zero) spmv++;
            reduce_publish<2>(P.U.rb, 0, acc, lds);
        }
        ok = ok && grid_sync(gs, gen);
        double rn;
        {
            double o[2];
            reduce_collect<2>(P.U.rb, 0, gridDim.x, o, lds);
            rn = sqrt(o[0]);
            if (ref == 0) bn = sqrt(o[1]);
        }
        if (rn <= P.inner_rtol * bn || rn == 0.0 || ref == P.max_refine) break;
        const double inv = 1.0 / rn;
        for (int i = gtid; i < n; i += gsize) P.cg_r[i] = (float)(P.rr[i] * inv);
        total += cg32(P, lam, gs, gen, ok, lds);
        for (int i = gtid; i < n; i += gsize) {
            double xi = zero ? 0.0 : x[i];
            x[i] = xi + rn * (double)P.cg_e[i];
        }
        if (L.nh > 0) {
            ok = ok && grid_sync(gs, gen);
            constrain_v(L, x, gtid, gsize);
        }
        ok = ok && grid_sync(gs, gen);
    }
    return total;
}

// --------------------------------------------------------- persistent Uzawa

__global__ void __launch_bounds__(MCL_BLOCK) k_uzawa_mixed(MixedParams P)
{
    __shared__ dd lds[4][MCL_BLOCK / MCL_WAVE];
    const UzawaParams &U = P.U;
    const LevelDev &L = U.L;
    const float *lam = P.lam;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    const int neq = L.neq, nel = L.nel;
    unsigned gen = __hip_atomic_load(U.gs.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    bool ok = true;
    long long spmv = 0, inner = 0;
    for (int e = gtid; e < nel; e += gsize) {
        double s = 0.0;
        int lev = L.elev[e];
        for (int t = 0; t < 8; t++) {
            int r = L.gt_row[t * nel + e];
            if (r < 0) break;
            double g = L.gtab[lev * 8 + L.gt_p[t * nel + e]];
            s += g * g / L.Kdiag[r];
        }
        U.BAB[e] = s;
    }
    // momentum defect of the initial guess: Ah = F - G p - K d
    for (int i = gtid; i < neq; i += gsize) {
        double u = ell_row(L.Kcol, L.Kval, L.Kncol[i], neq, i, U.d);
        U.Ahv[i] = L.F[i] - g_row(L, i, U.p) - u;
        U.u1[i] = 0.0;
    }
    spmv++;
    ok = ok && grid_sync(U.gs, gen);
    inner += inner_mixed(P, lam, U.Ahv, U.u1, true, U.gs, gen, ok, lds, spmv);
    for (int i = gtid; i < neq; i += gsize) U.d[i] += U.u1[i];
    ok = ok && grid_sync(U.gs, gen);
    double *r0 = U.r0, *r1 = U.r1, *r2 = U.r2, *z0 = U.z0, *z1 = U.z1, *s1 = U.s1, *s2 = U.s2;
    for (int e = gtid; e < nel; e += gsize) r1[e] = gt_elem(L, e, U.d);
    double dv = 1.0, dp = 1.0;
    int cycle = 0;
    while (ok && (dv > U.citcom_accu || (dp > U.citcom_accu && U.check_v_only == 0)) &&
           cycle < U.max_citcom_cycle) {
        {
            dd acc[2] = {dd_zero(), dd_zero()};
            for (int e = gtid; e < nel; e += gsize) {
                double z = U.schur ? r1[e] / U.BAB[e] : U.BAB[e] * r1[e];
                z1[e] = z;
                acc[0] = dd_add_prod(acc[0], r1[e], z);
                if (cycle > 0) acc[1] = dd_add_prod(acc[1], r0[e], z0[e]);
            }
            reduce_publish<2>(U.rb, 4, acc, lds);
        }
        ok = ok && grid_sync(U.gs, gen);
        double r1z1, delta = 0.0;
        {
            double o[2];
            reduce_collect<2>(U.rb, 4, gridDim.x, o, lds);
            r1z1 = o[0];
            if (cycle > 0) delta = r1z1 / o[1];
        }
        for (int e = gtid; e < nel; e += gsize) s2[e] = cycle == 0 ? z1[e] : z1[e] + delta * s1[e];
        ok = ok && grid_sync(U.gs, gen);
        for (int i = gtid; i < neq; i += gsize) {
            U.Ahv[i] = g_row(L, i, s2);
            U.u1[i] = 0.0;
        }
        ok = ok && grid_sync(U.gs, gen);
        inner += inner_mixed(P, lam, U.Ahv, U.u1, true, U.gs, gen, ok, lds, spmv);
        {
            dd acc[1] = {dd_zero()};
            for (int e = gtid; e < nel; e += gsize) {
                double a = gt_elem(L, e, U.u1);
                U.Ahp[e] = a;
                acc[0] = dd_add_prod(acc[0], s2[e], a);
            }
            reduce_publish<1>(U.rb, 6, acc, lds);
        }
        ok = ok && grid_sync(U.gs, gen);
        double alpha;
        {
            double o[1];
            reduce_collect<1>(U.rb, 6, gridDim.x, o, lds);
            alpha = r1z1 / o[0];
        }
        {
            dd acc[4] = {dd_zero(), dd_zero(), dd_zero(), dd_zero()};
            for (int e = gtid; e < nel; e += gsize) {
                r2[e] = r1[e] - alpha * U.Ahp[e];
                double pe = U.p[e] + alpha * s2[e];
                U.p[e] = pe;
                acc[0] = dd_add_prod(acc[0], s2[e], s2[e]);
                acc[1] = dd_add_prod(acc[1], pe, pe);
            }
            for (int i = gtid; i < neq; i += gsize) {
                double u = U.u1[i];
                double di = U.d[i] - alpha * u;
                U.d[i] = di;
                acc[2] = dd_add_prod(acc[2], u, u);
                acc[3] = dd_add_prod(acc[3], di, di);
            }
            reduce_publish<4>(U.rb, 8, acc, lds);
        }
        ok = ok && grid_sync(U.gs, gen);
        {
            double o[4];
            reduce_collect<4>(U.rb, 8, gridDim.x, o, lds);
            dp = alpha * sqrt(o[0] / (1.0e-32 + o[1]));
            dv = alpha * sqrt(o[2] / (1.0e-32 + o[3]));
        }
        cycle++;
        double *sh;
        sh = s1; s1 = s2; s2 = sh;
        sh = r0; r0 = r1; r1 = r2; r2 = sh;
        sh = z0; z0 = z1; z1 = sh;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        U.stats->cycles = cycle;
        U.stats->dv = dv;
        U.stats->dp = dp;
        U.stats->inner = inner;
        U.stats->spmv = spmv;
        U.stats->aborted = ok ? 0 : 1;
        if (!(dv == dv) || !(dp == dp)) U.stats->nan_flag = 1;
    }
}
