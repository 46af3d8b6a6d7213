// Mixed-precision Stokes solve of libmcl_hip.so (gfx950).
//
// The Uzawa outer iteration (Solver.c:14-249) runs in FP64 exactly as in the
// FP64 path; only the inner velocity solves K u = b change: FP64 conjugate
// gradients (FP64 ELL products, FP64 recurrences, compensated dot products)
// preconditioned by an FP32 geometric-multigrid V-cycle over the quadtree
// levels (damped-Jacobi smoothing, restriction P^T, prolongation P from the
// reference's interp maps, dense inverse on the coarsest grid); Jacobi when
// there is a single level.  The preconditioner is symmetric (equal pre- and
// post-sweeps, R = P^T), so plain CG applies; FP32 only perturbs it.
// Levels small enough for one workgroup are processed by workgroup 0 alone with
// workgroup barriers; larger levels use the whole grid and grid barriers.
#include "mcl_kernels.h"
#include "mcl_ops.h"

// Optional phase timers (build with -DMCL_PHASE_TIMERS=1): thread 0 of
// workgroup 0 charges the 100 MHz wall clock since its previous stamp to a
// phase; grid barriers keep every workgroup in step, so workgroup 0's split
// is the grid's.  Phases: 0 d'/q pass, 1 x/r update (+ start), 2 V-cycle down
// (grid levels), 7 restriction into the LDS levels, 8 LDS levels down,
// 9 coarsest solve, 3 LDS levels up, 4 V-cycle up (grid levels), 5 last
// sweep + r.z, 6 Uzawa outer work.
#if MCL_PHASE_TIMERS
__shared__ long long mcl_ph[MCL_PHASES + 1];
#define MCL_PHASE(k)                                                           \
    do {                                                                       \
        if (blockIdx.x == 0 && threadIdx.x == 0) {                             \
            long long t_ = (long long)wall_clock64();                          \
            mcl_ph[k] += t_ - mcl_ph[MCL_PHASES];                              \
            mcl_ph[MCL_PHASES] = t_;                                           \
        }                                                                      \
    } while (0)
#else
#define MCL_PHASE(k) \
    do {             \
    } while (0)
#endif

// ------------------------------------------------------------ set-up kernels

// FP32 copy of an FP64 ELL matrix, 1/diag, and the Gershgorin bound of D^-1 K
// (max over rows of sum_j |K_ij| / K_ii), written as float bits with atomicMax
// (positive floats order like their bit patterns).  One launch covers every
// multigrid level: each workgroup takes MCL_BLOCK rows of one level.
__global__ void __launch_bounds__(MCL_BLOCK) k_prepare_fp32(const PrepLevel *__restrict__ lv, int nlev)
{
    int lk = 0;
    while (lk + 1 < nlev && (int)blockIdx.x >= lv[lk + 1].blk0) lk++;
    const PrepLevel &A = lv[lk];
    const int i = ((int)blockIdx.x - A.blk0) * MCL_BLOCK + (int)threadIdx.x;
    float m = 0.f;
    if (i < A.n) {
        double s = 0.0;
        for (int k = 0; k < A.W; k++) {
            const double v = A.val[k * A.n + i];
            A.val32[k * A.n + i] = (float)v;
            s += fabs(v);
        }
        const double d = A.diag[i];
        A.dinv[i] = (float)(1.0 / d);
        m = (float)(s / d);
    }
    for (int o = MCL_WAVE / 2; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, MCL_WAVE));
    if ((threadIdx.x & (MCL_WAVE - 1)) == 0) atomicMax(A.bound, __float_as_uint(m));
}

// Augmented Lagrangian of the mixed solve: K_al = K + gamma G W G^T with W =
// diag(w), w_e = 1 / BAB_e of the plain K (BAB_e = sum_i G_ie^2 / K_ii,
// Solver.c:54-64, the Uzawa preconditioner's diagonal).  G W G^T couples the
// velocity rows of one element, so K_al keeps K's ELL pattern.  The term
// vanishes on divergence-free velocities, so (u, p) solves K u + G p = F,
// G^T u = 0 exactly when (u, p - gamma W G^T u) solves the augmented system;
// its Schur complement S_al = G^T K_al^-1 G has S_al^-1 = S^-1 + gamma W^-1,
// a narrower spectrum relative to BAB, so the Uzawa iteration needs fewer
// steps (its inner solves somewhat more).

// w_e, one thread per element of every level
__global__ void __launch_bounds__(MCL_BLOCK) k_al_weights(const AlLevel *__restrict__ lv, int nlev)
{
    int lk = 0;
    while (lk + 1 < nlev && (int)blockIdx.x >= lv[lk + 1].eblk0) lk++;
    const AlLevel &A = lv[lk];
    const LevelDev &L = A.L;
    const int e = ((int)blockIdx.x - A.eblk0) * MCL_BLOCK + (int)threadIdx.x;
    if (e >= L.nel) return;
    double s = 0.0;
    const int lev = L.elev[e];
    for (int t = 0; t < 8; t++) {
        const int r = L.gt_row[t * L.nel + e];
        if (r < 0) break;
        const double g = L.gtab[lev * 8 + L.gt_p[t * L.nel + e]];
        s += g * g / L.Kdiag[r];
    }
    A.w[e] = s > 0.0 ? 1.0 / s : 0.0;
}

// K_al rows, one thread per row of every level: the row of K, then for each
// element at the row's node (ascending) and each of its free dofs q the term
// gamma w_e G_Pe G_qe in the slot of q (Kslot); diag = the diagonal slot
__global__ void __launch_bounds__(MCL_BLOCK) k_al_rows(const AlLevel *__restrict__ lv, int nlev)
{
    int lk = 0;
    while (lk + 1 < nlev && (int)blockIdx.x >= lv[lk + 1].rblk0) lk++;
    const AlLevel &A = lv[lk];
    const LevelDev &L = A.L;
    const int n = L.neq;
    const int P = ((int)blockIdx.x - A.rblk0) * MCL_BLOCK + (int)threadIdx.x;
    if (P >= n) return;
    for (int s = 0; s < L.WK; s++) A.val[s * n + P] = L.Kval[s * n + P];
    const double gamma = *A.gamma;
    if (gamma == 0.0) {
        A.diag[P] = L.Kdiag[P];
        return;
    }
    const int nd = L.veq_node[P], ip = L.veq_dof[P];
    int sd = -1;
    for (int k = 0; k < L.p2n_cnt[nd]; k++) {
        const int e = L.p2n_el[k * L.nno + nd], a = L.p2n_a[k * L.nno + nd];
        const double *g = L.gtab + L.elev[e] * 8;
        const double c = gamma * A.w[e] * g[2 * a + ip];
        for (int q = 0; q < 8; q++) {
            const int s = L.Kslot[(k * 8 + q) * n + P];
            if (s < 0) continue;
            if (q == 2 * a + ip) sd = s;
            A.val[s * n + P] += c * g[q];
        }
    }
    A.diag[P] = sd >= 0 ? A.val[sd * n + P] : L.Kdiag[P];
}

// Warm start of a mixed solve by linear extrapolation in model time: with
// x1 the last solution (in d, p; at time t1) and x0 the one before (in the
// history array h; at t0), the new solve at time t starts from
// x1 + r (x1 - x0), r = (t - t1) / (t1 - t0), as long as 0 < r <= rmax; the
// velocity changes smoothly from step to step (at 256x256, Ra 1e6, step 21:
// initial velocity error 4.8e-2 from x1 alone, 2.5e-3 extrapolated).  Then
// h = x1.  Only from the (after + 1)-th solve since the history was reset:
// a cold start's first steps are a fast transient (the flow accelerates from
// rest, dt shrinks by up to half a step to the next) where the straight line
// through the last two solutions often overshoots (measured at 256x256,
// Ra 1e6: steps 4 and 7 took 14 % and 40 % longer with it; on 96 seeded
// instances of sides 32-512, Ra up to 1e7, their steps 2-4 took 0.43, 0.28
// and 0.26 s longer in all).  (Quadratic extrapolation from three solutions starts closer
// still, 1.3e-4, but the Q1-P0 checkerboard pressure mode, which the Uzawa
// iteration barely corrects, then grows from step to step: from step ~40 the
// Uzawa counts oscillate with period 3 between 5 and 130.)
__global__ void __launch_bounds__(MCL_BLOCK) k_extrapolate(int nv, int np, double *__restrict__ d,
                                                           double *__restrict__ p, double *__restrict__ h,
                                                           const double *__restrict__ ws,
                                                           const double *__restrict__ model_time, int after,
                                                           double rmax)
{
    const double t = *model_time, t1 = ws[1], t0 = ws[2];
    const double r = ws[0] >= (double)max(after, 2) && t1 > t0 ? (t - t1) / (t1 - t0) : 0.0;
    const bool ex = r > 0.0 && r <= rmax;
    const int n = nv + np, gs = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gs) {
        double *x = i < nv ? d + i : p + (i - nv);
        const double x1 = *x;
        if (ex) *x = x1 + r * (x1 - h[i]);
        h[i] = x1;
    }
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
// Jacobi sweeps on a level alternate between two buffers: the iterate after k
// sweeps from x = 0 is sweep_buf(A, k).  A V-cycle applies `pre` sweeps on the
// way down and `pre` more on the way up (equal counts and R = P^T keep the
// preconditioner symmetric), so every level's result is sweep_buf(A, 2 pre).

// Rows a thread visits on a level: begin, end, stride.
struct Ex {
    int b, e, s;
};

// Contiguous row ownership of the V-cycle's grid levels: workgroup g owns rows
// [g R, (g+1) R) with R = ceil(n / gridDim.x).  The equations are numbered
// node line by node line, so a workgroup's rows are about one line of the
// mesh and it exchanges data with the few workgroups owning the lines next to
// it (the host's neighbour lists use the same formula).
__device__ __forceinline__ Ex own_rows(int n)
{
    const int R = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    const int b0 = (int)blockIdx.x * R;
    return Ex{b0 + (int)threadIdx.x, min(b0 + R, n), (int)blockDim.x};
}

// all rows of a level, by one workgroup (the LDS levels)
__device__ __forceinline__ Ex wg_rows(int n) { return Ex{(int)threadIdx.x, n, (int)blockDim.x}; }

// Address spaces of the V-cycle's data: grid levels live in global memory
// (AS 1), the coarse levels' vectors and matrices in LDS (AS 3).  Typed
// pointers make the loads global_load / ds_read instead of flat accesses that
// wait on both memory counters.  The level descriptors are read from the
// workgroup's LDS copy; the pointers they hold are cast to AS 1.
#define MCL_G __attribute__((address_space(1)))
#define MCL_L __attribute__((address_space(3)))
template <typename T>
__device__ __forceinline__ MCL_G T *gp(T *p)
{
    return (MCL_G T *)p;
}

// Vector accesses of the V-cycle.  CO = true for vectors of the grid levels
// between neighbour waits: agent-scope atomic loads and stores (coherent
// across the XCDs' L2 caches, see nbr_sync_coherent); CO = false for LDS
// vectors and plain global accesses.
#ifndef MCL_VCYCLE_COHERENT
#define MCL_VCYCLE_COHERENT 1
#endif
constexpr bool kCO = MCL_VCYCLE_COHERENT != 0;

template <bool CO, typename Ptr>
__device__ __forceinline__ float vld(Ptr p)
{
    static_assert(sizeof(*p) == sizeof(float), "vld reads float vectors (vldd: double)");
    if constexpr (CO) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}

template <bool CO, typename Ptr>
__device__ __forceinline__ void vst(Ptr p, float v)
{
    static_assert(sizeof(*p) == sizeof(float), "vst writes float vectors (vstd: double)");
    if constexpr (CO) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

template <bool CO, typename Ptr>
__device__ __forceinline__ double vldd(Ptr p)
{
    static_assert(sizeof(*p) == sizeof(double), "vldd reads double vectors");
    if constexpr (CO) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}

template <bool CO, typename Ptr>
__device__ __forceinline__ void vstd(Ptr p, double v)
{
    static_assert(sizeof(*p) == sizeof(double), "vstd writes double vectors");
    if constexpr (CO) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

// A grid level as the level operations see it.
struct GLevel {
    int n, WK, rw;
    const MCL_G int *col;
    const MCL_G float *val, *dinv;
    const MCL_G int *pn, *pcol, *rcol;
    const MCL_G float *rval;
    MCL_G float *x, *b, *r, *t;
    int fuse_up, fuse_down, deep_up, deep_down;
    const MCL_G int *hu_ptr, *hu_row, *hd_ptr, *hd_row, *h2_ptr, *h2_row, *th_ptr, *th_row, *uh_ptr, *uh_row;
    const MCL_G unsigned short *kl, *hm, *mt, *mf;
    const MCL_G short *rl;
};

// A coarse level as seen by one workgroup: vectors x, t, b (3 n floats at
// lds_off), 1/diag(K) (n floats at lds_off / 3) and the FP32 K with 16-bit
// columns (WK n entries at moff) in LDS, copied from HBM once per launch; the
// transfer maps stay in HBM.
struct LdsLevel {
    int n, WK, rw;
    const MCL_L unsigned short *col;
    const MCL_L float *val, *dinv;
    const MCL_G int *pn, *pcol, *rcol;
    const MCL_G float *rval;
    MCL_L float *x, *b, *r, *t;
};

// the workgroup's LDS blocks: level descriptors, Gershgorin bounds (both read
// in every phase, so kept out of the grid barriers' cache invalidations), and
// the coarse levels' vectors and matrices
struct MgLds {
    const MCL_L MgLevelDev *lev;
    const MCL_L float *lam;
    MCL_L float *vec, *dinv, *val;
    MCL_L unsigned short *col;
    MCL_L float *halo;       // [MCL_HALO_LDS] buffer of the fused phases
};

template <typename LV>
__device__ __forceinline__ auto sweep_buf(const LV &A, int k)
{
    return (k & 1) ? A.x : A.t;
}

__device__ __forceinline__ GLevel glevel(const MCL_L MgLevelDev *G)
{
    GLevel A;
    A.n = G->n;
    A.WK = G->WK;
    A.rw = G->rw;
    A.col = gp(G->col);
    A.val = gp(G->val);
    A.dinv = gp(G->dinv);
    A.pn = gp(G->pn);
    A.pcol = gp(G->pcol);
    A.rcol = gp(G->rcol);
    A.rval = gp(G->rval);
    A.x = gp(G->x);
    A.b = gp(G->b);
    A.r = gp(G->r);
    A.t = gp(G->t);
    A.fuse_up = G->fuse_up;
    A.fuse_down = G->fuse_down;
    A.deep_up = G->deep_up;
    A.deep_down = G->deep_down;
    A.hu_ptr = gp(G->hu_ptr);
    A.hu_row = gp(G->hu_row);
    A.hd_ptr = gp(G->hd_ptr);
    A.hd_row = gp(G->hd_row);
    A.h2_ptr = gp(G->h2_ptr);
    A.h2_row = gp(G->h2_row);
    A.th_ptr = gp(G->th_ptr);
    A.th_row = gp(G->th_row);
    A.uh_ptr = gp(G->uh_ptr);
    A.uh_row = gp(G->uh_row);
    A.kl = gp(G->kl);
    A.hm = gp(G->hm);
    A.mt = gp(G->mt);
    A.mf = gp(G->mf);
    A.rl = gp(G->rl);
    return A;
}

// the residual reuses the sweep buffer that is free after the pre-sweeps
__device__ __forceinline__ LdsLevel lds_level(const MgLds &S, int l, int pre)
{
    const MCL_L MgLevelDev *G = S.lev + l;
    LdsLevel A;
    A.n = G->n;
    A.WK = G->WK;
    A.rw = G->rw;
    A.col = S.col + G->moff;
    A.val = S.val + G->moff;
    A.dinv = S.dinv + G->lds_off / 3;
    A.pn = gp(G->pn);
    A.pcol = gp(G->pcol);
    A.rcol = gp(G->rcol);
    A.rval = gp(G->rval);
    A.x = S.vec + G->lds_off;
    A.t = A.x + A.n;
    A.b = A.t + A.n;
    A.r = sweep_buf(A, pre - 1);
    return A;
}

// Row i of an FP32 ELL matrix (WK slots, slot-major over n rows, columns col)
// times x.  Rows of width 18 (the bilinear-quad velocity stencil) issue all
// their loads at once, other widths nine entries per round trip; the sum keeps
// the slot order.
template <bool CO, typename CP, typename VP, typename XP>
__device__ __forceinline__ float ell32c(int WK, int n, CP col, VP val, int i, XP x)
{
    float s = 0.f;
    if (WK == 18) {
        int c[18];
        float v[18], g[18];
#pragma unroll
        for (int u = 0; u < 18; u++) {
            c[u] = col[u * n + i];
            v[u] = val[u * n + i];
        }
#pragma unroll
        for (int u = 0; u < 18; u++) g[u] = vld<CO>(x + c[u]);
#pragma unroll
        for (int u = 0; u < 18; u++) s += v[u] * g[u];
        return s;
    }
    int k = 0;
    for (; k + 9 <= WK; k += 9) {
        int c[9];
        float v[9], g[9];
#pragma unroll
        for (int u = 0; u < 9; u++) {
            c[u] = col[(k + u) * n + i];
            v[u] = val[(k + u) * n + i];
        }
#pragma unroll
        for (int u = 0; u < 9; u++) g[u] = vld<CO>(x + c[u]);
#pragma unroll
        for (int u = 0; u < 9; u++) s += v[u] * g[u];
    }
    for (; k < WK; k++) s += val[k * n + i] * vld<CO>(x + col[k * n + i]);
    return s;
}

template <bool CO, typename LV, typename XP>
__device__ __forceinline__ float ell32(const LV &A, int i, XP x)
{
    return ell32c<CO>(A.WK, A.n, A.col, A.val, i, x);
}

// sweep k >= 2: x_k = x_{k-1} + w D^-1 (b - K x_{k-1})
template <bool CO, typename LV>
__device__ __forceinline__ void jacobi(const LV &A, int k, float w, Ex ex)
{
    const auto in = sweep_buf(A, k - 1);
    const auto out = sweep_buf(A, k);
    for (int i = ex.b; i < ex.e; i += ex.s)
        vst<CO>(out + i, vld<CO>(in + i) + w * A.dinv[i] * (vld<CO>(A.b + i) - ell32<CO>(A, i, in)));
}

// r = b - K x on a level
template <bool CO, typename LV, typename XP>
__device__ __forceinline__ void residual(const LV &A, XP x, Ex ex)
{
    for (int i = ex.b; i < ex.e; i += ex.s) vst<CO>(A.r + i, vld<CO>(A.b + i) - ell32<CO>(A, i, x));
}

// Coarse right-hand side b_c = sum_k R_ck r_u (R = P^T, arrays on the coarse
// level), fused with the coarse level's first sweep x_c = w_c D^-1 b_c unless
// the coarse level is solved directly.  CF: access mode of the fine level's
// vectors, CC: of the coarse level's.  Nine entries per round trip (all map
// loads, then all gathers); the sum keeps the slot order.
template <bool CF, bool CC, typename LC, typename LF>
__device__ __forceinline__ void restrict_to(const LC &C, const LF &F, bool sweep, float wc, Ex ex)
{
    for (int c = ex.b; c < ex.e; c += ex.s) {
        float s = 0.f;
        for (int k0 = 0; k0 < C.rw; k0 += 9) {
            int u[9];
            float v[9], g[9];
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const bool in = k0 + k < C.rw;
                u[k] = in ? C.rcol[(k0 + k) * C.n + c] : -1;
                v[k] = in ? C.rval[(k0 + k) * C.n + c] : 0.f;
            }
#pragma unroll
            for (int k = 0; k < 9; k++) g[k] = u[k] >= 0 ? vld<CF>(F.r + u[k]) : 0.f;
#pragma unroll
            for (int k = 0; k < 9; k++)
                if (u[k] >= 0) s += v[k] * g[k];
        }
        vst<CC>(C.b + c, s);
        if (sweep) vst<CC>(C.x + c, wc * C.dinv[c] * s);
    }
}

// x_f += P x_c: fine row u takes the mean of its pn[u] coarse rows (pcol has
// four slots, -1 past pn[u] and for fixed coarse rows; all loads issued at once)
template <bool CF, bool CC, typename LF, typename XF, typename XC>
__device__ __forceinline__ void prolong_add(const LF &F, XF xf, XC xc, Ex ex)
{
    for (int u = ex.b; u < ex.e; u += ex.s) {
        const int m = F.pn[u];
        int l[4];
        float g[4];
#pragma unroll
        for (int k = 0; k < 4; k++) l[k] = F.pcol[k * F.n + u];
#pragma unroll
        for (int k = 0; k < 4; k++) g[k] = (k < m && l[k] >= 0) ? vld<CC>(xc + l[k]) : 0.f;
        if (m == 0) continue;
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < m && l[k] >= 0) s += g[k];
        vst<CF>(xf + u, vld<CF>(xf + u) + s / (float)m);
    }
}

// Restriction of the residual held in LDS (rF, indexed through C.rl) to the
// workgroup's coarse rows, with the coarse first sweep (restrict_to's order).
template <typename LB>
__device__ __forceinline__ void restrict_from_lds(const GLevel &C, LB rF, float wc)
{
    const Ex ex = own_rows(C.n);
    for (int c = ex.b; c < ex.e; c += ex.s) {
        float s = 0.f;
        for (int k0 = 0; k0 < C.rw; k0 += 9) {
            int l[9];
            float v[9];
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const bool in = k0 + k < C.rw;
                l[k] = in ? C.rl[(k0 + k) * C.n + c] : -1;
                v[k] = in ? C.rval[(k0 + k) * C.n + c] : 0.f;
            }
#pragma unroll
            for (int k = 0; k < 9; k++)
                if (l[k] >= 0) s += v[k] * rF[l[k]];
        }
        vst<kCO>(C.b + c, s);
        vst<kCO>(C.x + c, wc * C.dinv[c] * s);
    }
}

// Fused down-leg phase of grid level F (coarse grid level C): the residual
// r = b - K x_pre on the fine rows this workgroup's coarse rows restrict from
// (its own and a halo, into LDS), then b_c = R r and the coarse first sweep
// x_c = w_c D_c^-1 b_c on its coarse rows.  Same operations in the same order
// as residual + restrict_to, with one wait fewer.
template <typename LB>
__device__ __forceinline__ void fused_down(const GLevel &F, const GLevel &C, LB buf, float wc, int pre)
{
    const int h0 = F.hd_ptr[blockIdx.x], nh = F.hd_ptr[blockIdx.x + 1] - h0;
    const auto x = sweep_buf(F, pre);
    for (int t = threadIdx.x; t < nh; t += blockDim.x) {
        const int u = F.hd_row[h0 + t];
        buf[t] = vld<kCO>(F.b + u) - ell32<kCO>(F, u, x);
    }
    __syncthreads();
    restrict_from_lds(C, buf, wc);
}

// Fused up-leg phase of grid level F: x' = x_pre + P x_c on the workgroup's
// rows and its halo (into LDS; x_c the coarse level's result, CC its access
// mode), then the first post-sweep x_{pre+1} = x' + w D^-1 (b - K x') on its
// rows with K's columns mapped into the LDS copy.  Same operations in the
// same order as prolong_add + jacobi, with one wait fewer.
template <bool CC, typename XC, typename LB>
__device__ __forceinline__ void fused_up(const GLevel &F, XC xc, LB buf, float w, int pre)
{
    const int R = (F.n + (int)gridDim.x - 1) / (int)gridDim.x, b0 = (int)blockIdx.x * R;
    const int no = max(0, min(R, F.n - b0));
    const int h0 = F.hu_ptr[blockIdx.x], nh = F.hu_ptr[blockIdx.x + 1] - h0;
    const auto xin = sweep_buf(F, pre);
    for (int t = threadIdx.x; t < R + nh; t += blockDim.x) {
        if (t >= no && t < R) continue;
        const int u = t < R ? b0 + t : F.hu_row[h0 + t - R];
        const int m = F.pn[u];
        int lc[4];
        float g[4];
#pragma unroll
        for (int k = 0; k < 4; k++) lc[k] = F.pcol[k * F.n + u];
#pragma unroll
        for (int k = 0; k < 4; k++) g[k] = (k < m && lc[k] >= 0) ? vld<CC>(xc + lc[k]) : 0.f;
        float v = vld<kCO>(xin + u);
        if (m != 0) {
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (k < m && lc[k] >= 0) s += g[k];
            v = v + s / (float)m;
        }
        buf[t] = v;
    }
    __syncthreads();
    const auto out = sweep_buf(F, pre + 1);
    for (int i = b0 + (int)threadIdx.x; i < b0 + no; i += blockDim.x)
        vst<kCO>(out + i,
                 buf[i - b0] + w * F.dinv[i] * (vld<kCO>(F.b + i) - ell32c<false>(F.WK, F.n, F.kl, F.val, i, buf)));
}

// Row `row` of F's FP32 K times a vector held in LDS: the columns come from a
// local map (slot k of local row t at map[k nloc + t]), the values from the
// matrix; the sum in slot order, as ell32.
template <typename MP, typename XP>
__device__ __forceinline__ float ell32_map(const GLevel &F, MP map, int nloc, int t, int row, XP x)
{
    float s = 0.f;
    if (F.WK == 18) {
        int c[18];
        float v[18], g[18];
#pragma unroll
        for (int u = 0; u < 18; u++) {
            c[u] = map[u * nloc + t];
            v[u] = F.val[u * F.n + row];
        }
#pragma unroll
        for (int u = 0; u < 18; u++) g[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < 18; u++) s += v[u] * g[u];
        return s;
    }
    int k = 0;
    for (; k + 9 <= F.WK; k += 9) {
        int c[9];
        float v[9], g[9];
#pragma unroll
        for (int u = 0; u < 9; u++) {
            c[u] = map[(k + u) * nloc + t];
            v[u] = F.val[(k + u) * F.n + row];
        }
#pragma unroll
        for (int u = 0; u < 9; u++) g[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < 9; u++) s += v[u] * g[u];
    }
    for (; k < F.WK; k++) s += F.val[k * F.n + row] * x[map[k * nloc + t]];
    return s;
}

// A workgroup's rows O = [b0, b0 + no) of a grid level (R = own_rows' block).
struct Own {
    int R, b0, no;
};

__device__ __forceinline__ Own own_block(int n)
{
    const int R = (n + (int)gridDim.x - 1) / (int)gridDim.x, b0 = (int)blockIdx.x * R;
    return Own{R, b0, max(0, min(R, n - b0))};
}

// ------------------------------------ one-row-per-thread two-sweep phases
// On the coarse grid levels every list of a workgroup fits its MCL_BLOCK
// threads, so a thread holds at most one row of each.  These forms issue all
// loads of a phase that do not wait on another workgroup at its start -- the
// K rows' local columns and values, 1/diag, right-hand sides, transfer maps --
// right after the row indices, so the phase costs three dependent memory
// round trips instead of two per list; a K row needed by two sweeps is loaded
// once.  Same operations in the same order as the general forms below.

// a K row of width 18 in registers: its columns as local indices and values
struct KRow18 {
    int c[18];
    float v[18];
};

__device__ __forceinline__ void krow_load(KRow18 &k, const MCL_G unsigned short *map, int ms, int mi,
                                          const GLevel &F, int row)
{
#pragma unroll
    for (int u = 0; u < 18; u++) {
        k.c[u] = map[u * ms + mi];
        k.v[u] = F.val[u * F.n + row];
    }
}

template <typename XP>
__device__ __forceinline__ float krow_dot(const KRow18 &k, XP x)
{
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < 18; u++) s += k.v[u] * x[k.c[u]];
    return s;
}

// deep_down for lists of at most MCL_BLOCK rows (false: not applicable)
template <typename LB>
__device__ __forceinline__ bool deep_down1(const GLevel &F, const GLevel &C, LB buf, float wf, float wc)
{
    const Own o = own_block(F.n), oc = own_block(C.n);
    const int g = blockIdx.x;
    const int t0 = F.th_ptr[g], nth = F.th_ptr[g + 1] - t0;
    const int u0 = F.uh_ptr[g], nuh = F.uh_ptr[g + 1] - u0;
    const int f0 = F.hd_ptr[g], nf = F.hd_ptr[g + 1] - f0;
    const int nT = o.R + nth, nU = nT + nuh;
    if (F.WK != 18 || C.rw > 9 || nU > MCL_BLOCK || nf > MCL_BLOCK || oc.R > MCL_BLOCK) return false;
    const int t = threadIdx.x;
    const auto xU = buf;
    const auto xT = buf + nU;
    const auto rF = xT + nT;
    // this thread's row of U (T is its prefix) and of F
    const bool inU = t < nU && (t >= o.R || t < o.no), inT = inU && t < nT, inF = t < nf;
    const int ru = inU ? (t < o.R ? o.b0 + t : t < nT ? F.th_row[t0 + t - o.R] : F.uh_row[u0 + t - nT]) : 0;
    const int rf = inF ? F.hd_row[f0 + t] : 0;
    // every load that does not wait on another workgroup
    KRow18 kt, kf;
    if (inT) krow_load(kt, F.mt + (size_t)((long long)g * o.R + t0) * F.WK, nT, t, F, ru);
    if (inF) krow_load(kf, F.mf + (size_t)f0 * F.WK, nf, t, F, rf);
    const float dT = inT ? F.dinv[ru] : 0.f;
    const float bT = inT ? vld<kCO>(F.b + ru) : 0.f;
    const float bF = inF ? vld<kCO>(F.b + rf) : 0.f;
    const float xu = inU ? vld<kCO>(sweep_buf(F, 1) + ru) : 0.f;
    const int c = oc.b0 + t;
    const bool inC = t < oc.no;
    int l[9];
    float rv[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const bool in = inC && k < C.rw;
        l[k] = in ? C.rl[k * C.n + c] : -1;
        rv[k] = in ? C.rval[k * C.n + c] : 0.f;
    }
    const float dC = inC ? C.dinv[c] : 0.f;
    if (inU) xU[t] = xu;
    __syncthreads();
    if (inT) {
        const float v = xU[t] + wf * dT * (bT - krow_dot(kt, xU));
        xT[t] = v;
        if (t < o.R) vst<kCO>(sweep_buf(F, 2) + ru, v);
    }
    __syncthreads();
    if (inF) rF[t] = bF - krow_dot(kf, xT);
    __syncthreads();
    if (inC) {
        float sc = 0.f;
#pragma unroll
        for (int k = 0; k < 9; k++)
            if (l[k] >= 0) sc += rv[k] * rF[l[k]];
        vst<kCO>(C.b + c, sc);
        vst<kCO>(C.x + c, wc * dC * sc);
    }
    return true;
}

// Two-sweep down-leg phase of grid level F (pre = 2) above a coarse grid
// level C: x_1 on U into LDS, sweep 2 on T (stored for the workgroup's own
// rows), the residual on F, then b_c = R r and the coarse first sweep.  The
// operations and their order are those of jacobi, residual and restrict_to.
template <typename LB>
__device__ __forceinline__ void deep_down(const GLevel &F, const GLevel &C, LB buf, float wf, float wc)
{
    if (deep_down1(F, C, buf, wf, wc)) return;
    const Own o = own_block(F.n);
    const int g = blockIdx.x;
    const int t0 = F.th_ptr[g], nth = F.th_ptr[g + 1] - t0;
    const int u0 = F.uh_ptr[g], nuh = F.uh_ptr[g + 1] - u0;
    const int f0 = F.hd_ptr[g], nf = F.hd_ptr[g + 1] - f0;
    const int nT = o.R + nth, nU = nT + nuh;
    const auto xU = buf;
    const auto xT = buf + nU;
    const auto rF = xT + nT;
    auto urow = [&](int t) { return t < o.R ? o.b0 + t : t < nT ? F.th_row[t0 + t - o.R] : F.uh_row[u0 + t - nT]; };
    const auto x1 = sweep_buf(F, 1);
    for (int t = threadIdx.x; t < nU; t += blockDim.x)
        if (t >= o.R || t < o.no) xU[t] = vld<kCO>(x1 + urow(t));
    __syncthreads();
    const auto x2 = sweep_buf(F, 2);
    const MCL_G unsigned short *mt = F.mt + (size_t)((long long)g * o.R + t0) * F.WK;
    for (int t = threadIdx.x; t < nT; t += blockDim.x) {
        if (t < o.R && t >= o.no) continue;
        const int row = urow(t);
        const float v = xU[t] + wf * F.dinv[row] * (vld<kCO>(F.b + row) - ell32_map(F, mt, nT, t, row, xU));
        xT[t] = v;
        if (t < o.R) vst<kCO>(x2 + row, v);
    }
    __syncthreads();
    const MCL_G unsigned short *mf = F.mf + (size_t)f0 * F.WK;
    for (int t = threadIdx.x; t < nf; t += blockDim.x) {
        const int row = F.hd_row[f0 + t];
        rF[t] = vld<kCO>(F.b + row) - ell32_map(F, mf, nf, t, row, xT);
    }
    __syncthreads();
    restrict_from_lds(C, rF, wc);
}

// x on V = [O | H1 | H2] rows; a local row t of W = [O | H1] times K from V
// (the map: kl for O rows, hm for H1 rows)
struct Halo2 {
    Own o;
    int h10, nh1, h20, nh2, nW, nV;
};

__device__ __forceinline__ Halo2 halo2(const GLevel &F)
{
    Halo2 h;
    h.o = own_block(F.n);
    const int g = blockIdx.x;
    h.h10 = F.hu_ptr[g];
    h.nh1 = F.hu_ptr[g + 1] - h.h10;
    h.h20 = F.h2_ptr[g];
    h.nh2 = F.h2_ptr[g + 1] - h.h20;
    h.nW = h.o.R + h.nh1;
    h.nV = h.nW + h.nh2;
    return h;
}

__device__ __forceinline__ int vrow(const GLevel &F, const Halo2 &h, int t)
{
    return t < h.o.R ? h.o.b0 + t : t < h.nW ? F.hu_row[h.h10 + t - h.o.R] : F.h2_row[h.h20 + t - h.nW];
}

template <typename XP>
__device__ __forceinline__ float kw_row(const GLevel &F, const Halo2 &h, int t, int row, XP xV)
{
    if (t < h.o.R) return ell32c<false>(F.WK, F.n, F.kl, F.val, row, xV);
    return ell32_map(F, F.hm + (size_t)h.h10 * F.WK, h.nh1, t - h.o.R, row, xV);
}

// The thread's row of V (-1: none), and its K row when it is a row of W:
// O rows' columns through kl, H1 rows' through hm (both index V, whose
// prefix W is; an O row's kl columns also index W alone).
__device__ __forceinline__ int v_row1(const GLevel &F, const Halo2 &h, int t, KRow18 &kw, float &dW, float &bW)
{
    const bool inV = t < h.nV && (t >= h.o.R || t < h.o.no);
    const int rv = inV ? vrow(F, h, t) : -1;
    dW = 0.f;
    bW = 0.f;
    if (inV && t < h.nW) {
        if (t < h.o.R) krow_load(kw, F.kl, F.n, rv, F, rv);
        else krow_load(kw, F.hm + (size_t)h.h10 * F.WK, h.nh1, t - h.o.R, F, rv);
        dW = F.dinv[rv];
        bW = vld<kCO>(F.b + rv);
    }
    return rv;
}

// deep_down_last for lists of at most MCL_BLOCK rows (false: not applicable)
template <typename LB>
__device__ __forceinline__ bool deep_down_last1(const GLevel &F, const Halo2 &h, LB buf, float wf)
{
    if (F.WK != 18 || h.nV > MCL_BLOCK) return false;
    const int t = threadIdx.x;
    const auto xV = buf;
    const auto xW = buf + h.nV;
    KRow18 kw;
    float dW, bW;
    const int rv = v_row1(F, h, t, kw, dW, bW);
    const float xv = rv >= 0 ? vld<kCO>(sweep_buf(F, 1) + rv) : 0.f;
    if (rv >= 0) xV[t] = xv;
    __syncthreads();
    if (rv >= 0 && t < h.nW) {
        const float v = xV[t] + wf * dW * (bW - krow_dot(kw, xV));
        xW[t] = v;
        if (t < h.o.R) vst<kCO>(sweep_buf(F, 2) + rv, v);
    }
    __syncthreads();
    if (t < h.o.no) vst<kCO>(F.r + rv, bW - krow_dot(kw, xW));
    return true;
}

// Two-sweep down-leg phase of the grid level just above the LDS levels
// (pre = 2): x_1 on V, sweep 2 on W (stored for the workgroup's own rows),
// the residual on its own rows (F.r, which every workgroup then restricts
// into its LDS).
template <typename LB>
__device__ __forceinline__ void deep_down_last(const GLevel &F, LB buf, float wf)
{
    const Halo2 h = halo2(F);
    if (deep_down_last1(F, h, buf, wf)) return;
    const auto xV = buf;
    const auto xW = buf + h.nV;
    const auto x1 = sweep_buf(F, 1);
    for (int t = threadIdx.x; t < h.nV; t += blockDim.x)
        if (t >= h.o.R || t < h.o.no) xV[t] = vld<kCO>(x1 + vrow(F, h, t));
    __syncthreads();
    const auto x2 = sweep_buf(F, 2);
    for (int t = threadIdx.x; t < h.nW; t += blockDim.x) {
        if (t < h.o.R && t >= h.o.no) continue;
        const int row = vrow(F, h, t);
        const float v = xV[t] + wf * F.dinv[row] * (vld<kCO>(F.b + row) - kw_row(F, h, t, row, xV));
        xW[t] = v;
        if (t < h.o.R) vst<kCO>(x2 + row, v);
    }
    __syncthreads();
    for (int i = h.o.b0 + (int)threadIdx.x; i < h.o.b0 + h.o.no; i += blockDim.x)
        vst<kCO>(F.r + i, vld<kCO>(F.b + i) - ell32c<false>(F.WK, F.n, F.kl, F.val, i, xW));
}

// Two-sweep up-leg phase of grid level F (pre = 2): x' = x_2 + P x_c on V
// (x_c the coarse result, CC its access mode), sweep 3 on W, sweep 4 on the
// workgroup's own rows, written to F.x (the buffer of x_1/x_3: its
// neighbours still read x_2 in this phase).  On the finest level sweep 4 is
// z, zeroed on hanging equations, with sum r_i z_i published to `slot`.  The
// operations and their order are those of prolong_add, jacobi and
// last_sweep_rz.
// deep_up for lists of at most MCL_BLOCK rows (false: not applicable)
template <bool CC, typename XC, typename LB>
__device__ __forceinline__ bool deep_up1(const GLevel &F, const Halo2 &h, XC xc, LB buf, float w, bool finest,
                                         const MixedParams &P, const double *__restrict__ r, int slot,
                                         dd (*lds)[MCL_BLOCK / MCL_WAVE])
{
    if (F.WK != 18 || h.nV > MCL_BLOCK) return false;
    const int t = threadIdx.x;
    const auto xV = buf;
    const auto xW = buf + h.nV;
    KRow18 kw;
    float dW, bW;
    const int rv = v_row1(F, h, t, kw, dW, bW);
    const int u = rv >= 0 ? rv : 0;
    const int m = rv >= 0 ? F.pn[u] : 0;
    int lc[4];
    float gv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) lc[k] = rv >= 0 ? F.pcol[k * F.n + u] : -1;
    const float x2v = rv >= 0 ? vld<kCO>(sweep_buf(F, 2) + u) : 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) gv[k] = (k < m && lc[k] >= 0) ? vld<CC>(xc + lc[k]) : 0.f;
    if (rv >= 0) {
        float v = x2v;
        if (m != 0) {
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (k < m && lc[k] >= 0) s += gv[k];
            v = v + s / (float)m;
        }
        xV[t] = v;
    }
    __syncthreads();
    if (rv >= 0 && t < h.nW) xW[t] = xV[t] + w * dW * (bW - krow_dot(kw, xV));
    __syncthreads();
    dd acc[1] = {dd_zero()};
    if (t < h.o.no) {
        float z = xW[t] + w * dW * (bW - krow_dot(kw, xW));
        if (finest) {
            const LevelDev &L = P.U.L;
            if (L.nh > 0 && L.hp_n[rv] != 0) z = 0.f;  // z lives in the constrained space
            acc[0] = dd_add_prod(acc[0], r[rv], (double)z);
        }
        vst<kCO>(F.x + rv, z);
    }
    if (finest) reduce_publish<1>(P.U.rb, slot, acc, lds);
    return true;
}

template <bool CC, typename XC, typename LB>
__device__ __forceinline__ void deep_up(const GLevel &F, XC xc, LB buf, float w, bool finest, const MixedParams &P,
                                        const double *__restrict__ r, int slot, dd (*lds)[MCL_BLOCK / MCL_WAVE])
{
    const Halo2 h = halo2(F);
    if (deep_up1<CC>(F, h, xc, buf, w, finest, P, r, slot, lds)) return;
    const auto xV = buf;
    const auto xW = buf + h.nV;
    const auto x2 = sweep_buf(F, 2);
    for (int t = threadIdx.x; t < h.nV; t += blockDim.x) {
        if (t < h.o.R && t >= h.o.no) continue;
        const int u = vrow(F, h, t);
        const int m = F.pn[u];
        int lc[4];
        float gv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) lc[k] = F.pcol[k * F.n + u];
#pragma unroll
        for (int k = 0; k < 4; k++) gv[k] = (k < m && lc[k] >= 0) ? vld<CC>(xc + lc[k]) : 0.f;
        float v = vld<kCO>(x2 + u);
        if (m != 0) {
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (k < m && lc[k] >= 0) s += gv[k];
            v = v + s / (float)m;
        }
        xV[t] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < h.nW; t += blockDim.x) {
        if (t < h.o.R && t >= h.o.no) continue;
        const int row = vrow(F, h, t);
        xW[t] = xV[t] + w * F.dinv[row] * (vld<kCO>(F.b + row) - kw_row(F, h, t, row, xV));
    }
    __syncthreads();
    const LevelDev &L = P.U.L;
    dd acc[1] = {dd_zero()};
    for (int i = h.o.b0 + (int)threadIdx.x; i < h.o.b0 + h.o.no; i += blockDim.x) {
        float z = xW[i - h.o.b0] + w * F.dinv[i] * (vld<kCO>(F.b + i) - ell32c<false>(F.WK, F.n, F.kl, F.val, i, xW));
        if (finest) {
            if (L.nh > 0 && L.hp_n[i] != 0) z = 0.f;  // z lives in the constrained space
            acc[0] = dd_add_prod(acc[0], r[i], (double)z);
        }
        vst<kCO>(F.x + i, z);
    }
    if (finest) reduce_publish<1>(P.U.rb, slot, acc, lds);
}

// Grid-wide wait between phases whose cross-workgroup data is accessed only
// with agent-scope atomics (no cache maintenance), or a fenced grid barrier
// when the V-cycle runs with plain accesses.
__device__ __forceinline__ bool csync(const GridSync &gs, unsigned &gen, bool ok)
{
    if constexpr (kCO) return ok && grid_sync_coherent(gs, gen);
    else return ok && grid_sync(gs, gen);
}

// Wait between two stencil phases of the V-cycle: only for the workgroups in
// this workgroup's neighbour list (thread t polls neighbour t; nbr < 0: none),
// or for the whole grid when the lists are off.
__device__ __forceinline__ bool vsync(const GridSync &gs, unsigned &gen, bool ok, int nbr, bool use_nbr)
{
#if MCL_WAIT_TIMERS
    // split of the neighbour waits: work before the wait (12), the wait (13)
    MCL_PHASE(12);
    bool r;
    if (!use_nbr) r = ok && grid_sync(gs, gen);
    else if constexpr (kCO) r = ok && nbr_sync_coherent(gs, gen, nbr);
    else r = ok && nbr_sync(gs, gen, nbr);
    MCL_PHASE(13);
    return r;
#endif
    if (!use_nbr) return ok && grid_sync(gs, gen);
    if constexpr (kCO) return ok && nbr_sync_coherent(gs, gen, nbr);
    else return ok && nbr_sync(gs, gen, nbr);
}

// Copy of the LDS levels' matrices into the workgroup's LDS block.
__device__ __forceinline__ void lds_load_levels(const MixedParams &P, const MgLds &S)
{
    for (int l = 0; l + 1 < P.nmg; l++) {
        const MCL_L MgLevelDev *G = S.lev + l;
        if (!G->block) continue;
        const int m = G->WK * G->n, moff = G->moff, doff = G->lds_off / 3, n = G->n;
        const MCL_G float *val = gp(G->val), *dinv = gp(G->dinv);
        const MCL_G unsigned short *col = gp(G->col16);
        for (int e = threadIdx.x; e < m; e += blockDim.x) {
            S.val[moff + e] = val[e];
            S.col[moff + e] = col[e];
        }
        for (int i = threadIdx.x; i < n; i += blockDim.x) S.dinv[doff + i] = dinv[i];
    }
    __syncthreads();
}

// Coarsest level: x = K0^-1 b by the dense FP64 inverse, four lanes per row
// (MCL_BLOCK >= 4 n threads; n <= 64).
__device__ __forceinline__ void coarse_solve(const MixedParams &P, const LdsLevel &C)
{
    const int i = threadIdx.x >> 2, q = threadIdx.x & 3;
    const MCL_G double *cinv = gp(P.cinv);
    double s = 0.0;
    if (i < C.n)
        for (int j = q; j < C.n; j += 4) s += cinv[i * C.n + j] * (double)C.b[j];
    s += __shfl_xor(s, 1, MCL_WAVE);
    s += __shfl_xor(s, 2, MCL_WAVE);
    if (q == 0 && i < C.n) sweep_buf(C, 2 * P.pre)[i] = (float)s;
}

// Finest level's last post-sweep z = x_{2 pre}, fused with sum r_i z_i
// (published to reduction slot `slot`).
__device__ __forceinline__ void last_sweep_rz(const MixedParams &P, const MgLds &mgl, const double *__restrict__ r,
                                              int slot, dd (*lds)[MCL_BLOCK / MCL_WAVE])
{
    const GLevel A = glevel(mgl.lev + P.nmg - 1);
    const float w = P.omega / mgl.lam[P.nmg - 1];
    const int kf = 2 * P.pre;
    const auto in = sweep_buf(A, kf - 1);
    const auto out = sweep_buf(A, kf);
    dd acc[1] = {dd_zero()};
    const Ex ex = own_rows(A.n);
    const LevelDev &L = P.U.L;
    for (int i = ex.b; i < ex.e; i += ex.s) {
        float zi = vld<kCO>(in + i) + w * A.dinv[i] * (vld<kCO>(A.b + i) - ell32<kCO>(A, i, in));
        if (L.nh > 0 && L.hp_n[i] != 0) zi = 0.f;  // z lives in the constrained space
        vst<kCO>(out + i, zi);
        acc[0] = dd_add_prod(acc[0], r[i], (double)zi);
    }
    reduce_publish<1>(P.U.rb, slot, acc, lds);
}

// Whether a grid level runs its two-sweep phases (needs pre = 2 and the lists).
__device__ __forceinline__ bool deep_up_on(const MixedParams &P, const GLevel &A) { return P.pre == 2 && A.deep_up; }
__device__ __forceinline__ bool deep_down_on(const MixedParams &P, const GLevel &A)
{
    return P.pre == 2 && A.deep_down;
}

// Where a grid level's V-cycle result (sweep 2 pre) lies.
__device__ __forceinline__ MCL_G float *level_result(const MixedParams &P, const GLevel &A)
{
    return deep_up_on(P, A) ? A.x : sweep_buf(A, 2 * P.pre);
}

// The V-cycle after the finest level's right-hand side and first sweep are in
// place and synchronised, through the finest level's last post-sweep z with
// sum r_i z_i published to reduction slot `slot` (the caller waits on the
// grid before reading it or z).
//  * grid levels (lb+1 .. top): rows spread over the grid, neighbour waits;
//    with pre = 2 each level's down leg (sweep 2, residual, restriction) and
//    up leg (prolongation, sweeps 3 and 4) run as one phase each over the
//    workgroup's halos (deep_down / deep_up), else as the one-sweep fused
//    phases or the plain sequence;
//  * coarse levels (0 .. lb, `block` set): every workgroup runs them itself,
//    identically, with its vectors and matrices in LDS and workgroup barriers
//    only, so no workgroup waits on another and level lb+1 prolongates from
//    its own LDS.
__device__ __forceinline__ void vcycle_rest(const MixedParams &P, const MgLds &mgl, const GridSync &gs, unsigned &gen,
                                            bool &ok, int nbr, const double *__restrict__ r, int slot,
                                            dd (*lds)[MCL_BLOCK / MCL_WAVE])
{
    const int top = P.nmg - 1, pre = P.pre, kf = 2 * P.pre;
    const bool un = P.use_nbr != 0;
    const MCL_L float *lam = mgl.lam;
    int lb = 0;
    for (int l = 0; l < top; l++)
        if (mgl.lev[l].block) lb = l;
    for (int l = top; l > lb; l--) {
        const GLevel A = glevel(mgl.lev + l);
        const Ex ex = own_rows(A.n);
        const float w = P.omega / lam[l];
        if (deep_down_on(P, A)) {
            if (l - 1 > lb) {
                deep_down(A, glevel(mgl.lev + l - 1), mgl.halo, w, P.omega / lam[l - 1]);
                ok = vsync(gs, gen, ok, nbr, un);
            } else {
                // every workgroup restricts the whole residual into its LDS
                deep_down_last(A, mgl.halo, w);
                ok = csync(gs, gen, ok);
            }
            continue;
        }
        for (int k = 2; k <= pre; k++) {
            jacobi<kCO>(A, k, w, ex);
            ok = vsync(gs, gen, ok, nbr, un);
        }
        if (l - 1 > lb) {
            const GLevel C = glevel(mgl.lev + l - 1);
            if (A.fuse_down) {
                fused_down(A, C, mgl.halo, P.omega / lam[l - 1], pre);
            } else {
                residual<kCO>(A, sweep_buf(A, pre), ex);
                ok = vsync(gs, gen, ok, nbr, un);
                restrict_to<kCO, kCO>(C, A, true, P.omega / lam[l - 1], own_rows(C.n));
            }
            ok = vsync(gs, gen, ok, nbr, un);
        } else {
            // every workgroup restricts the whole residual into its LDS
            residual<kCO>(A, sweep_buf(A, pre), ex);
            ok = csync(gs, gen, ok);
        }
    }
    MCL_PHASE(2);
    {
        const LdsLevel B = lds_level(mgl, lb, pre);
        restrict_to<kCO, false>(B, glevel(mgl.lev + lb + 1), lb > 0, P.omega / lam[lb], wg_rows(B.n));
        __syncthreads();
        MCL_PHASE(7);
        for (int l = lb; l > 0; l--) {
            const LdsLevel A = lds_level(mgl, l, pre);
            const Ex ex = wg_rows(A.n);
            const float w = P.omega / lam[l];
            for (int k = 2; k <= pre; k++) {
                jacobi<false>(A, k, w, ex);
                __syncthreads();
            }
            residual<false>(A, sweep_buf(A, pre), ex);
            __syncthreads();
            const LdsLevel C = lds_level(mgl, l - 1, pre);
            restrict_to<false, false>(C, A, l - 1 > 0, P.omega / lam[l - 1], wg_rows(C.n));
            __syncthreads();
        }
        MCL_PHASE(8);
        coarse_solve(P, lds_level(mgl, 0, pre));
        __syncthreads();
        MCL_PHASE(9);
        for (int l = 1; l <= lb; l++) {
            const LdsLevel A = lds_level(mgl, l, pre);
            const LdsLevel C = lds_level(mgl, l - 1, pre);
            const Ex ex = wg_rows(A.n);
            prolong_add<false, false>(A, sweep_buf(A, pre), sweep_buf(C, kf), ex);
            __syncthreads();
            const float w = P.omega / lam[l];
            for (int k = pre + 1; k <= kf; k++) {
                jacobi<false>(A, k, w, ex);
                __syncthreads();
            }
        }
    }
    MCL_PHASE(3);
    for (int l = lb + 1; l <= top; l++) {
        const GLevel A = glevel(mgl.lev + l);
        const Ex ex = own_rows(A.n);
        const float w = P.omega / lam[l];
        const bool finest = l == top;
        if (deep_up_on(P, A)) {
            if (l - 1 == lb) {
                const LdsLevel C = lds_level(mgl, lb, pre);
                deep_up<false>(A, sweep_buf(C, kf), mgl.halo, w, finest, P, r, slot, lds);
            } else {
                deep_up<kCO>(A, level_result(P, glevel(mgl.lev + l - 1)), mgl.halo, w, finest, P, r, slot, lds);
            }
            if (!finest) ok = vsync(gs, gen, ok, nbr, un);
            continue;
        }
        const int kend = finest ? kf - 1 : kf;
        // the fused phase makes sweep pre + 1 (the finest level's last sweep
        // is last_sweep_rz's, so there it needs pre >= 2)
        const bool fuse = A.fuse_up && kend >= pre + 1;
        if (l - 1 == lb) {
            const LdsLevel C = lds_level(mgl, lb, pre);
            if (fuse) fused_up<false>(A, sweep_buf(C, kf), mgl.halo, w, pre);
            else prolong_add<kCO, false>(A, sweep_buf(A, pre), sweep_buf(C, kf), ex);
        } else {
            const auto xc = level_result(P, glevel(mgl.lev + l - 1));
            if (fuse) fused_up<kCO>(A, xc, mgl.halo, w, pre);
            else prolong_add<kCO, kCO>(A, sweep_buf(A, pre), xc, ex);
        }
        ok = vsync(gs, gen, ok, nbr, un);
        for (int k = fuse ? pre + 2 : pre + 1; k <= kend; k++) {
            jacobi<kCO>(A, k, w, ex);
            ok = vsync(gs, gen, ok, nbr, un);
        }
        if (finest) last_sweep_rz(P, mgl, r, slot, lds);
    }
    MCL_PHASE(4);
}

// ------------------------------------------- hanging nodes (Galerkin form)
// With hanging nodes the inner solve works on the constrained space u = C u_f
// (a hanging value is the weighted mean of its free parents): it solves
// C^T K C u_f = C^T b by CG on vectors that are zero on hanging equations and
// expands the result at the end.  (The FP64 path keeps the reference's
// overwrite after each update instead, Solver.c:1368-1437.)

// (C v)_j for a vector v that is zero on hanging equations
template <typename F>
__device__ __forceinline__ double c_expand(const LevelDev &L, int j, F v)
{
    const int np = L.hp_n[j];
    if (np == 0) return v(j);
    double s = 0.0;
    for (int k = 0; k < np; k++) s += L.hp_w[k * L.neq + j] * v(L.hp_i[k * L.neq + j]);
    return s;
}

// row i of K (C v)
template <typename F>
__device__ __forceinline__ double k_row_c(const LevelDev &L, int i, F v)
{
    const int nc = L.Kncol[i], n = L.neq;
    double s = 0.0;
    for (int k = 0; k < nc; k++) s += L.Kval[k * n + i] * c_expand(L, L.Kcol[k * n + i], v);
    return s;
}

// (C^T y)_i for a free equation i: y_i plus the weighted rows of its children
template <typename F>
__device__ __forceinline__ double ct_gather(const LevelDev &L, int i, double yi, F child_value)
{
    for (int t = L.ch_ptr[i]; t < L.ch_ptr[i + 1]; t++) yi += L.ch_w[t] * child_value(L.ch_i[t]);
    return yi;
}

// ------------------------------------------------- FP64 CG, FP32 V-cycle
// Inner solve K x = b from x = 0 by FP64 conjugate gradients preconditioned
// with the FP32 multigrid V-cycle (Jacobi when there is a single level): K x
// products, residuals and recurrences stay in FP64 with compensated dot
// products; only z = M r is FP32.  Stops at |b - K x|^2 <= max(rtol^2 |b|^2,
// floor2), floor2 = ff_coef ff.  Per iteration:
//   (1) d' = z + beta d and q = K d' in one pass (q gathers z + beta d),
//       with d'.q                                              | barrier
//   (2) x += alpha d', r -= alpha q, r.r, finest b = r and the
//       first Jacobi sweep                                     | barrier
//   (3) rest of the V-cycle, last sweep fused with r.z          | barrier
// The caller writes b on this thread's rows (own_rows(neq)) in the phase
// that calls the solve, so the start needs no barrier before it (with
// hanging nodes the start reads b across rows: the caller waits on the grid
// first).  With collect_ff the caller has published |F|^2 to slot 5 in that
// phase and ff receives it.  x is stored with agent-scope stores and is
// complete for agent-scope loads on return, which follows a barrier (with
// hanging nodes, a fenced one after the expansion u = C u_f).
// Reduction slots: 0,1 (start), 2 (r.z), 15 (d.q), 13,14 (r.r, r.z).
// Returns the iteration count.
__device__ __forceinline__ int inner_mixed(const MixedParams &P, const double *__restrict__ b,
                                           double *__restrict__ x, const GridSync &gs, unsigned &gen, bool &ok,
                                           dd (*lds)[MCL_BLOCK / MCL_WAVE], const MgLds &mgl, double ff_coef,
                                           double &ff, bool collect_ff, int nbr, double rtol)
{
    const LevelDev &L = P.U.L;
    const bool use_mg = P.nmg > 1;
    const GLevel A = glevel(mgl.lev + P.nmg - 1);
    const float wtop = use_mg ? P.omega / mgl.lam[P.nmg - 1] : 0.f;
    const MCL_G float *zf = use_mg ? level_result(P, A) : A.x;
    const int n = L.neq;
    // rows of this thread: the same as on the V-cycle's finest level, so x, r
    // and q stay private to their thread; d and z are read by neighbours
    // (agent-scope accesses), and between the phases of an iteration the grid
    // waits without cache maintenance (grid_sync_coherent)
    const Ex ex = own_rows(n);
    double *__restrict__ r = P.rr, *__restrict__ q = P.U.cg_q;
    MCL_G double *dold = gp(P.U.cg_d), *dnew = gp(P.U.cg_s);
    MCL_G double *xg = gp(x);
    // r = b, then M r
    {
        dd acc[2] = {dd_zero(), dd_zero()};
        for (int i = ex.b; i < ex.e; i += ex.s) {
            double ri = b[i];
            vstd<kCO>(xg + i, 0.0);
            if (L.nh > 0) {
                // r = C^T b (the solve starts from x = 0 whenever there are hanging nodes)
                ri = L.hp_n[i] != 0 ? 0.0 : ct_gather(L, i, b[i], [&](int h) { return b[h]; });
            }
            r[i] = ri;
            acc[0] = dd_add_prod(acc[0], ri, ri);
            if (use_mg) {
                float bi = (float)ri;
                vst<kCO>(A.b + i, bi);
                vst<kCO>(A.x + i, wtop * A.dinv[i] * bi);
            } else {
                float zi = (float)((double)A.dinv[i] * ri);
                vst<kCO>(A.x + i, zi);
                acc[1] = dd_add_prod(acc[1], ri, (double)zi);
            }
        }
        reduce_publish<2>(P.U.rb, 0, acc, lds);
    }
    ok = csync(gs, gen, ok);
    double bn2, delta;
    {
        double o[2];
        reduce_collect<2>(P.U.rb, 0, gridDim.x, o, lds);
        bn2 = o[0];
        delta = o[1];
        if (collect_ff) {
            double f[1];
            reduce_collect<1>(P.U.rb, 5, gridDim.x, f, lds);
            ff = f[0];
        }
    }
    MCL_PHASE(1);
    const double stop2 = fmax(rtol * rtol * bn2, ff_coef * ff);
    if (!(bn2 > stop2)) return 0;   // x = 0 (also on hanging equations)
    // rz: the latest r.z (with Jacobi it came with the reduction above)
    double beta = 0.0, rz = delta;
    bool first = true;
    int it = 0;
    while (ok && it < P.max_inner) {
        // z = M r: the V-cycle's one call site, so the kernel holds one copy
        if (use_mg) {
            vcycle_rest(P, mgl, gs, gen, ok, nbr, r, 2, lds);
            ok = csync(gs, gen, ok);
            double o[1];
            reduce_collect<1>(P.U.rb, 2, gridDim.x, o, lds);
            rz = o[0];
            MCL_PHASE(5);
        }
        if (!first) beta = rz / delta;
        delta = rz;
        double dq;
        {
            dd acc[1] = {dd_zero()};
            if (L.nh > 0) {
                // q = C^T K C d', d' = z + beta d (zero on hanging equations)
                auto dv = [&](int j) {
                    double v = (double)vld<kCO>(zf + j);
                    if (!first) v += beta * vldd<kCO>(dold + j);
                    return v;
                };
                for (int i = ex.b; i < ex.e; i += ex.s) {
                    double qi = 0.0, di = 0.0;
                    if (L.hp_n[i] == 0) {
                        di = dv(i);
                        qi = ct_gather(L, i, k_row_c(L, i, dv), [&](int h) { return k_row_c(L, h, dv); });
                    }
                    vstd<kCO>(dnew + i, di);
                    q[i] = qi;
                    acc[0] = dd_add_prod(acc[0], di, qi);
                }
            } else
            for (int i = ex.b; i < ex.e; i += ex.s) {
                const int nc = L.Kncol[i];
                double s = 0.0;
                int k = 0;
                if (L.WK == 18) {
                    // the whole width-18 row in one round trip (padding
                    // slots hold column i and value 0, adding exact zeros)
                    int c[18];
                    double v[18];
#pragma unroll
                    for (int u = 0; u < 18; u++) {
                        c[u] = L.Kcol[u * n + i];
                        v[u] = L.Kval[u * n + i];
                    }
                    float zj[18];
                    double dj[18];
#pragma unroll
                    for (int u = 0; u < 18; u++) {
                        zj[u] = vld<kCO>(zf + c[u]);
                        dj[u] = first ? 0.0 : vldd<kCO>(dold + c[u]);
                    }
#pragma unroll
                    for (int u = 0; u < 18; u++) {
                        double d = (double)zj[u];
                        if (!first) d += beta * dj[u];
                        s += v[u] * d;
                    }
                    k = nc;
                }
                for (; k + 9 <= nc; k += 9) {
                    int c[9];
                    double v[9];
#pragma unroll
                    for (int u = 0; u < 9; u++) {
                        c[u] = L.Kcol[(k + u) * n + i];
                        v[u] = L.Kval[(k + u) * n + i];
                    }
                    float zj[9];
                    double dj[9];
#pragma unroll
                    for (int u = 0; u < 9; u++) {
                        zj[u] = vld<kCO>(zf + c[u]);
                        dj[u] = first ? 0.0 : vldd<kCO>(dold + c[u]);
                    }
#pragma unroll
                    for (int u = 0; u < 9; u++) {
                        double d = (double)zj[u];
                        if (!first) d += beta * dj[u];
                        s += v[u] * d;
                    }
                }
                for (; k < nc; k++) {
                    const int j = L.Kcol[k * n + i];
                    double dj = (double)vld<kCO>(zf + j);
                    if (!first) dj += beta * vldd<kCO>(dold + j);
                    s += L.Kval[k * n + i] * dj;
                }
                double di = (double)vld<kCO>(zf + i);
                if (!first) di += beta * vldd<kCO>(dold + i);
                vstd<kCO>(dnew + i, di);
                q[i] = s;
                acc[0] = dd_add_prod(acc[0], di, s);
            }
            reduce_publish<1>(P.U.rb, 15, acc, lds);
            ok = csync(gs, gen, ok);
            double o[1];
            reduce_collect<1>(P.U.rb, 15, gridDim.x, o, lds);
            dq = o[0];
            MCL_PHASE(0);
        }
        if (!(dq > 0.0)) break;
        const double alpha = delta / dq;
        double rr;
        rz = 0.0;
        {
            dd acc[2] = {dd_zero(), dd_zero()};
            for (int i = ex.b; i < ex.e; i += ex.s) {
                vstd<kCO>(xg + i, vldd<kCO>(xg + i) + alpha * vldd<kCO>(dnew + i));
                double ri = r[i] - alpha * q[i];
                r[i] = ri;
                acc[0] = dd_add_prod(acc[0], ri, ri);
                if (use_mg) {
                    float bi = (float)ri;
                    vst<kCO>(A.b + i, bi);
                    vst<kCO>(A.x + i, wtop * A.dinv[i] * bi);
                } else {
                    float zi = (float)((double)A.dinv[i] * ri);
                    vst<kCO>(A.x + i, zi);
                    acc[1] = dd_add_prod(acc[1], ri, (double)zi);
                }
            }
            reduce_publish<2>(P.U.rb, 13, acc, lds);
            ok = csync(gs, gen, ok);
            double o[2];
            reduce_collect<2>(P.U.rb, 13, gridDim.x, o, lds);
            rr = o[0];
            rz = o[1];
            MCL_PHASE(1);
        }
        it++;
        if (rr <= stop2) break;
        first = false;
        MCL_G double *t = dold;
        dold = dnew;
        dnew = t;
    }
    // every exit follows a barrier that ends the phase of the last x update
    if (L.nh > 0) {
        // u = C u_f (hanging values from their parents across rows)
        ok = ok && grid_sync(gs, gen);
        for (int i = ex.b; i < ex.e; i += ex.s)
            if (L.hp_n[i] != 0) vstd<kCO>(xg + i, c_expand(L, i, [&](int j) { return vldd<kCO>(xg + j); }));
        ok = ok && grid_sync(gs, gen);
    }
    return it;
}

// --------------------------------------------------------- persistent Uzawa

__device__ __forceinline__ void uzawa_mixed(const MixedParams &P)
{
    __shared__ dd lds[MCL_RED_ROWS][MCL_BLOCK / MCL_WAVE];
    __shared__ float mg_vec[3 * MCL_MG_LDS_ROWS], mg_dinv[MCL_MG_LDS_ROWS], mg_val[MCL_MG_LDS_NNZ];
    __shared__ unsigned short mg_col[MCL_MG_LDS_NNZ];
    // level descriptors and Gershgorin bounds (read in every phase; LDS keeps
    // them out of the grid barriers' cache invalidations)
    __shared__ MgLevelDev mg_lev[MCL_MG_MAX_LEVELS];
    __shared__ float mg_lam[MCL_MG_MAX_LEVELS];
    __shared__ float mg_halo[MCL_HALO_LDS];
    {
        const int nw = P.nmg * (int)(sizeof(MgLevelDev) / sizeof(int));
        for (int e = threadIdx.x; e < nw; e += blockDim.x) ((int *)mg_lev)[e] = ((const int *)P.mg)[e];
        for (int e = threadIdx.x; e < P.nmg; e += blockDim.x) mg_lam[e] = P.lam[e];
        __syncthreads();
    }
    const MgLds mgl{(const MCL_L MgLevelDev *)mg_lev, (const MCL_L float *)mg_lam, (MCL_L float *)mg_vec,
                    (MCL_L float *)mg_dinv, (MCL_L float *)mg_val, (MCL_L unsigned short *)mg_col,
                    (MCL_L float *)mg_halo};
    if (P.nmg > 1) lds_load_levels(P, mgl);
    const UzawaParams &U = P.U;
    const LevelDev &L = U.L;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    const int neq = L.neq, nel = L.nel;
    // this solve's augmented-Lagrangian weight (read before the first wait;
    // workgroup 0 rewrites it after the last)
    const double al_g = P.al_state ? *P.al_state : 0.0;
    unsigned gen = __hip_atomic_load(U.gs.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    bool ok = true;
    long long spmv = 0, inner = 0;
    // this thread's neighbour for the V-cycle's point-to-point waits
    int nbr = -1;
    if (P.use_nbr) {
        const int n0 = P.nbr_ptr[blockIdx.x], n1 = P.nbr_ptr[blockIdx.x + 1];
        if ((int)threadIdx.x < n1 - n0) nbr = P.nbr_idx[n0 + threadIdx.x];
    }
#if MCL_PHASE_TIMERS
    if (threadIdx.x == 0) {
        for (int k = 0; k < MCL_PHASES; k++) mcl_ph[k] = 0;
        mcl_ph[MCL_PHASES] = (long long)wall_clock64();
    }
#endif
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
    // One inner solve call site for all three kinds of inner solve (the
    // kernel then holds one copy of the inner solve and its V-cycle):
    //   stage 0: the first solve, for the momentum defect of the initial guess;
    //   stage 1: the Uzawa iterations' solves K u1 = G s2;
    //   stage 2: with relaxed inner accuracy, the final velocity correction.
    // The inner solve's right-hand side is written on the rows the solve's
    // threads own (own_rows), in the phase that starts the solve.  Element
    // vectors are read by the thread that writes them (the element loops use
    // the grid stride), except z and s, which G gathers across workgroups,
    // and u1, which the solve stores and G^T gathers: those three use
    // agent-scope accesses, and the waits between their phases need no cache
    // maintenance.
    double *r0 = U.r0, *r1 = U.r1, *r2 = U.r2, *z0 = U.z0, *z1 = U.z1, *s1 = U.s1, *s2 = U.s2;
    MCL_G double *u1 = gp(U.u1);
    const Ex rows = own_rows(neq);
    double dv = 1.0, dp = 1.0, ff = 0.0, r1z1 = 0.0, r0z0 = 0.0;
    int cycle = 0, stage = 0;
    while (ok) {
        double rtol = P.inner_rtol, ffc = 0.0;
        if (stage != 1) {
            // momentum defect Ah = F - G p - K d; |F|^2 gives the solve an
            // absolute floor inner_rtol |F| (a warm start's small defect is
            // not reduced by another factor inner_rtol).  After the Uzawa
            // iterations d and p were written by their owning threads with
            // plain stores and are now gathered across workgroups: a fenced
            // barrier first.
            if (stage == 2) ok = ok && grid_sync(U.gs, gen);
            dd acc[1] = {dd_zero()};
            for (int i = rows.b; i < rows.e; i += rows.s) {
                double u = ell_row(L.Kcol, L.Kval, L.Kncol[i], neq, i, U.d);
                double f = L.F[i];
                U.Ahv[i] = f - g_row(L, i, U.p) - u;
                acc[0] = dd_add_prod(acc[0], f, f);
            }
            spmv++;
            if (stage == 0) reduce_publish<1>(U.rb, 5, acc, lds);
            ffc = P.inner_rtol * P.inner_rtol;
        } else {
            // Uzawa iteration `cycle`: the search direction s2 = z1 + delta
            // s1 and the inner right-hand side G s2 = G z1 + delta G s1, both
            // from z1 and s1 (complete since the last barrier)
            const double delta = cycle > 0 ? r1z1 / r0z0 : 0.0;
            for (int e = gtid; e < nel; e += gsize) {
                const double z = vldd<kCO>(gp(z1) + e);
                vstd<kCO>(gp(s2) + e, cycle == 0 ? z : z + delta * vldd<kCO>(gp(s1) + e));
            }
            for (int i = rows.b; i < rows.e; i += rows.s) {
                double a = g_row_ld(L, i, [&](int e) { return vldd<kCO>(gp(z1) + e); });
                if (cycle > 0) a += delta * g_row_ld(L, i, [&](int e) { return vldd<kCO>(gp(s1) + e); });
                U.Ahv[i] = a;
            }
            // relaxed inner accuracy (P.relax > 0): the update this solve
            // feeds is about dv |d| in size, so its defect matters in
            // proportion; the tolerance grows as relax / dv up to relax_max,
            // and the final correction (stage 2) restores the momentum residual
            if (P.relax > 0.0) rtol = fmin(P.relax_max, fmax(P.inner_rtol, P.relax / fmax(dv, 1e-300)));
        }
        MCL_PHASE(6);
        if (L.nh > 0) ok = ok && grid_sync(U.gs, gen);   // the solve's start reads b across rows
        const int it = inner_mixed(P, U.Ahv, U.u1, U.gs, gen, ok, lds, mgl, ffc, ff, stage == 0, nbr, rtol);
        inner += it;
        spmv += it;
        if (stage != 1) {
            for (int i = gtid; i < neq; i += gsize) U.d[i] += vldd<kCO>(u1 + i);
            ok = ok && grid_sync(U.gs, gen);
            if (stage == 2) break;
            if (P.probe) {
                // initial Schur residual r1 = G^T K^-1 (F - G p) of this
                // (d, p): the quantity the Uzawa iteration has to reduce
                dd acc[1] = {dd_zero()};
                for (int e = gtid; e < nel; e += gsize) {
                    double r = gt_elem(L, e, U.d);
                    acc[0] = dd_add_prod(acc[0], r, r);
                }
                reduce_publish<1>(U.rb, 4, acc, lds);
                ok = ok && grid_sync(U.gs, gen);
                double o[1];
                reduce_collect<1>(U.rb, 4, gridDim.x, o, lds);
                if (blockIdx.x == 0 && threadIdx.x == 0) {
                    U.stats->probe_r1 = o[0];
                    U.stats->cycles = 0;
                    U.stats->inner = inner;
                    U.stats->spmv = spmv;
                    U.stats->aborted = ok ? 0 : 1;
                }
                return;
            }
            // Schur residual r1 = G^T d, z1 = M_S r1 and r1.z1 for the first
            // Uzawa iteration
            {
                dd acc[1] = {dd_zero()};
                for (int e = gtid; e < nel; e += gsize) {
                    const double r = gt_elem(L, e, U.d);
                    const double z = U.schur ? r / U.BAB[e] : U.BAB[e] * r;
                    r1[e] = r;
                    vstd<kCO>(gp(z1) + e, z);
                    acc[0] = dd_add_prod(acc[0], r, z);
                }
                reduce_publish<1>(U.rb, 4, acc, lds);
            }
            ok = csync(U.gs, gen, ok);
            {
                double o[1];
                reduce_collect<1>(U.rb, 4, gridDim.x, o, lds);
                r1z1 = o[0];
            }
            stage = 1;
        } else {
            {
                dd acc[1] = {dd_zero()};
                for (int e = gtid; e < nel; e += gsize) {
                    double a = gt_elem_ld(L, e, [&](int r) { return vldd<kCO>(u1 + r); });
                    U.Ahp[e] = a;
                    acc[0] = dd_add_prod(acc[0], vldd<kCO>(gp(s2) + e), a);
                }
                reduce_publish<1>(U.rb, 6, acc, lds);
            }
            ok = csync(U.gs, gen, ok);
            double alpha;
            {
                double o[1];
                reduce_collect<1>(U.rb, 6, gridDim.x, o, lds);
                alpha = r1z1 / o[0];
            }
            {
                // updates, their norms, and the next iteration's z = M_S r
                // with r.z (into z0's buffer, which the rotation makes z1)
                dd acc[5] = {dd_zero(), dd_zero(), dd_zero(), dd_zero(), dd_zero()};
                for (int e = gtid; e < nel; e += gsize) {
                    const double rn = r1[e] - alpha * U.Ahp[e];
                    r2[e] = rn;
                    const double se = vldd<kCO>(gp(s2) + e);
                    double pe = U.p[e] + alpha * se;
                    U.p[e] = pe;
                    acc[0] = dd_add_prod(acc[0], se, se);
                    acc[1] = dd_add_prod(acc[1], pe, pe);
                    const double z = U.schur ? rn / U.BAB[e] : U.BAB[e] * rn;
                    vstd<kCO>(gp(z0) + e, z);
                    acc[4] = dd_add_prod(acc[4], rn, z);
                }
                for (int i = gtid; i < neq; i += gsize) {
                    double u = vldd<kCO>(u1 + i);
                    double di = U.d[i] - alpha * u;
                    U.d[i] = di;
                    acc[2] = dd_add_prod(acc[2], u, u);
                    acc[3] = dd_add_prod(acc[3], di, di);
                }
                reduce_publish<5>(U.rb, 8, acc, lds);
            }
            ok = csync(U.gs, gen, ok);
            {
                double o[5];
                reduce_collect<5>(U.rb, 8, gridDim.x, o, lds);
                dp = alpha * sqrt(o[0] / (1.0e-32 + o[1]));
                dv = alpha * sqrt(o[2] / (1.0e-32 + o[3]));
                r0z0 = r1z1;
                r1z1 = o[4];
            }
            cycle++;
            double *sh;
            sh = s1; s1 = s2; s2 = sh;
            sh = r0; r0 = r1; r1 = r2; r2 = sh;
            sh = z0; z0 = z1; z1 = sh;
        }
        // the Uzawa loop's test (Solver.c:131), then the final correction
        if (!((dv > U.citcom_accu || (dp > U.citcom_accu && U.check_v_only == 0)) &&
              cycle < U.max_citcom_cycle)) {
            if (P.relax > 0.0 && cycle > 0) stage = 2;
            else break;
        }
    }
    if (al_g > 0.0 && ok) {
        // the plain system's pressure p + gamma W G^T d (d was written by
        // its owners with plain stores: a fenced barrier first)
        ok = grid_sync(U.gs, gen);
        if (ok)
            for (int e = gtid; e < nel; e += gsize) U.p[e] += al_g * P.al_w[e] * gt_elem(L, e, U.d);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (P.al_state) *P.al_state = cycle >= (al_g > 0.0 ? P.al_keep : P.al_start) ? P.al_gamma : 0.0;
        if (P.ws && ok) {
            P.ws[2] = P.ws[1];
            P.ws[1] = *P.model_time;
            P.ws[0] += 1.0;
        }
        U.stats->cycles = cycle;
        U.stats->dv = dv;
        U.stats->dp = dp;
        U.stats->inner = inner;
        U.stats->spmv = spmv;
        U.stats->aborted = ok ? 0 : 1;
        if (!(dv == dv) || !(dp == dp)) U.stats->nan_flag = 1;
    }
    MCL_PHASE(6);
#if MCL_PHASE_TIMERS
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int k = 0; k < MCL_PHASES; k++) U.stats->phase[k] = mcl_ph[k];
#endif
}

// the solve, and the guess gate's probe (same body, P.probe = 1), as two
// kernels so that their launches are told apart in traces
__global__ void __launch_bounds__(MCL_BLOCK) k_uzawa_mixed(MixedParams P) { uzawa_mixed(P); }
__global__ void __launch_bounds__(MCL_BLOCK) k_uzawa_probe(MixedParams P) { uzawa_mixed(P); }
