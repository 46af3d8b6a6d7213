// FP64 kernels of libmcl_hip.so (gfx950).
//
// This translation unit is compiled with -ffp-contract=off: every operation
// on the parity path rounds like the reference's x86-64 build (no FMA), and
// every sum runs in the reference's order (element-ascending assembly, sorted
// row products), so assembly, sparse products and vector updates reproduce the
// reference bit for bit.  Only the dot products differ: they are compensated
// parallel reductions (the reference sums sequentially).
//
// Work decomposition (one HIP thread per item, grid-stride):
//   nodes     viscosity, buoyancy, nodal velocity, stress averaging
//   rows      velocity / heat equations: assembly of one matrix row from the
//             row's node-adjacent elements (atomic-free, element-ascending),
//             ELL products, G p, CG vector updates
//   elements  BAB, G^T d, pressure-space updates, SUPG tau, element stress
// The Uzawa solve and the heat step each run as one persistent kernel launched
// cooperatively; phases are separated by grid barriers (mcl_device.h).
#include "mcl_kernels.h"
#include "mcl_ops.h"

__constant__ ShapeTables c_shape;

extern "C" hipError_t mcl_upload_shape(const ShapeTables *s)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(c_shape), s, sizeof(ShapeTables));
}

// ------------------------------------------------------------------ physics

// Viscosity.c:3-24: eta = exp(A (refT - T)), times lm_jump below z = 0.77.
__global__ void k_viscosity(int nno, const double *__restrict__ T, const double *__restrict__ X,
                            double A, double refT, double lm_jump, double *__restrict__ vis)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nno; i += gridDim.x * blockDim.x) {
        double v = exp(A * (refT - T[i]));
        if (X[nno + i] < 0.77) v *= lm_jump;
        vis[i] = v;
    }
}

// Buoyancy.c:3-22 with get_node_horiz_avg (Parallel.c:354-437): one thread per
// node row k walks the row's contribution list in the reference's order
// (elements ascending; within an element the interior points, then the top
// row or right column points, then the corner node).  Each entry adds
// N_a(xi, eta) * data[node_a] for a = 0..3 separately, as the reference does.
// Row averages in two passes.  k_buoy_terms evaluates every term of every
// row's ordered contribution list in parallel (the products are the
// reference's, independent of the running sum); k_buoy_rows then adds each
// row's terms in list order, one thread per row, with its loads unrolled, so
// the sum rounds exactly as the reference's sequential loop.  Entries with a
// negative element contribute one term (Ra T of the last node) and three
// +0.0 terms, which leave the sum unchanged.
__global__ void k_buoy_terms(int nent, int nno, const int *__restrict__ ent_e, const signed char *__restrict__ ent_xi,
                             const signed char *__restrict__ ent_eta, const int *__restrict__ ienl, int nel,
                             const double *__restrict__ T, double Ra, double *__restrict__ terms)
{
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nent; t += gridDim.x * blockDim.x) {
        int e = ent_e[t];
        double v0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
        if (e < 0) {
            v0 = Ra * T[nno - 1];
        } else {
            double xi = (double)ent_xi[t], eta = (double)ent_eta[t];
            // Na values of Shape_function.c:180-197
            double n0 = 0.25 * (1.0 - xi) * (1.0 - eta);
            double n1 = 0.25 * (1.0 - xi) * (1.0 + eta);
            double n2 = 0.25 * (1.0 + xi) * (1.0 + eta);
            double n3 = 0.25 * (1.0 + xi) * (1.0 - eta);
            v0 = n0 * (Ra * T[ienl[e]]);
            v1 = n1 * (Ra * T[ienl[nel + e]]);
            v2 = n2 * (Ra * T[ienl[2 * nel + e]]);
            v3 = n3 * (Ra * T[ienl[3 * nel + e]]);
        }
        terms[4 * (size_t)t] = v0;
        terms[4 * (size_t)t + 1] = v1;
        terms[4 * (size_t)t + 2] = v2;
        terms[4 * (size_t)t + 3] = v3;
    }
}

__global__ void k_buoy_rows(int gnoz, int gnox, const int *__restrict__ ptr, const double *__restrict__ terms,
                            double *__restrict__ avg)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= gnoz) return;
    const int t0 = 4 * ptr[k], t1 = 4 * ptr[k + 1];
    double s = 0.0;
    int t = t0;
    for (; t + 8 <= t1; t += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = terms[t + u];
#pragma unroll
        for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; t < t1; t++) s += terms[t];
    avg[k] = s / gnox;
}

__global__ void k_buoy_apply(int nno, int gnoz, const int *__restrict__ node, const double *__restrict__ T,
                             double Ra, const double *__restrict__ avg, double *__restrict__ buoy)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nno; i += gridDim.x * blockDim.x) {
        int nd = node[i];
        int ix = nd / gnoz, iz = nd - ix * gnoz;
        buoy[i] = Ra * T[i] - avg[iz];
    }
}

// ------------------------------------------------------------ Stokes assembly

// Upper-triangle element stiffness entry of get_element_K (Build_matrix.c:
// 192-271) for row dof (a,i) and column dof (b,j) with 2a+i <= 2b+j; eta_g is
// the linearly interpolated viscosity at the four Gauss points.
__device__ __forceinline__ double ke_upper(int a, int i, int b, int j, const double eta_g[4], double dxdxi,
                                           double dzdeta, double jac)
{
    const double dzdxi = 0.0, dxdeta = 0.0;
    double K = 0.0;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        double eta = eta_g[g];
        double D00 = 2.0 * eta, D11 = 2.0 * eta, D22 = eta;
        const double D01 = 0.0, D02 = 0.0, D10 = 0.0, D12 = 0.0, D20 = 0.0, D21 = 0.0;
        double ax = c_shape.dxi[a][g] * dzdeta - c_shape.deta[a][g] * dzdxi;
        double az = c_shape.deta[a][g] * dxdxi - c_shape.dxi[a][g] * dxdeta;
        double bx = c_shape.dxi[b][g] * dzdeta - c_shape.deta[b][g] * dzdxi;
        double bz = c_shape.deta[b][g] * dxdxi - c_shape.dxi[b][g] * dxdeta;
        if (i == 0 && j == 0)
            K += ((ax * D00 + az * D20) * bx + (ax * D02 + az * D22) * bz);
        else if (i == 1 && j == 0)
            K += ((az * D10 + ax * D20) * bx + (az * D12 + ax * D22) * bz);
        else if (i == 0 && j == 1)
            K += ((ax * D01 + az * D21) * bz + (ax * D02 + az * D22) * bx);
        else
            K += ((az * D11 + ax * D21) * bz + (az * D12 + ax * D22) * bx);
    }
    return K / jac;
}

// Element entry K_e[p][q] after the reference's mirror step: the value of the
// upper entry (min(p,q), max(p,q)).
__device__ __forceinline__ double ke_entry(int p, int q, const double eta_g[4], double dxdxi, double dzdeta,
                                           double jac)
{
    int lo = p < q ? p : q, hi = p < q ? q : p;
    return ke_upper(lo >> 1, lo & 1, hi >> 1, hi & 1, eta_g, dxdxi, dzdeta, jac);
}

__device__ __forceinline__ void elem_eta(const LevelDev &L, int e, const double *__restrict__ vis, double eta_g[4])
{
    double vn[4];
#pragma unroll
    for (int d = 0; d < 4; d++) vn[d] = vis[L.ienl[d * L.nel + e]];
#pragma unroll
    for (int g = 0; g < 4; g++) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 4; d++) s += c_shape.N[d][g] * vn[d];
        eta_g[g] = s;
    }
}

// One thread per velocity row P: K row slots (element-ascending sums, the
// order of K_matrix, Build_matrix.c:42-115), F[P] (all boundary-value
// subtractions first, then the element loads of F_matrix, :117-140), Kdiag.
// The thread owns its row, so the slots accumulate without atomics: in the
// thread's own LDS column for rows of at most 18 slots (written to HBM once at
// the end), else in place in HBM; the additions to each slot come in the same
// order either way.  One launch assembles every level: each workgroup takes
// MCL_BLOCK rows of one level (lv[k].blk0 gives the first workgroup of level
// k), so the short levels run side by side instead of one launch each.
#define MCL_ASM_LDS_SLOTS 18
__global__ void __launch_bounds__(MCL_BLOCK) k_assemble_K(const AsmLevel *__restrict__ lv, int nlev)
{
    __shared__ double kacc[MCL_ASM_LDS_SLOTS * MCL_BLOCK];
    double *acc = kacc + threadIdx.x;  // slot u at acc[u * MCL_BLOCK]
    int lk = 0;
    while (lk + 1 < nlev && (int)blockIdx.x >= lv[lk + 1].blk0) lk++;
    const LevelDev &L = lv[lk].L;
    const double *__restrict__ vis = lv[lk].vis;
    const double *__restrict__ buoy = lv[lk].buoy;
    const bool inl = L.WK <= MCL_ASM_LDS_SLOTS;
    {
        const int P = ((int)blockIdx.x - lv[lk].blk0) * MCL_BLOCK + (int)threadIdx.x;
        if (P >= L.neq) return;
        const int n = L.veq_node[P], ip = L.veq_dof[P];
        const int nadj = L.p2n_cnt[n];
        const int nc = L.Kncol[P];
        if (inl)
            for (int u = 0; u < L.WK; u++) acc[u * MCL_BLOCK] = 0.0;
        else
            for (int s = 0; s < L.WK; s++) L.Kval[s * L.neq + P] = 0.0;
        double F = 0.0;
        for (int k = 0; k < nadj; k++) {
            int e = L.p2n_el[k * L.nno + n], a = L.p2n_a[k * L.nno + n];
            int lev = L.elev[e];
            double dxdxi = 0.5 * L.lvl_dx[lev], dzdeta = 0.5 * L.lvl_dz[lev];
            double jac = dxdxi * dzdeta - 0.0 * 0.0;
            double eta_g[4];
            elem_eta(L, e, vis, eta_g);
            int p = 2 * a + ip;
            for (int j = 0; j < 2; j++)
                for (int b = 0; b < 4; b++) {
                    int q = 2 * b + j;
                    double kv = ke_entry(p, q, eta_g, dxdxi, dzdeta, jac);
                    int slot = L.Kslot[(k * 8 + q) * L.neq + P];
                    if (slot >= 0) {
                        if (inl) acc[slot * MCL_BLOCK] += kv;
                        else L.Kval[slot * L.neq + P] += kv;
                    }
                    int nb = L.ienl[b * L.nel + e];
                    if (L.v_bc[j * L.nno + nb] == 'V') F -= kv * L.v_bc_val[j * L.nno + nb];
                }
        }
        // element loads, get_element_F (Build_matrix.c:273-344)
        for (int k = 0; k < nadj; k++) {
            int e = L.p2n_el[k * L.nno + n], a = L.p2n_a[k * L.nno + n];
            int lev = L.elev[e];
            double dxdxi = 0.5 * L.lvl_dx[lev], dzdeta = 0.5 * L.lvl_dz[lev];
            double jac = dxdxi * dzdeta - 0.0 * 0.0;
            int idx[4];
#pragma unroll
            for (int d = 0; d < 4; d++) idx[d] = L.ienl[d * L.nel + e];
            double Fe = 0.0;
            if (ip == 1) {
                double fg[4];
#pragma unroll
                for (int g = 0; g < 4; g++) {
                    double s = 0.0;
#pragma unroll
                    for (int d = 0; d < 4; d++) s += c_shape.N[d][g] * buoy[idx[d]];
                    fg[g] = s;
                }
#pragma unroll
                for (int g = 0; g < 4; g++) Fe += c_shape.N[a][g] * fg[g];
                Fe *= jac;
            }
            int nat = 0;
#pragma unroll
            for (int d = 0; d < 4; d++)
                if (L.v_bc[ip * L.nno + idx[d]] == 'S') nat += d;
            if (nat == 3) nat = 1;
            else if (nat == 1 || nat == 5) nat = 2;
            else if (nat == 0 || nat == 2) nat = 0;
            if (nat != 0) {
                double Lx = L.X[ip * L.nno + idx[2]] - L.X[ip * L.nno + idx[0]];
                Fe += 0.5 * Lx * L.v_bc_val[ip * L.nno + idx[a]];
            }
            F += Fe;
        }
        double diag = 0.0;
        if (inl) {
            for (int u = 0; u < L.WK; u++) {
                const double v = acc[u * MCL_BLOCK];
                L.Kval[u * L.neq + P] = v;
                if (u < nc && L.Kcol[u * L.neq + P] == P) diag = v;
            }
        } else {
            for (int s = 0; s < nc; s++)
                if (L.Kcol[s * L.neq + P] == P) diag = L.Kval[s * L.neq + P];
        }
        L.F[P] = F;
        L.Kdiag[P] = diag;
    }
}

// sum_i x_i y_i in index order, rounded after every product and every
// addition exactly as the reference's sequential loops (global_v_prod /
// global_p_prod, Array_work.c:287-341, at one rank).  Every workgroup computes
// it itself: the products in parallel into LDS, the additions in order by one
// thread, so all workgroups obtain the identical value without a barrier.
// The vectors must be complete (called after a grid barrier).
#define MCL_SEQ_CHUNK 2048
__device__ double seq_dot(const double *__restrict__ x, const double *__restrict__ y, int n)
{
    __shared__ double buf[MCL_SEQ_CHUNK];
    __shared__ double res;
    double s = 0.0;
    for (int base = 0; base < n; base += MCL_SEQ_CHUNK) {
        const int m = min(MCL_SEQ_CHUNK, n - base);
        __syncthreads();
        for (int t = threadIdx.x; t < m; t += blockDim.x) buf[t] = x[base + t] * y[base + t];
        __syncthreads();
        if (threadIdx.x == 0) {
            // loads batched sixteen at a time, additions strictly in order
            int t = 0;
            for (; t + 16 <= m; t += 16) {
                double v[16];
#pragma unroll
                for (int u = 0; u < 16; u++) v[u] = buf[t + u];
#pragma unroll
                for (int u = 0; u < 16; u++) s = s + v[u];
            }
            for (; t < m; t++) s = s + buf[t];
        }
    }
    if (threadIdx.x == 0) res = s;
    __syncthreads();
    return res;
}

// Conj_Grad (Solver.c:1589-1797) / Conj_Grad_heat (Update_temperature.c:
// 1140-1354) inside a persistent kernel.  All workgroups execute the same
// control flow: the scalars come from reduce_collect, identical everywhere.
// kind 0: velocity (constraint on velocity dofs), 1: heat.
// Returns the iteration count; *ok is cleared if a barrier aborted.
__device__ int pcg_persistent(const LevelDev &L, const EllMat &A, const double *__restrict__ b,
                              double *__restrict__ x, bool x_zero, double accu, int max_cycle, int precond,
                              int kind, CgVecs v, const GridSync &gs, unsigned &gen, const RedBuf &rb,
                              dd (*lds)[MCL_BLOCK / MCL_WAVE], double *ldsm, bool &ok, long long &spmv,
                              bool exact)
{
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    const int n = A.n;
    {
        dd acc[1] = {dd_zero()};
        for (int i = gtid; i < n; i += gsize) {
            double ax = x_zero ? 0.0 : ell_row(A.col, A.val, A.ncol[i], n, i, x);
            double r = b[i] - ax;
            double dv = precond ? r / ell_diag(A, i) : r;
            v.r[i] = r;
            v.d[i] = dv;
            acc[0] = dd_add_prod(acc[0], r, dv);
        }
        if (!x_zero) spmv++;
        reduce_publish<1>(rb, 0, acc, lds);
    }
    ok = ok && grid_sync(gs, gen);
    double delta_new;
    {
        double o[1];
        reduce_collect<1>(rb, 0, gridDim.x, o, lds);
        delta_new = exact ? seq_dot(v.r, v.d, n) : o[0];
    }
    const double delta0 = delta_new;
    int cycle = 0;
    while (ok && cycle < max_cycle && delta_new > accu * delta0) {
        // q = K d, d.q
        {
            dd acc[1] = {dd_zero()};
            for (int i = gtid; i < n; i += gsize) {
                double q = ell_row(A.col, A.val, A.ncol[i], n, i, v.d);
                v.q[i] = q;
                acc[0] = dd_add_prod(acc[0], v.d[i], q);
            }
            spmv++;
            reduce_publish<1>(rb, 1, acc, lds);
        }
        ok = ok && grid_sync(gs, gen);
        double t1;
        {
            double o[1];
            reduce_collect<1>(rb, 1, gridDim.x, o, lds);
            t1 = exact ? seq_dot(v.d, v.q, n) : o[0];
        }
        const double alpha = delta_new / t1;
        for (int i = gtid; i < n; i += gsize) x[i] += alpha * v.d[i];
        if (L.nh > 0) {
            ok = ok && grid_sync(gs, gen);
            if (kind == 0) constrain_v(L, x, gtid, gsize);
            else constrain_T(L, x, gtid, gsize);
        }
        const bool recompute = (cycle % 50) == 0;
        if (recompute) ok = ok && grid_sync(gs, gen);
        double delta_old = delta_new;
        {
            dd acc[1] = {dd_zero()};
            double mx = 0.0;
            for (int i = gtid; i < n; i += gsize) {
                double r;
                if (recompute) r = b[i] - ell_row(A.col, A.val, A.ncol[i], n, i, x);
                else r = v.r[i] - alpha * v.q[i];
                v.r[i] = r;
                double s = precond ? r / ell_diag(A, i) : r;
                v.s[i] = s;
                acc[0] = dd_add_prod(acc[0], r, s);
                mx = fmax(mx, fabs(alpha * r));
            }
            if (recompute) spmv++;
            reduce_publish<1>(rb, 2, acc, lds);
            double bm = block_max(mx, ldsm);
            if (threadIdx.x == 0) rb.pmax[0 * MCL_MAX_WG + blockIdx.x] = bm;
        }
        ok = ok && grid_sync(gs, gen);
        double res;
        {
            double o[1];
            reduce_collect<1>(rb, 2, gridDim.x, o, lds);
            delta_new = exact ? seq_dot(v.r, v.s, n) : o[0];
            double m = 0.0;
            for (int bb = threadIdx.x; bb < (int)gridDim.x; bb += MCL_BLOCK) m = fmax(m, rb.pmax[bb]);
            res = block_max(m, ldsm);
        }
        const double beta = delta_new / delta_old;
        for (int i = gtid; i < n; i += gsize) v.d[i] = v.s[i] + beta * v.d[i];
        cycle++;
        bool brk = precond ? (res < accu) : (kind == 0 ? res < accu : (res < accu && cycle > 10));
        ok = ok && grid_sync(gs, gen);
        if (brk) break;
    }
    return cycle;
}

// ------------------------------------------------ stationary inner solvers
// The reference's solver="jc", "gs" and "mg" inner solves (Solver.c:82-89),
// with its arithmetic: every row sum in ascending-column order with the
// diagonal left out (prod += a_ij x_j, x_i = (b_i - prod) / a_ii), the stop on
// max |x - x_old| < accu, no hanging-node constraint.

// jacobi (Solver.c:1800-1864) over the whole grid: sweep k reads one buffer
// and writes the other (x_old needs no copy); the sweep maxima alternate
// between two reduction slots, so a workgroup writing a slot two sweeps later
// cannot overtake a slower one still reading it.
__device__ int jacobi_grid(const EllMat &A, const double *__restrict__ b, double *x, double *xt, double accu,
                           int max_cycle, const GridSync &gs, unsigned &gen, const RedBuf &rb, double *ldsm, bool &ok,
                           long long &spmv)
{
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    const int n = A.n;
    double *src = x, *dst = xt;
    int cycle = 0;
    while (ok && cycle < max_cycle) {
        double mx = 0.0;
        for (int i = gtid; i < n; i += gsize) {
            double prod = 0.0, diag = 1.0;
            const int nc = A.ncol[i];
            for (int j = 0; j < nc; j++) {
                const int c = A.col[j * n + i];
                const double v = A.val[j * n + i];
                if (c != i) prod += v * src[c];
                else diag = v;
            }
            const double xn = (b[i] - prod) / diag;
            dst[i] = xn;
            mx = fmax(mx, fabs(xn - src[i]));
        }
        spmv++;
        const int slot = 4 + (cycle & 1);
        double bm = block_max(mx, ldsm);
        if (threadIdx.x == 0) rb.pmax[slot * MCL_MAX_WG + blockIdx.x] = bm;
        ok = ok && grid_sync(gs, gen);
        double m = 0.0;
        for (int bb = threadIdx.x; bb < (int)gridDim.x; bb += MCL_BLOCK) m = fmax(m, rb.pmax[slot * MCL_MAX_WG + bb]);
        const double diff = block_max(m, ldsm);
        double *t = src;
        src = dst;
        dst = t;
        cycle++;
        if (diff < accu) break;
    }
    if (src != x) {
        for (int i = gtid; i < n; i += gsize) x[i] = src[i];
        ok = ok && grid_sync(gs, gen);
    }
    return cycle;
}

// gauss_seidel (Solver.c:1866-1928) in one workgroup: each sweep walks the
// level's wavefront groups in order, the rows of a group in parallel.
__device__ int gs_block(const GsLevelDev &A, const double *b, double *x, double *xt, double accu, int max_cycle,
                        double *ldsm, long long &spmv)
{
    const int t = threadIdx.x, n = A.n;
    int cycle = 0;
    while (cycle < max_cycle) {
        for (int i = t; i < n; i += blockDim.x) xt[i] = x[i];
        __syncthreads();
        for (int g = 0; g < A.nsched; g++) {
            const int k1 = A.sptr[g + 1];
            for (int k = A.sptr[g] + t; k < k1; k += blockDim.x) {
                const int i = A.srows[k];
                double prod = 0.0, diag = 1.0;
                const int nc = A.ncol[i];
                for (int j = 0; j < nc; j++) {
                    const int c = A.col[j * n + i];
                    const double v = A.val[j * n + i];
                    if (c != i) prod += v * x[c];
                    else diag = v;
                }
                x[i] = (b[i] - prod) / diag;
            }
            __syncthreads();
        }
        spmv++;
        cycle++;
        double mx = 0.0;
        for (int i = t; i < n; i += blockDim.x) mx = fmax(mx, fabs(x[i] - xt[i]));
        if (block_max(mx, ldsm) < accu) break;
    }
    return cycle;
}

// multigrid + MG (Solver.c:668-769) in one workgroup, levels coarsest first
// (lv[0] = min_level, lv[nl-1] = the finest).  Below the finest the iterate
// is the level's err vector itself, so the prolongation overwrites the
// smoothed iterate before it is added, as in the reference.  err keeps its
// values across V-cycles and is zeroed once per call (initial_error).
__device__ int mg_block(const GsLevelDev *lv, int nl, const double *b, double *x, double *xt_cycle, double *xt,
                        double accu, int mg_cycle, int fi, int ci, double *ldsm, long long &spmv)
{
    const int t = threadIdx.x, top = nl - 1;
    for (int l = 0; l < nl; l++)
        for (int i = t; i < lv[l].n; i += blockDim.x) lv[l].err[i] = 0.0;
    __syncthreads();
    const int n = lv[top].n;
    int cycle = 0;
    bool conv = false;
    while (!conv && cycle < mg_cycle) {
        for (int i = t; i < n; i += blockDim.x) xt_cycle[i] = x[i];
        __syncthreads();
        for (int l = top; l >= 1; l--) {
            const GsLevelDev &A = lv[l];
            const double *rhs = l == top ? b : A.rhs;
            double *xl = l == top ? x : A.err;
            gs_block(A, rhs, xl, xt, accu, fi, ldsm, spmv);
            // get_residual (Solver.c:1291-1316): res = b, then res -= a_ij x_j in column order
            for (int i = t; i < A.n; i += blockDim.x) {
                double r = rhs[i];
                const int nc = A.ncol[i];
                for (int j = 0; j < nc; j++) r -= A.val[j * A.n + i] * xl[A.col[j * A.n + i]];
                A.resu[i] = r;
            }
            __syncthreads();
            // project_res (Solver.c:770-832): injection
            for (int k = t; k < A.nproj; k += blockDim.x) lv[l - 1].rhs[A.proj_l[k]] = A.resu[A.proj_u[k]];
            __syncthreads();
        }
        gs_block(lv[0], lv[0].rhs, lv[0].err, xt, accu, 1000, ldsm, spmv);
        for (int l = 1; l <= top; l++) {
            const GsLevelDev &A = lv[l];
            const double *rhs = l == top ? b : A.rhs;
            double *xl = l == top ? x : A.err;
            const double *ec = lv[l - 1].err;
            // interp_error (Solver.c:1009-1143, local part)
            for (int u = t; u < A.n; u += blockDim.x) {
                const int m = A.pn[u];
                if (m > 0) {
                    double v = 0.0;
                    for (int a = 0; a < m; a++) {
                        const int c = A.pcol[a * A.n + u];
                        if (c != -1) v += ec[c];
                    }
                    A.err[u] = v / m;
                }
            }
            __syncthreads();
            for (int i = t; i < A.n; i += blockDim.x) xl[i] += A.err[i];
            __syncthreads();
            gs_block(A, rhs, xl, xt, accu, ci, ldsm, spmv);
        }
        double mx = 0.0;
        for (int i = t; i < n; i += blockDim.x) mx = fmax(mx, fabs(x[i] - xt_cycle[i]));
        conv = block_max(mx, ldsm) < accu;
        cycle++;
    }
    return cycle;
}

// The inner solve of citcom_solver (Solver.c:82-89) for the chosen solver.
__device__ int inner_solve(const UzawaParams &P, const LevelDev &L, const EllMat &K, const double *b, double *x,
                           bool x_zero, CgVecs cv, unsigned &gen, dd (*lds)[MCL_BLOCK / MCL_WAVE], double *ldsm,
                           bool &ok, long long &spmv)
{
    switch (P.solver) {
    case MCL_INNER_JC:
        return jacobi_grid(K, b, x, P.xt, P.accuracy, 1000, P.gs, gen, P.rb, ldsm, ok, spmv);
    case MCL_INNER_GS: {
        const int c = gs_block(P.gsl[P.ngsl - 1], b, x, P.xt, P.accuracy, 1000, ldsm, spmv);
        ok = ok && grid_sync(P.gs, gen);
        return c;
    }
    case MCL_INNER_MG: {
        const int c = mg_block(P.gsl, P.ngsl, b, x, P.xt2, P.xt, P.accuracy, P.mg_cycle, P.mg_finest_iteration,
                               P.mg_coarse_iteration, ldsm, spmv);
        ok = ok && grid_sync(P.gs, gen);
        return c;
    }
    default:
        return pcg_persistent(L, K, b, x, x_zero, P.accuracy, P.max_cg_cycle, P.cg_precondition, 0, cv, P.gs, gen,
                              P.rb, lds, ldsm, ok, spmv, P.exact_dots != 0);
    }
}

// citcom_solver (Solver.c:14-249) as one persistent kernel.
__global__ void __launch_bounds__(MCL_BLOCK) k_uzawa_fp64(UzawaParams P)
{
    __shared__ dd lds[MCL_RED_ROWS][MCL_BLOCK / MCL_WAVE];
    __shared__ double ldsm[MCL_BLOCK / MCL_WAVE];
    const LevelDev &L = P.L;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    const int neq = L.neq, nel = L.nel;
    if (blockIdx.x == 0 && threadIdx.x == 0) P.stats->t_begin = wall_clock64();
    unsigned gen = __hip_atomic_load(P.gs.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    bool ok = true;
    long long spmv = 0, inner = 0;
    EllMat K{L.Kcol, L.Kncol, L.Kval, L.Kdiag, nullptr, neq};
    CgVecs cv{P.cg_r, P.cg_s, P.cg_d, P.cg_q};
    // Preconditioner data, Solver.c:54-64: BAB[e] = sum_i G_ie^2 / Kdiag_i
    for (int e = gtid; e < nel; e += gsize) {
        double s = 0.0;
        int lev = L.elev[e];
        for (int t = 0; t < 8; t++) {
            int r = L.gt_row[t * nel + e];
            if (r < 0) break;
            double g = L.gtab[lev * 8 + L.gt_p[t * nel + e]];
            s += g * g / L.Kdiag[r];
        }
        P.BAB[e] = s;
    }
    // u1 = K d ; Ah = F - G p - u1   (:67-77)
    for (int i = gtid; i < neq; i += gsize) {
        double u = ell_row(L.Kcol, L.Kval, L.Kncol[i], neq, i, P.d);
        double gp = g_row(L, i, P.p);
        P.u1[i] = u;
        P.Ahv[i] = L.F[i] - gp - u;
    }
    spmv++;
    ok = ok && grid_sync(P.gs, gen);
    inner += inner_solve(P, L, K, P.Ahv, P.u1, false, cv, gen, lds, ldsm, ok, spmv);
    for (int i = gtid; i < neq; i += gsize) P.d[i] += P.u1[i];
    ok = ok && grid_sync(P.gs, gen);
    double *r0 = P.r0, *r1 = P.r1, *r2 = P.r2, *z0 = P.z0, *z1 = P.z1, *s1 = P.s1, *s2 = P.s2;
    for (int e = gtid; e < nel; e += gsize) r1[e] = gt_elem(L, e, P.d);
    double dv = 1.0, dp = 1.0;
    int cycle = 0;
    while (ok && (dv > P.citcom_accu || (dp > P.citcom_accu && P.check_v_only == 0)) &&
           cycle < P.max_citcom_cycle) {
        // z1 = BAB r1 (or r1 / BAB); r1.z1, r0.z0
        {
            dd acc[2] = {dd_zero(), dd_zero()};
            for (int e = gtid; e < nel; e += gsize) {
                double z = P.schur ? r1[e] / P.BAB[e] : P.BAB[e] * r1[e];
                z1[e] = z;
                acc[0] = dd_add_prod(acc[0], r1[e], z);
                if (cycle > 0) acc[1] = dd_add_prod(acc[1], r0[e], z0[e]);
            }
            reduce_publish<2>(P.rb, 4, acc, lds);
        }
        ok = ok && grid_sync(P.gs, gen);
        double r1z1, delta = 0.0;
        {
            double o[2];
            reduce_collect<2>(P.rb, 4, gridDim.x, o, lds);
            if (P.exact_dots) {
                o[0] = seq_dot(r1, z1, nel);
                if (cycle > 0) o[1] = seq_dot(r0, z0, nel);
            }
            r1z1 = o[0];
            if (cycle > 0) delta = r1z1 / o[1];
        }
        if (!(r1z1 == r1z1)) P.stats->nan_flag = 1;
        for (int e = gtid; e < nel; e += gsize) s2[e] = cycle == 0 ? z1[e] : z1[e] + delta * s1[e];
        ok = ok && grid_sync(P.gs, gen);
        // Ah = G s2 ; u1 = 0 ; K u1 = Ah
        for (int i = gtid; i < neq; i += gsize) {
            P.Ahv[i] = g_row(L, i, s2);
            P.u1[i] = 0.0;
        }
        ok = ok && grid_sync(P.gs, gen);
        inner += inner_solve(P, L, K, P.Ahv, P.u1, true, cv, gen, lds, ldsm, ok, spmv);
        // Ah = G^T u1 ; alpha = r1.z1 / s2.Ah
        {
            dd acc[1] = {dd_zero()};
            for (int e = gtid; e < nel; e += gsize) {
                double a = gt_elem(L, e, P.u1);
                P.Ahp[e] = a;
                acc[0] = dd_add_prod(acc[0], s2[e], a);
            }
            reduce_publish<1>(P.rb, 6, acc, lds);
        }
        ok = ok && grid_sync(P.gs, gen);
        double alpha;
        {
            double o[1];
            reduce_collect<1>(P.rb, 6, gridDim.x, o, lds);
            if (P.exact_dots) o[0] = seq_dot(s2, P.Ahp, nel);
            alpha = r1z1 / o[0];
        }
        {
            dd acc[4] = {dd_zero(), dd_zero(), dd_zero(), dd_zero()};
            for (int e = gtid; e < nel; e += gsize) {
                r2[e] = r1[e] - alpha * P.Ahp[e];
                double pe = P.p[e] + alpha * s2[e];
                P.p[e] = pe;
                acc[0] = dd_add_prod(acc[0], s2[e], s2[e]);
                acc[1] = dd_add_prod(acc[1], pe, pe);
            }
            for (int i = gtid; i < neq; i += gsize) {
                double u = P.u1[i];
                double di = P.d[i] - alpha * u;
                P.d[i] = di;
                acc[2] = dd_add_prod(acc[2], u, u);
                acc[3] = dd_add_prod(acc[3], di, di);
            }
            reduce_publish<4>(P.rb, 8, acc, lds);
        }
        ok = ok && grid_sync(P.gs, gen);
        {
            double o[4];
            reduce_collect<4>(P.rb, 8, gridDim.x, o, lds);
            if (P.exact_dots) {
                o[0] = seq_dot(s2, s2, nel);
                o[1] = seq_dot(P.p, P.p, nel);
                o[2] = seq_dot(P.u1, P.u1, neq);
                o[3] = seq_dot(P.d, P.d, neq);
            }
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
        P.stats->cycles = cycle;
        P.stats->dv = dv;
        P.stats->dp = dp;
        P.stats->inner = inner;
        P.stats->spmv = spmv;
        P.stats->aborted = ok ? 0 : 1;
        if (!(dv == dv) || !(dp == dp)) P.stats->nan_flag = 1;
        P.stats->t_end = wall_clock64();
    }
}

// Zero n 32-bit words (a kernel rather than hipMemsetAsync, so that a caller
// may capture it in a graph).
__global__ void k_zero_words(unsigned *p, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// v_to_nodes (Solver.c:1319-1343)
__global__ void k_v_to_nodes(LevelDev L, const double *__restrict__ d, double *__restrict__ V)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L.nno; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 2; k++) {
            int j = L.id[k * L.nno + i];
            V[k * L.nno + i] = j >= 0 ? d[j] : L.v_bc_val[k * L.nno + i];
        }
}

// compute_stress element part (Grain_size_evolution.c:154-209)
__global__ void k_stress_elem(LevelDev L, const double *__restrict__ vis,
                              const double *__restrict__ V, double *__restrict__ es, double *__restrict__ area)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L.nel; e += gridDim.x * blockDim.x) {
        int lev = L.elev[e];
        double dxdxi = 0.5 * L.lvl_dx[lev], dzdeta = 0.5 * L.lvl_dz[lev];
        double jac = dxdxi * dzdeta - 0.0 * 0.0;
        double eta = 0.0, u11 = 0.0, u22 = 0.0, u12 = 0.0;
        for (int a = 0; a < 4; a++) eta += 0.25 * vis[L.ienl[a * L.nel + e]];
        for (int a = 0; a < 4; a++) {
            int k = L.ienl[a * L.nel + e];
            double ax = c_shape.dxic[a] * dzdeta - c_shape.detac[a] * 0.0;
            double az = c_shape.detac[a] * dxdxi - c_shape.dxic[a] * 0.0;
            u11 += ax * V[k];
            u22 += az * V[L.nno + k];
            u12 += az * V[k];
        }
        u11 /= jac;
        u22 /= jac;
        u12 /= jac;
        double u21 = u12;
        double sr = u11 * u11 + u22 * u22 + u12 * u21 * 2.0;
        sr = sqrt(0.5 * sr);
        es[e] = 2.0 * eta * sr;
        int n0 = L.ienl[e], n2 = L.ienl[2 * L.nel + e];
        area[e] = (L.X[n2] - L.X[n0]) * (L.X[L.nno + n2] - L.X[L.nno + n0]);
    }
}

// p_to_nodes (Solver.c:1345-1366): area-weighted mean over the node's elements.
__global__ void k_stress_node(LevelDev L, const double *__restrict__ es, const double *__restrict__ area,
                              double *__restrict__ stress)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L.nno; i += gridDim.x * blockDim.x) {
        double s = 0.0, t = 0.0;
        for (int k = 0; k < L.p2n_cnt[i]; k++) {
            int e = L.p2n_el[k * L.nno + i];
            s += area[e] * es[e];
            t += area[e];
        }
        stress[i] = s / t;
    }
}

// True residual norms after a solve: |F - K d - G p|^2, |G^T d|^2, |F|^2 as
// partial sums in reduction slots 12..14 (k_collect3 finishes them).
// Residual norms after a solve: |F - K d - G p|^2, |G^T d|^2 and |F|^2.
// galerkin != 0 (the mixed path on meshes with hanging nodes, which solves
// C^T K C u_f = C^T b): the momentum residual and F are measured in the
// constrained space, C^T (F - K d - G p) and C^T F, rows of hanging
// equations folded into their free parents.
__global__ void k_residuals(LevelDev L, RedBuf rb, const double *__restrict__ d, const double *__restrict__ p,
                            int galerkin)
{
    __shared__ dd lds[MCL_RED_ROWS][MCL_BLOCK / MCL_WAVE];
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    const bool gal = galerkin && L.nh > 0;
    dd acc[3] = {dd_zero(), dd_zero(), dd_zero()};
    for (int i = gtid; i < L.neq; i += gsize) {
        if (gal && L.hp_n[i] != 0) continue;
        double r = L.F[i] - ell_row(L.Kcol, L.Kval, L.Kncol[i], L.neq, i, d) - g_row(L, i, p);
        double f = L.F[i];
        if (gal) {
            for (int t = L.ch_ptr[i]; t < L.ch_ptr[i + 1]; t++) {
                const int h = L.ch_i[t];
                const double w = L.ch_w[t];
                r += w * (L.F[h] - ell_row(L.Kcol, L.Kval, L.Kncol[h], L.neq, h, d) - g_row(L, h, p));
                f += w * L.F[h];
            }
        }
        acc[0] = dd_add_prod(acc[0], r, r);
        acc[2] = dd_add_prod(acc[2], f, f);
    }
    for (int e = gtid; e < L.nel; e += gsize) {
        double c = gt_elem(L, e, d);
        acc[1] = dd_add_prod(acc[1], c, c);
    }
    reduce_publish<3>(rb, 12, acc, lds);
}

__global__ void k_collect3(RedBuf rb, int slot0, int nblocks, double *out)
{
    __shared__ dd lds[MCL_RED_ROWS][MCL_BLOCK / MCL_WAVE];
    double o[3];
    reduce_collect<3>(rb, slot0, nblocks, o, lds);
    if (threadIdx.x == 0)
        for (int r = 0; r < 3; r++) out[r] = o[r];
}

// ------------------------------------------------------------------ heat step

// SUPG tau, calculate_artificial_diffusivity (Update_temperature.c:241-291),
// and each element's advective time-step bound (advection_dt, :135-177, which
// uses the sizes L of the last element of the diffusive loop for every element).
__device__ __forceinline__ void elem_center_u(const LevelDev &L, int e, const double *__restrict__ V, double u[2])
{
    for (int d = 0; d < 2; d++) {
        double s = 0.0;
        for (int a = 0; a < 4; a++) s += 0.25 * V[d * L.nno + L.ienl[a * L.nel + e]];
        u[d] = s;
    }
}

// Heat row assembly (K_matrix_heat :293-423 with get_element_K_heat,
// F_matrix_heat :425-508, C_matrix_heat :510-594): one thread per heat row,
// element-ascending sums accumulated in place in the row's own slots.
__device__ void heat_assemble_row(const LevelDev &L, const HeatParams &H, int P)
{
    const int n = L.heq_node[P];
    const int nadj = L.p2n_cnt[n];
    const int nT = L.neqT;
    for (int s = 0; s < L.WT; s++) {
        H.hK[s * nT + P] = 0.0;
        H.hC[s * nT + P] = 0.0;
    }
    double F = 0.0;
    for (int k = 0; k < nadj; k++) {
        int e = L.p2n_el[k * L.nno + n], a = L.p2n_a[k * L.nno + n];
        int lev = L.elev[e];
        double dxdxi = 0.5 * L.lvl_dx[lev], dzdeta = 0.5 * L.lvl_dz[lev];
        double jac = dxdxi * dzdeta - 0.0 * 0.0;
        double u1 = 0.0, u2 = 0.0;
        for (int c = 0; c < 4; c++) {
            int kn = L.ienl[c * L.nel + e];
            u1 += H.V[kn] * 0.25;
            u2 += H.V[L.nno + kn] * 0.25;
        }
        double t0 = H.tau[e], t1 = H.tau[L.nel + e];
        for (int b = 0; b < 4; b++) {
            double kv = 0.0, cv = 0.0;
            for (int g = 0; g < 4; g++) {
                double ax = c_shape.dxi[a][g] * dzdeta - c_shape.deta[a][g] * 0.0;
                double az = c_shape.deta[a][g] * dxdxi - c_shape.dxi[a][g] * 0.0;
                double bx = c_shape.dxi[b][g] * dzdeta - c_shape.deta[b][g] * 0.0;
                double bz = c_shape.deta[b][g] * dxdxi - c_shape.dxi[b][g] * 0.0;
                double nh = t0 * ax + t1 * az;
                double tt1 = u1 * c_shape.N[a][g] * bx + u2 * c_shape.N[a][g] * bz;
                double tt2 = u1 * nh * bx + u2 * nh * bz;
                double tt3 = ax * 1.0 * bx + ax * 0.0 * bz + az * 0.0 * bx + az * 1.0 * bz;
                kv += tt1 + (tt2 + tt3) / jac;
                cv += (c_shape.N[a][g] + nh) * c_shape.N[b][g];
            }
            cv *= jac;
            int slot = L.Tslot[(k * 4 + b) * nT + P];
            if (slot >= 0) {
                H.hK[slot * nT + P] += kv;
                H.hC[slot * nT + P] += cv;
            }
            int nb = L.ienl[b * L.nel + e];
            if (L.T_bc[nb] == 'T') F -= kv * L.T_bc_val[nb];
        }
    }
    for (int k = 0; k < nadj; k++) {
        int e = L.p2n_el[k * L.nno + n], a = L.p2n_a[k * L.nno + n];
        int lev = L.elev[e];
        double dxdxi = 0.5 * L.lvl_dx[lev], dzdeta = 0.5 * L.lvl_dz[lev];
        double jac = dxdxi * dzdeta - 0.0 * 0.0;
        double t0 = H.tau[e], t1 = H.tau[L.nel + e];
        int nat = 0;
        for (int c = 0; c < 4; c++)
            if (L.T_bc[L.ienl[c * L.nel + e]] == 'F') nat += c;
        if (nat == 3) nat = 1;
        else if (nat == 1 || nat == 5) nat = 2;
        else if (nat == 0 || nat == 2) nat = 0;
        double f = 0.0;
        for (int g = 0; g < 4; g++) {
            double ax = c_shape.dxi[a][g] * dzdeta - c_shape.deta[a][g] * 0.0;
            double az = c_shape.deta[a][g] * dxdxi - c_shape.dxi[a][g] * 0.0;
            double nh = t0 * ax + t1 * az;
            f += (c_shape.N[a][g] + nh) * H.Q;
        }
        f *= jac;
        if (nat != 0) f -= 0.5 * L.lvl_dz[lev] * L.T_bc_val[L.ienl[a * L.nel + e]];
        F += f;
    }
    H.hF[P] = F;
}

// update_temperature (Update_temperature.c:3-133) as one persistent kernel.
__global__ void __launch_bounds__(MCL_BLOCK) k_heat_fp64(HeatParams H)
{
    __shared__ dd lds[MCL_RED_ROWS][MCL_BLOCK / MCL_WAVE];
    __shared__ double ldsm[MCL_BLOCK / MCL_WAVE];
    const LevelDev &L = H.L;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    const int n = L.neqT, nel = L.nel;
    unsigned gen = __hip_atomic_load(H.gs.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    bool ok = true;
    long long spmv = 0;
    // tau per element and the advective dt bound
    double dtmin = 1e10;
    for (int e = gtid; e < nel; e += gsize) {
        int lev = L.elev[e];
        double Lx = L.lvl_dx[lev], Lz = L.lvl_dz[lev];
        double u[2];
        elem_center_u(L, e, H.V, u);
        double al[2], xi[2];
        double Ld[2] = {Lx, Lz};
        for (int d = 0; d < 2; d++) {
            al[d] = 0.5 * u[d] * Ld[d] / 1.0;
            xi[d] = fabs(al[d]) < H.accuracy ? 0.0 : cosh(al[d]) / sinh(al[d]) - 1.0 / al[d];
        }
        double us = (u[0] * u[0] + u[1] * u[1]);
        double kh = (xi[0] * u[0] * Ld[0] + xi[1] * u[1] * Ld[1]) * 0.5;
        H.tau[e] = kh * u[0] / us;
        H.tau[nel + e] = kh * u[1] / us;
        // advection_dt with the sizes of the last element (H.Llast)
        double uu = fabs(u[0]) / H.Llast0 + fabs(u[1]) / H.Llast1;
        double step = 0.5 / uu;
        dtmin = step < dtmin ? step : dtmin;
    }
    {
        double bm = block_min(dtmin, ldsm);
        if (threadIdx.x == 0) H.rb.pmax[1 * MCL_MAX_WG + blockIdx.x] = bm;
    }
    ok = ok && grid_sync(H.gs, gen);
    for (int P = gtid; P < n; P += gsize) heat_assemble_row(L, H, P);
    double dt;
    {
        double m = 1e10;
        for (int b = threadIdx.x; b < (int)gridDim.x; b += MCL_BLOCK) {
            double v = H.rb.pmax[1 * MCL_MAX_WG + b];
            m = v < m ? v : m;
        }
        dt = block_min(m, ldsm);
        dt = dt < H.diff_dt ? dt : H.diff_dt;
        dt = dt * 0.75;
    }
    ok = ok && grid_sync(H.gs, gen);
    EllMat Kh{L.Tcol, L.Tncol, H.hK, nullptr, L.Tdiag_slot, n};
    EllMat Ch{L.Tcol, L.Tncol, H.hC, nullptr, L.Tdiag_slot, n};
    EllMat Ah{L.Tcol, L.Tncol, H.hA, nullptr, L.Tdiag_slot, n};
    CgVecs cv{H.cg_r, H.cg_s, H.cg_d, H.cg_q};
    int init_cycles = 0;
    if (H.first_call) {
        // initial_T_dot (:179-239)
        for (int i = gtid; i < n; i += gsize) H.hd[i] = H.T[L.heq_node[i]];
        ok = ok && grid_sync(H.gs, gen);
        for (int i = gtid; i < n; i += gsize) {
            H.B[i] = H.hF[i] - ell_row(L.Tcol, H.hK, L.Tncol[i], n, i, H.hd);
            H.hdd[i] = 0.0;
        }
        ok = ok && grid_sync(H.gs, gen);
        init_cycles = pcg_persistent(L, Ch, H.B, H.hdd, true, 1e-3, 400, H.cg_precondition, 1, cv, H.gs, gen,
                                     H.rb, lds, ldsm, ok, spmv, H.exact_dots != 0);
    }
    const double td = H.theta * dt;
    for (int i = gtid; i < n; i += gsize) {
        double ct = ell_row(L.Tcol, H.hC, L.Tncol[i], n, i, H.hd);
        H.d0[i] = H.hd[i];
        H.B[i] = H.hF[i] + ct / td;
        for (int s = 0; s < L.Tncol[i]; s++) H.hA[s * n + i] = H.hC[s * n + i] / td + H.hK[s * n + i];
    }
    ok = ok && grid_sync(H.gs, gen);
    int cycles = pcg_persistent(L, Ah, H.B, H.hd, false, 1e-3, 1000, H.cg_precondition, 1, cv, H.gs, gen, H.rb, lds,
                                ldsm, ok, spmv, H.exact_dots != 0);
    for (int i = gtid; i < L.nno; i += gsize) {
        int k = L.heat_id[i];
        H.T[i] = k >= 0 ? H.hd[k] : L.T_bc_val[i];
    }
    for (int i = gtid; i < n; i += gsize)
        H.hdd[i] = 1.0 / (H.theta * dt) * (H.hd[i] - H.d0[i]) - (1.0 - H.theta) / H.theta * H.hdd[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        H.stats->cycles = cycles;
        H.stats->init_cycles = init_cycles;
        H.stats->dt = dt;
        H.stats->aborted = ok ? 0 : 1;
        *H.model_time += dt;
    }
}
