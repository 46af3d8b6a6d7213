// Host side of libmcl_hip.so: the C-ABI declared in include/mcl.h.
//
// mcl_mesh_upload turns the reference's mesh arrays (one view per quadtree
// level) into the device layout of LevelDev: local connectivity, node ->
// element adjacency, sorted row patterns with per-(element, dof) slot maps for
// atomic-free row assembly, element -> row lists for G^T, the buoyancy row
// contribution lists, and the static time-step terms.  All integer work here
// runs once per mesh.  The per-step entry points only launch kernels on the
// context's stream; nothing is copied between host and device unless the
// caller asks for it (mcl_set_field / mcl_get_field).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mcl.h"
#include "mcl_kernels.h"
#include "mcl_net.h"

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct Level {
    // host copies of the views needed later
    int L = 0, lev = 0, nel = 0, nno = 0, neq = 0, neqT = 0, gnox = 0, gnoz = 0;
    std::vector<int> Mcode, ienl, node, id, heat_id;
    std::vector<signed char> v_bc, T_bc;
    std::vector<double> v_bc_val, T_bc_val, X;
    std::vector<int> Kncol_h, Kcol_h, Tncol_h, Tcol_h;   // patterns (slot-major)
    int WK = 0, WT = 0;
    double Llast0 = 0, Llast1 = 0, diff_dt = 0;
    LevelDev dev{};
    // device fields
    double *T = nullptr, *vis = nullptr, *buoy = nullptr, *V = nullptr, *d = nullptr, *p = nullptr;
    double *stress = nullptr, *es = nullptr, *area = nullptr, *avg = nullptr;
    double *bterms = nullptr;  // [4][entries] terms of the buoyancy row averages
    int bnent = 0;
    // multigrid (mixed precision): prolongation maps of this level (fine side)
    std::vector<int> interp_n, interp_eq;
    int *c2f = nullptr;                       // coarse node -> finest local node
    float *val32 = nullptr, *dinv32 = nullptr, *mx = nullptr, *mb = nullptr, *mr = nullptr, *mt = nullptr;
    int *pn = nullptr, *pcol = nullptr, *rcol = nullptr;
    float *rval = nullptr;
    int rw = 0;
    std::vector<int> rcol_h;                  // host copy of rcol ([rw][neq], -1 padded)
    // injection into the level below (the reference's project maps, solver="mg")
    std::vector<int> proj_l, proj_u;
    // Jacobi / Gauss-Seidel / multigrid inner solvers: wavefront schedule and level vectors
    int nsched = 0;
    int *sptr = nullptr, *srows = nullptr, *dproj_l = nullptr, *dproj_u = nullptr;
    double *gerr = nullptr, *grhs = nullptr, *gres = nullptr;
    // augmented Lagrangian of the mixed solve (K_al, its diagonal, w)
    double *kval_al = nullptr, *kdiag_al = nullptr, *al_w = nullptr;
    // buoyancy row lists
    int *brow_ptr = nullptr, *brow_e = nullptr;
    signed char *brow_xi = nullptr, *brow_eta = nullptr;
    std::vector<void *> owned;
};

}  // namespace

struct mcl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int nlev = 0;
    const AsmLevel *asm_tab = nullptr;  // [nlev] all-levels assembly table (device)
    int asm_blocks = 0;                 // workgroups of its launch
    const PrepLevel *prep_tab = nullptr;  // [prep_n] FP32 preparation table of the mixed solve (device)
    int prep_n = 0, prep_blocks = 0;
    double al_gamma = 0.0;               // > 0: the mixed solve runs on K_al (al_tab: [prep_n] levels)
    double *al_state = nullptr;          // device: the next solve's weight (al_gamma or 0)
    // warm-start extrapolation of the mixed solve (MixedParams::ws): history
    // and previous solution on the device; extrap_done: this step's
    // extrapolation already ran (before the learned guess's gate)
    double *ws = nullptr, *hist = nullptr;
    int extrap_after = 5;                // solves after a reset before it extrapolates
    bool extrap_on = true, extrap_done = false;
    const AlLevel *al_tab = nullptr;
    int al_eblocks = 0, al_rblocks = 0;
    std::vector<Level> levels;
    mcl_control ctl{};
    bool have_ctl = false;
    bool first_heat = true;
    bool been_here_dp = false;
    int max_coop_blocks = 0;
    int block_cap = 0;               // > 0: most workgroups of any persistent launch (set by a batch)
    int num_cu = 0;
    int wall_khz = 0;                // rate of the device wall clock (wall_clock64)
    // solver work space (finest level)
    GridSync gs{};
    RedBuf rb{};
    DevStats *stats = nullptr;       // device
    double *work = nullptr;          // device scratch vectors
    size_t work_doubles = 0;
    double *model_time = nullptr;    // device scalar
    double *resid = nullptr;         // device [3]
    // mixed precision: multigrid level table, smoother bounds, coarse inverse, FP32 CG vectors
    MgLevelDev *mg_dev = nullptr;      // level descriptors of the mixed solve
    std::vector<MgLevelDev> mg_host;   // host copy (the halo lists are added per workgroup count)
    std::vector<void *> halo_bufs;     // device allocations of the halo lists
    unsigned *lam_bits = nullptr;
    double *cinv = nullptr;
    double *rr = nullptr;
    bool have_mg = false;
    bool mixed_ready = false;        // FP32 hierarchy matches the assembled K
    int mg_lb = 0;                   // highest multigrid level kept in LDS
    int nbr_nb = 0;                  // workgroup count the neighbour lists were built for
    int nbr_ok = 0;                  // lists usable (every list within MCL_BLOCK)
    int *nbr_ptr = nullptr, *nbr_idx = nullptr;
    // the reference's stationary inner solvers (jc, gs, mg)
    const GsLevelDev *gs_tab = nullptr;  // [nlev] (device)
    bool gs_mg_ok = false;           // transfer maps of every level present (solver="mg")
    double *xt = nullptr, *xt2 = nullptr;
    double *gd = nullptr, *gp = nullptr;  // learned-guess candidate (d, p)
    // learned guess: per-row partial sums and the sample scale [2 H + 1]
    double *net_rows = nullptr;
    int net_rows_h = 0;
    // heat work (finest level)
    double *tau = nullptr, *hK = nullptr, *hC = nullptr, *hA = nullptr, *hF = nullptr;
    double *hd = nullptr, *hdd = nullptr, *hd0 = nullptr, *hB = nullptr;
    std::vector<void *> owned;
};

static int fail(mcl_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, call)                                                                                   \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) return fail(ctx, MCL_E_DEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                              \
    } while (0)

template <typename T>
static int dev_alloc(mcl_ctx *c, std::vector<void *> &owner, T **out, size_t count)
{
    void *p = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(c, MCL_E_DEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    hipMemset(p, 0, bytes);
    owner.push_back(p);
    *out = (T *)p;
    return 0;
}

template <typename T>
static int dev_upload(mcl_ctx *c, std::vector<void *> &owner, const T **out, const std::vector<T> &h)
{
    T *p = nullptr;
    int rc = dev_alloc(c, owner, &p, h.size());
    if (rc) return rc;
    if (!h.empty()) {
        hipError_t e = hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(c, MCL_E_DEVICE, "hipMemcpy: %s", hipGetErrorString(e));
    }
    *out = p;
    return 0;
}

static void free_all(std::vector<void *> &v)
{
    for (void *p : v) hipFree(p);
    v.clear();
}

// Shape tables computed with the expressions of Shape_function.c:17-121.
static void shape_tables(ShapeTables *s)
{
    const double pts[4][2] = {{-1.0, -1.0}, {-1.0, 1.0}, {1.0, 1.0}, {1.0, -1.0}};
    volatile double beta = 1.0 / std::sqrt(3.0);
    for (int g = 0; g < 4; g++) {
        volatile double xi = pts[g][0] * beta, eta = pts[g][1] * beta;
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
    s->dxic[0] = -0.25; s->dxic[1] = -0.25; s->dxic[2] = 0.25; s->dxic[3] = 0.25;
    s->detac[0] = -0.25; s->detac[1] = 0.25; s->detac[2] = 0.25; s->detac[3] = -0.25;
}

// ------------------------------------------------------------ static scratch
// Work space of mcl_solve_velocity_pressure_noalloc: the FP64 Uzawa solve's
// grid-barrier words, statistics, residual norms, reduction partials and work
// vectors, in one file-scope device array per GPU (each GPU the library runs
// on holds its copy until the process ends).  Host code never uses the array's
// name as an address: each device's copy is located once, with that device
// current, by hipGetSymbolAddress at mcl_create, and only that pointer is
// handed to kernels.  Every call on a GPU shares its copy, so allocation-free
// solves that overlap on one GPU race.
#define MCL_SCRATCH_BYTES (48ull << 20)
#define MCL_SCRATCH_DEVICES 64
static __device__ __align__(256) unsigned char scratch[MCL_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static unsigned char *g_scratch[MCL_SCRATCH_DEVICES];
static hipError_t g_scratch_err[MCL_SCRATCH_DEVICES];

// Locate the current device's copy (once per device); call with `device` current.
static void scratch_locate(int device)
{
    if (device < 0 || device >= MCL_SCRATCH_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    if (e == hipSuccess) g_scratch[device] = (unsigned char *)p;
}

extern "C" int mcl_abi_version(void) { return MCL_ABI_VERSION; }

extern "C" const char *mcl_last_error(const mcl_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int mcl_create(int device, mcl_ctx **out)
{
    if (!out) return MCL_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return MCL_E_DEVICE;
    mcl_ctx *c = new mcl_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return MCL_E_DEVICE; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return MCL_E_DEVICE; }
    c->num_cu = prop.multiProcessorCount;
    if (hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess) c->wall_khz = 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return MCL_E_DEVICE; }
    scratch_locate(device);
    ShapeTables st;
    shape_tables(&st);
    if (mcl_upload_shape(&st) != hipSuccess) { delete c; return MCL_E_DEVICE; }
    int per_cu_u = 0, per_cu_h = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_u, (const void *)k_uzawa_fp64, MCL_BLOCK, 0);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_h, (const void *)k_heat_fp64, MCL_BLOCK, 0);
    int per_cu_m = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_m, (const void *)k_uzawa_mixed, MCL_BLOCK, 0);
    int per_cu = std::min(std::min(per_cu_u, per_cu_h), per_cu_m);
    c->max_coop_blocks = std::min(std::max(1, per_cu) * c->num_cu, MCL_MAX_WG);
    // grid barrier and reduction buffers
    unsigned *sync = nullptr;
    if (dev_alloc(c, c->owned, &sync, MCL_SYNC_WORDS)) { delete c; return MCL_E_DEVICE; }
    c->gs.gen = sync;
    c->gs.abort = sync + 32;
    c->gs.arrive = sync + 64;
    if (dev_alloc(c, c->owned, &c->rb.part, (size_t)MCL_RED_SLOTS * MCL_MAX_WG) ||
        dev_alloc(c, c->owned, &c->rb.pmax, (size_t)MCL_RED_SLOTS * MCL_MAX_WG) ||
        dev_alloc(c, c->owned, &c->stats, 1) || dev_alloc(c, c->owned, &c->model_time, 1) ||
        dev_alloc(c, c->owned, &c->resid, 4)) {
        free_all(c->owned);
        delete c;
        return MCL_E_DEVICE;
    }
    *out = c;
    return 0;
}

extern "C" void mcl_destroy(mcl_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto &L : c->levels) free_all(L.owned);
    free_all(c->owned);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

// Local node index of a global node id (node[] is sorted ascending).
static int local_of(const std::vector<int> &node, int g)
{
    auto it = std::lower_bound(node.begin(), node.end(), g);
    if (it == node.end() || *it != g) return -1;
    return (int)(it - node.begin());
}

static int build_level(mcl_ctx *c, const mcl_mesh_view &v, Level &Lv)
{
    const int nel = v.nel, nno = v.nno, neq = v.neq, neqT = v.neqT;
    if (nel <= 0 || nno <= 0 || neq < 0 || neqT < 0 || v.L <= 0 || v.L > MCL_MAX_LEVELS)
        return fail(c, MCL_E_ARG, "level %d: bad sizes", v.lev);
    if (!v.Mcode || !v.ien || !v.node || !v.id || !v.lm || !v.heat_id || !v.heat_lm || !v.v_bc || !v.v_bc_val ||
        !v.T_bc || !v.T_bc_val || !v.X || !v.dx || !v.dz)
        return fail(c, MCL_E_ARG, "level %d: missing array", v.lev);
    Lv.L = v.L; Lv.lev = v.lev; Lv.nel = nel; Lv.nno = nno; Lv.neq = neq; Lv.neqT = neqT;
    Lv.gnox = v.gnox; Lv.gnoz = v.gnoz;
    Lv.Mcode.assign(v.Mcode, v.Mcode + nel);
    Lv.node.assign(v.node, v.node + nno);
    for (int i = 1; i < nno; i++)
        if (Lv.node[i] <= Lv.node[i - 1]) return fail(c, MCL_E_ARG, "level %d: node list not sorted", v.lev);
    Lv.id.assign(v.id, v.id + 2 * (size_t)nno);
    Lv.heat_id.assign(v.heat_id, v.heat_id + nno);
    Lv.v_bc.assign(v.v_bc, v.v_bc + 2 * (size_t)nno);
    Lv.v_bc_val.assign(v.v_bc_val, v.v_bc_val + 2 * (size_t)nno);
    Lv.T_bc.assign(v.T_bc, v.T_bc + nno);
    Lv.T_bc_val.assign(v.T_bc_val, v.T_bc_val + nno);
    Lv.X.assign(v.X, v.X + 2 * (size_t)nno);

    std::vector<int> elev(nel), ienl(4 * (size_t)nel);
    for (int e = 0; e < nel; e++) {
        elev[e] = v.Mcode[e] & 15;
        if (elev[e] >= v.L) return fail(c, MCL_E_ARG, "element %d: level %d >= L", e, elev[e]);
        for (int a = 0; a < 4; a++) {
            int li = local_of(Lv.node, v.ien[a * nel + e]);
            if (li < 0) return fail(c, MCL_E_ARG, "element %d corner %d: node not in node list", e, a);
            ienl[a * (size_t)nel + e] = li;
        }
    }
    Lv.ienl = ienl;
    // node -> elements (ascending), corner index
    std::vector<int> cnt(nno, 0), p2n_el(4 * (size_t)nno, -1), p2n_a(4 * (size_t)nno, -1);
    for (int e = 0; e < nel; e++)
        for (int a = 0; a < 4; a++) {
            int n = ienl[a * (size_t)nel + e];
            if (cnt[n] >= 4) return fail(c, MCL_E_ARG, "node %d in more than 4 elements", n);
            p2n_el[cnt[n] * (size_t)nno + n] = e;
            p2n_a[cnt[n] * (size_t)nno + n] = a;
            cnt[n]++;
        }
    // equations -> node, dof (ids numbered in node order)
    std::vector<int> veq_node(std::max(neq, 1), -1), veq_dof(std::max(neq, 1), -1), heq_node(std::max(neqT, 1), -1);
    for (int i = 0; i < nno; i++) {
        for (int d = 0; d < 2; d++) {
            int q = v.id[d * nno + i];
            if (q >= neq) return fail(c, MCL_E_ARG, "id out of range");
            if (q >= 0) { veq_node[q] = i; veq_dof[q] = d; }
        }
        int t = v.heat_id[i];
        if (t >= neqT) return fail(c, MCL_E_ARG, "heat id out of range");
        if (t >= 0) heq_node[t] = i;
    }
    for (int q = 0; q < neq; q++)
        if (veq_node[q] < 0) return fail(c, MCL_E_ARG, "velocity equation %d has no node", q);
    for (int q = 0; q < neqT; q++)
        if (heq_node[q] < 0) return fail(c, MCL_E_ARG, "heat equation %d has no node", q);
    // lm consistency with id
    for (int e = 0; e < nel; e++)
        for (int d = 0; d < 2; d++)
            for (int a = 0; a < 4; a++)
                if (v.lm[(d * 4 + a) * (size_t)nel + e] != v.id[d * nno + ienl[a * (size_t)nel + e]])
                    return fail(c, MCL_E_ARG, "lm inconsistent with id at element %d", e);
    // velocity row patterns: sorted unique columns over the node's elements
    std::vector<std::vector<int>> rows(neq);
    int WK = 0;
    for (int P = 0; P < neq; P++) {
        int n = veq_node[P];
        auto &r = rows[P];
        for (int k = 0; k < cnt[n]; k++) {
            int e = p2n_el[k * (size_t)nno + n];
            for (int q = 0; q < 8; q++) {
                int Q = v.lm[((q & 1) * 4 + (q >> 1)) * (size_t)nel + e];
                if (Q >= 0) r.push_back(Q);
            }
        }
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        WK = std::max<int>(WK, (int)r.size());
    }
    if (WK > 127) return fail(c, MCL_E_ARG, "velocity row too long (%d)", WK);
    std::vector<int> Kcol((size_t)std::max(WK, 1) * std::max(neq, 1), 0), Kncol(std::max(neq, 1), 0);
    std::vector<signed char> Kslot(32 * (size_t)std::max(neq, 1), -1);
    for (int P = 0; P < neq; P++) {
        auto &r = rows[P];
        Kncol[P] = (int)r.size();
        for (size_t s = 0; s < r.size(); s++) Kcol[s * neq + P] = r[s];
        for (size_t s = r.size(); s < (size_t)WK; s++) Kcol[s * neq + P] = P;  // padding reads x[P] * 0
        int n = veq_node[P];
        for (int k = 0; k < cnt[n]; k++) {
            int e = p2n_el[k * (size_t)nno + n];
            for (int q = 0; q < 8; q++) {
                int Q = v.lm[((q & 1) * 4 + (q >> 1)) * (size_t)nel + e];
                if (Q < 0) continue;
                int s = (int)(std::lower_bound(r.begin(), r.end(), Q) - r.begin());
                Kslot[(k * 8 + q) * (size_t)neq + P] = (signed char)s;
            }
        }
    }
    Lv.WK = WK;
    Lv.Kncol_h = Kncol;
    Lv.Kcol_h = Kcol;
    // element -> free rows ascending (G^T)
    std::vector<int> gt_row(8 * (size_t)nel, -1);
    std::vector<signed char> gt_p(8 * (size_t)nel, 0);
    for (int e = 0; e < nel; e++) {
        std::vector<std::pair<int, int>> rp;
        for (int d = 0; d < 2; d++)
            for (int a = 0; a < 4; a++) {
                int P = v.lm[(d * 4 + a) * (size_t)nel + e];
                if (P >= 0) rp.push_back({P, 2 * a + d});
            }
        std::sort(rp.begin(), rp.end());
        for (size_t t = 0; t < rp.size(); t++) {
            gt_row[t * nel + e] = rp[t].first;
            gt_p[t * nel + e] = (signed char)rp[t].second;
        }
    }
    // heat row patterns
    std::vector<std::vector<int>> hrows(neqT);
    int WT = 0;
    for (int P = 0; P < neqT; P++) {
        int n = heq_node[P];
        auto &r = hrows[P];
        for (int k = 0; k < cnt[n]; k++) {
            int e = p2n_el[k * (size_t)nno + n];
            for (int b = 0; b < 4; b++) {
                int Q = v.heat_lm[b * (size_t)nel + e];
                if (Q >= 0) r.push_back(Q);
            }
        }
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        WT = std::max<int>(WT, (int)r.size());
    }
    std::vector<int> Tcol((size_t)std::max(WT, 1) * std::max(neqT, 1), 0), Tncol(std::max(neqT, 1), 0),
        Tdiag(std::max(neqT, 1), 0);
    std::vector<signed char> Tslot(16 * (size_t)std::max(neqT, 1), -1);
    for (int P = 0; P < neqT; P++) {
        auto &r = hrows[P];
        Tncol[P] = (int)r.size();
        for (size_t s = 0; s < r.size(); s++) Tcol[s * neqT + P] = r[s];
        for (size_t s = r.size(); s < (size_t)WT; s++) Tcol[s * neqT + P] = P;
        auto it = std::lower_bound(r.begin(), r.end(), P);
        if (it == r.end() || *it != P) return fail(c, MCL_E_ARG, "heat row %d has no diagonal", P);
        Tdiag[P] = (int)(it - r.begin());
        int n = heq_node[P];
        for (int k = 0; k < cnt[n]; k++) {
            int e = p2n_el[k * (size_t)nno + n];
            for (int b = 0; b < 4; b++) {
                int Q = v.heat_lm[b * (size_t)nel + e];
                if (Q < 0) continue;
                Tslot[(k * 4 + b) * (size_t)neqT + P] =
                    (signed char)(std::lower_bound(r.begin(), r.end(), Q) - r.begin());
            }
        }
    }
    Lv.WT = WT;
    Lv.Tncol_h = Tncol;
    Lv.Tcol_h = Tcol;
    if (v.nproj > 0 && v.proj_l && v.proj_u) {
        Lv.proj_l.assign(v.proj_l, v.proj_l + v.nproj);
        Lv.proj_u.assign(v.proj_u, v.proj_u + v.nproj);
        for (int k = 0; k < v.nproj; k++)
            if (Lv.proj_u[k] < 0 || Lv.proj_u[k] >= neq)
                return fail(c, MCL_E_ARG, "proj_u[%d] = %d outside the level's %d equations", k, Lv.proj_u[k], neq);
    }
    if (v.interp_n && v.interp_eq) {
        Lv.interp_n.assign(v.interp_n, v.interp_n + neq);
        Lv.interp_eq.assign(v.interp_eq, v.interp_eq + 4 * (size_t)neq);
        for (int u = 0; u < neq; u++)
            if (Lv.interp_n[u] < 0 || Lv.interp_n[u] > 4)
                return fail(c, MCL_E_ARG, "interp_n[%d] = %d outside 0..4", u, Lv.interp_n[u]);
    }
    // buoyancy row-average contribution lists, in the reference's order
    // (get_node_horiz_avg, Parallel.c:354-437)
    std::vector<std::vector<int>> be(v.gnoz);
    std::vector<std::vector<signed char>> bxi(v.gnoz), beta(v.gnoz);
    for (int e = 0; e < nel; e++) {
        int n0 = v.ien[0 * nel + e], n2 = v.ien[2 * nel + e];
        int ix0 = n0 / v.gnoz, iz0 = n0 - ix0 * v.gnoz, ix1 = n2 / v.gnoz, iz1 = n2 - ix1 * v.gnoz;
        for (int i = ix0; i < ix1; i++) {
            int xi = (2 * i - ix0 - ix1) / (ix1 - ix0);
            for (int k = iz0; k < iz1; k++) {
                int eta = (2 * k - iz0 - iz1) / (iz1 - iz0);
                be[k].push_back(e); bxi[k].push_back((signed char)xi); beta[k].push_back((signed char)eta);
            }
        }
        if (iz1 == v.gnoz - 1)
            for (int i = ix0; i < ix1; i++) {
                int xi = (2 * i - ix0 - ix1) / (ix1 - ix0);
                be[iz1].push_back(e); bxi[iz1].push_back((signed char)xi); beta[iz1].push_back(1);
            }
        if (ix1 == v.gnox - 1)
            for (int k = iz0; k < iz1; k++) {
                int eta = (2 * k - iz0 - iz1) / (iz1 - iz0);
                be[k].push_back(e); bxi[k].push_back(1); beta[k].push_back((signed char)eta);
            }
        if (n2 == v.gnox * v.gnoz - 1) {
            be[v.gnoz - 1].push_back(-1); bxi[v.gnoz - 1].push_back(0); beta[v.gnoz - 1].push_back(0);
        }
    }
    std::vector<int> bptr(v.gnoz + 1, 0), bent;
    std::vector<signed char> bx, bz;
    for (int k = 0; k < v.gnoz; k++) {
        bent.insert(bent.end(), be[k].begin(), be[k].end());
        bx.insert(bx.end(), bxi[k].begin(), bxi[k].end());
        bz.insert(bz.end(), beta[k].begin(), beta[k].end());
        bptr[k + 1] = (int)bent.size();
    }
    // static dt terms of advection_dt (Update_temperature.c:135-160)
    double diff = 1e10, L0 = 0.0, L1 = 0.0;
    for (int e = 0; e < nel; e++) {
        L0 = v.dx[elev[e]];
        L1 = v.dz[elev[e]];
        double t0 = L0 * L0, t1 = L1 * L1;
        if (diff > t0) diff = t0;
        if (diff > t1) diff = t1;
    }
    Lv.Llast0 = L0; Lv.Llast1 = L1; Lv.diff_dt = diff;
    // level tables: sizes and element gradient vectors (get_element_G)
    ShapeTables sh;
    shape_tables(&sh);
    std::vector<double> ldx(MCL_MAX_LEVELS, 0.0), ldz(MCL_MAX_LEVELS, 0.0), gtab(MCL_MAX_LEVELS * 8, 0.0);
    for (int l = 0; l < v.L; l++) {
        ldx[l] = v.dx[l];
        ldz[l] = v.dz[l];
        volatile double dxdxi = 0.5 * v.dx[l], dzdeta = 0.5 * v.dz[l];
        const double zero = 0.0;
        for (int a = 0; a < 4; a++) {
            volatile double g0 = 0.0, g1 = 0.0;
            for (int gs = 0; gs < 4; gs++) {
                volatile double t = sh.dxi[a][gs] * dzdeta;
                volatile double u = sh.deta[a][gs] * zero;
                g0 = g0 - (t - u);
            }
            for (int gs = 0; gs < 4; gs++) {
                volatile double t = sh.deta[a][gs] * dxdxi;
                volatile double u = sh.dxi[a][gs] * zero;
                g1 = g1 - (t - u);
            }
            gtab[l * 8 + 2 * a] = g0;
            gtab[l * 8 + 2 * a + 1] = g1;
        }
    }
    // hanging nodes: parallel when no parent is itself a hanging node
    int nh = v.nh > 0 ? v.nh : 0, h_serial = 0;
    std::vector<int> h_ni(std::max(nh, 1), 0), h_ci(2 * (size_t)std::max(nh, 1), 0);
    if (nh > 0) {
        if (!v.h_ni || !v.h_ci) return fail(c, MCL_E_ARG, "hanging arrays missing");
        std::vector<char> is_h(nno, 0);
        for (int k = 0; k < nh; k++) {
            h_ni[k] = v.h_ni[k];
            h_ci[k] = v.h_ci[k];
            h_ci[nh + k] = v.h_ci[nh + k];
            if (h_ni[k] >= 0) is_h[h_ni[k]] = 1;
        }
        for (int k = 0; k < nh; k++)
            if (is_h[h_ci[k]] || is_h[h_ci[nh + k]]) h_serial = 1;
    }

    LevelDev &D = Lv.dev;
    std::vector<void *> &o = Lv.owned;
    D.L = v.L; D.lev = v.lev; D.nel = nel; D.nno = nno; D.neq = neq; D.neqT = neqT; D.gnox = v.gnox; D.gnoz = v.gnoz;
    std::vector<int> lm(v.lm, v.lm + 8 * (size_t)nel);
    std::vector<int> nodev(Lv.node), idv(Lv.id), hid(Lv.heat_id);
    int rc = 0;
    rc |= dev_upload(c, o, &D.lvl_dx, ldx);
    rc |= dev_upload(c, o, &D.lvl_dz, ldz);
    rc |= dev_upload(c, o, &D.gtab, gtab);
    rc |= dev_upload(c, o, &D.ienl, ienl);
    rc |= dev_upload(c, o, &D.elev, elev);
    rc |= dev_upload(c, o, &D.lm, lm);
    rc |= dev_upload(c, o, &D.gt_row, gt_row);
    rc |= dev_upload(c, o, &D.gt_p, gt_p);
    rc |= dev_upload(c, o, &D.node, nodev);
    rc |= dev_upload(c, o, &D.p2n_cnt, cnt);
    rc |= dev_upload(c, o, &D.p2n_el, p2n_el);
    rc |= dev_upload(c, o, &D.p2n_a, p2n_a);
    rc |= dev_upload(c, o, &D.id, idv);
    rc |= dev_upload(c, o, &D.heat_id, hid);
    rc |= dev_upload(c, o, &D.v_bc, Lv.v_bc);
    rc |= dev_upload(c, o, &D.v_bc_val, Lv.v_bc_val);
    rc |= dev_upload(c, o, &D.T_bc, Lv.T_bc);
    rc |= dev_upload(c, o, &D.T_bc_val, Lv.T_bc_val);
    rc |= dev_upload(c, o, &D.X, Lv.X);
    rc |= dev_upload(c, o, &D.veq_node, veq_node);
    rc |= dev_upload(c, o, &D.veq_dof, veq_dof);
    D.WK = WK;
    rc |= dev_upload(c, o, &D.Kcol, Kcol);
    rc |= dev_upload(c, o, &D.Kncol, Kncol);
    rc |= dev_upload(c, o, &D.Kslot, Kslot);
    rc |= dev_alloc(c, o, &D.Kval, (size_t)std::max(WK, 1) * std::max(neq, 1));
    rc |= dev_alloc(c, o, &D.Kdiag, std::max(neq, 1));
    rc |= dev_alloc(c, o, &D.F, std::max(neq, 1));
    rc |= dev_upload(c, o, &D.heq_node, heq_node);
    D.WT = WT;
    rc |= dev_upload(c, o, &D.Tcol, Tcol);
    rc |= dev_upload(c, o, &D.Tncol, Tncol);
    rc |= dev_upload(c, o, &D.Tslot, Tslot);
    rc |= dev_upload(c, o, &D.Tdiag_slot, Tdiag);
    D.nh = nh;
    D.h_serial = h_serial;
    {
        // Galerkin constraint maps (see LevelDev): expand each hanging
        // equation to weighted free parents, then the transpose
        std::vector<int> hidx(nno, -1);
        for (int k = 0; k < nh; k++)
            if (h_ni[k] >= 0) hidx[h_ni[k]] = k;
        const size_t nq = std::max(neq, 1);
        std::vector<int> hpn(nq, 0), hpi(4 * nq, 0);
        std::vector<double> hpw(4 * nq, 0.0);
        std::vector<std::vector<std::pair<int, double>>> ch(nq);
        bool over = false;
        for (int k = 0; k < nh; k++) {
            for (int dd = 0; dd < 2; dd++) {
                int eq = Lv.id[dd * nno + h_ni[k]];
                if (eq < 0) continue;
                std::vector<std::pair<int, double>> par;
                std::vector<std::pair<int, double>> todo = {{h_ci[k], 0.5}, {h_ci[nh + k], 0.5}};
                int guard = 0;
                while (!todo.empty() && guard++ < 64) {
                    auto [node, w] = todo.back();
                    todo.pop_back();
                    if (hidx[node] >= 0) {
                        int kk = hidx[node];
                        todo.push_back({h_ci[kk], 0.5 * w});
                        todo.push_back({h_ci[nh + kk], 0.5 * w});
                    } else {
                        int pe = Lv.id[dd * nno + node];
                        if (pe >= 0) par.push_back({pe, w});
                    }
                }
                if (par.size() > 4) { over = true; continue; }
                hpn[eq] = (int)par.size();
                if (par.empty()) hpn[eq] = -1;  // all parents fixed: the value is 0
                for (size_t t = 0; t < par.size(); t++) {
                    hpi[t * nq + eq] = par[t].first;
                    hpw[t * nq + eq] = par[t].second;
                    ch[par[t].first].push_back({eq, par[t].second});
                }
            }
        }
        if (over) return fail(c, MCL_E_UNSUPPORTED, "hanging node with more than 4 free parents");
        std::vector<int> cptr(nq + 1, 0), ci;
        std::vector<double> cw;
        for (size_t i = 0; i < nq; i++) {
            for (auto &pw : ch[i]) { ci.push_back(pw.first); cw.push_back(pw.second); }
            cptr[i + 1] = (int)ci.size();
        }
        if (ci.empty()) { ci.push_back(0); cw.push_back(0.0); }
        int rc2 = 0;
        rc2 |= dev_upload(c, Lv.owned, &D.hp_n, hpn);
        rc2 |= dev_upload(c, Lv.owned, &D.hp_i, hpi);
        rc2 |= dev_upload(c, Lv.owned, &D.hp_w, hpw);
        rc2 |= dev_upload(c, Lv.owned, &D.ch_ptr, cptr);
        rc2 |= dev_upload(c, Lv.owned, &D.ch_i, ci);
        rc2 |= dev_upload(c, Lv.owned, &D.ch_w, cw);
        if (rc2) return MCL_E_DEVICE;
    }
    rc |= dev_upload(c, o, &D.h_ni, h_ni);
    rc |= dev_upload(c, o, &D.h_ci, h_ci);
    const int *cp;
    const signed char *sp;
    rc |= dev_upload(c, o, &cp, bptr); Lv.brow_ptr = (int *)cp;
    rc |= dev_upload(c, o, &cp, bent); Lv.brow_e = (int *)cp;
    rc |= dev_upload(c, o, &sp, bx); Lv.brow_xi = (signed char *)sp;
    rc |= dev_upload(c, o, &sp, bz); Lv.brow_eta = (signed char *)sp;
    Lv.bnent = (int)bent.size();
    rc |= dev_alloc(c, o, &Lv.bterms, 4 * std::max<size_t>(bent.size(), 1));
    rc |= dev_alloc(c, o, &Lv.T, nno);
    rc |= dev_alloc(c, o, &Lv.vis, nno);
    rc |= dev_alloc(c, o, &Lv.buoy, nno);
    rc |= dev_alloc(c, o, &Lv.V, 2 * (size_t)nno);
    rc |= dev_alloc(c, o, &Lv.d, std::max(neq, 1));
    rc |= dev_alloc(c, o, &Lv.p, nel);
    rc |= dev_alloc(c, o, &Lv.stress, nno);
    rc |= dev_alloc(c, o, &Lv.es, nel);
    rc |= dev_alloc(c, o, &Lv.area, nel);
    rc |= dev_alloc(c, o, &Lv.avg, v.gnoz);
    return rc ? (c->err.empty() ? fail(c, MCL_E_DEVICE, "device allocation failed") : MCL_E_DEVICE) : 0;
}

// Multigrid data for the mixed-precision inner solver: per level an FP32 copy
// of K, prolongation from the level below (the reference's interp maps), the
// restriction R = P^T stored on the coarse level, level vectors, and the map of
// coarse nodes to finest nodes used to inject the temperature.
static int block_rows_threshold()
{
    const char *e = getenv("MCL_MG_BLOCK_ROWS");
    return e ? atoi(e) : 1024;
}

static unsigned coop_blocks(mcl_ctx *c, long long n, int rows_per_block, const char *knob);
static int build_neighbours(mcl_ctx *c, int nb);
// the mixed solve: one workgroup per 128 rows up to one per CU (256 from
// 128x128 up; 64 at 64x64, 16 at 32x32), the fastest counts measured there
#define MCL_MIXED_ROWS_PER_BLOCK 128

static int build_multigrid(mcl_ctx *c)
{
    c->have_mg = false;
    const int nlev = c->nlev;
    Level &Fn = c->levels.back();
    std::vector<MgLevelDev> mg(nlev);
    for (int k = 0; k < nlev; k++) {
        Level &Lv = c->levels[k];
        std::vector<void *> &o = Lv.owned;
        int n = std::max(Lv.neq, 1);
        int rc = 0;
        rc |= dev_alloc(c, o, &Lv.val32, (size_t)std::max(Lv.WK, 1) * n);
        rc |= dev_alloc(c, o, &Lv.dinv32, n);
        rc |= dev_alloc(c, o, &Lv.mx, n);
        rc |= dev_alloc(c, o, &Lv.mb, n);
        rc |= dev_alloc(c, o, &Lv.mr, n);
        rc |= dev_alloc(c, o, &Lv.mt, n);
        if (rc) return MCL_E_DEVICE;
        MgLevelDev &M = mg[k];
        M.n = Lv.neq;
        M.WK = Lv.WK;
        M.block = 0;
        M.lds_off = 0;
        M.col = Lv.dev.Kcol;
        M.ncol = Lv.dev.Kncol;
        M.val = Lv.val32;
        M.dinv = Lv.dinv32;
        M.val64 = Lv.dev.Kval;
        M.diag64 = Lv.dev.Kdiag;
        M.x = Lv.mx;
        M.b = Lv.mb;
        M.r = Lv.mr;
        M.t = Lv.mt;
        // the coarse node -> finest node map
        if (k < nlev - 1) {
            std::vector<int> c2f(Lv.nno);
            for (int i = 0; i < Lv.nno; i++) {
                c2f[i] = local_of(Fn.node, Lv.node[i]);
                if (c2f[i] < 0) return fail(c, MCL_E_ARG, "level %d node %d not on the finest level", Lv.lev, i);
            }
            const int *cp;
            if (dev_upload(c, o, &cp, c2f)) return MCL_E_DEVICE;
            Lv.c2f = (int *)cp;
        }
    }
    // coarse levels run per workgroup from LDS: level 0 (the dense solve) and
    // the levels above it while each has at most block_rows_threshold() rows,
    // their rows sum to at most MCL_MG_LDS_ROWS, their ELL entries to at most
    // MCL_MG_LDS_NNZ, and the level is not the finest
    {
        int off = 0, nnz = 0;
        for (int k = 0; k < nlev - 1; k++) {
            Level &Lv = c->levels[k];
            const int n = Lv.neq, m = Lv.WK * n;
            if (k > 0 && n > block_rows_threshold()) break;
            if (off / 3 + n > MCL_MG_LDS_ROWS || nnz + m > MCL_MG_LDS_NNZ || n > 65535) break;
            std::vector<unsigned short> col16(std::max(m, 1), 0);
            for (int e = 0; e < m; e++) col16[e] = (unsigned short)Lv.Kcol_h[e];
            const unsigned short *cp;
            if (dev_upload(c, Lv.owned, &cp, col16)) return MCL_E_DEVICE;
            mg[k].col16 = cp;
            mg[k].block = 1;
            mg[k].lds_off = off;
            mg[k].moff = nnz;
            off += 3 * n;
            nnz += m;
            c->mg_lb = k;
        }
    }
    c->nbr_nb = 0;
    bool maps = nlev > 1;
    for (int k = 1; k < nlev; k++)
        if (c->levels[k].interp_n.empty()) maps = false;
    if (maps) {
        for (int k = 1; k < nlev; k++) {
            Level &F = c->levels[k], &C = c->levels[k - 1];
            const int nf = F.neq, nc = C.neq;
            // prolongation
            std::vector<int> pcol(4 * (size_t)std::max(nf, 1), -1);
            for (int a = 0; a < 4; a++)
                for (int u = 0; u < nf; u++)
                    if (F.interp_n[u] > a) pcol[a * (size_t)nf + u] = F.interp_eq[a * (size_t)nf + u];
            const int *cp;
            if (dev_upload(c, F.owned, &cp, F.interp_n)) return MCL_E_DEVICE;
            F.pn = (int *)cp;
            if (dev_upload(c, F.owned, &cp, pcol)) return MCL_E_DEVICE;
            F.pcol = (int *)cp;
            mg[k].pn = F.pn;
            mg[k].pcol = F.pcol;
            // restriction = transpose, rows of the coarse level
            std::vector<std::vector<std::pair<int, float>>> rows(std::max(nc, 1));
            for (int u = 0; u < nf; u++) {
                int m = F.interp_n[u];
                for (int a = 0; a < m; a++) {
                    int l = F.interp_eq[a * (size_t)nf + u];
                    if (l >= 0) rows[l].push_back({u, 1.0f / (float)m});
                }
            }
            int rw = 1;
            for (auto &r : rows) rw = std::max<int>(rw, (int)r.size());
            std::vector<int> rcol((size_t)rw * std::max(nc, 1), -1);
            std::vector<float> rval((size_t)rw * std::max(nc, 1), 0.f);
            for (int l = 0; l < nc; l++)
                for (size_t t = 0; t < rows[l].size(); t++) {
                    rcol[t * nc + l] = rows[l][t].first;
                    rval[t * nc + l] = rows[l][t].second;
                }
            const float *fp;
            if (dev_upload(c, C.owned, &cp, rcol)) return MCL_E_DEVICE;
            C.rcol = (int *)cp;
            C.rcol_h = rcol;
            if (dev_upload(c, C.owned, &fp, rval)) return MCL_E_DEVICE;
            C.rval = (float *)fp;
            C.rw = rw;
            mg[k - 1].rw = rw;
            mg[k - 1].rcol = C.rcol;
            mg[k - 1].rval = C.rval;
        }
        if (c->levels[0].neq > 64) maps = false;  // the coarse solve holds a dense inverse of at most 64 rows
    }
    for (void *p : {(void *)c->mg_dev, (void *)c->lam_bits, (void *)c->cinv, (void *)c->rr}) {
        if (p) {
            (void)hipFree(p);
            c->owned.erase(std::remove(c->owned.begin(), c->owned.end(), p), c->owned.end());
        }
    }
    int rc = 0;
    // deeper hierarchies than the kernel's LDS copy holds use the Jacobi
    // preconditioner (a larger mg_min_level shortens the hierarchy)
    if (nlev > MCL_MG_MAX_LEVELS) maps = false;
    c->mg_dev = nullptr;
    if (maps) {
        rc |= dev_alloc(c, c->owned, &c->mg_dev, nlev);
        if (!rc) rc |= hipMemcpy(c->mg_dev, mg.data(), nlev * sizeof(MgLevelDev), hipMemcpyHostToDevice) != hipSuccess;
        c->mg_host = mg;
    } else {
        // single level: Jacobi preconditioner from the finest entry's FP32 1/diag
        rc |= dev_alloc(c, c->owned, &c->mg_dev, 1);
        MgLevelDev top = mg[nlev - 1];
        top.block = 0;
        if (!rc) rc |= hipMemcpy(c->mg_dev, &top, sizeof(MgLevelDev), hipMemcpyHostToDevice) != hipSuccess;
        c->mg_host.assign(1, top);
    }
    rc |= dev_alloc(c, c->owned, &c->lam_bits, MCL_MAX_LEVELS);
    rc |= dev_alloc(c, c->owned, &c->cinv, 64 * 64);
    rc |= dev_alloc(c, c->owned, &c->rr, std::max(Fn.neq, 1));
    if (rc) return fail(c, MCL_E_DEVICE, "multigrid allocation failed");
    c->have_mg = maps;
    c->mixed_ready = false;
    // the mixed solve's neighbour and halo lists for its default workgroup
    // count, built here with the rest of the mesh set-up (a different count,
    // MCL_MIXED_BLOCKS, rebuilds them at the first solve)
    if (maps) {
        Level &T = c->levels.back();
        rc = build_neighbours(c, (int)coop_blocks(c, std::max(T.neq, T.nel), MCL_MIXED_ROWS_PER_BLOCK,
                                                  "MCL_MIXED_BLOCKS"));
        if (rc) return rc;
    }
    return 0;
}

// Wavefront schedule of a Gauss-Seidel sweep over rows 0..n-1 in order: row i
// must be updated after every lower-numbered row it is coupled to (it reads
// their new values) and before every higher-numbered one (they read its old
// value), so group(i) = 1 + max group of its coupled lower rows; rows of one
// group are independent.  The coupling is taken in both directions, so an
// unsymmetric pattern is handled too.
static void gs_schedule(const Level &Lv, std::vector<int> &ptr, std::vector<int> &rows)
{
    const int n = Lv.neq;
    std::vector<std::vector<int>> lower(n);
    for (int i = 0; i < n; i++)
        for (int s = 0; s < Lv.Kncol_h[i]; s++) {
            const int j = Lv.Kcol_h[(size_t)s * n + i];
            if (j < i) lower[i].push_back(j);
            else if (j > i) lower[j].push_back(i);
        }
    std::vector<int> grp(n, 0);
    int ng = 0;
    for (int i = 0; i < n; i++) {
        int g = 0;
        for (int j : lower[i]) g = std::max(g, grp[j] + 1);
        grp[i] = g;
        ng = std::max(ng, g + 1);
    }
    ptr.assign(ng + 1, 0);
    for (int i = 0; i < n; i++) ptr[grp[i] + 1]++;
    for (int g = 0; g < ng; g++) ptr[g + 1] += ptr[g];
    rows.assign(std::max(n, 1), 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; i++) rows[fill[grp[i]]++] = i;   // ascending within a group
}

// Data of the reference's stationary inner solvers (solver = jc, gs, mg):
// per level the wavefront schedule, the injection maps and three vectors,
// the level table, and two finest-level vectors.
static int build_stationary(mcl_ctx *c)
{
    const int nlev = c->nlev;
    std::vector<GsLevelDev> tab(nlev);
    bool mg_ok = nlev >= 2;
    for (int k = 0; k < nlev; k++) {
        Level &Lv = c->levels[k];
        std::vector<int> ptr, rows;
        gs_schedule(Lv, ptr, rows);
        const int *cp;
        if (dev_upload(c, Lv.owned, &cp, ptr)) return MCL_E_DEVICE;
        Lv.sptr = (int *)cp;
        if (dev_upload(c, Lv.owned, &cp, rows)) return MCL_E_DEVICE;
        Lv.srows = (int *)cp;
        Lv.nsched = (int)ptr.size() - 1;
        const int n = std::max(Lv.neq, 1);
        if (dev_alloc(c, Lv.owned, &Lv.gerr, n) || dev_alloc(c, Lv.owned, &Lv.grhs, n) ||
            dev_alloc(c, Lv.owned, &Lv.gres, n))
            return MCL_E_DEVICE;
        GsLevelDev &G = tab[k];
        G.n = Lv.neq;
        G.WK = Lv.WK;
        G.col = Lv.dev.Kcol;
        G.ncol = Lv.dev.Kncol;
        G.val = Lv.dev.Kval;
        G.nsched = Lv.nsched;
        G.sptr = Lv.sptr;
        G.srows = Lv.srows;
        G.err = Lv.gerr;
        G.rhs = Lv.grhs;
        G.resu = Lv.gres;
        if (k > 0) {
            Level &C = c->levels[k - 1];
            // every equation of the level below must receive an injected residual
            std::vector<char> hit(std::max(C.neq, 1), 0);
            for (size_t q = 0; q < Lv.proj_l.size(); q++) {
                const int l = Lv.proj_l[q];
                if (l < 0 || l >= C.neq) return fail(c, MCL_E_ARG, "proj_l[%zu] = %d outside level %d", q, l, C.lev);
                hit[l] = 1;
            }
            bool covered = !Lv.proj_l.empty();
            for (int l = 0; l < C.neq && covered; l++) covered = hit[l] != 0;
            if (!covered || Lv.interp_n.empty() || !Lv.pn) mg_ok = false;
            G.pn = Lv.pn;
            G.pcol = Lv.pcol;
            G.nproj = (int)Lv.proj_l.size();
            if (G.nproj > 0) {
                if (dev_upload(c, Lv.owned, &cp, Lv.proj_l)) return MCL_E_DEVICE;
                Lv.dproj_l = (int *)cp;
                if (dev_upload(c, Lv.owned, &cp, Lv.proj_u)) return MCL_E_DEVICE;
                Lv.dproj_u = (int *)cp;
            }
            G.proj_l = Lv.dproj_l;
            G.proj_u = Lv.dproj_u;
        }
    }
    const GsLevelDev *tp;
    if (dev_upload(c, c->levels.back().owned, &tp, tab)) return MCL_E_DEVICE;
    c->gs_tab = tp;
    c->gs_mg_ok = mg_ok;
    for (void *p : {(void *)c->xt, (void *)c->xt2}) {
        if (p) {
            (void)hipFree(p);
            c->owned.erase(std::remove(c->owned.begin(), c->owned.end(), p), c->owned.end());
        }
    }
    c->xt = c->xt2 = nullptr;
    const size_t nv = std::max(c->levels.back().neq, 1);
    if (dev_alloc(c, c->owned, &c->xt, nv) || dev_alloc(c, c->owned, &c->xt2, nv)) return MCL_E_DEVICE;
    return 0;
}

extern "C" int mcl_mesh_upload(mcl_ctx *c, const mcl_mesh_view *levels, int nlev)
{
    if (!c) return MCL_E_ARG;
    if (!levels || nlev <= 0 || nlev > MCL_MAX_LEVELS) return fail(c, MCL_E_ARG, "bad level list");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto &L : c->levels) free_all(L.owned);
    c->levels.clear();
    c->levels.resize(nlev);
    for (int l = 0; l < nlev; l++) {
        int rc = build_level(c, levels[l], c->levels[l]);
        if (rc) {
            for (auto &L : c->levels) free_all(L.owned);
            c->levels.clear();
            c->nlev = 0;
            return rc;
        }
    }
    c->nlev = nlev;
    {
        int rc = build_multigrid(c);
        if (!rc) rc = build_stationary(c);
        if (rc) return rc;
    }
    {
        // all-levels assembly table (owned by the finest level, so a new
        // upload frees it with the levels)
        std::vector<AsmLevel> tab(nlev);
        int blk = 0;
        for (int l = 0; l < nlev; l++) {
            Level &Lv = c->levels[l];
            tab[l].L = Lv.dev;
            tab[l].vis = Lv.vis;
            tab[l].buoy = Lv.buoy;
            tab[l].blk0 = blk;
            blk += (Lv.neq + MCL_BLOCK - 1) / MCL_BLOCK;
        }
        const AsmLevel *tp = nullptr;
        int rc = dev_upload(c, c->levels[nlev - 1].owned, &tp, tab);
        if (rc) return rc;
        c->asm_tab = tp;
        c->asm_blocks = blk;
    }
    {
        // warm-start history of the mixed solve (MCL_EXTRAPOLATE=n:
        // extrapolate once n solves have followed a reset, default 5; 0
        // turns it off); zeroed, so a new mesh starts without history
        Level &Fv = c->levels.back();
        const char *e = getenv("MCL_EXTRAPOLATE");
        c->extrap_after = e ? atoi(e) : 5;
        c->extrap_on = c->extrap_after > 0;
        c->extrap_done = false;
        int rc = dev_alloc(c, Fv.owned, &c->ws, 3);
        rc |= dev_alloc(c, Fv.owned, &c->hist, std::max<size_t>((size_t)Fv.neq + Fv.nel, 1));
        if (rc) return MCL_E_DEVICE;
    }
    {
        // augmented Lagrangian of the mixed solve's levels (MCL_AL_GAMMA,
        // default 3; 0 switches it off): on meshes without hanging nodes
        // whose fixed velocities are zero (else F would need gamma G W G^T
        // v_bc as well).  A solve runs with it while the Uzawa counts are
        // high: the first solve on a mesh does, later ones if the previous
        // solve took at least MCL_AL_KEEP (100) Uzawa iterations with it or
        // MCL_AL_START (200) without.  (Measured at 256x256, Ra 1e6, contrast
        // 1e3: the cold first solve needs 155 Uzawa and 536 inner iterations
        // instead of 326 and 681, 25 % less time; once the flow settles and
        // the counts drop below ~100 the inner solves' extra iterations cost
        // more than the Uzawa iterations saved.)
        const int nm = c->have_mg ? nlev : 1;
        const char *e = getenv("MCL_AL_GAMMA");
        double gamma = e ? atof(e) : 3.0;
        const Level &Fv = c->levels.back();
        for (int l = 0; l < nlev && gamma > 0; l++)
            if (c->levels[l].dev.nh > 0) gamma = 0;
        for (size_t q = 0; q < Fv.v_bc.size() && gamma > 0; q++)
            if (Fv.v_bc[q] == 'V' && Fv.v_bc_val[q] != 0.0) gamma = 0;
        c->al_gamma = gamma > 0 ? gamma : 0.0;
        c->al_state = nullptr;
        c->al_tab = nullptr;
        c->al_eblocks = c->al_rblocks = 0;
        if (c->al_gamma > 0) {
            std::vector<AlLevel> tab(nm);
            int eb = 0, rb = 0;
            for (int m = 0; m < nm; m++) {
                Level &Lv = c->have_mg ? c->levels[m] : c->levels.back();
                int rc = dev_alloc(c, Lv.owned, &Lv.kval_al, (size_t)std::max(Lv.WK, 1) * std::max(Lv.neq, 1));
                rc |= dev_alloc(c, Lv.owned, &Lv.kdiag_al, std::max(Lv.neq, 1));
                rc |= dev_alloc(c, Lv.owned, &Lv.al_w, std::max(Lv.nel, 1));
                if (rc) return MCL_E_DEVICE;
                tab[m].L = Lv.dev;
                tab[m].val = Lv.kval_al;
                tab[m].diag = Lv.kdiag_al;
                tab[m].w = Lv.al_w;
                tab[m].eblk0 = eb;
                tab[m].rblk0 = rb;
                eb += (Lv.nel + MCL_BLOCK - 1) / MCL_BLOCK;
                rb += (Lv.neq + MCL_BLOCK - 1) / MCL_BLOCK;
            }
            const AlLevel *tp = nullptr;
            const double *sp = nullptr;
            int rc = dev_upload(c, c->levels[nlev - 1].owned, &sp, std::vector<double>{c->al_gamma});
            if (rc) return rc;
            c->al_state = (double *)sp;
            for (auto &t : tab) t.gamma = sp;
            rc = dev_upload(c, c->levels[nlev - 1].owned, &tp, tab);
            if (rc) return rc;
            c->al_tab = tp;
            c->al_eblocks = eb;
            c->al_rblocks = rb;
        }
    }
    {
        // FP32 copies and Gershgorin bounds of the mixed solve's levels (all
        // multigrid levels, or the finest alone without multigrid; bound m
        // in lam_bits[m]), from K_al with the augmented Lagrangian
        const int nm = c->have_mg ? nlev : 1;
        std::vector<PrepLevel> tab(nm);
        int blk = 0;
        for (int m = 0; m < nm; m++) {
            Level &Lv = c->have_mg ? c->levels[m] : c->levels.back();
            tab[m].n = Lv.neq;
            tab[m].W = Lv.WK;
            tab[m].val = c->al_gamma > 0 ? Lv.kval_al : Lv.dev.Kval;
            tab[m].diag = c->al_gamma > 0 ? Lv.kdiag_al : Lv.dev.Kdiag;
            tab[m].val32 = Lv.val32;
            tab[m].dinv = Lv.dinv32;
            tab[m].bound = c->lam_bits + m;
            tab[m].blk0 = blk;
            blk += (Lv.neq + MCL_BLOCK - 1) / MCL_BLOCK;
        }
        const PrepLevel *tp = nullptr;
        int rc = dev_upload(c, c->levels[nlev - 1].owned, &tp, tab);
        if (rc) return rc;
        c->prep_tab = tp;
        c->prep_n = nm;
        c->prep_blocks = blk;
    }
    // finest-level solver work space
    Level &F = c->levels[nlev - 1];
    for (void *p : {(void *)c->work, (void *)c->tau, (void *)c->hK, (void *)c->hC, (void *)c->hA, (void *)c->hF,
                    (void *)c->hd, (void *)c->hdd, (void *)c->hd0, (void *)c->hB}) {
        if (p) {
            hipFree(p);
            c->owned.erase(std::remove(c->owned.begin(), c->owned.end(), p), c->owned.end());
        }
    }
    size_t nv = std::max(F.neq, 1), np = F.nel, nt = std::max(F.neqT, 1);
    c->work_doubles = 6 * nv + 9 * np + 4 * std::max(nv, nt);
    int rc = dev_alloc(c, c->owned, &c->work, c->work_doubles);
    size_t wt = (size_t)std::max(F.WT, 1) * nt;
    rc |= dev_alloc(c, c->owned, &c->tau, 2 * np);
    rc |= dev_alloc(c, c->owned, &c->hK, wt);
    rc |= dev_alloc(c, c->owned, &c->hC, wt);
    rc |= dev_alloc(c, c->owned, &c->hA, wt);
    rc |= dev_alloc(c, c->owned, &c->hF, nt);
    rc |= dev_alloc(c, c->owned, &c->hd, nt);
    rc |= dev_alloc(c, c->owned, &c->hdd, nt);
    rc |= dev_alloc(c, c->owned, &c->hd0, nt);
    rc |= dev_alloc(c, c->owned, &c->hB, nt);
    if (rc) return MCL_E_DEVICE;
    c->first_heat = true;
    c->been_here_dp = false;
    return 0;
}

extern "C" int mcl_set_control(mcl_ctx *c, const mcl_control *ctl)
{
    if (!c || !ctl) return MCL_E_ARG;
    if (ctl->solver < 0 || ctl->solver > 3) return fail(c, MCL_E_ARG, "solver %d", ctl->solver);
    if (ctl->solver == MCL_SOLVER_MG && (ctl->mg_cycle < 0 || ctl->mg_finest_iteration < 0 ||
                                         ctl->mg_coarse_iteration < 0))
        return fail(c, MCL_E_ARG, "negative multigrid cycle or iteration count");
    if (ctl->precision != MCL_PREC_FP64 && ctl->precision != MCL_PREC_MIXED)
        return fail(c, MCL_E_ARG, "precision %d", ctl->precision);
    if (ctl->theta == 0.0) return fail(c, MCL_E_ARG, "theta = 0 divides by zero in the heat step");
    c->ctl = *ctl;
    c->have_ctl = true;
    return 0;
}

static Level *level_of(mcl_ctx *c, int lev)
{
    for (auto &L : c->levels)
        if (L.lev == lev) return &L;
    return nullptr;
}

static int field_ptr(mcl_ctx *c, Level *L, int field, double **ptr, size_t *count)
{
    bool finest = L == &c->levels.back();
    switch (field) {
    case MCL_FIELD_T: *ptr = L->T; *count = L->nno; return 0;
    case MCL_FIELD_VIS: *ptr = L->vis; *count = L->nno; return 0;
    case MCL_FIELD_BUOYANCY: *ptr = L->buoy; *count = L->nno; return 0;
    case MCL_FIELD_V: *ptr = L->V; *count = 2 * (size_t)L->nno; return 0;
    case MCL_FIELD_D: *ptr = L->d; *count = L->neq; return 0;
    case MCL_FIELD_P: *ptr = L->p; *count = L->nel; return 0;
    case MCL_FIELD_F: *ptr = L->dev.F; *count = L->neq; return 0;
    case MCL_FIELD_KDIAG: *ptr = L->dev.Kdiag; *count = L->neq; return 0;
    case MCL_FIELD_STRESS: *ptr = L->stress; *count = L->nno; return 0;
    }
    if (!finest) return fail(c, MCL_E_ARG, "heat fields exist on the finest level only");
    switch (field) {
    case MCL_FIELD_HEAT_D: *ptr = c->hd; *count = L->neqT; return 0;
    case MCL_FIELD_HEAT_DDOT: *ptr = c->hdd; *count = L->neqT; return 0;
    case MCL_FIELD_HEAT_F: *ptr = c->hF; *count = L->neqT; return 0;
    case MCL_FIELD_TAU: *ptr = c->tau; *count = 2 * (size_t)L->nel; return 0;
    }
    return fail(c, MCL_E_ARG, "unknown field %d", field);
}

extern "C" int mcl_set_field(mcl_ctx *c, int lev, int field, const double *src, size_t count)
{
    if (!c || !src) return MCL_E_ARG;
    Level *L = level_of(c, lev);
    if (!L) return fail(c, MCL_E_STATE, "level %d not uploaded", lev);
    double *p;
    size_t n;
    int rc = field_ptr(c, L, field, &p, &n);
    if (rc) return rc;
    if (count != n) return fail(c, MCL_E_ARG, "field %d: %zu values, expected %zu", field, count, n);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(p, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (field == MCL_FIELD_HEAT_D) c->first_heat = false;
    if (c->ws && L == &c->levels.back() && (field == MCL_FIELD_T || field == MCL_FIELD_D || field == MCL_FIELD_P))
        HIPCHK(c, hipMemsetAsync(c->ws, 0, sizeof(double), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mcl_get_field(mcl_ctx *c, int lev, int field, double *dst, size_t count)
{
    if (!c || !dst) return MCL_E_ARG;
    Level *L = level_of(c, lev);
    if (!L) return fail(c, MCL_E_STATE, "level %d not uploaded", lev);
    double *p;
    size_t n;
    int rc = field_ptr(c, L, field, &p, &n);
    if (rc) return rc;
    if (count != n) return fail(c, MCL_E_ARG, "field %d: %zu values, expected %zu", field, count, n);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(dst, p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mcl_get_matrix(mcl_ctx *c, int lev, int matrix, int width, int *ncol, int *col, double *val)
{
    if (!c || !ncol || !col || !val || width <= 0) return MCL_E_ARG;
    Level *L = level_of(c, lev);
    if (!L) return fail(c, MCL_E_STATE, "level %d not uploaded", lev);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int n, W;
    const std::vector<int> *pn, *pc;
    const double *dv;
    bool finest = L == &c->levels.back();
    if (matrix == MCL_MATRIX_K) { n = L->neq; W = L->WK; pn = &L->Kncol_h; pc = &L->Kcol_h; dv = L->dev.Kval; }
    else if (matrix == MCL_MATRIX_HEAT_K || matrix == MCL_MATRIX_HEAT_C) {
        if (!finest) return fail(c, MCL_E_ARG, "heat matrices exist on the finest level only");
        n = L->neqT; W = L->WT; pn = &L->Tncol_h; pc = &L->Tcol_h;
        dv = matrix == MCL_MATRIX_HEAT_K ? c->hK : c->hC;
    } else if (matrix == MCL_MATRIX_G) {
        // G is applied from element tables; rebuild its rows (columns = elements ascending)
        std::vector<double> gt(MCL_MAX_LEVELS * 8);
        HIPCHK(c, hipMemcpy(gt.data(), L->dev.gtab, gt.size() * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<int> cnt(L->nno), el(4 * (size_t)L->nno), aa(4 * (size_t)L->nno);
        HIPCHK(c, hipMemcpy(cnt.data(), L->dev.p2n_cnt, cnt.size() * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(el.data(), L->dev.p2n_el, el.size() * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(aa.data(), L->dev.p2n_a, aa.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int i = 0; i < L->nno; i++)
            for (int d = 0; d < 2; d++) {
                int P = L->id[d * L->nno + i];
                if (P < 0) continue;
                if (cnt[i] > width) return fail(c, MCL_E_ARG, "width too small");
                ncol[P] = cnt[i];
                for (int k = 0; k < cnt[i]; k++) {
                    int e = el[k * (size_t)L->nno + i], a = aa[k * (size_t)L->nno + i];
                    col[(size_t)P * width + k] = e;
                    val[(size_t)P * width + k] = gt[(L->Mcode[e] & 15) * 8 + 2 * a + d];
                }
            }
        return 0;
    } else return fail(c, MCL_E_ARG, "unknown matrix %d", matrix);
    if (W > width) return fail(c, MCL_E_ARG, "width %d < row length %d", width, W);
    std::vector<double> v((size_t)std::max(W, 1) * std::max(n, 1));
    HIPCHK(c, hipMemcpy(v.data(), dv, v.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        ncol[i] = (*pn)[i];
        for (int s = 0; s < (*pn)[i]; s++) {
            col[(size_t)i * width + s] = (*pc)[(size_t)s * n + i];
            val[(size_t)i * width + s] = v[(size_t)s * n + i];
        }
    }
    return 0;
}

extern "C" int mcl_set_model_time(mcl_ctx *c, double t)
{
    if (!c) return MCL_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->model_time, &t, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mcl_get_model_time(mcl_ctx *c, double *t)
{
    if (!c || !t) return MCL_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(t, c->model_time, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mcl_synchronize(mcl_ctx *c)
{
    if (!c) return MCL_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mcl_set_launch_cap(mcl_ctx *c, int max_workgroups)
{
    if (!c || max_workgroups < 0) return MCL_E_ARG;
    c->block_cap = max_workgroups;
    return 0;
}

static int ready(mcl_ctx *c)
{
    if (!c) return MCL_E_ARG;
    if (c->nlev == 0) return fail(c, MCL_E_STATE, "no mesh uploaded");
    if (!c->have_ctl) return fail(c, MCL_E_STATE, "no control set");
    return 0;
}

static unsigned grid_for(mcl_ctx *c, long long n)
{
    long long b = (n + MCL_BLOCK - 1) / MCL_BLOCK;
    b = std::max<long long>(1, std::min<long long>(b, 4LL * c->num_cu));
    return (unsigned)b;
}

extern "C" int mcl_viscosity(mcl_ctx *c)
{
    int rc = ready(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    for (auto &L : c->levels) {
        hipLaunchKernelGGL(k_viscosity, dim3(grid_for(c, L.nno)), dim3(MCL_BLOCK), 0, c->stream, L.nno, L.T,
                           L.dev.X, c->ctl.A, c->ctl.refT, c->ctl.lm_jump, L.vis);
        HIPCHK(c, hipGetLastError());
    }
    return 0;
}

static int inject_temperature(mcl_ctx *c)
{
    Level &F = c->levels.back();
    for (int k = 0; k + 1 < c->nlev; k++) {
        Level &Lv = c->levels[k];
        hipLaunchKernelGGL(k_inject, dim3(grid_for(c, Lv.nno)), dim3(MCL_BLOCK), 0, c->stream, Lv.nno,
                           (const int *)Lv.c2f, (const double *)F.T, Lv.T);
        HIPCHK(c, hipGetLastError());
    }
    return 0;
}

extern "C" int mcl_buoyancy_force(mcl_ctx *c)
{
    int rc = ready(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    // The mixed path rediscretises the coarse levels from the current
    // temperature (the reference keeps the initial one there).
    if (c->ctl.precision == MCL_PREC_MIXED && c->nlev > 1) {
        rc = inject_temperature(c);
        if (rc) return rc;
    }
    for (auto &L : c->levels) {
        // the mixed path's coarse levels carry only K (their F is not used)
        if (c->ctl.precision == MCL_PREC_MIXED && &L != &c->levels.back()) continue;
        if (L.bnent > 0)
            hipLaunchKernelGGL(k_buoy_terms, dim3(grid_for(c, L.bnent)), dim3(MCL_BLOCK), 0, c->stream, L.bnent, L.nno,
                               L.brow_e, L.brow_xi, L.brow_eta, L.dev.ienl, L.nel, L.T, c->ctl.Ra, L.bterms);
        hipLaunchKernelGGL(k_buoy_rows, dim3((L.gnoz + MCL_BLOCK - 1) / MCL_BLOCK), dim3(MCL_BLOCK), 0, c->stream,
                           L.gnoz, L.gnox, L.brow_ptr, (const double *)L.bterms, L.avg);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(k_buoy_apply, dim3(grid_for(c, L.nno)), dim3(MCL_BLOCK), 0, c->stream, L.nno, L.gnoz,
                           L.dev.node, L.T, c->ctl.Ra, L.avg, L.buoy);
        HIPCHK(c, hipGetLastError());
    }
    return 0;
}

static int mixed_prepare(mcl_ctx *c, bool by_kernel);

extern "C" int mcl_velocity_pressure_KGF_matrix(mcl_ctx *c)
{
    int rc = ready(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->asm_blocks > 0) {
        hipLaunchKernelGGL(k_assemble_K, dim3(c->asm_blocks), dim3(MCL_BLOCK), 0, c->stream,
                           (const AsmLevel *)c->asm_tab, c->nlev);
        HIPCHK(c, hipGetLastError());
    }
    c->mixed_ready = false;
    if (c->ctl.precision == MCL_PREC_MIXED) return mixed_prepare(c, false);
    return 0;
}

// Workgroups of a persistent solve: one per rows_per_block rows, at most what
// fits co-resident; the environment variable `knob` (MCL_FP64_BLOCKS for the
// FP64 Uzawa and heat kernels, MCL_MIXED_BLOCKS for the mixed solve), if set,
// replaces the count (still capped at co-residency).  The FP64 solvers take two rows per thread
// (their phases are grid barriers, whose cost grows with the workgroup
// count); the mixed solve takes at most 128 rows per workgroup (its V-cycle
// waits are mostly point-to-point, and more workgroups keep more loads in
// flight: at 256x256 and above every CU has one workgroup).
#define MCL_FP64_ROWS_PER_THREAD 2
static unsigned coop_blocks(mcl_ctx *c, long long n, int rows_per_block, const char *knob)
{
    const long long per = rows_per_block;
    long long b = (n + per - 1) / per;
    const char *e = getenv(knob);
    if (e && atoi(e) > 0) b = atoi(e);
    long long cap = c->max_coop_blocks;
    if (c->block_cap > 0) cap = std::min<long long>(cap, c->block_cap);
    return (unsigned)std::max<long long>(1, std::min<long long>(b, cap));
}

// Device buffers a Stokes solve works in besides the mesh and the fields:
// grid-barrier words, reduction partials, the kernel's statistics, the
// residual norms and the Uzawa / CG work vectors (6 n_v + 9 n_p doubles).
// The context's own (allocated at set-up), or carved from the static scratch
// array by the allocation-free solve.
struct SolveWork {
    GridSync gs;
    RedBuf rb;
    DevStats *stats;
    double *resid;
    double *work;
};

static SolveWork ctx_work(mcl_ctx *c) { return SolveWork{c->gs, c->rb, c->stats, c->resid, c->work}; }

// Zero `bytes` (a multiple of 4) on the context's stream: hipMemsetAsync for
// the allocating entry points, the zero-fill kernel for the allocation-free one.
static int zero_async(mcl_ctx *c, void *p, size_t bytes, bool by_kernel)
{
    if (!by_kernel) {
        HIPCHK(c, hipMemsetAsync(p, 0, bytes, c->stream));
        return 0;
    }
    const long long n = (long long)(bytes / 4);
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<long long>(std::max(1LL, (n + MCL_BLOCK - 1) / MCL_BLOCK),
                                                                         4LL * c->num_cu)),
                       dim3(MCL_BLOCK), 0, c->stream, (unsigned *)p, n);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// Launch a persistent kernel: its workgroups wait on each other, so all of
// them must be resident at once, and the grid never exceeds what the occupancy
// query says fits (coop_blocks).  A cooperative launch guarantees residency;
// the runtime refuses it (hipErrorIllegalState) while another process holds
// the device's cooperative-launch resources (a second process on the same
// GPU), and then the same grid is launched plainly: with the CUs free every
// workgroup is dispatched at once, and a workgroup kept waiting longer than the
// barriers' spin limit ends the solve with MCL_E_DEVICE instead of a hang.
static hipError_t launch_persistent(mcl_ctx *c, const void *fn, unsigned nb, void **args, bool plain = false)
{
    // a batch's captured graphs launch plainly: their grids are capped so that
    // the persistent kernels of all its lanes fit on the device together
    if (plain) return hipLaunchKernel(fn, dim3(nb), dim3(MCL_BLOCK), args, 0, c->stream);
    hipError_t e = hipLaunchCooperativeKernel(fn, dim3(nb), dim3(MCL_BLOCK), args, 0, c->stream);
    if (e == hipErrorIllegalState) {
        (void)hipGetLastError();
        e = hipLaunchKernel(fn, dim3(nb), dim3(MCL_BLOCK), args, 0, c->stream);
    }
    return e;
}

static int reset_sync_if_aborted(mcl_ctx *c, const DevStats &st, const SolveWork &W, bool by_kernel)
{
    if (st.aborted) {
        int rc = zero_async(c, W.gs.gen, MCL_SYNC_WORDS * sizeof(unsigned), by_kernel);
        if (rc) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return fail(c, MCL_E_DEVICE, "grid barrier timed out");
    }
    return 0;
}

// Uzawa kernel parameters for the finest level's (d, p) and the work space.
static UzawaParams uzawa_params(mcl_ctx *c, double accuracy, const SolveWork &W)
{
    Level &L = c->levels.back();
    const mcl_control &k = c->ctl;
    UzawaParams P{};
    P.L = L.dev;
    P.gs = W.gs;
    P.rb = W.rb;
    P.stats = W.stats;
    P.d = L.d;
    P.p = L.p;
    size_t nv = std::max(L.neq, 1), np = L.nel;
    double *w = W.work;
    P.u1 = w; w += nv;
    P.Ahv = w; w += nv;
    P.cg_r = w; w += nv;
    P.cg_s = w; w += nv;
    P.cg_d = w; w += nv;
    P.cg_q = w; w += nv;
    P.Ahp = w; w += np;
    P.BAB = w; w += np;
    P.r0 = w; w += np;
    P.r1 = w; w += np;
    P.r2 = w; w += np;
    P.z0 = w; w += np;
    P.z1 = w; w += np;
    P.s1 = w; w += np;
    P.s2 = w; w += np;
    P.accuracy = accuracy;
    P.citcom_accu = k.citcom_accu;
    P.max_citcom_cycle = k.max_citcom_cycle;
    P.max_cg_cycle = k.max_cg_cycle;
    P.cg_precondition = k.cg_precondition;
    P.check_v_only = k.check_v_only;
    P.schur = k.schur;
    P.exact_dots = k.exact_dots;
    P.gs.nblocks = coop_blocks(c, std::max(L.neq, L.nel), MCL_FP64_ROWS_PER_THREAD * MCL_BLOCK, "MCL_FP64_BLOCKS");
    P.solver = k.solver;
    P.mg_cycle = k.mg_cycle;
    P.mg_finest_iteration = k.mg_finest_iteration;
    P.mg_coarse_iteration = k.mg_coarse_iteration;
    P.xt = c->xt;
    P.xt2 = c->xt2;
    P.gsl = c->gs_tab;
    P.ngsl = c->nlev;
    // Gauss-Seidel sweeps and the multigrid V-cycle are sequential in the
    // row order: the whole solve runs in one workgroup
    if (k.solver == MCL_SOLVER_GS || k.solver == MCL_SOLVER_MG) P.gs.nblocks = 1;
    return P;
}

// FP32 copies of K and smoother bounds per multigrid level and the dense
// inverse of the coarsest K: once per assembly (mixed precision).
static int mixed_prepare(mcl_ctx *c, bool by_kernel)
{
    int rc = zero_async(c, c->lam_bits, MCL_MAX_LEVELS * sizeof(unsigned), by_kernel);
    if (rc) return rc;
    if (c->al_gamma > 0) {
        hipLaunchKernelGGL(k_al_weights, dim3(c->al_eblocks), dim3(MCL_BLOCK), 0, c->stream, c->al_tab, c->prep_n);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(k_al_rows, dim3(c->al_rblocks), dim3(MCL_BLOCK), 0, c->stream, c->al_tab, c->prep_n);
        HIPCHK(c, hipGetLastError());
    }
    if (c->prep_blocks > 0) {
        hipLaunchKernelGGL(k_prepare_fp32, dim3(c->prep_blocks), dim3(MCL_BLOCK), 0, c->stream,
                           (const PrepLevel *)c->prep_tab, c->prep_n);
        HIPCHK(c, hipGetLastError());
    }
    if (c->have_mg) {
        Level &C0 = c->levels[0];
        hipLaunchKernelGGL(k_coarse_inverse, dim3(1), dim3(MCL_BLOCK), 0, c->stream, C0.neq, C0.WK, C0.dev.Kcol,
                           C0.dev.Kncol, (const double *)(c->al_gamma > 0 ? C0.kval_al : C0.dev.Kval), c->cinv);
        HIPCHK(c, hipGetLastError());
    }
    c->mixed_ready = true;
    return 0;
}

// Halo lists of the fused V-cycle phases for nb workgroups (kernel:
// fused_up / fused_down, deep_up / deep_down), their extra data dependencies
// reported through add(a, b), and the level descriptors re-uploaded.  On grid
// level l a workgroup owns rows O = [b0, e0) (b0 = g R, R = ceil(n / nb)).
// Local row lists keep an inner set as a prefix of an outer one, so a local
// index into the inner set is valid in the outer set:
//   up:   W = [O | H1], H1 = K columns of O outside O;
//         V = [W | H2], H2 = K columns of W outside W;
//   down: F = the fine rows that the workgroup's coarse rows restrict from;
//         T = [O | Th], Th = K columns of F outside O;
//         U = [T | Uh], Uh = K columns of T outside T.
// (O is indexed b0 .. b0 + R - 1 even where the last workgroup owns fewer.)
// A phase stays in its unfused form if some workgroup's lists would not fit
// the LDS buffer (MCL_HALO_LDS floats); MCL_MG_FUSE=0 switches fusion off,
// MCL_MG_FUSE=1 keeps the one-sweep fusion only.
template <typename T>
static int halo_upload(mcl_ctx *c, const T **out, const std::vector<T> &h)
{
    if (dev_upload(c, c->owned, out, h.empty() ? std::vector<T>(1) : h)) return 1;
    c->halo_bufs.push_back(c->owned.back());
    return 0;
}

template <typename Add>
static int build_halos(mcl_ctx *c, int nb, Add &add)
{
    const int top = c->nlev - 1, lb = c->mg_lb;
    if ((int)c->mg_host.size() != c->nlev) return 0;
    // descriptors without lists first (no kernel may still read the old ones)
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (MgLevelDev &M : c->mg_host) {
        M.fuse_up = M.fuse_down = M.deep_up = M.deep_down = 0;
        M.hu_ptr = M.hu_row = M.hd_ptr = M.hd_row = M.h2_ptr = M.h2_row = M.th_ptr = M.th_row = nullptr;
        M.uh_ptr = M.uh_row = nullptr;
        M.kl = M.hm = M.mt = M.mf = nullptr;
        M.rl = nullptr;
    }
    HIPCHK(c, hipMemcpy(c->mg_dev, c->mg_host.data(), c->mg_host.size() * sizeof(MgLevelDev), hipMemcpyHostToDevice));
    for (void *q : c->halo_bufs) {
        (void)hipFree(q);
        c->owned.erase(std::remove(c->owned.begin(), c->owned.end(), q), c->owned.end());
    }
    c->halo_bufs.clear();
    const char *fe = getenv("MCL_MG_FUSE");
    const int fuse_level = fe ? atoi(fe) : 2;
    if (fuse_level <= 0) return 0;
    const long long cap = MCL_HALO_LDS;
    std::vector<MgLevelDev> mg = c->mg_host;
    auto sorted_unique = [](std::vector<int> &v) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    };
    auto pos = [](const std::vector<int> &v, int j) {
        auto it = std::lower_bound(v.begin(), v.end(), j);
        return (it != v.end() && *it == j) ? (int)(it - v.begin()) : -1;
    };
    for (int l = lb + 1; l <= top; l++) {
        MgLevelDev &M = mg[l];
        Level &A = c->levels[l];
        Level &C = c->levels[l - 1];
        const int n = A.neq, WK = A.WK, nc = C.neq, rw = C.rw;
        if (n == 0) continue;
        const int R = (n + nb - 1) / nb, Rc = (nc + nb - 1) / nb;
        const bool coarse_grid = l - 1 > lb;
        auto kcol = [&](int i, int s) { return A.Kcol_h[(size_t)s * n + i]; };
        auto parents = [&](int g, int j) {
            if (!coarse_grid) return;
            for (int a = 0; a < A.interp_n[j]; a++) {
                const int cc = A.interp_eq[(size_t)a * n + j];
                if (cc >= 0) add(g, cc / Rc);
            }
        };
        std::vector<int> h1_ptr(nb + 1, 0), h1_row, h2_ptr(nb + 1, 0), h2_row;
        std::vector<unsigned short> kl((size_t)WK * n, 0), hm;
        std::vector<int> f_ptr(nb + 1, 0), f_row, th_ptr(nb + 1, 0), th_row, uh_ptr(nb + 1, 0), uh_row;
        std::vector<unsigned short> mt, mf;
        std::vector<short> rl((size_t)std::max(rw, 1) * std::max(nc, 1), -1);
        bool up1 = true, up2 = fuse_level >= 2, dn1 = coarse_grid && (int)C.rcol_h.size() == rw * nc && nc > 0;
        bool dn2 = dn1 && fuse_level >= 2;
        std::vector<int> H1, H2, F, Th, Uh;
        for (int g = 0; g < nb; g++) {
            const int b0 = g * R, e0 = std::min(b0 + R, n);
            auto own = [&](int j) { return j >= b0 && j < e0; };
            // up: H1, H2
            H1.clear();
            for (int s = 0; s < WK; s++)
                for (int i = b0; i < e0; i++)
                    if (!own(kcol(i, s))) H1.push_back(kcol(i, s));
            sorted_unique(H1);
            H2.clear();
            for (int s = 0; s < WK; s++)
                for (int j : H1) {
                    const int q = kcol(j, s);
                    if (!own(q) && pos(H1, q) < 0) H2.push_back(q);
                }
            sorted_unique(H2);
            const long long nW = R + (long long)H1.size(), nV = nW + (long long)H2.size();
            if (nW > cap) up1 = false;
            if (nV + nW > cap) up2 = false;
            // local index in V (-1: missing, which the construction excludes;
            // checked so that no LDS index can leave its list)
            auto locV = [&](int j) {
                if (own(j)) return j - b0;
                const int p1 = pos(H1, j);
                if (p1 >= 0) return R + p1;
                const int p2 = pos(H2, j);
                return p2 >= 0 ? R + (int)H1.size() + p2 : -1;
            };
            for (int s = 0; s < WK; s++)
                for (int i = b0; i < e0; i++) {
                    const int q = locV(kcol(i, s));
                    if (q < 0 || q >= nW) up1 = up2 = false;
                    kl[(size_t)s * n + i] = (unsigned short)std::max(q, 0);
                }
            // hm: H1 rows -> V, slot-major within the workgroup's block
            for (int s = 0; s < WK; s++)
                for (int j : H1) {
                    const int q = locV(kcol(j, s));
                    if (q < 0 || q >= nV) up2 = false;
                    hm.push_back((unsigned short)std::max(q, 0));
                }
            h1_row.insert(h1_row.end(), H1.begin(), H1.end());
            h1_ptr[g + 1] = (int)h1_row.size();
            h2_row.insert(h2_row.end(), H2.begin(), H2.end());
            h2_ptr[g + 1] = (int)h2_row.size();
            // up dependencies: the coarse parents of every row x' is formed on
            // (the level's own vectors were written before the grid barrier
            // that precedes the LDS levels)
            for (int i = b0; i < e0; i++) parents(g, i);
            for (int j : H1) parents(g, j);
            if (up2)
                for (int j : H2) parents(g, j);
            if (!coarse_grid) {
                // the down phase above the LDS levels uses T = W, U = V and
                // reads x_1 on V, written in the previous phase
                for (int j : H1) add(g, j / R);
                for (int j : H2) add(g, j / R);
                continue;
            }
            // down: F, Th, Uh
            const int c0 = g * Rc, c1 = std::min(c0 + Rc, nc);
            F.clear();
            for (int k = 0; k < rw; k++)
                for (int cc = c0; cc < c1; cc++) {
                    const int u = C.rcol_h[(size_t)k * nc + cc];
                    if (u >= 0) F.push_back(u);
                }
            sorted_unique(F);
            if ((long long)F.size() > cap) dn1 = false;
            for (int k = 0; k < rw; k++)
                for (int cc = c0; cc < c1; cc++) {
                    const int u = C.rcol_h[(size_t)k * nc + cc];
                    if (u >= 0) rl[(size_t)k * nc + cc] = (short)pos(F, u);
                }
            Th.clear();
            for (int s = 0; s < WK; s++)
                for (int u : F)
                    if (!own(kcol(u, s))) Th.push_back(kcol(u, s));
            sorted_unique(Th);
            Uh.clear();
            for (int s = 0; s < WK; s++) {
                for (int i = b0; i < e0; i++) {
                    const int q = kcol(i, s);
                    if (!own(q) && pos(Th, q) < 0) Uh.push_back(q);
                }
                for (int j : Th) {
                    const int q = kcol(j, s);
                    if (!own(q) && pos(Th, q) < 0) Uh.push_back(q);
                }
            }
            sorted_unique(Uh);
            const long long nT = R + (long long)Th.size(), nU = nT + (long long)Uh.size();
            if (nU + nT + (long long)F.size() > cap) dn2 = false;
            auto locU = [&](int j) {
                if (own(j)) return j - b0;
                const int pt = pos(Th, j);
                if (pt >= 0) return R + pt;
                const int pu = pos(Uh, j);
                return pu >= 0 ? R + (int)Th.size() + pu : -1;
            };
            // mt: T rows (O padded to R, then Th) -> U, slot-major in the block
            for (int s = 0; s < WK; s++) {
                for (int t = 0; t < R; t++) {
                    const int i = b0 + t;
                    const int q = i < e0 ? locU(kcol(i, s)) : 0;
                    if (q < 0 || q >= nU) dn2 = false;
                    mt.push_back((unsigned short)std::max(q, 0));
                }
                for (int j : Th) {
                    const int q = locU(kcol(j, s));
                    if (q < 0 || q >= nU) dn2 = false;
                    mt.push_back((unsigned short)std::max(q, 0));
                }
            }
            // mf: F rows -> T (K columns of F lie in T by construction)
            for (int s = 0; s < WK; s++)
                for (int u : F) {
                    const int q = locU(kcol(u, s));
                    if (q < 0 || q >= nT) dn2 = false;
                    mf.push_back((unsigned short)std::max(q, 0));
                }
            f_row.insert(f_row.end(), F.begin(), F.end());
            f_ptr[g + 1] = (int)f_row.size();
            th_row.insert(th_row.end(), Th.begin(), Th.end());
            th_ptr[g + 1] = (int)th_row.size();
            uh_row.insert(uh_row.end(), Uh.begin(), Uh.end());
            uh_ptr[g + 1] = (int)uh_row.size();
            // down dependencies: the residual reads b and x on F's K columns
            // (one-sweep form); the two-sweep form reads x_1 on U
            for (int u : F) {
                add(g, u / R);
                for (int s2 = 0; s2 < WK; s2++) add(g, kcol(u, s2) / R);
            }
            if (dn2)
                for (int j : Uh) add(g, j / R);
        }
        if (up1) {
            if (halo_upload(c, &M.hu_ptr, h1_ptr) || halo_upload(c, &M.hu_row, h1_row) || halo_upload(c, &M.kl, kl))
                return MCL_E_DEVICE;
            M.fuse_up = 1;
            if (up2) {
                if (halo_upload(c, &M.h2_ptr, h2_ptr) || halo_upload(c, &M.h2_row, h2_row) ||
                    halo_upload(c, &M.hm, hm))
                    return MCL_E_DEVICE;
                M.deep_up = 1;
                if (!coarse_grid) M.deep_down = 1;   // T = W, U = V
            }
        }
        if (dn1) {
            if (halo_upload(c, &M.hd_ptr, f_ptr) || halo_upload(c, &M.hd_row, f_row) ||
                halo_upload(c, &mg[l - 1].rl, rl))
                return MCL_E_DEVICE;
            M.fuse_down = 1;
            if (dn2) {
                if (halo_upload(c, &M.th_ptr, th_ptr) || halo_upload(c, &M.th_row, th_row) ||
                    halo_upload(c, &M.uh_ptr, uh_ptr) || halo_upload(c, &M.uh_row, uh_row) ||
                    halo_upload(c, &M.mt, mt) || halo_upload(c, &M.mf, mf))
                    return MCL_E_DEVICE;
                M.deep_down = 1;
            }
        }
    }
    HIPCHK(c, hipMemcpy(c->mg_dev, mg.data(), mg.size() * sizeof(MgLevelDev), hipMemcpyHostToDevice));
    c->mg_host = mg;
    return 0;
}

// Neighbour lists for the point-to-point waits of the mixed V-cycle.  With
// the contiguous ownership owner(l, i) = i / ceil(n_l / nb) of the kernel, a
// workgroup's neighbours are the workgroups that own a column of one of its
// rows on any grid level (all ELL slots), or a row coupled to one of its rows
// by the prolongation between two grid levels (restriction is its transpose);
// the relation is made symmetric, so every reader of a workgroup's rows is
// its neighbour.  Lists longer than MCL_BLOCK turn the waits into grid
// barriers.
static int build_neighbours(mcl_ctx *c, int nb)
{
    const int top = c->nlev - 1, lb = c->mg_lb;
    const int words = (nb + 63) / 64;
    std::vector<uint64_t> bits((size_t)nb * words, 0);
    auto add = [&](int a, int b) {
        if (a == b) return;
        bits[(size_t)a * words + (b >> 6)] |= 1ull << (b & 63);
        bits[(size_t)b * words + (a >> 6)] |= 1ull << (a & 63);
    };
    for (int l = lb + 1; l <= top; l++) {
        Level &A = c->levels[l];
        const int n = A.neq;
        if (n == 0) continue;
        const int R = (n + nb - 1) / nb;
        for (int s = 0; s < A.WK; s++)
            for (int i = 0; i < n; i++) add(i / R, A.Kcol_h[(size_t)s * n + i] / R);
        if (l - 1 > lb) {
            Level &C = c->levels[l - 1];
            const int Rc = (C.neq + nb - 1) / nb;
            for (int a = 0; a < 4; a++)
                for (int u = 0; u < n; u++) {
                    if (A.interp_n[u] <= a) continue;
                    int cc = A.interp_eq[(size_t)a * n + u];
                    if (cc >= 0) add(u / R, cc / Rc);
                }
        }
    }
    c->nbr_nb = -1;
    c->nbr_ok = 0;
    int rc = build_halos(c, nb, add);
    if (rc) return rc;
    std::vector<int> ptr(nb + 1, 0), idx;
    int maxdeg = 0;
    for (int b = 0; b < nb; b++) {
        for (int w = 0; w < words; w++) {
            uint64_t m = bits[(size_t)b * words + w];
            while (m) {
                int t = __builtin_ctzll(m);
                idx.push_back(w * 64 + t);
                m &= m - 1;
            }
        }
        ptr[b + 1] = (int)idx.size();
        maxdeg = std::max(maxdeg, ptr[b + 1] - ptr[b]);
    }
    // the old lists go first: until the new ones are in place no solve may
    // use either (nbr_nb = -1 matches no workgroup count)
    c->nbr_nb = -1;
    c->nbr_ok = 0;
    for (void *q : {(void *)c->nbr_ptr, (void *)c->nbr_idx}) {
        if (q) {
            (void)hipFree(q);
            c->owned.erase(std::remove(c->owned.begin(), c->owned.end(), q), c->owned.end());
        }
    }
    c->nbr_ptr = c->nbr_idx = nullptr;
    const int *cp;
    if (dev_upload(c, c->owned, &cp, ptr)) return MCL_E_DEVICE;
    c->nbr_ptr = (int *)cp;
    if (idx.empty()) idx.push_back(0);
    if (dev_upload(c, c->owned, &cp, idx)) return MCL_E_DEVICE;
    c->nbr_idx = (int *)cp;
    c->nbr_nb = nb;
    const char *e = getenv("MCL_NBR_SYNC");
    c->nbr_ok = (maxdeg <= MCL_BLOCK && !(e && atoi(e) == 0)) ? 1 : 0;
    return 0;
}

static int mixed_params(mcl_ctx *c, const UzawaParams &P, MixedParams &M)
{
    const mcl_control &k = c->ctl;
    M = MixedParams{};
    M.U = P;
    M.U.gs.nblocks = coop_blocks(c, std::max(P.L.neq, P.L.nel), MCL_MIXED_ROWS_PER_BLOCK, "MCL_MIXED_BLOCKS");
    M.mg = c->mg_dev;
    M.lam = (const float *)c->lam_bits;
    M.nmg = c->have_mg ? c->nlev : 1;
    M.cinv = c->cinv;
    M.nc = c->levels[0].neq;
    M.rr = c->rr;
    M.inner_rtol = k.inner_rtol > 0 ? k.inner_rtol : 3e-7;
    {
        // relaxed inner accuracy in the Uzawa loop, then one correction solve
        // (measured at 256x256 over steps 0-19 against converged FP64: relax
        // 3e-8 gave 2.3x less solve time than none with Vrms within 8e-8;
        // 1e-7 another 10 % with Vrms within 1.5e-7; 3e-7 8.6e-7; 1e-6
        // 3.2e-6, past the 1e-6 bar.  With the augmented velocity block in
        // the high-count solves: 1e-7 5.0e-8, 2e-7 2.0e-7 and 5 % less
        // time, 3e-7 3.0e-7 and 7 %); MCL_INNER_RELAX=0 switches it off
        const char *e = getenv("MCL_INNER_RELAX"), *m = getenv("MCL_INNER_RELAX_MAX");
        M.relax = e ? atof(e) : 2e-7;
        M.relax_max = m && atof(m) > 0 ? atof(m) : 1e-1;
    }
    M.max_inner = 1000;
    M.probe = 0;
    M.use_nbr = 0;
    const int nb = (int)M.U.gs.nblocks;
    if (c->have_mg && nb != c->nbr_nb) {
        // halo and neighbour lists for this workgroup count; on failure both
        // are left marked unusable and the solve is not launched
        int rc = build_neighbours(c, nb);
        if (rc) return rc;
    }
    if (c->have_mg && c->nbr_nb == nb) {
        M.nbr_ptr = c->nbr_ptr;
        M.nbr_idx = c->nbr_idx;
        M.use_nbr = c->nbr_ok;
    }
    const char *sw = getenv("MCL_MG_SWEEPS");
    M.pre = sw && atoi(sw) > 0 ? atoi(sw) : 2;
    M.omega = 4.0f / 3.0f;
    if (c->extrap_on && k.warm_start) {
        M.ws = c->ws;
        M.model_time = c->model_time;
    }
    if (c->al_gamma > 0) {
        Level &F = c->levels.back();
        M.U.L.Kval = F.kval_al;
        M.U.L.Kdiag = F.kdiag_al;
        M.al_state = c->al_state;
        M.al_w = F.al_w;
        M.al_gamma = c->al_gamma;
        const char *ek = getenv("MCL_AL_KEEP"), *es = getenv("MCL_AL_START");
        M.al_keep = ek ? atoi(ek) : 100;
        M.al_start = es ? atoi(es) : 200;
    }
    return 0;
}

// Two HIP events around a launch, destroyed on every return path.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    int create(mcl_ctx *c)
    {
        HIPCHK(c, hipEventCreate(&a));
        HIPCHK(c, hipEventCreate(&b));
        return 0;
    }
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// How a solve or heat step is enqueued: NORMAL (the allocating entry points:
// hipMemsetAsync, cooperative launch, HIP events), NOALLOC (nothing
// allocated: buffers cleared by the zero-fill kernel, no events) or GRAPH
// (captured into a batch's graph: zero-fill kernel, plain launch, no events,
// no host synchronisation).
enum EnqMode { ENQ_NORMAL, ENQ_NOALLOC, ENQ_GRAPH };

// The mixed solve's warm start from the last two solutions (k_extrapolate),
// once per step: before the learned guess's gate if there is one, else
// before the solve.
static int extrapolate(mcl_ctx *c)
{
    if (c->ctl.precision != MCL_PREC_MIXED || !c->ctl.warm_start || !c->extrap_on || c->extrap_done) return 0;
    Level &L = c->levels.back();
    const long long n = (long long)L.neq + L.nel;
    hipLaunchKernelGGL(k_extrapolate, dim3(grid_for(c, n)), dim3(MCL_BLOCK), 0, c->stream, L.neq, L.nel, L.d, L.p,
                       c->hist, (const double *)c->ws, (const double *)c->model_time, c->extrap_after, 2.0);
    HIPCHK(c, hipGetLastError());
    c->extrap_done = true;
    return 0;
}

// The device work of solve_velocity_pressure (Solver.c:3-12) on work space W,
// up to the residual norms in W.resid; `accuracy` is the inner tolerance.
static int solve_enqueue(mcl_ctx *c, const SolveWork &W, EnqMode mode, EventPair *ev, double accuracy)
{
    Level &L = c->levels.back();
    const mcl_control &k = c->ctl;
    const bool by_kernel = mode != ENQ_NORMAL, plain = mode == ENQ_GRAPH;
    int rc;
    // initial_d_p (Solver.c:251-278): d = 0 on every level; p = 0 unless a
    // previous call exists, AMR is off and zero_P is off.
    if (!k.warm_start) {
        for (auto &Lv : c->levels) {
            rc = zero_async(c, Lv.d, std::max(Lv.neq, 1) * sizeof(double), by_kernel);
            if (rc) return rc;
            if (!(c->been_here_dp && k.AMR == 0 && k.zero_P == 0)) {
                rc = zero_async(c, Lv.p, Lv.nel * sizeof(double), by_kernel);
                if (rc) return rc;
            }
        }
    }
    c->been_here_dp = true;
    if (mode == ENQ_NOALLOC) {
        // barrier words (the kernel starts from generation 0)
        rc = zero_async(c, W.gs.gen, MCL_SYNC_WORDS * sizeof(unsigned), true);
        if (rc) return rc;
    }
    rc = zero_async(c, W.stats, sizeof(DevStats), by_kernel);
    if (rc) return rc;
    UzawaParams P = uzawa_params(c, accuracy, W);
    if (k.precision == MCL_PREC_MIXED) {
        if (!c->mixed_ready || mode == ENQ_GRAPH) {
            rc = mixed_prepare(c, by_kernel);
            if (rc) return rc;
        }
        MixedParams M;
        rc = mixed_params(c, P, M);
        if (rc) return rc;
        rc = extrapolate(c);
        if (rc) return rc;
        c->extrap_done = false;
        void *margs[] = {&M};
        if (ev) HIPCHK(c, hipEventRecord(ev->a, c->stream));
        HIPCHK(c, launch_persistent(c, (const void *)k_uzawa_mixed, M.U.gs.nblocks, margs, plain));
    } else {
        void *args[] = {&P};
        if (ev) HIPCHK(c, hipEventRecord(ev->a, c->stream));
        HIPCHK(c, launch_persistent(c, (const void *)k_uzawa_fp64, P.gs.nblocks, args, plain));
    }
    if (ev) HIPCHK(c, hipEventRecord(ev->b, c->stream));
    // nodal velocity and stress (Solver.c:8-9)
    hipLaunchKernelGGL(k_v_to_nodes, dim3(grid_for(c, L.nno)), dim3(MCL_BLOCK), 0, c->stream, L.dev,
                       (const double *)L.d, L.V);
    hipLaunchKernelGGL(k_stress_elem, dim3(grid_for(c, L.nel)), dim3(MCL_BLOCK), 0, c->stream, L.dev,
                       (const double *)L.vis, (const double *)L.V, L.es, L.area);
    hipLaunchKernelGGL(k_stress_node, dim3(grid_for(c, L.nno)), dim3(MCL_BLOCK), 0, c->stream, L.dev,
                       (const double *)L.es, (const double *)L.area, L.stress);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(k_residuals, dim3(grid_for(c, L.neq)), dim3(MCL_BLOCK), 0, c->stream, L.dev, W.rb,
                       (const double *)L.d, (const double *)L.p, k.precision == MCL_PREC_MIXED ? 1 : 0);
    hipLaunchKernelGGL(k_collect3, dim3(1), dim3(MCL_BLOCK), 0, c->stream, W.rb, 12, (int)grid_for(c, L.neq),
                       W.resid);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// Statistics of a finished solve from the kernel's block and the residual norms.
static void solve_stats(mcl_ctx *c, const DevStats &st, const double *r, double accuracy, float ms,
                        mcl_stokes_stats *out)
{
    const mcl_control &k = c->ctl;
    if (!out) return;
    out->uzawa_cycles = st.cycles;
    out->dv = st.dv;
    out->dp = st.dp;
    out->converged = (st.dv <= k.citcom_accu && (k.check_v_only || st.dp <= k.citcom_accu)) ? 1 : 0;
    out->inner_iterations = st.inner;
    out->spmv = st.spmv;
    out->accuracy_used = accuracy;
    double fn = std::sqrt(r[2]);
    out->res_momentum = fn > 0 ? std::sqrt(r[0]) / fn : std::sqrt(r[0]);
    out->res_continuity = fn > 0 ? std::sqrt(r[1]) / fn : std::sqrt(r[1]);
    out->device_ms = ms;
}

static int solve_checks(mcl_ctx *c)
{
    const mcl_control &k = c->ctl;
    if (k.precision == MCL_PREC_FP64 && k.solver == MCL_SOLVER_MG) {
        // Solver.c:678-683
        if (c->nlev < 2) return fail(c, MCL_E_ARG, "for MG solver, min_level needs to be smaller than max_level-1");
        if (!c->gs_mg_ok)
            return fail(c, MCL_E_ARG, "solver mg needs the interp and project maps of every level above the coarsest");
    }
    return 0;
}

// solve_velocity_pressure (Solver.c:3-12) on work space W, synchronously.
// noalloc: the allocation-free form (FP64 only; no events, device time from
// the kernel's clock stamps; every buffer cleared by the zero-fill kernel).
static int solve_run(mcl_ctx *c, mcl_stokes_stats *out, const SolveWork &W, bool noalloc)
{
    Level &L = c->levels.back();
    const mcl_control &k = c->ctl;
    int rc = solve_checks(c);
    if (rc) return rc;
    double accuracy = k.accuracy;
    if (k.auto_accuracy) {
        // Solver.c:41-51: accuracy = citcom_accu * sqrt(F.F / neq)
        hipLaunchKernelGGL(k_residuals, dim3(grid_for(c, L.neq)), dim3(MCL_BLOCK), 0, c->stream, L.dev, W.rb,
                           (const double *)L.d, (const double *)L.p, 0);
        hipLaunchKernelGGL(k_collect3, dim3(1), dim3(MCL_BLOCK), 0, c->stream, W.rb, 12,
                           (int)grid_for(c, L.neq), W.resid);
        double r[3];
        HIPCHK(c, hipMemcpyAsync(r, W.resid, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        accuracy = k.citcom_accu * std::sqrt(r[2] / L.neq);
    }
    EventPair ev;
    if (!noalloc) {
        rc = ev.create(c);
        if (rc) return rc;
    }
    rc = solve_enqueue(c, W, noalloc ? ENQ_NOALLOC : ENQ_NORMAL, noalloc ? nullptr : &ev, accuracy);
    if (rc) return rc;
    DevStats st;
    double r[3];
    HIPCHK(c, hipMemcpyAsync(&st, W.stats, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(r, W.resid, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    if (!noalloc) hipEventElapsedTime(&ms, ev.a, ev.b);
    else if (c->wall_khz > 0) ms = (float)((double)(st.t_end - st.t_begin) / c->wall_khz);
    rc = reset_sync_if_aborted(c, st, W, noalloc);
    if (rc) return rc;
#if MCL_PHASE_TIMERS
    {
        int khz = 0;
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device);
        fprintf(stderr, "[mcl phases ms]");
        for (int q = 0; q < MCL_PHASES; q++) fprintf(stderr, " %.2f", khz > 0 ? st.phase[q] / (double)khz : 0.0);
        fprintf(stderr, "\n");
    }
#endif
    solve_stats(c, st, r, accuracy, ms, out);
    if (st.nan_flag) return fail(c, MCL_E_NUMERIC, "NaN in the Stokes solve");
    return 0;
}

extern "C" int mcl_solve_velocity_pressure(mcl_ctx *c, mcl_stokes_stats *out)
{
    int rc = ready(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    return solve_run(c, out, ctx_work(c), false);
}

// Byte offsets of the allocation-free solve's buffers in the scratch array
// (each 256-aligned); total = bytes needed.
struct ScratchLayout {
    size_t sync, stats, resid, part, pmax, work, total;
};

static ScratchLayout scratch_layout(size_t nv, size_t np)
{
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    ScratchLayout S{};
    size_t o = 0;
    S.sync = o; o = up(o + MCL_SYNC_WORDS * sizeof(unsigned));
    S.stats = o; o = up(o + sizeof(DevStats));
    S.resid = o; o = up(o + 4 * sizeof(double));
    S.part = o; o = up(o + (size_t)MCL_RED_SLOTS * MCL_MAX_WG * sizeof(dd));
    S.pmax = o; o = up(o + (size_t)MCL_RED_SLOTS * MCL_MAX_WG * sizeof(double));
    S.work = o; o = up(o + (6 * nv + 9 * np) * sizeof(double));
    S.total = o;
    return S;
}

extern "C" int mcl_solve_velocity_pressure_noalloc(mcl_ctx *c, mcl_stokes_stats *out)
{
    int rc = ready(c);
    if (rc) return rc;
    if (c->ctl.precision != MCL_PREC_FP64)
        return fail(c, MCL_E_UNSUPPORTED, "the allocation-free solve runs the FP64 Uzawa kernel only (precision 0)");
    if (c->ctl.solver != MCL_SOLVER_CG)
        return fail(c, MCL_E_UNSUPPORTED, "the allocation-free solve supports the conjugate-gradient inner solver only "
                                          "(solver cg)");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->device < 0 || c->device >= MCL_SCRATCH_DEVICES || !g_scratch[c->device])
        return fail(c, MCL_E_DEVICE, "static scratch of device %d not located: %s", c->device,
                    hipGetErrorString(c->device >= 0 && c->device < MCL_SCRATCH_DEVICES ? g_scratch_err[c->device]
                                                                                        : hipErrorInvalidDevice));
    Level &L = c->levels.back();
    const size_t nv = std::max(L.neq, 1), np = L.nel;
    const ScratchLayout S = scratch_layout(nv, np);
    if (S.total > MCL_SCRATCH_BYTES)
        return fail(c, MCL_E_ARG,
                    "mesh too large for the allocation-free solve: %zu velocity and %zu pressure unknowns need %zu "
                    "bytes of scratch, the maximum is %llu",
                    nv, np, S.total, (unsigned long long)MCL_SCRATCH_BYTES);
    unsigned char *base = g_scratch[c->device];
    SolveWork W{};
    W.gs.gen = (unsigned *)(base + S.sync);
    W.gs.abort = W.gs.gen + 32;
    W.gs.arrive = W.gs.gen + 64;
    W.stats = (DevStats *)(base + S.stats);
    W.resid = (double *)(base + S.resid);
    W.rb.part = (dd *)(base + S.part);
    W.rb.pmax = (double *)(base + S.pmax);
    W.work = (double *)(base + S.work);
    return solve_run(c, out, W, true);
}

// The device work of update_temperature (Update_temperature.c:3-133).
static int heat_enqueue(mcl_ctx *c, EnqMode mode, EventPair *ev)
{
    Level &L = c->levels.back();
    const mcl_control &k = c->ctl;
    int rc = zero_async(c, c->stats, sizeof(DevStats), mode != ENQ_NORMAL);
    if (rc) return rc;
    HeatParams H{};
    H.L = L.dev;
    H.gs = c->gs;
    H.rb = c->rb;
    H.stats = c->stats;
    H.V = L.V;
    H.T = L.T;
    H.tau = c->tau;
    H.hK = c->hK;
    H.hC = c->hC;
    H.hA = c->hA;
    H.hF = c->hF;
    H.hd = c->hd;
    H.hdd = c->hdd;
    H.d0 = c->hd0;
    H.B = c->hB;
    double *w = c->work;
    size_t nt = std::max(L.neqT, 1);
    H.cg_r = w; w += nt;
    H.cg_s = w; w += nt;
    H.cg_d = w; w += nt;
    H.cg_q = w; w += nt;
    H.model_time = c->model_time;
    H.accuracy = k.accuracy;
    H.theta = k.theta;
    H.Q = k.Q;
    H.Llast0 = L.Llast0;
    H.Llast1 = L.Llast1;
    H.diff_dt = L.diff_dt;
    H.first_call = c->first_heat ? 1 : 0;
    H.cg_precondition = k.cg_precondition;
    H.exact_dots = k.exact_dots;
    H.gs.nblocks = coop_blocks(c, std::max(L.neqT, L.nel), MCL_FP64_ROWS_PER_THREAD * MCL_BLOCK, "MCL_FP64_BLOCKS");
    if (ev) HIPCHK(c, hipEventRecord(ev->a, c->stream));
    void *args[] = {&H};
    HIPCHK(c, launch_persistent(c, (const void *)k_heat_fp64, H.gs.nblocks, args, mode == ENQ_GRAPH));
    if (ev) HIPCHK(c, hipEventRecord(ev->b, c->stream));
    return 0;
}

extern "C" int mcl_update_temperature(mcl_ctx *c, mcl_heat_stats *out)
{
    int rc = ready(c);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    EventPair ev;
    rc = ev.create(c);
    if (rc) return rc;
    rc = heat_enqueue(c, ENQ_NORMAL, &ev);
    if (rc) return rc;
    DevStats st;
    double mt;
    HIPCHK(c, hipMemcpyAsync(&st, c->stats, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&mt, c->model_time, sizeof mt, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    hipEventElapsedTime(&ms, ev.a, ev.b);
    rc = reset_sync_if_aborted(c, st, ctx_work(c), false);
    if (rc) return rc;
    c->first_heat = false;
    if (out) {
        out->cg_cycles = st.cycles;
        out->init_cg_cycles = st.init_cycles;
        out->dt = st.dt;
        out->model_time = mt;
        out->device_ms = ms;
    }
    return 0;
}

// ------------------------------------------------------------ batches
// Many independent instances advanced together on one device.  Each
// instance's time step after its first -- buoyancy, viscosity, assembly (and
// the FP32 multigrid hierarchy), the Stokes solve, nodal velocity and stress,
// residual norms, the heat step, and the copies of its statistics into pinned
// host memory -- is captured once into a HIP graph.  A round replays every
// instance's graph on one of `lanes` streams (instances dealt largest first
// to the least loaded lane, so the lanes carry similar work) and waits for
// all; the lanes run concurrently and fill the device that one small
// instance leaves mostly idle.  Every persistent launch of the batch's
// instances is capped at max_coop_blocks / lanes workgroups, so the
// persistent kernels of all lanes fit on the device together.
namespace {
struct HostSlot {
    DevStats solve;
    double resid[3];
    DevStats heat;
    double model_time;
};
}  // namespace

struct mcl_batch {
    int device = 0;
    std::vector<mcl_ctx *> ctx;
    std::vector<int> lane;               // lane of each instance
    std::vector<hipStream_t> streams;    // one per lane
    std::vector<hipGraphExec_t> exec;
    std::vector<hipGraph_t> graph;
    std::vector<hipEvent_t> ev0, ev1;    // around each instance's graph, on its lane
    std::vector<double> accuracy;
    std::vector<int> prev_cap;           // each instance's launch cap before the batch
    HostSlot *slots = nullptr;           // pinned, one per instance
    std::string err;
};

static int batch_fail(mcl_batch *b, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

extern "C" void mcl_batch_destroy(mcl_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    for (hipStream_t s : b->streams) (void)hipStreamSynchronize(s);
    for (auto e : b->exec) if (e) (void)hipGraphExecDestroy(e);
    for (auto g : b->graph) if (g) (void)hipGraphDestroy(g);
    for (auto e : b->ev0) if (e) (void)hipEventDestroy(e);
    for (auto e : b->ev1) if (e) (void)hipEventDestroy(e);
    for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
    if (b->slots) (void)hipHostFree(b->slots);
    for (size_t i = 0; i < b->prev_cap.size() && i < b->ctx.size(); i++) b->ctx[i]->block_cap = b->prev_cap[i];
    delete b;
}

extern "C" const char *mcl_batch_last_error(const mcl_batch *b) { return b ? b->err.c_str() : "null batch"; }

// Capture one instance's step into a graph (its own stream in capture mode).
static int batch_capture(mcl_batch *b, int i)
{
    mcl_ctx *c = b->ctx[i];
    HostSlot *hs = b->slots + i;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed) != hipSuccess)
        return batch_fail(b, MCL_E_DEVICE, "instance %d: stream capture could not begin", i);
    int rc = mcl_buoyancy_force(c);
    if (!rc) rc = mcl_viscosity(c);
    if (!rc && c->asm_blocks > 0) {
        hipLaunchKernelGGL(k_assemble_K, dim3(c->asm_blocks), dim3(MCL_BLOCK), 0, c->stream,
                           (const AsmLevel *)c->asm_tab, c->nlev);
        if (hipGetLastError() != hipSuccess) rc = fail(c, MCL_E_DEVICE, "assembly launch");
    }
    // (the mixed path's FP32 hierarchy is rebuilt inside solve_enqueue)
    if (!rc) rc = solve_enqueue(c, ctx_work(c), ENQ_GRAPH, nullptr, b->accuracy[i]);
    if (!rc && hipMemcpyAsync(&hs->solve, c->stats, sizeof(DevStats), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = fail(c, MCL_E_DEVICE, "stats copy");
    if (!rc && hipMemcpyAsync(hs->resid, c->resid, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = fail(c, MCL_E_DEVICE, "residual copy");
    if (!rc) rc = heat_enqueue(c, ENQ_GRAPH, nullptr);
    if (!rc && hipMemcpyAsync(&hs->heat, c->stats, sizeof(DevStats), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = fail(c, MCL_E_DEVICE, "heat stats copy");
    if (!rc && hipMemcpyAsync(&hs->model_time, c->model_time, sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
                   hipSuccess)
        rc = fail(c, MCL_E_DEVICE, "model time copy");
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return batch_fail(b, rc, "instance %d: %s", i, c->err.c_str());
    }
    if (e != hipSuccess || !g) return batch_fail(b, MCL_E_DEVICE, "instance %d: capture: %s", i, hipGetErrorString(e));
    b->graph[i] = g;
    e = hipGraphInstantiate(&b->exec[i], g, nullptr, nullptr, 0);
    if (e != hipSuccess) return batch_fail(b, MCL_E_DEVICE, "instance %d: instantiate: %s", i, hipGetErrorString(e));
    return 0;
}

extern "C" int mcl_batch_create(mcl_ctx *const *ctxs, int n, int lanes, mcl_batch **out)
{
    if (!out) return MCL_E_ARG;
    *out = nullptr;
    if (!ctxs || n <= 0 || lanes <= 0) return MCL_E_ARG;
    mcl_batch *b = new mcl_batch();
    b->device = ctxs[0] ? ctxs[0]->device : 0;
    for (int i = 0; i < n; i++) {
        mcl_ctx *c = ctxs[i];
        int rc = ready(c);
        if (rc) { delete b; return rc; }
        if (c->device != b->device) { delete b; return MCL_E_ARG; }
        if (c->first_heat || !c->been_here_dp) {
            // the first step (initial_T_dot, initial_d_p's first call) stays eager
            c->err = "a batch starts after each instance's first step";
            delete b;
            return MCL_E_STATE;
        }
        if (c->ctl.auto_accuracy) {
            c->err = "auto_accuracy needs a host read-back in every solve: not available in a batch";
            delete b;
            return MCL_E_UNSUPPORTED;
        }
        if (solve_checks(c)) { delete b; return MCL_E_ARG; }
        b->ctx.push_back(c);
    }
    int rc = 0;
    if (hipSetDevice(b->device) != hipSuccess) { delete b; return MCL_E_DEVICE; }
    lanes = std::min(lanes, n);
    b->lane.assign(n, 0);
    b->graph.assign(n, nullptr);
    b->exec.assign(n, nullptr);
    b->ev0.assign(n, nullptr);
    b->ev1.assign(n, nullptr);
    b->accuracy.assign(n, 0.0);
    for (mcl_ctx *c : b->ctx) b->prev_cap.push_back(c->block_cap);
    // lanes: instances dealt largest first to the lane with the fewest rows
    {
        std::vector<int> order(n);
        for (int i = 0; i < n; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return b->ctx[x]->levels.back().neq > b->ctx[y]->levels.back().neq;
        });
        std::vector<long long> load(lanes, 0);
        for (int i : order) {
            int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            b->lane[i] = l;
            load[l] += b->ctx[i]->levels.back().neq;
        }
    }
    b->streams.assign(lanes, nullptr);
    for (int l = 0; l < lanes && !rc; l++)
        if (hipStreamCreateWithFlags(&b->streams[l], hipStreamNonBlocking) != hipSuccess)
            rc = batch_fail(b, MCL_E_DEVICE, "lane stream");
    if (!rc && hipHostMalloc((void **)&b->slots, n * sizeof(HostSlot), hipHostMallocDefault) != hipSuccess)
        rc = batch_fail(b, MCL_E_DEVICE, "pinned statistics");
    for (int i = 0; i < n && !rc; i++) {
        mcl_ctx *c = b->ctx[i];
        if (hipEventCreate(&b->ev0[i]) != hipSuccess || hipEventCreate(&b->ev1[i]) != hipSuccess) {
            rc = batch_fail(b, MCL_E_DEVICE, "events");
            break;
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = batch_fail(b, MCL_E_DEVICE, "instance %d stream", i);
            break;
        }
        const int share = std::max(1, c->max_coop_blocks / lanes);
        c->block_cap = c->block_cap > 0 ? std::min(c->block_cap, share) : share;
        b->accuracy[i] = c->ctl.accuracy;
        // the mixed solve's halo and neighbour lists for the capped workgroup
        // count are built now: their set-up allocates and may not be captured
        if (c->ctl.precision == MCL_PREC_MIXED && c->have_mg) {
            Level &T = c->levels.back();
            const int nb = (int)coop_blocks(c, std::max(T.neq, T.nel), MCL_MIXED_ROWS_PER_BLOCK, "MCL_MIXED_BLOCKS");
            if (nb != c->nbr_nb && build_neighbours(c, nb)) {
                rc = batch_fail(b, MCL_E_DEVICE, "instance %d: %s", i, c->err.c_str());
                break;
            }
        }
        rc = batch_capture(b, i);
    }
    if (rc) {
        int code = rc;
        std::string msg = b->err;
        mcl_batch_destroy(b);
        (void)msg;
        return code;
    }
    *out = b;
    return 0;
}

// One time step of every instance whose skip[i] is 0 (skip may be NULL):
// graph replays on the lanes, then a wait for all; statistics of instance i
// in stokes[i], heat[i] and its step's device time (events around its graph
// on its lane) in step_ms[i] (each output may be NULL).
extern "C" int mcl_batch_step(mcl_batch *b, const int *skip, mcl_stokes_stats *stokes, mcl_heat_stats *heat,
                              double *step_ms)
{
    if (!b) return MCL_E_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return batch_fail(b, MCL_E_DEVICE, "device");
    const int n = (int)b->ctx.size();
    for (int i = 0; i < n; i++) {
        if (skip && skip[i]) continue;
        hipStream_t s = b->streams[b->lane[i]];
        if (hipEventRecord(b->ev0[i], s) != hipSuccess || hipGraphLaunch(b->exec[i], s) != hipSuccess ||
            hipEventRecord(b->ev1[i], s) != hipSuccess)
            return batch_fail(b, MCL_E_DEVICE, "instance %d: graph launch", i);
    }
    for (hipStream_t s : b->streams)
        if (hipStreamSynchronize(s) != hipSuccess) return batch_fail(b, MCL_E_DEVICE, "lane synchronise");
    int rc = 0;
    for (int i = 0; i < n; i++) {
        if (skip && skip[i]) continue;
        mcl_ctx *c = b->ctx[i];
        const HostSlot &hs = b->slots[i];
        float ms = 0.f;
        hipEventElapsedTime(&ms, b->ev0[i], b->ev1[i]);
        if (step_ms) step_ms[i] = ms;
        if (hs.solve.aborted || hs.heat.aborted) {
            (void)zero_async(c, c->gs.gen, MCL_SYNC_WORDS * sizeof(unsigned), false);
            (void)hipStreamSynchronize(c->stream);
            rc = batch_fail(b, MCL_E_DEVICE, "instance %d: grid barrier timed out", i);
            continue;
        }
        solve_stats(c, hs.solve, hs.resid, b->accuracy[i], 0.f, stokes ? stokes + i : nullptr);
        if (heat) {
            heat[i].cg_cycles = hs.heat.cycles;
            heat[i].init_cg_cycles = hs.heat.init_cycles;
            heat[i].dt = hs.heat.dt;
            heat[i].model_time = hs.model_time;
            heat[i].device_ms = 0.0;
        }
        if (hs.solve.nan_flag && !rc) rc = batch_fail(b, MCL_E_NUMERIC, "instance %d: NaN in the Stokes solve", i);
    }
    return rc;
}

// ------------------------------------------------------------ learned guess

// Element-centre samples of the finest (uniform) level on an H x W image with
// rows = z and columns = x: inputs (T, log10 eta), targets (u_x, u_z, p) / Ra.
__global__ void k_net_sample(LevelDev L, int W, const double *__restrict__ T, const double *__restrict__ vis,
                             const double *__restrict__ V, const double *__restrict__ p, double inv_Ra,
                             float *__restrict__ x, float *__restrict__ t)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L.nel; e += gridDim.x * blockDim.x) {
        int g = L.node[L.ienl[e]];
        int ix = g / L.gnoz, iz = g - ix * L.gnoz;
        int pix = iz * W + ix;
        double tc = 0.0, lc = 0.0, uc = 0.0, wc = 0.0;
        for (int a = 0; a < 4; a++) {
            int k = L.ienl[a * L.nel + e];
            tc += 0.25 * T[k];
            lc += 0.25 * log10(vis[k]);
            uc += 0.25 * V[k];
            wc += 0.25 * V[L.nno + k];
        }
        x[2 * pix] = (float)tc;
        x[2 * pix + 1] = (float)lc;
        if (t) {
            t[3 * pix] = (float)(uc * inv_Ra);
            t[3 * pix + 1] = (float)(wc * inv_Ra);
            t[3 * pix + 2] = (float)(p[e] * inv_Ra);
        }
    }
}

// Sample normalisation.  Stokes flow is linear in the buoyancy Ra (T - row
// mean of T) for a fixed viscosity, so the network sees the temperature
// anomaly scaled to unit RMS, a = (T - rowmean) / s, and learns the flow per
// unit anomaly: targets (u, w, p - mean p) / (Ra s).  The guess is rescaled by
// Ra s.  Pass 1 (one workgroup per image row): subtract the row mean of T and
// accumulate the row's sum of squares and pressure sum.  Pass 2: every
// workgroup sums the H row partials in the same order, so all get the same s
// and pressure mean; it scales its pixels, and workgroup 0 stores s.
__global__ void k_net_norm_rows(int H, int W, float *__restrict__ x, const float *__restrict__ t,
                                double *__restrict__ rows)
{
    __shared__ double red[2][MCL_BLOCK / MCL_WAVE];
    const int iz = blockIdx.x;
    if (iz >= H) return;
    double st = 0.0, sp = 0.0;
    for (int ix = threadIdx.x; ix < W; ix += blockDim.x) {
        st += (double)x[2 * (iz * W + ix)];
        if (t) sp += (double)t[3 * (iz * W + ix) + 2];
    }
    for (int o = MCL_WAVE / 2; o > 0; o >>= 1) {
        st += __shfl_xor(st, o, MCL_WAVE);
        sp += __shfl_xor(sp, o, MCL_WAVE);
    }
    const int wv = threadIdx.x / MCL_WAVE, ln = threadIdx.x % MCL_WAVE;
    if (ln == 0) { red[0][wv] = st; red[1][wv] = sp; }
    __syncthreads();
    st = 0.0; sp = 0.0;
    for (int k = 0; k < (int)(blockDim.x / MCL_WAVE); k++) { st += red[0][k]; sp += red[1][k]; }
    __syncthreads();
    const double mean = st / W;
    double sq = 0.0;
    for (int ix = threadIdx.x; ix < W; ix += blockDim.x) {
        float a = (float)((double)x[2 * (iz * W + ix)] - mean);
        x[2 * (iz * W + ix)] = a;
        sq += (double)a * a;
    }
    for (int o = MCL_WAVE / 2; o > 0; o >>= 1) sq += __shfl_xor(sq, o, MCL_WAVE);
    if (ln == 0) red[0][wv] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = 0.0;
        for (int k = 0; k < (int)(blockDim.x / MCL_WAVE); k++) q += red[0][k];
        rows[iz] = q;
        rows[H + iz] = sp;
    }
}

__global__ void k_net_norm_apply(int H, int W, float *__restrict__ x, float *__restrict__ t,
                                 double *__restrict__ rows)
{
    double q = 0.0, sp = 0.0;
    for (int k = 0; k < H; k++) { q += rows[k]; sp += rows[H + k]; }
    const double n = (double)H * W;
    double sc = sqrt(q / n);
    if (!(sc > 1e-300)) sc = 1e-300;
    const double pm = sp / n;
    const float inv = (float)(1.0 / sc);
    for (int pix = blockIdx.x * blockDim.x + threadIdx.x; pix < H * W; pix += gridDim.x * blockDim.x) {
        x[2 * pix] *= inv;
        if (t) {
            t[3 * pix] *= inv;
            t[3 * pix + 1] *= inv;
            t[3 * pix + 2] = (float)(((double)t[3 * pix + 2] - pm) / sc);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) rows[2 * H] = sc;
}

// Network output -> (d, p): pressure per element, nodal velocity = mean of the
// adjacent element-centre values, written to the free velocity equations.
__global__ void k_net_unpack_p(LevelDev L, int W, const float *__restrict__ y, double scale0,
                               const double *__restrict__ sc, double *__restrict__ p)
{
    const double scale = scale0 * sc[0];
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L.nel; e += gridDim.x * blockDim.x) {
        int g = L.node[L.ienl[e]];
        int ix = g / L.gnoz, iz = g - ix * L.gnoz;
        p[e] = scale * (double)y[3 * (iz * W + ix) + 2];
    }
}

__global__ void k_net_unpack_d(LevelDev L, int W, const float *__restrict__ y, double su, double sw,
                               const double *__restrict__ sc, double *__restrict__ d)
{
    const double scale[2] = {su * sc[0], sw * sc[0]};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L.nno; i += gridDim.x * blockDim.x) {
        int cnt = L.p2n_cnt[i];
        for (int k = 0; k < 2; k++) {
            int eq = L.id[k * L.nno + i];
            if (eq < 0) continue;
            double s = 0.0;
            for (int j = 0; j < cnt; j++) {
                int e = L.p2n_el[j * L.nno + i];
                int g = L.node[L.ienl[e]];
                int ix = g / L.gnoz, iz = g - ix * L.gnoz;
                s += (double)y[3 * (iz * W + ix) + k];
            }
            d[eq] = scale[k] * s / cnt;
        }
    }
}

static int net_check(mcl_ctx *c, mcl_net *net, int *H, int *W, int *B)
{
    int rc = ready(c);
    if (rc) return rc;
    if (!net) return fail(c, MCL_E_ARG, "null network");
    mcl_net_dims(net, H, W, B);
    Level &L = c->levels.back();
    if ((long long)(*H) * (*W) != L.nel || L.gnoz - 1 != *H || L.gnox - 1 != *W)
        return fail(c, MCL_E_ARG, "network grid %dx%d does not match the uniform %dx%d mesh", *W, *H, L.gnox - 1,
                    L.gnoz - 1);
    for (int m : L.Mcode)
        if ((m & 15) != L.L - 1) return fail(c, MCL_E_UNSUPPORTED, "learned guess needs a uniform mesh");
    return 0;
}

// sample the fields onto the network grid and normalise them (see above)
static int net_sample(mcl_ctx *c, int H, int W, float *x, float *t)
{
    Level &L = c->levels.back();
    if (c->net_rows_h != H) {
        if (c->net_rows) {
            (void)hipFree(c->net_rows);
            c->owned.erase(std::remove(c->owned.begin(), c->owned.end(), (void *)c->net_rows), c->owned.end());
            c->net_rows = nullptr;
        }
        if (dev_alloc(c, c->owned, &c->net_rows, 2 * (size_t)H + 1)) return MCL_E_DEVICE;
        c->net_rows_h = H;
    }
    hipLaunchKernelGGL(k_net_sample, dim3(grid_for(c, L.nel)), dim3(MCL_BLOCK), 0, c->stream, L.dev, W,
                       (const double *)L.T, (const double *)L.vis, (const double *)L.V, (const double *)L.p,
                       1.0 / c->ctl.Ra, x, t);
    hipLaunchKernelGGL(k_net_norm_rows, dim3(H), dim3(MCL_BLOCK), 0, c->stream, H, W, x, (const float *)t,
                       c->net_rows);
    hipLaunchKernelGGL(k_net_norm_apply, dim3(grid_for(c, (long long)H * W)), dim3(MCL_BLOCK), 0, c->stream, H, W,
                       x, t, c->net_rows);
    HIPCHK(c, hipGetLastError());
    return 0;
}

extern "C" int mcl_net_capture(mcl_ctx *c, mcl_net *net)
{
    int H, W, B;
    int rc = net_check(c, net, &H, &W, &B);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    Level &L = c->levels.back();
    float *x = mcl_net_in(net), *t = mcl_net_tgt(net);  // batch slot 0 as staging
    HIPCHK(c, hipStreamSynchronize(mcl_net_stream(net)));
    rc = net_sample(c, H, W, x, t);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)L;
    return mcl_net_add_snapshot(net, x, t) ? fail(c, MCL_E_DEVICE, "snapshot: %s", mcl_net_last_error(net)) : 0;
}

// Initial Schur residual |G^T K^-1 (F - G p)| of the candidate (d, p): a
// probe launch of the mixed Uzawa kernel runs the first inner solve (d becomes
// K^-1 (F - G p), starting from the candidate's d) and stops.
static int uzawa_probe(mcl_ctx *c, double *d, double *p, double *r1)
{
    HIPCHK(c, hipMemsetAsync(c->stats, 0, sizeof(DevStats), c->stream));
    UzawaParams P = uzawa_params(c, c->ctl.accuracy, ctx_work(c));
    P.d = d;
    P.p = p;
    MixedParams M;
    int rc = mixed_params(c, P, M);
    if (rc) return rc;
    M.probe = 1;
    {
        // the probe only ranks two candidates: a looser inner solve suffices
        // (the solve that follows corrects d to full accuracy)
        const char *e = getenv("MCL_PROBE_RTOL");
        M.inner_rtol = e && atof(e) > 0 ? atof(e) : 1e-3;
    }
    void *margs[] = {&M};
    HIPCHK(c, launch_persistent(c, (const void *)k_uzawa_probe, M.U.gs.nblocks, margs));
    DevStats st;
    HIPCHK(c, hipMemcpyAsync(&st, c->stats, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    rc = reset_sync_if_aborted(c, st, ctx_work(c), false);
    if (rc) return rc;
    *r1 = std::sqrt(st.probe_r1);
    return 0;
}

extern "C" int mcl_stokes_guess(mcl_ctx *c, mcl_net *net, int *accepted, double *res_net, double *res_prev)
{
    int H, W, B;
    int rc = net_check(c, net, &H, &W, &B);
    if (rc) return rc;
    if (c->ctl.precision != MCL_PREC_MIXED || !c->ctl.warm_start)
        return fail(c, MCL_E_UNSUPPORTED, "the learned guess needs precision=1 and warm_start=1");
    HIPCHK(c, hipSetDevice(c->device));
    Level &L = c->levels.back();
    if (!c->mixed_ready) {
        rc = mixed_prepare(c, false);
        if (rc) return rc;
    }
    if (!c->gd) {
        if (dev_alloc(c, c->owned, &c->gd, std::max(L.neq, 1)) || dev_alloc(c, c->owned, &c->gp, L.nel))
            return MCL_E_DEVICE;
    }
    float *x = mcl_net_in(net);
    HIPCHK(c, hipStreamSynchronize(mcl_net_stream(net)));
    rc = net_sample(c, H, W, x, nullptr);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (mcl_net_forward_device(net)) return fail(c, MCL_E_DEVICE, "network: %s", mcl_net_last_error(net));
    HIPCHK(c, hipStreamSynchronize(mcl_net_stream(net)));
    // candidate from the network; equations the network does not set keep the current d
    const double Ra = c->ctl.Ra;
    HIPCHK(c, hipMemcpyAsync(c->gd, L.d, std::max(L.neq, 1) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_net_unpack_d, dim3(grid_for(c, L.nno)), dim3(MCL_BLOCK), 0, c->stream, L.dev, W,
                       (const float *)mcl_net_out(net), Ra * mcl_net_out_scale(net, 0), Ra * mcl_net_out_scale(net, 1),
                       (const double *)(c->net_rows + 2 * H), c->gd);
    hipLaunchKernelGGL(k_net_unpack_p, dim3(grid_for(c, L.nel)), dim3(MCL_BLOCK), 0, c->stream, L.dev, W,
                       (const float *)mcl_net_out(net), Ra * mcl_net_out_scale(net, 2),
                       (const double *)(c->net_rows + 2 * H), c->gp);
    HIPCHK(c, hipGetLastError());
    // gate: the initial Schur residual of each candidate (the current fields
    // are the previous step's solution, extrapolated when there are two, or
    // zero on a cold start)
    rc = extrapolate(c);
    if (rc) return rc;
    double rp = 0, rn = 0;
    rc = uzawa_probe(c, L.d, L.p, &rp);
    if (rc) return rc;
    rc = uzawa_probe(c, c->gd, c->gp, &rn);
    if (rc) return rc;
    int acc = rn < rp ? 1 : 0;
    if (acc) {
        HIPCHK(c, hipMemcpyAsync(L.d, c->gd, std::max(L.neq, 1) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(L.p, c->gp, L.nel * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        // the fields no longer hold the last solution: no extrapolation
        // until two solutions follow
        HIPCHK(c, hipMemsetAsync(c->ws, 0, sizeof(double), c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (accepted) *accepted = acc;
    if (res_net) *res_net = rn;
    if (res_prev) *res_prev = rp;
    return 0;
}
