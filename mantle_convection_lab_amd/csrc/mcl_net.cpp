// Learned initial guess for the Stokes solve: host side of the U-Net.
//
// The network maps two fields resampled to element centres of a uniform
// mesh -- temperature and log10 viscosity -- to three: u_x, u_z and p at the
// element centres, in units of Ra (Stokes flow is linear in the buoyancy Ra T,
// so the solution at any Ra is Ra times the solution at Ra = 1) and scaled by a
// dataset constant.  Architecture: 4 levels, two 3x3 convolutions with ReLU per
// level (32, 64, 128, 256 channels), 2x2 average pooling down, nearest 2x
// upsampling up with the encoder activation concatenated, a 1x1 head.
// Training data are snapshots of the solver's own fields captured on the
// device; training runs on the device with the kernels of mcl_unet.hip (bf16
// MFMA products, FP32 accumulation and master weights, Adam).  The guess is
// accepted only if its momentum residual is smaller than that of the current
// (previous-step) guess.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mcl.h"
#include "mcl_net.h"

// kernels of mcl_unet.hip
struct ConvShape {
    int B, H, W;
    int Cin, Cout;
    int cs_in, ci_off;
    int cs_out, co_off;
};
__global__ void k_conv3_fwd(ConvShape s, const uint16_t *in, const uint16_t *Wf, const float *bias, uint16_t *out,
                            int relu);
__global__ void k_conv3_dgrad(ConvShape s, const uint16_t *dout, const uint16_t *relu_y, const uint16_t *Wd,
                              uint16_t *din, int accum);
__global__ void k_conv3_wgrad(ConvShape s, const uint16_t *dout, const uint16_t *relu_y, const uint16_t *in,
                              float *dW, int pix_per_split);
__global__ void k_conv3_mm(ConvShape s, int dgrad, const uint16_t *X, const uint16_t *mask, const uint16_t *Wt,
                           const float *bias, uint16_t *Y, int flag);
__global__ void k_conv3_mm_n32(ConvShape s, int dgrad, const uint16_t *X, const uint16_t *mask, const uint16_t *Wt,
                               const float *bias, uint16_t *Y, int flag);
__global__ void k_conv3_wgrad_mm(ConvShape s, const uint16_t *dout, const uint16_t *relu_y, const uint16_t *in,
                                 float *dW, float *db, int tiles_per_split);
__global__ void k_bias_grad(int M, int Cout, int cs, int coff, const uint16_t *dout, const uint16_t *relu_y,
                            float *db);
__global__ void k_pool_fwd(int B, int H, int W, int C, const uint16_t *in, int cs_in, int ci_off, uint16_t *out,
                           int cs_out, int co_off);
__global__ void k_pool_bwd(int B, int H, int W, int C, const uint16_t *dout, int cs_out, int co_off, uint16_t *din,
                           int cs_in, int ci_off, int accum);
__global__ void k_up_fwd(int B, int H, int W, int C, const uint16_t *in, int cs_in, int ci_off, uint16_t *out,
                         int cs_out, int co_off);
__global__ void k_up_bwd(int B, int H, int W, int C, const uint16_t *dout, int cs_out, int co_off, uint16_t *din,
                         int cs_in, int ci_off);
__global__ void k_head_fwd(int M, const uint16_t *x, int cs, const float *Wh, const float *bh, float *y);
__global__ void k_head_bwd(int M, const uint16_t *x, int cs, const float *Wh, const float *y, const float *tgt,
                           float scale, uint16_t *dx, float *dWh, float *dbh, float *loss);
__global__ void k_adam(int n, float *w, const float *g, float *m1, float *m2, float lr, float b1, float b2, float eps,
                       float c1, float c2);
__global__ void k_weights_bf16(int Cout, int Cin, const float *w, uint16_t *Wf, uint16_t *Wd);
__global__ void k_f32_to_bf16(size_t n, const float *x, uint16_t *y);
__global__ void k_bf16_to_f32(size_t n, const uint16_t *x, float *y);

namespace {

struct Act {           // one NHWC bf16 activation buffer and its gradient
    int H = 0, W = 0, C = 0;
    uint16_t *y = nullptr, *g = nullptr;
};

struct Conv {          // one 3x3 convolution layer
    int Cin = 0, Cout = 0;
    size_t nw = 0;     // Cout*9*Cin
    float *w = nullptr, *b = nullptr, *dw = nullptr, *db = nullptr;
    float *m1 = nullptr, *m2 = nullptr, *bm1 = nullptr, *bm2 = nullptr;
    uint16_t *Wf = nullptr, *Wd = nullptr;
};

}  // namespace

struct mcl_net {
    int H = 0, W = 0, B = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> owned;
    std::string err;
    // activations (encoder, concat buffers, decoder)
    Act x0, e1a, cat1, p1, e2a, cat2, p2, e3a, cat3, p3, b1, b2, d3a, d3b, d2a, d2b, d1a, d1b;
    Conv enc1a, enc1b, enc2a, enc2b, enc3a, enc3b, bota, botb, dec3a, dec3b, dec2a, dec2b, dec1a, dec1b;
    // head (1x1, 32 -> 3, FP32)
    float *wh = nullptr, *bh = nullptr, *dwh = nullptr, *dbh = nullptr, *hm1 = nullptr, *hm2 = nullptr,
          *hbm1 = nullptr, *hbm2 = nullptr;
    float *y = nullptr, *tgt = nullptr, *xin = nullptr, *loss = nullptr;
    long long adam_t = 0;
    float out_scale[3] = {1.0f, 1.0f, 1.0f};   // per output channel (RMS of the training targets)
    bool scaled = false;
    // dataset of snapshots on the device: inputs [N][H][W][2], targets [N][H][W][3]
    std::vector<float *> ds_x, ds_t;
};

#define NCHK(call)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            n->err = std::string(#call) + ": " + hipGetErrorString(e_);                   \
            return MCL_E_DEVICE;                                                         \
        }                                                                                \
    } while (0)

template <typename T>
static int nalloc(mcl_net *n, T **p, size_t count)
{
    void *q = nullptr;
    NCHK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    NCHK(hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T)));
    n->owned.push_back(q);
    *p = (T *)q;
    return 0;
}

static int make_act(mcl_net *n, Act &a, int H, int W, int C)
{
    a.H = H;
    a.W = W;
    a.C = C;
    size_t cnt = (size_t)n->B * H * W * C;
    if (nalloc(n, &a.y, cnt) || nalloc(n, &a.g, cnt)) return MCL_E_DEVICE;
    return 0;
}

static int make_conv(mcl_net *n, Conv &c, int Cin, int Cout, std::mt19937_64 &rng)
{
    c.Cin = Cin;
    c.Cout = Cout;
    c.nw = (size_t)Cout * 9 * Cin;
    if (nalloc(n, &c.w, c.nw) || nalloc(n, &c.b, Cout) || nalloc(n, &c.dw, c.nw) || nalloc(n, &c.db, Cout) ||
        nalloc(n, &c.m1, c.nw) || nalloc(n, &c.m2, c.nw) || nalloc(n, &c.bm1, Cout) || nalloc(n, &c.bm2, Cout) ||
        nalloc(n, &c.Wf, c.nw) || nalloc(n, &c.Wd, c.nw))
        return MCL_E_DEVICE;
    // He initialisation for ReLU layers
    std::normal_distribution<float> nd(0.f, std::sqrt(2.0f / (9.0f * (float)Cin)));
    std::vector<float> w(c.nw);
    for (auto &v : w) v = nd(rng);
    NCHK(hipMemcpy(c.w, w.data(), c.nw * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

static unsigned gblocks(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 8192); }

// The LDS-staged kernels (k_conv3_mm: 64 pixels x 64 channels a workgroup;
// k_conv3_mm_n32: 128 x 32, for N <= 32 output channels) apply when MT
// consecutive pixels form a rectangle of one image (TW = min(W, MT) divides
// both W and MT, TH = MT / TW divides H) and the summed channels and the
// staged buffer's channel stride and offset are multiples of 8 (16-byte
// loads).  Returns MT, or 0 for the unstaged kernels; MCL_CONV_DIRECT=1
// keeps the unstaged kernels (for comparisons).
static int conv_staged(const ConvShape &s, int N, int Kc, int xst, int xof)
{
    const char *e = getenv("MCL_CONV_DIRECT");
    if (e && atoi(e)) return 0;
    if (Kc % 8 || xst % 8 || xof % 8) return 0;
    const int MT = N <= 32 ? 128 : 64;
    const int TW = s.W < MT ? s.W : MT;
    if (TW <= 0 || s.W % TW || MT % TW || s.H % (MT / TW)) return 0;
    return MT;
}

// one launch of the staged kernel of tile MT (64 or 128) over M pixels, N channels
static void conv_mm(mcl_net *n, const ConvShape &s, int MT, int M, int N, int dgrad, const uint16_t *X,
                    const uint16_t *mask, const uint16_t *Wt, const float *bias, uint16_t *Y, int flag)
{
    if (MT == 128)
        hipLaunchKernelGGL(k_conv3_mm_n32, dim3(M / 128, (N + 31) / 32), dim3(256), 0, n->stream, s, dgrad, X, mask,
                           Wt, bias, Y, flag);
    else
        hipLaunchKernelGGL(k_conv3_mm, dim3(M / 64, (N + 63) / 64), dim3(256), 0, n->stream, s, dgrad, X, mask, Wt,
                           bias, Y, flag);
}

static void conv_fwd(mcl_net *n, Conv &c, const Act &in, int ci_off, Act &out, int co_off, bool relu)
{
    ConvShape s{n->B, in.H, in.W, c.Cin, c.Cout, in.C, ci_off, out.C, co_off};
    int M = n->B * in.H * in.W;
    const int MT = conv_staged(s, c.Cout, c.Cin, in.C, ci_off);
    if (MT)
        conv_mm(n, s, MT, M, c.Cout, 0, (const uint16_t *)in.y, nullptr, (const uint16_t *)c.Wf, (const float *)c.b,
                out.y, relu ? 1 : 0);
    else
        hipLaunchKernelGGL(k_conv3_fwd, dim3((M + 63) / 64, (c.Cout + 63) / 64), dim3(256), 0, n->stream, s,
                           (const uint16_t *)in.y, (const uint16_t *)c.Wf, (const float *)c.b, out.y, relu ? 1 : 0);
}

// Backward of a ReLU convolution: weight and bias gradients, and the data
// gradient into in.g (skipped when din is false).
static void conv_bwd(mcl_net *n, Conv &c, const Act &in, int ci_off, Act &out, int co_off, bool din, bool accum)
{
    ConvShape s{n->B, in.H, in.W, c.Cin, c.Cout, in.C, ci_off, out.C, co_off};
    int M = n->B * in.H * in.W;
    const uint16_t *mask = out.y;
    const int TW = in.W < 64 ? in.W : 64;
    if (conv_staged(s, 64, c.Cin, in.C, ci_off) == 64 && TW % 8 == 0 && out.C % 8 == 0 && co_off % 8 == 0) {
        // staged weight gradient (and bias gradient): workgroups of 64 x 32 x
        // 9 taps, about 512 in all, each summing at least 8 tiles of 64 pixels
        const int tiles = M / 64, nbn = (c.Cout + 63) / 64, nbc = (c.Cin + 31) / 32;
        int splits = std::max(1, std::min(tiles / 8, 512 / (nbn * nbc)));
        const int tps = (tiles + splits - 1) / splits;
        splits = (tiles + tps - 1) / tps;
        hipLaunchKernelGGL(k_conv3_wgrad_mm, dim3(nbn, nbc, splits), dim3(256), 0, n->stream, s,
                           (const uint16_t *)out.g, mask, (const uint16_t *)in.y, c.dw, c.db, tps);
    } else {
        int splits = std::max(1, std::min(256, M / 4096));
        int pps = ((M + splits - 1) / splits + 15) / 16 * 16;
        splits = (M + pps - 1) / pps;
        dim3 gw((c.Cout + 63) / 64, (9 * c.Cin + 63) / 64, splits);
        hipLaunchKernelGGL(k_conv3_wgrad, gw, dim3(256), 0, n->stream, s, (const uint16_t *)out.g, mask,
                           (const uint16_t *)in.y, c.dw, pps);
        hipLaunchKernelGGL(k_bias_grad, dim3(std::min(1024, (M + 63) / 64)), dim3(256), 0, n->stream, M, c.Cout,
                           out.C, co_off, (const uint16_t *)out.g, mask, c.db);
    }
    if (din) {
        dim3 gd((M + 63) / 64, (c.Cin + 63) / 64);
        const int MT = conv_staged(s, c.Cin, c.Cout, out.C, co_off);
        if (MT)
            conv_mm(n, s, MT, M, c.Cin, 1, (const uint16_t *)out.g, mask, (const uint16_t *)c.Wd, nullptr, in.g,
                    accum ? 1 : 0);
        else
            hipLaunchKernelGGL(k_conv3_dgrad, gd, dim3(256), 0, n->stream, s, (const uint16_t *)out.g, mask,
                               (const uint16_t *)c.Wd, in.g, accum ? 1 : 0);
    }
}

static void refresh_bf16(mcl_net *n, Conv &c)
{
    hipLaunchKernelGGL(k_weights_bf16, dim3(gblocks(c.nw)), dim3(256), 0, n->stream, c.Cout, c.Cin, (const float *)c.w,
                       c.Wf, c.Wd);
}

static Conv *convs(mcl_net *n, int i)
{
    Conv *all[] = {&n->enc1a, &n->enc1b, &n->enc2a, &n->enc2b, &n->enc3a, &n->enc3b, &n->bota,
                   &n->botb,  &n->dec3a, &n->dec3b, &n->dec2a, &n->dec2b, &n->dec1a, &n->dec1b};
    return i < 14 ? all[i] : nullptr;
}

int mcl_net_create(int H, int W, int batch, unsigned long long seed, mcl_net **out)
{
    if (!out || H % 8 || W % 8 || H < 16 || W < 16 || batch < 1) return MCL_E_ARG;
    mcl_net *n = new mcl_net();
    n->H = H;
    n->W = W;
    n->B = batch;
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        return MCL_E_DEVICE;
    }
    std::mt19937_64 rng(seed);
    int rc = 0;
    rc |= make_act(n, n->x0, H, W, 16);
    rc |= make_act(n, n->e1a, H, W, 32);
    rc |= make_act(n, n->cat1, H, W, 96);
    rc |= make_act(n, n->p1, H / 2, W / 2, 32);
    rc |= make_act(n, n->e2a, H / 2, W / 2, 64);
    rc |= make_act(n, n->cat2, H / 2, W / 2, 192);
    rc |= make_act(n, n->p2, H / 4, W / 4, 64);
    rc |= make_act(n, n->e3a, H / 4, W / 4, 128);
    rc |= make_act(n, n->cat3, H / 4, W / 4, 384);
    rc |= make_act(n, n->p3, H / 8, W / 8, 128);
    rc |= make_act(n, n->b1, H / 8, W / 8, 256);
    rc |= make_act(n, n->b2, H / 8, W / 8, 256);
    rc |= make_act(n, n->d3a, H / 4, W / 4, 128);
    rc |= make_act(n, n->d3b, H / 4, W / 4, 128);
    rc |= make_act(n, n->d2a, H / 2, W / 2, 64);
    rc |= make_act(n, n->d2b, H / 2, W / 2, 64);
    rc |= make_act(n, n->d1a, H, W, 32);
    rc |= make_act(n, n->d1b, H, W, 32);
    rc |= make_conv(n, n->enc1a, 16, 32, rng);
    rc |= make_conv(n, n->enc1b, 32, 32, rng);
    rc |= make_conv(n, n->enc2a, 32, 64, rng);
    rc |= make_conv(n, n->enc2b, 64, 64, rng);
    rc |= make_conv(n, n->enc3a, 64, 128, rng);
    rc |= make_conv(n, n->enc3b, 128, 128, rng);
    rc |= make_conv(n, n->bota, 128, 256, rng);
    rc |= make_conv(n, n->botb, 256, 256, rng);
    rc |= make_conv(n, n->dec3a, 384, 128, rng);
    rc |= make_conv(n, n->dec3b, 128, 128, rng);
    rc |= make_conv(n, n->dec2a, 192, 64, rng);
    rc |= make_conv(n, n->dec2b, 64, 64, rng);
    rc |= make_conv(n, n->dec1a, 96, 32, rng);
    rc |= make_conv(n, n->dec1b, 32, 32, rng);
    size_t M = (size_t)batch * H * W;
    rc |= nalloc(n, &n->wh, 3 * 32) | nalloc(n, &n->bh, 3) | nalloc(n, &n->dwh, 3 * 32) | nalloc(n, &n->dbh, 3);
    rc |= nalloc(n, &n->hm1, 3 * 32) | nalloc(n, &n->hm2, 3 * 32) | nalloc(n, &n->hbm1, 3) | nalloc(n, &n->hbm2, 3);
    rc |= nalloc(n, &n->y, 3 * M) | nalloc(n, &n->tgt, 3 * M) | nalloc(n, &n->xin, 2 * M) | nalloc(n, &n->loss, 4);
    if (rc) {
        mcl_net_destroy(n);
        return MCL_E_DEVICE;
    }
    std::normal_distribution<float> nd(0.f, std::sqrt(1.0f / 32.0f));
    std::vector<float> wh(96);
    for (auto &v : wh) v = nd(rng);
    if (hipMemcpy(n->wh, wh.data(), 96 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        mcl_net_destroy(n);
        return MCL_E_DEVICE;
    }
    for (int i = 0; i < 14; i++) refresh_bf16(n, *convs(n, i));
    *out = n;
    return 0;
}

void mcl_net_destroy(mcl_net *n)
{
    if (!n) return;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (void *p : n->owned) (void)hipFree(p);
    for (float *p : n->ds_x) (void)hipFree(p);
    for (float *p : n->ds_t) (void)hipFree(p);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

const char *mcl_net_last_error(const mcl_net *n) { return n ? n->err.c_str() : "null network"; }

// x0 (bf16, 16 channels, 2 used) from FP32 [B][H][W][2] on the device
__global__ void k_pack_input(size_t M, const float *__restrict__ x, uint16_t *__restrict__ x0)
{
    for (size_t m = blockIdx.x * (size_t)blockDim.x + threadIdx.x; m < M; m += (size_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < 16; c++) {
            float v = c < 2 ? x[m * 2 + c] : 0.f;
            unsigned u = __float_as_uint(v);
            u += 0x7fffu + ((u >> 16) & 1u);
            x0[m * 16 + c] = (uint16_t)(u >> 16);
        }
    }
}

static void forward(mcl_net *n)
{
    const int B = n->B;
    size_t M = (size_t)B * n->H * n->W;
    hipLaunchKernelGGL(k_pack_input, dim3(gblocks(M)), dim3(256), 0, n->stream, M, (const float *)n->xin, n->x0.y);
    conv_fwd(n, n->enc1a, n->x0, 0, n->e1a, 0, true);
    conv_fwd(n, n->enc1b, n->e1a, 0, n->cat1, 64, true);
    hipLaunchKernelGGL(k_pool_fwd, dim3(gblocks(M / 4 * 32)), dim3(256), 0, n->stream, B, n->H, n->W, 32,
                       (const uint16_t *)n->cat1.y, 96, 64, n->p1.y, 32, 0);
    conv_fwd(n, n->enc2a, n->p1, 0, n->e2a, 0, true);
    conv_fwd(n, n->enc2b, n->e2a, 0, n->cat2, 128, true);
    hipLaunchKernelGGL(k_pool_fwd, dim3(gblocks(M / 16 * 64)), dim3(256), 0, n->stream, B, n->H / 2, n->W / 2, 64,
                       (const uint16_t *)n->cat2.y, 192, 128, n->p2.y, 64, 0);
    conv_fwd(n, n->enc3a, n->p2, 0, n->e3a, 0, true);
    conv_fwd(n, n->enc3b, n->e3a, 0, n->cat3, 256, true);
    hipLaunchKernelGGL(k_pool_fwd, dim3(gblocks(M / 64 * 128)), dim3(256), 0, n->stream, B, n->H / 4, n->W / 4, 128,
                       (const uint16_t *)n->cat3.y, 384, 256, n->p3.y, 128, 0);
    conv_fwd(n, n->bota, n->p3, 0, n->b1, 0, true);
    conv_fwd(n, n->botb, n->b1, 0, n->b2, 0, true);
    hipLaunchKernelGGL(k_up_fwd, dim3(gblocks(M / 16 * 256)), dim3(256), 0, n->stream, B, n->H / 8, n->W / 8, 256,
                       (const uint16_t *)n->b2.y, 256, 0, n->cat3.y, 384, 0);
    conv_fwd(n, n->dec3a, n->cat3, 0, n->d3a, 0, true);
    conv_fwd(n, n->dec3b, n->d3a, 0, n->d3b, 0, true);
    hipLaunchKernelGGL(k_up_fwd, dim3(gblocks(M / 4 * 128)), dim3(256), 0, n->stream, B, n->H / 4, n->W / 4, 128,
                       (const uint16_t *)n->d3b.y, 128, 0, n->cat2.y, 192, 0);
    conv_fwd(n, n->dec2a, n->cat2, 0, n->d2a, 0, true);
    conv_fwd(n, n->dec2b, n->d2a, 0, n->d2b, 0, true);
    hipLaunchKernelGGL(k_up_fwd, dim3(gblocks(M * 64)), dim3(256), 0, n->stream, B, n->H / 2, n->W / 2, 64,
                       (const uint16_t *)n->d2b.y, 64, 0, n->cat1.y, 96, 0);
    conv_fwd(n, n->dec1a, n->cat1, 0, n->d1a, 0, true);
    conv_fwd(n, n->dec1b, n->d1a, 0, n->d1b, 0, true);
    hipLaunchKernelGGL(k_head_fwd, dim3(gblocks(M)), dim3(256), 0, n->stream, (int)M, (const uint16_t *)n->d1b.y,
                       32, (const float *)n->wh, (const float *)n->bh, n->y);
}

static void zero_grads(mcl_net *n)
{
    for (int i = 0; i < 14; i++) {
        Conv &c = *convs(n, i);
        (void)hipMemsetAsync(c.dw, 0, c.nw * sizeof(float), n->stream);
        (void)hipMemsetAsync(c.db, 0, c.Cout * sizeof(float), n->stream);
    }
    (void)hipMemsetAsync(n->dwh, 0, 96 * sizeof(float), n->stream);
    (void)hipMemsetAsync(n->dbh, 0, 3 * sizeof(float), n->stream);
    (void)hipMemsetAsync(n->loss, 0, sizeof(float), n->stream);
}

static void backward(mcl_net *n)
{
    const int B = n->B;
    const int M = B * n->H * n->W;
    float scale = 2.0f / (3.0f * (float)M);
    hipLaunchKernelGGL(k_head_bwd, dim3(std::min(1024, (M + 255) / 256)), dim3(256), 0, n->stream, M,
                       (const uint16_t *)n->d1b.y, 32, (const float *)n->wh, (const float *)n->y,
                       (const float *)n->tgt, scale, n->d1b.g, n->dwh, n->dbh, n->loss);
    conv_bwd(n, n->dec1b, n->d1a, 0, n->d1b, 0, true, false);
    conv_bwd(n, n->dec1a, n->cat1, 0, n->d1a, 0, true, false);
    hipLaunchKernelGGL(k_up_bwd, dim3(gblocks((size_t)M / 4 * 64)), dim3(256), 0, n->stream, B, n->H / 2, n->W / 2, 64,
                       (const uint16_t *)n->cat1.g, 96, 0, n->d2b.g, 64, 0);
    conv_bwd(n, n->dec2b, n->d2a, 0, n->d2b, 0, true, false);
    conv_bwd(n, n->dec2a, n->cat2, 0, n->d2a, 0, true, false);
    hipLaunchKernelGGL(k_up_bwd, dim3(gblocks((size_t)M / 16 * 128)), dim3(256), 0, n->stream, B, n->H / 4, n->W / 4,
                       128, (const uint16_t *)n->cat2.g, 192, 0, n->d3b.g, 128, 0);
    conv_bwd(n, n->dec3b, n->d3a, 0, n->d3b, 0, true, false);
    conv_bwd(n, n->dec3a, n->cat3, 0, n->d3a, 0, true, false);
    hipLaunchKernelGGL(k_up_bwd, dim3(gblocks((size_t)M / 64 * 256)), dim3(256), 0, n->stream, B, n->H / 8, n->W / 8,
                       256, (const uint16_t *)n->cat3.g, 384, 0, n->b2.g, 256, 0);
    conv_bwd(n, n->botb, n->b1, 0, n->b2, 0, true, false);
    conv_bwd(n, n->bota, n->p3, 0, n->b1, 0, true, false);
    hipLaunchKernelGGL(k_pool_bwd, dim3(gblocks((size_t)M / 16 * 128)), dim3(256), 0, n->stream, B, n->H / 4, n->W / 4,
                       128, (const uint16_t *)n->p3.g, 128, 0, n->cat3.g, 384, 256, 1);
    conv_bwd(n, n->enc3b, n->e3a, 0, n->cat3, 256, true, false);
    conv_bwd(n, n->enc3a, n->p2, 0, n->e3a, 0, true, false);
    hipLaunchKernelGGL(k_pool_bwd, dim3(gblocks((size_t)M / 4 * 64)), dim3(256), 0, n->stream, B, n->H / 2, n->W / 2,
                       64, (const uint16_t *)n->p2.g, 64, 0, n->cat2.g, 192, 128, 1);
    conv_bwd(n, n->enc2b, n->e2a, 0, n->cat2, 128, true, false);
    conv_bwd(n, n->enc2a, n->p1, 0, n->e2a, 0, true, false);
    hipLaunchKernelGGL(k_pool_bwd, dim3(gblocks((size_t)M * 32)), dim3(256), 0, n->stream, B, n->H, n->W, 32,
                       (const uint16_t *)n->p1.g, 32, 0, n->cat1.g, 96, 64, 1);
    conv_bwd(n, n->enc1b, n->e1a, 0, n->cat1, 64, true, false);
    conv_bwd(n, n->enc1a, n->x0, 0, n->e1a, 0, false, false);
}

static void adam_step(mcl_net *n, float lr)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    n->adam_t++;
    float c1 = 1.0f - std::pow(b1, (float)n->adam_t), c2 = 1.0f - std::pow(b2, (float)n->adam_t);
    for (int i = 0; i < 14; i++) {
        Conv &c = *convs(n, i);
        hipLaunchKernelGGL(k_adam, dim3(gblocks(c.nw)), dim3(256), 0, n->stream, (int)c.nw, c.w, (const float *)c.dw,
                           c.m1, c.m2, lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, n->stream, c.Cout, c.b, (const float *)c.db, c.bm1, c.bm2,
                           lr, b1, b2, eps, c1, c2);
        refresh_bf16(n, c);
    }
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, n->stream, 96, n->wh, (const float *)n->dwh, n->hm1, n->hm2, lr,
                       b1, b2, eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, n->stream, 3, n->bh, (const float *)n->dbh, n->hbm1, n->hbm2, lr,
                       b1, b2, eps, c1, c2);
}

int mcl_net_forward(mcl_net *n, const float *x_host, float *y_host)
{
    if (!n || !x_host || !y_host) return MCL_E_ARG;
    size_t M = (size_t)n->B * n->H * n->W;
    NCHK(hipMemcpyAsync(n->xin, x_host, 2 * M * sizeof(float), hipMemcpyHostToDevice, n->stream));
    forward(n);
    NCHK(hipGetLastError());
    NCHK(hipMemcpyAsync(y_host, n->y, 3 * M * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    NCHK(hipStreamSynchronize(n->stream));
    return 0;
}

int mcl_net_train_step(mcl_net *n, const float *x_host, const float *t_host, float lr, float *loss)
{
    if (!n || !x_host || !t_host) return MCL_E_ARG;
    size_t M = (size_t)n->B * n->H * n->W;
    NCHK(hipMemcpyAsync(n->xin, x_host, 2 * M * sizeof(float), hipMemcpyHostToDevice, n->stream));
    NCHK(hipMemcpyAsync(n->tgt, t_host, 3 * M * sizeof(float), hipMemcpyHostToDevice, n->stream));
    zero_grads(n);
    forward(n);
    backward(n);
    if (lr > 0.f) adam_step(n, lr);
    NCHK(hipGetLastError());
    float l = 0.f;
    NCHK(hipMemcpyAsync(&l, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    NCHK(hipStreamSynchronize(n->stream));
    if (loss) *loss = l / (3.0f * (float)M);
    return 0;
}

// Parameters in a flat FP32 vector: for each conv in order w[Cout][3][3][Cin]
// then b[Cout]; then the head w[3][32], b[3].
long long mcl_net_num_params(const mcl_net *n)
{
    long long s = 0;
    for (int i = 0; i < 14; i++) {
        Conv &c = *convs(const_cast<mcl_net *>(n), i);
        s += (long long)c.nw + c.Cout;
    }
    return s + 96 + 3;
}

static int copy_params(mcl_net *n, float *host, bool get, bool grads)
{
    size_t off = 0;
    for (int i = 0; i < 14; i++) {
        Conv &c = *convs(n, i);
        float *w = grads ? c.dw : c.w, *b = grads ? c.db : c.b;
        if (get) {
            NCHK(hipMemcpy(host + off, w, c.nw * sizeof(float), hipMemcpyDeviceToHost));
            NCHK(hipMemcpy(host + off + c.nw, b, c.Cout * sizeof(float), hipMemcpyDeviceToHost));
        } else {
            NCHK(hipMemcpy(w, host + off, c.nw * sizeof(float), hipMemcpyHostToDevice));
            NCHK(hipMemcpy(b, host + off + c.nw, c.Cout * sizeof(float), hipMemcpyHostToDevice));
        }
        off += c.nw + c.Cout;
    }
    float *wh = grads ? n->dwh : n->wh, *bh = grads ? n->dbh : n->bh;
    if (get) {
        NCHK(hipMemcpy(host + off, wh, 96 * sizeof(float), hipMemcpyDeviceToHost));
        NCHK(hipMemcpy(host + off + 96, bh, 3 * sizeof(float), hipMemcpyDeviceToHost));
    } else {
        NCHK(hipMemcpy(wh, host + off, 96 * sizeof(float), hipMemcpyHostToDevice));
        NCHK(hipMemcpy(bh, host + off + 96, 3 * sizeof(float), hipMemcpyHostToDevice));
        for (int i = 0; i < 14; i++) refresh_bf16(n, *convs(n, i));
        NCHK(hipStreamSynchronize(n->stream));
    }
    return 0;
}

int mcl_net_get_params(mcl_net *n, float *host) { return n && host ? copy_params(n, host, true, false) : MCL_E_ARG; }
int mcl_net_set_params(mcl_net *n, const float *host)
{
    return n && host ? copy_params(n, const_cast<float *>(host), false, false) : MCL_E_ARG;
}
int mcl_net_get_grads(mcl_net *n, float *host) { return n && host ? copy_params(n, host, true, true) : MCL_E_ARG; }

// ------------------------------------------------- internal interface, data
float *mcl_net_in(mcl_net *n) { return n->xin; }
float *mcl_net_out(mcl_net *n) { return n->y; }
float *mcl_net_tgt(mcl_net *n) { return n->tgt; }
hipStream_t mcl_net_stream(mcl_net *n) { return n->stream; }
float mcl_net_out_scale(const mcl_net *n, int ch) { return (ch >= 0 && ch < 3) ? n->out_scale[ch] : 0.f; }

int mcl_net_dims(const mcl_net *n, int *H, int *W, int *B)
{
    if (!n) return MCL_E_ARG;
    *H = n->H;
    *W = n->W;
    *B = n->B;
    return 0;
}

int mcl_net_forward_device(mcl_net *n)
{
    forward(n);
    NCHK(hipGetLastError());
    return 0;
}

// sum of squares of each of the 3 channels of npix pixels [npix][3] (order of
// accumulation irrelevant: it only sets the output scales)
__global__ void k_sumsq3(size_t npix, const float *__restrict__ x, double *__restrict__ out)
{
    double s[3] = {0.0, 0.0, 0.0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; c++) s[c] += (double)x[3 * i + c] * x[3 * i + c];
    for (int c = 0; c < 3; c++) {
        for (int o = 32; o > 0; o >>= 1) s[c] += __shfl_xor(s[c], o, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(out + c, s[c]);
    }
}

__global__ void k_scale3(size_t npix, float a0, float a1, float a2, float *__restrict__ x)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x) {
        x[3 * i] *= a0;
        x[3 * i + 1] *= a1;
        x[3 * i + 2] *= a2;
    }
}

// Mirror one sample in x (column ix <-> W-1-ix) in place: inputs [H][W][2],
// targets [H][W][3] with the horizontal velocity negated (the box, its
// boundary conditions and gravity are symmetric under x -> L - x).
__global__ void k_flip_x(int H, int W, float *__restrict__ x, float *__restrict__ t)
{
    const int half = W / 2;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < H * half; q += gridDim.x * blockDim.x) {
        const int iz = q / half, ix = q - iz * half;
        const int a = iz * W + ix, b = iz * W + (W - 1 - ix);
        for (int c = 0; c < 2; c++) {
            float u = x[2 * a + c], v = x[2 * b + c];
            x[2 * a + c] = v;
            x[2 * b + c] = u;
        }
        for (int c = 0; c < 3; c++) {
            float u = t[3 * a + c], v = t[3 * b + c];
            float sg = c == 0 ? -1.f : 1.f;
            t[3 * a + c] = sg * v;
            t[3 * b + c] = sg * u;
        }
    }
}

int mcl_net_add_snapshot(mcl_net *n, const float *x_dev, const float *t_dev)
{
    size_t P = (size_t)n->H * n->W;
    float *x = nullptr, *t = nullptr;
    NCHK(hipMalloc(&x, 2 * P * sizeof(float)));
    NCHK(hipMalloc(&t, 3 * P * sizeof(float)));
    NCHK(hipMemcpyAsync(x, x_dev, 2 * P * sizeof(float), hipMemcpyDeviceToDevice, n->stream));
    NCHK(hipMemcpyAsync(t, t_dev, 3 * P * sizeof(float), hipMemcpyDeviceToDevice, n->stream));
    NCHK(hipStreamSynchronize(n->stream));
    n->ds_x.push_back(x);
    n->ds_t.push_back(t);
    return 0;
}

long long mcl_net_dataset_size(const mcl_net *n) { return n ? (long long)n->ds_x.size() : 0; }

int mcl_net_train(mcl_net *n, int epochs, float lr, unsigned long long seed, float *losses)
{
    if (!n || epochs < 0) return MCL_E_ARG;
    const size_t N = n->ds_x.size();
    if (N < (size_t)n->B) {
        n->err = "dataset smaller than one batch";
        return MCL_E_STATE;
    }
    const size_t P = (size_t)n->H * n->W;
    // output scales: RMS of each target channel over the dataset (set once)
    if (!n->scaled) {
        double *acc = nullptr;
        NCHK(hipMalloc(&acc, 3 * sizeof(double)));
        NCHK(hipMemsetAsync(acc, 0, 3 * sizeof(double), n->stream));
        for (size_t i = 0; i < N; i++)
            hipLaunchKernelGGL(k_sumsq3, dim3(gblocks(P)), dim3(256), 0, n->stream, P, (const float *)n->ds_t[i], acc);
        double sq[3] = {0.0, 0.0, 0.0};
        NCHK(hipMemcpyAsync(sq, acc, 3 * sizeof(double), hipMemcpyDeviceToHost, n->stream));
        NCHK(hipStreamSynchronize(n->stream));
        (void)hipFree(acc);
        for (int c = 0; c < 3; c++) {
            double rms = std::sqrt(sq[c] / ((double)P * N));
            n->out_scale[c] = rms > 0 ? (float)rms : 1.0f;
        }
        n->scaled = true;
    }
    std::mt19937_64 rng(seed);
    std::vector<size_t> order(N);
    for (size_t i = 0; i < N; i++) order[i] = i;
    const size_t nb = N / n->B;
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(order.begin(), order.end(), rng);
        double sum = 0.0;
        for (size_t b = 0; b < nb; b++) {
            for (int k = 0; k < n->B; k++) {
                size_t s = order[b * n->B + k];
                NCHK(hipMemcpyAsync(n->xin + k * 2 * P, n->ds_x[s], 2 * P * sizeof(float), hipMemcpyDeviceToDevice,
                                    n->stream));
                NCHK(hipMemcpyAsync(n->tgt + k * 3 * P, n->ds_t[s], 3 * P * sizeof(float), hipMemcpyDeviceToDevice,
                                    n->stream));
                if (rng() & 1)
                    hipLaunchKernelGGL(k_flip_x, dim3(gblocks(P / 2)), dim3(256), 0, n->stream, n->H, n->W,
                                       n->xin + k * 2 * P, n->tgt + k * 3 * P);
            }
            hipLaunchKernelGGL(k_scale3, dim3(gblocks(P * n->B)), dim3(256), 0, n->stream, P * n->B,
                               1.0f / n->out_scale[0], 1.0f / n->out_scale[1], 1.0f / n->out_scale[2], n->tgt);
            zero_grads(n);
            forward(n);
            backward(n);
            adam_step(n, lr);
            NCHK(hipGetLastError());
            float l = 0.f;
            NCHK(hipMemcpyAsync(&l, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
            NCHK(hipStreamSynchronize(n->stream));
            sum += l / (3.0 * P * n->B);
        }
        if (losses) losses[ep] = (float)(sum / nb);
    }
    return 0;
}

double mcl_net_get_scale(const mcl_net *n, int channel)
{
    return (n && channel >= 0 && channel < 3) ? n->out_scale[channel] : 0.0;
}

int mcl_net_set_scale(mcl_net *n, int channel, double scale)
{
    if (!n || channel < 0 || channel > 2 || !(scale > 0)) return MCL_E_ARG;
    n->out_scale[channel] = (float)scale;
    n->scaled = true;
    return 0;
}
