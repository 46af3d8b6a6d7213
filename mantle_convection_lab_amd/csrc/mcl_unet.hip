// Learned initial guess: the U-Net's forward, backward and Adam kernels (gfx950).
//
// Activations are NHWC bf16 (stored as 16-bit words), weights are FP32 master
// copies plus bf16 copies in the two layouts the products need:
//   Wf[co][tap][ci]   forward      out[m,co]  = sum_{tap,ci} in[m+off(tap),ci] Wf[co][tap][ci]
//   Wd[ci][tap][co]   data grad    din[m,ci]  = sum_{tap,co} g[m-off(tap),co] Wd[ci][tap][co]
// and the weight gradient dW[co][tap][ci] = sum_m g[m,co] in[m+off(tap),ci].
// All three are implicit GEMMs on v_mfma_f32_16x16x16_bf16 with FP32
// accumulation.  A workgroup (4 waves) computes a 64 x 64 output tile; each
// wave a 32 x 32 block of four 16 x 16 MFMA tiles.  Operand lane layout of the
// instruction (K = 16): A[i][k] in lane i + 16*(k/4), element k%4; B[k][j] in
// lane j + 16*(k/4), element k%4; D[i][j] in lane j + 16*(i/4), element i%4.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((unsigned)h) << 16); }

// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct ConvShape {
    int B, H, W;          // batch and spatial size (same for input and output: 3x3, stride 1, pad 1)
    int Cin, Cout;        // channels of this convolution
    int cs_in, ci_off;    // channel stride / offset of the input buffer
    int cs_out, co_off;   // channel stride / offset of the output buffer
};

// Load 4 consecutive bf16 channels (8-byte aligned) or zeros.
__device__ __forceinline__ s4 load4(const uint16_t *p, bool valid)
{
    if (!valid) return s4{0, 0, 0, 0};
    uint2 v = *(const uint2 *)p;
    s4 r;
    r[0] = (short)(v.x & 0xffff);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xffff);
    r[3] = (short)(v.y >> 16);
    return r;
}

// ------------------------------------------------------------------ forward
// out[m, co] = relu?(bias[co] + sum in[m + off, ci] Wf[co][tap][ci]); requires
// Cin % 16 == 0 and Cout % 32 == 0, channel offsets/strides multiples of 4.
__global__ void __launch_bounds__(256) k_conv3_fwd(ConvShape s, const uint16_t *__restrict__ in,
                                                   const uint16_t *__restrict__ Wf, const float *__restrict__ bias,
                                                   uint16_t *__restrict__ out, int relu)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int M = s.B * s.H * s.W;
    const int m0 = blockIdx.x * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.y * 64 + (wave & 1) * 32;
    if (n0 >= s.Cout) return;  // whole wave idle (Cout = 32)
    f4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f4{0, 0, 0, 0};
    // this lane's A rows (pixels) for the two 16-row tiles
    int ph[2], pw[2], pn[2];
    bool pv[2];
    for (int a = 0; a < 2; a++) {
        int m = m0 + a * 16 + (lane & 15);
        pv[a] = m < M;
        int mm = pv[a] ? m : 0;
        pn[a] = mm / (s.H * s.W);
        int r = mm - pn[a] * s.H * s.W;
        ph[a] = r / s.W;
        pw[a] = r - ph[a] * s.W;
    }
    const int kq = (lane >> 4) * 4;  // this lane's 4 k values within a 16-chunk
    for (int tap = 0; tap < 9; tap++) {
        const int dh = tap / 3 - 1, dw = tap % 3 - 1;
        const uint16_t *ap[2];
        bool av[2];
        for (int a = 0; a < 2; a++) {
            int h = ph[a] + dh, w = pw[a] + dw;
            av[a] = pv[a] && h >= 0 && h < s.H && w >= 0 && w < s.W;
            ap[a] = in + ((size_t)((pn[a] * s.H + (av[a] ? h : 0)) * s.W + (av[a] ? w : 0))) * s.cs_in + s.ci_off;
        }
        for (int c0 = 0; c0 < s.Cin; c0 += 16) {
            s4 A[2], Bv[2];
            for (int a = 0; a < 2; a++) A[a] = load4(ap[a] + c0 + kq, av[a]);
            for (int b = 0; b < 2; b++) {
                int co = n0 + b * 16 + (lane & 15);
                Bv[b] = load4(Wf + ((size_t)co * 9 + tap) * s.Cin + c0 + kq, co < s.Cout);
            }
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[a], Bv[b], acc[a][b], 0, 0, 0);
        }
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            int co = n0 + b * 16 + (lane & 15);
            if (co >= s.Cout) continue;
            float bi = bias ? bias[co] : 0.f;
            for (int e = 0; e < 4; e++) {
                int m = m0 + a * 16 + (lane >> 4) * 4 + e;
                if (m >= M) continue;
                float v = acc[a][b][e] + bi;
                if (relu) v = fmaxf(v, 0.f);
                out[(size_t)m * s.cs_out + s.co_off + co] = f2bf(v);
            }
        }
}

// ------------------------------------------- LDS-staged forward / data grad
// Implicit-GEMM 3x3 convolution on v_mfma_f32_16x16x32_bf16 (gfx950) with the
// operands staged in LDS:
//   forward (dgrad = 0): Y[m, n] = relu?(bias[n] + sum_{tap,c} X[m + off(tap), c] Wt[n][tap][c])
//     with X = in, Wt = Wf, Y = out;
//   data grad (dgrad = 1): Y[m, n] (+)= sum_{tap,c} X[m - off(tap), c] Wt[n][tap][c]
//     with X = dout (times y > 0 when mask != NULL), Wt = Wd, Y = din.
// A workgroup (4 waves, each a 32 x 32 block of 2 x 2 MFMA tiles) computes MT
// pixels x NT channels: 64 x 64, or 128 x 32 for layers of at most 32 output
// channels (every wave busy).  Its MT pixels are a TH x TW rectangle of one
// image (TW = min(W, MT); the host checks that the shape allows it).  For
// every 32-channel chunk of the summed channels it stages the (TH + 2) x
// (TW + 2) patch of X and the chunk's weights of its NT output channels for
// all nine taps in LDS (rows padded by 16 bytes), then runs the nine taps, so
// a staged value feeds nine products instead of being fetched for each.
// Operand layout of the instruction (K = 32): A[i][k] in lane i + 16 (k / 8),
// element k % 8; B[k][j] in lane j + 16 (k / 8), element k % 8; D[i][j] in
// lane j + 16 (i / 4), element i % 4.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int CV_KC = 32;            // summed channels per stage
constexpr int CV_ROW = CV_KC + 8;    // bf16 words per staged pixel or weight row

__device__ __forceinline__ uint4 mask_bf16x8(uint4 g, uint4 y)
{
    unsigned *gv = (unsigned *)&g;
    const unsigned *yv = (const unsigned *)&y;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        unsigned lo = gv[q] & 0xffffu, hi = gv[q] >> 16;
        if (bf2f((uint16_t)(yv[q] & 0xffffu)) <= 0.f) lo = 0;
        if (bf2f((uint16_t)(yv[q] >> 16)) <= 0.f) hi = 0;
        gv[q] = lo | (hi << 16);
    }
    return g;
}

template <int MT, int NT>
__device__ __forceinline__ void conv3_mm(const ConvShape &s, int dgrad, const uint16_t *__restrict__ X,
                                         const uint16_t *__restrict__ mask, const uint16_t *__restrict__ Wt,
                                         const float *__restrict__ bias, uint16_t *__restrict__ Y, int flag)
{
    static_assert((MT / 32) * (NT / 32) == 4, "four waves of 32 x 32");
    constexpr int NW = NT / 32;
    __shared__ __attribute__((aligned(16))) uint16_t xs[3 * (MT + 2) * CV_ROW];
    __shared__ __attribute__((aligned(16))) uint16_t ws[9 * NT * CV_ROW];
    const int Kc = dgrad ? s.Cout : s.Cin, N = dgrad ? s.Cin : s.Cout;
    const int xst = dgrad ? s.cs_out : s.cs_in, xof = dgrad ? s.co_off : s.ci_off;
    const int yst = dgrad ? s.cs_in : s.cs_out, yof = dgrad ? s.ci_off : s.co_off;
    const int sg = dgrad ? -1 : 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / NW, wn = wave % NW;
    const int TW = s.W < MT ? s.W : MT, TH = MT / TW, PW = TW + 2, np = (TH + 2) * PW;
    const int m0 = blockIdx.x * MT, n0 = blockIdx.y * NT;
    const int img = m0 / (s.H * s.W), r0 = m0 - img * s.H * s.W;
    const int h0 = r0 / s.W, w0 = r0 - h0 * s.W;
    const bool active = n0 + wn * 32 < N;
    f4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f4{0, 0, 0, 0};
    // this lane's pixels (patch coordinates of the tap-free position)
    int py[2], px[2];
    for (int a = 0; a < 2; a++) {
        const int ml = wm * 32 + a * 16 + (lane & 15);
        py[a] = ml / TW + 1;
        px[a] = ml - (py[a] - 1) * TW + 1;
    }
    const int kq = 8 * (lane >> 4);
    for (int c0 = 0; c0 < Kc; c0 += CV_KC) {
        __syncthreads();   // the previous chunk's products are done with the LDS
        for (int t = threadIdx.x; t < np * 4; t += 256) {
            const int p = t >> 2, q = t & 3, c = c0 + 8 * q;
            const int hh = h0 - 1 + p / PW, ww = w0 - 1 + p % PW;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (c < Kc && hh >= 0 && hh < s.H && ww >= 0 && ww < s.W) {
                const size_t idx = ((size_t)(img * s.H + hh) * s.W + ww) * xst + xof + c;
                v = *(const uint4 *)(X + idx);
                if (mask) v = mask_bf16x8(v, *(const uint4 *)(mask + idx));
            }
            *(uint4 *)(xs + p * CV_ROW + 8 * q) = v;
        }
        for (int t = threadIdx.x; t < 9 * NT * 4; t += 256) {
            const int q = t & 3, row = t >> 2, nl = row % NT, tap = row / NT, c = c0 + 8 * q;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (c < Kc && n0 + nl < N) v = *(const uint4 *)(Wt + ((size_t)(n0 + nl) * 9 + tap) * Kc + c);
            *(uint4 *)(ws + (tap * NT + nl) * CV_ROW + 8 * q) = v;
        }
        __syncthreads();
        if (active) {
            for (int tap = 0; tap < 9; tap++) {
                const int dh = sg * (tap / 3 - 1), dw = sg * (tap % 3 - 1);
                bf8 A[2], B[2];
                for (int a = 0; a < 2; a++)
                    A[a] = *(const bf8 *)(xs + ((py[a] + dh) * PW + px[a] + dw) * CV_ROW + kq);
                for (int b = 0; b < 2; b++)
                    B[b] = *(const bf8 *)(ws + (tap * NT + wn * 32 + b * 16 + (lane & 15)) * CV_ROW + kq);
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a], B[b], acc[a][b], 0, 0, 0);
            }
        }
    }
    if (!active) return;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const int n = n0 + wn * 32 + b * 16 + (lane & 15);
            if (n >= N) continue;
            const float bi = (!dgrad && bias) ? bias[n] : 0.f;
            for (int e = 0; e < 4; e++) {
                const size_t idx = (size_t)(m0 + wm * 32 + a * 16 + (lane >> 4) * 4 + e) * yst + yof + n;
                float v = acc[a][b][e] + bi;
                if (!dgrad && flag) v = fmaxf(v, 0.f);
                if (dgrad && flag) v += bf2f(Y[idx]);
                Y[idx] = f2bf(v);
            }
        }
}

__global__ void __launch_bounds__(256) k_conv3_mm(ConvShape s, int dgrad, const uint16_t *__restrict__ X,
                                                  const uint16_t *__restrict__ mask, const uint16_t *__restrict__ Wt,
                                                  const float *__restrict__ bias, uint16_t *__restrict__ Y, int flag)
{
    conv3_mm<64, 64>(s, dgrad, X, mask, Wt, bias, Y, flag);
}

__global__ void __launch_bounds__(256) k_conv3_mm_n32(ConvShape s, int dgrad, const uint16_t *__restrict__ X,
                                                      const uint16_t *__restrict__ mask,
                                                      const uint16_t *__restrict__ Wt, const float *__restrict__ bias,
                                                      uint16_t *__restrict__ Y, int flag)
{
    conv3_mm<128, 32>(s, dgrad, X, mask, Wt, bias, Y, flag);
}

// ---------------------------------------------------------------- data grad
// din[m, ci] = sum_{tap,co} g[m - off(tap), co] Wd[ci][tap][co], g = dout * (y > 0)
// when relu_y != NULL.  Output channels ci of the input buffer (stride/offset of
// the forward input); accum adds into din.  Requires Cout % 16 == 0, Cin % 32 == 0.
__global__ void __launch_bounds__(256) k_conv3_dgrad(ConvShape s, const uint16_t *__restrict__ dout,
                                                     const uint16_t *__restrict__ relu_y,
                                                     const uint16_t *__restrict__ Wd, uint16_t *__restrict__ din,
                                                     int accum)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int M = s.B * s.H * s.W;
    const int m0 = blockIdx.x * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.y * 64 + (wave & 1) * 32;
    if (n0 >= s.Cin) return;
    f4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f4{0, 0, 0, 0};
    int ph[2], pw[2], pn[2];
    bool pv[2];
    for (int a = 0; a < 2; a++) {
        int m = m0 + a * 16 + (lane & 15);
        pv[a] = m < M;
        int mm = pv[a] ? m : 0;
        pn[a] = mm / (s.H * s.W);
        int r = mm - pn[a] * s.H * s.W;
        ph[a] = r / s.W;
        pw[a] = r - ph[a] * s.W;
    }
    const int kq = (lane >> 4) * 4;
    for (int tap = 0; tap < 9; tap++) {
        const int dh = tap / 3 - 1, dw = tap % 3 - 1;
        size_t off[2];
        bool av[2];
        for (int a = 0; a < 2; a++) {
            int h = ph[a] - dh, w = pw[a] - dw;
            av[a] = pv[a] && h >= 0 && h < s.H && w >= 0 && w < s.W;
            off[a] = ((size_t)((pn[a] * s.H + (av[a] ? h : 0)) * s.W + (av[a] ? w : 0))) * s.cs_out + s.co_off;
        }
        for (int c0 = 0; c0 < s.Cout; c0 += 16) {
            s4 A[2], Bv[2];
            for (int a = 0; a < 2; a++) {
                s4 g = load4(dout + off[a] + c0 + kq, av[a]);
                if (relu_y && av[a]) {
                    s4 y = load4(relu_y + off[a] + c0 + kq, true);
                    for (int e = 0; e < 4; e++)
                        if (bf2f((uint16_t)y[e]) <= 0.f) g[e] = 0;
                }
                A[a] = g;
            }
            for (int b = 0; b < 2; b++) {
                int ci = n0 + b * 16 + (lane & 15);
                Bv[b] = load4(Wd + ((size_t)ci * 9 + tap) * s.Cout + c0 + kq, ci < s.Cin);
            }
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[a], Bv[b], acc[a][b], 0, 0, 0);
        }
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            int ci = n0 + b * 16 + (lane & 15);
            if (ci >= s.Cin) continue;
            for (int e = 0; e < 4; e++) {
                int m = m0 + a * 16 + (lane >> 4) * 4 + e;
                if (m >= M) continue;
                size_t idx = (size_t)m * s.cs_in + s.ci_off + ci;
                float v = acc[a][b][e];
                if (accum) v += bf2f(din[idx]);
                din[idx] = f2bf(v);
            }
        }
}

// -------------------------------------------------------------- weight grad
// dW[co][tap][ci] += sum over a slice of pixels of g[m,co] in[m+off(tap),ci]
// (FP32 atomics; the pixel range is split over blockIdx.z).  GEMM rows co,
// columns (tap, ci) with ci fastest, K = pixels.
__global__ void __launch_bounds__(256) k_conv3_wgrad(ConvShape s, const uint16_t *__restrict__ dout,
                                                     const uint16_t *__restrict__ relu_y,
                                                     const uint16_t *__restrict__ in, float *__restrict__ dW,
                                                     int pix_per_split)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int M = s.B * s.H * s.W;
    const int NC = 9 * s.Cin;
    const int r0 = blockIdx.x * 64 + (wave >> 1) * 32;   // co
    const int c0 = blockIdx.y * 64 + (wave & 1) * 32;    // (tap, ci)
    if (r0 >= s.Cout || c0 >= NC) return;
    const int mbeg = blockIdx.z * pix_per_split;
    const int mend = min(M, mbeg + pix_per_split);
    f4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f4{0, 0, 0, 0};
    const int kq = (lane >> 4) * 4;
    int col_tap[2], col_ci[2];
    bool col_ok[2];
    for (int b = 0; b < 2; b++) {
        int c = c0 + b * 16 + (lane & 15);
        col_ok[b] = c < NC;
        int cc = col_ok[b] ? c : 0;
        col_tap[b] = cc / s.Cin;
        col_ci[b] = cc - col_tap[b] * s.Cin;
    }
    for (int mk = mbeg; mk < mend; mk += 16) {
        s4 A[2], Bv[2];
        for (int a = 0; a < 2; a++) {
            int co = r0 + a * 16 + (lane & 15);
            for (int e = 0; e < 4; e++) {
                int m = mk + kq + e;
                short v = 0;
                if (m < mend && co < s.Cout) {
                    size_t idx = (size_t)m * s.cs_out + s.co_off + co;
                    uint16_t g = dout[idx];
                    if (relu_y && bf2f(relu_y[idx]) <= 0.f) g = 0;
                    v = (short)g;
                }
                A[a][e] = v;
            }
        }
        for (int b = 0; b < 2; b++) {
            const int dh = col_tap[b] / 3 - 1, dw = col_tap[b] % 3 - 1;
            for (int e = 0; e < 4; e++) {
                int m = mk + kq + e;
                short v = 0;
                if (m < mend && col_ok[b]) {
                    int n = m / (s.H * s.W);
                    int r = m - n * s.H * s.W;
                    int h = r / s.W + dh, w = r % s.W + dw;
                    if (h >= 0 && h < s.H && w >= 0 && w < s.W)
                        v = (short)in[((size_t)((n * s.H + h) * s.W + w)) * s.cs_in + s.ci_off + col_ci[b]];
                }
                Bv[b][e] = v;
            }
        }
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[a], Bv[b], acc[a][b], 0, 0, 0);
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            int c = c0 + b * 16 + (lane & 15);
            if (c >= NC) continue;
            for (int e = 0; e < 4; e++) {
                int co = r0 + a * 16 + (lane >> 4) * 4 + e;
                if (co >= s.Cout) continue;
                atomicAdd(&dW[(size_t)co * NC + c], acc[a][b][e]);
            }
        }
}

// LDS-staged weight gradient on v_mfma_f32_16x16x32_bf16 (K = pixels):
// dW[n][tap][c] += sum_m G[m, n] X[m + off(tap), c], G = dout (y > 0 when
// relu_y != NULL), X = in.  A workgroup owns 64 output channels n, 32 input
// channels c and all nine taps (waves: 16 n each, 18 tiles of 16 x 16), and
// sums over tiles_per_split 64-pixel tiles (TH x TW rectangles of one image,
// as k_conv3_mm); the partial sums are added to dW with FP32 atomics, and the
// workgroups of the first channel block add the bias gradient db[n] =
// sum_m G[m, n] (the sums of their G fragments) when db != NULL.  Per tile it
// stages G transposed (channel-major, 64 pixels a row) and three
// copies of the X patch shifted by -1, 0, +1 columns, channel-major, so that
// every operand fragment -- 8 consecutive pixels of one channel -- is one
// 16-byte LDS read.  Requires TW % 8 == 0 (the host checks the shapes).
constexpr int WG_XP = 72;   // bf16 words per staged G row (64 pixels + pad)

__global__ void __launch_bounds__(256) k_conv3_wgrad_mm(ConvShape s, const uint16_t *__restrict__ dout,
                                                        const uint16_t *__restrict__ relu_y,
                                                        const uint16_t *__restrict__ in, float *__restrict__ dW,
                                                        float *__restrict__ db, int tiles_per_split)
{
    __shared__ __attribute__((aligned(16))) uint16_t gt[64 * WG_XP];
    __shared__ __attribute__((aligned(16))) uint16_t xs[3 * 32 * 3 * 72];   // largest: TW = 64, TH + 2 = 3
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int TW = s.W < 64 ? s.W : 64, TH = 64 / TW, PR = TH + 2, XP = TW + 8;
    const int n0 = blockIdx.x * 64, c0 = blockIdx.y * 32;
    const int ntiles = s.B * s.H * s.W / 64;
    const int tb = blockIdx.z * tiles_per_split, te = min(ntiles, tb + tiles_per_split);
    const bool active = n0 + 16 * wave < s.Cout;
    const bool bias = db != nullptr && blockIdx.y == 0;
    float bsum = 0.f;   // this lane's share of sum_m G[m, n] for n = n0 + 16 wave + lane % 16
    f4 acc[9][2];
    for (int t = 0; t < 9; t++)
        for (int b = 0; b < 2; b++) acc[t][b] = f4{0, 0, 0, 0};
    const int kq = 8 * (lane >> 4);
    for (int tile = tb; tile < te; tile++) {
        const int m0 = tile * 64;
        const int img = m0 / (s.H * s.W), r0 = m0 - img * s.H * s.W;
        const int h0 = r0 / s.W, w0 = r0 - h0 * s.W;
        __syncthreads();   // the previous tile's products are done with the LDS
        // G^T: 64 pixels x 64 channels (8 channels per 16-byte load)
        for (int t = threadIdx.x; t < 64 * 8; t += 256) {
            const int p = t >> 3, q = t & 7, n = n0 + 8 * q;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (n < s.Cout) {
                const size_t idx = (size_t)(m0 + p) * s.cs_out + s.co_off + n;
                v = *(const uint4 *)(dout + idx);
                if (relu_y) v = mask_bf16x8(v, *(const uint4 *)(relu_y + idx));
            }
            const uint16_t *e = (const uint16_t *)&v;
#pragma unroll
            for (int k = 0; k < 8; k++) gt[(8 * q + k) * WG_XP + p] = e[k];
        }
        // X patch (PR rows x TW + 2 columns) into the three shifted copies
        const int pw = TW + 2;
        for (int t = threadIdx.x; t < PR * pw * 4; t += 256) {
            const int pp = t >> 2, q = t & 3, c = c0 + 8 * q;
            const int prow = pp / pw, pcol = pp - prow * pw;
            const int hh = h0 - 1 + prow, ww = w0 - 1 + pcol;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (c < s.Cin && hh >= 0 && hh < s.H && ww >= 0 && ww < s.W)
                v = *(const uint4 *)(in + ((size_t)(img * s.H + hh) * s.W + ww) * s.cs_in + s.ci_off + c);
            const uint16_t *e = (const uint16_t *)&v;
            // copy sh holds X at column col + sh - 1 of the tile: col = pcol - sh
#pragma unroll
            for (int sh = 0; sh < 3; sh++) {
                const int col = pcol - sh;
                if (col < 0 || col >= TW) continue;
#pragma unroll
                for (int k = 0; k < 8; k++) xs[((sh * 32 + 8 * q + k) * PR + prow) * XP + col] = e[k];
            }
        }
        __syncthreads();
        if (!active) continue;
        for (int kk = 0; kk < 64; kk += 32) {
            const bf8 A = *(const bf8 *)(gt + (16 * wave + (lane & 15)) * WG_XP + kk + kq);
            if (bias) {
                const uint4 a4 = *(const uint4 *)(gt + (16 * wave + (lane & 15)) * WG_XP + kk + kq);
                const uint16_t *e = (const uint16_t *)&a4;
#pragma unroll
                for (int k = 0; k < 8; k++) bsum += bf2f(e[k]);
            }
            const int p = kk + kq, py = p / TW, px = p - py * TW;
            for (int tap = 0; tap < 9; tap++) {
                const int dh = tap / 3 - 1, sh = tap % 3;
                for (int b = 0; b < 2; b++) {
                    const bf8 B = *(const bf8 *)(xs + ((sh * 32 + 16 * b + (lane & 15)) * PR + py + dh + 1) * XP + px);
                    acc[tap][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[tap][b], 0, 0, 0);
                }
            }
        }
    }
    if (!active) return;
    if (bias) {
        bsum += __shfl_xor(bsum, 16, 64);
        bsum += __shfl_xor(bsum, 32, 64);
        const int n = n0 + 16 * wave + lane;
        if (lane < 16 && n < s.Cout) atomicAdd(&db[n], bsum);
    }
    const int NC = 9 * s.Cin;
    for (int tap = 0; tap < 9; tap++)
        for (int b = 0; b < 2; b++) {
            const int c = c0 + 16 * b + (lane & 15);
            if (c >= s.Cin) continue;
            for (int e = 0; e < 4; e++) {
                const int n = n0 + 16 * wave + (lane >> 4) * 4 + e;
                if (n < s.Cout) atomicAdd(&dW[(size_t)n * NC + tap * s.Cin + c], acc[tap][b][e]);
            }
        }
}

// db[co] += sum_m g[m, co]
__global__ void k_bias_grad(int M, int Cout, int cs, int coff, const uint16_t *__restrict__ dout,
                            const uint16_t *__restrict__ relu_y, float *__restrict__ db)
{
    int co = threadIdx.x % Cout;
    int lanes = blockDim.x / Cout;
    int sub = threadIdx.x / Cout;
    if (sub >= lanes) return;
    float s = 0.f;
    for (int m = blockIdx.x * lanes + sub; m < M; m += gridDim.x * lanes) {
        size_t idx = (size_t)m * cs + coff + co;
        float g = bf2f(dout[idx]);
        if (relu_y && bf2f(relu_y[idx]) <= 0.f) g = 0.f;
        s += g;
    }
    atomicAdd(&db[co], s);
}

// ------------------------------------------------------- pooling / upsampling
// 2x2 average pooling: out (B, H/2, W/2) from in (B, H, W), C channels.
__global__ void k_pool_fwd(int B, int H, int W, int C, const uint16_t *__restrict__ in, int cs_in, int ci_off,
                           uint16_t *__restrict__ out, int cs_out, int co_off)
{
    const int Ho = H / 2, Wo = W / 2;
    size_t total = (size_t)B * Ho * Wo * C;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        int c = t % C;
        size_t p = t / C;
        int w = p % Wo, h = (p / Wo) % Ho, n = p / ((size_t)Wo * Ho);
        float s = 0.f;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                s += bf2f(in[((size_t)((n * H + 2 * h + a) * W + 2 * w + b)) * cs_in + ci_off + c]);
        out[p * cs_out + co_off + c] = f2bf(0.25f * s);
    }
}

// din (+)= pool^T dout
__global__ void k_pool_bwd(int B, int H, int W, int C, const uint16_t *__restrict__ dout, int cs_out, int co_off,
                           uint16_t *__restrict__ din, int cs_in, int ci_off, int accum)
{
    const int Ho = H / 2, Wo = W / 2;
    size_t total = (size_t)B * H * W * C;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        int c = t % C;
        size_t p = t / C;
        int w = p % W, h = (p / W) % H, n = p / ((size_t)W * H);
        float g = 0.25f * bf2f(dout[((size_t)((n * Ho + h / 2) * Wo + w / 2)) * cs_out + co_off + c]);
        size_t idx = p * cs_in + ci_off + c;
        if (accum) g += bf2f(din[idx]);
        din[idx] = f2bf(g);
    }
}

// nearest 2x upsampling: out (B, 2H, 2W) from in (B, H, W); eight channels a
// thread (16-byte accesses) when C, the strides and the offsets allow it
__global__ void k_up_fwd(int B, int H, int W, int C, const uint16_t *__restrict__ in, int cs_in, int ci_off,
                         uint16_t *__restrict__ out, int cs_out, int co_off)
{
    const int Ho = 2 * H, Wo = 2 * W;
    if (C % 8 == 0 && cs_in % 8 == 0 && ci_off % 8 == 0 && cs_out % 8 == 0 && co_off % 8 == 0) {
        const int C8 = C / 8;
        size_t total8 = (size_t)B * Ho * Wo * C8;
        for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total8; t += (size_t)gridDim.x * blockDim.x) {
            int c = (int)(t % C8) * 8;
            size_t p = t / C8;
            int w = p % Wo, h = (p / Wo) % Ho, n = p / ((size_t)Wo * Ho);
            *(uint4 *)(out + p * cs_out + co_off + c) =
                *(const uint4 *)(in + ((size_t)((n * H + h / 2) * W + w / 2)) * cs_in + ci_off + c);
        }
        return;
    }
    size_t total = (size_t)B * Ho * Wo * C;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        int c = t % C;
        size_t p = t / C;
        int w = p % Wo, h = (p / Wo) % Ho, n = p / ((size_t)Wo * Ho);
        out[p * cs_out + co_off + c] = in[((size_t)((n * H + h / 2) * W + w / 2)) * cs_in + ci_off + c];
    }
}

__global__ void k_up_bwd(int B, int H, int W, int C, const uint16_t *__restrict__ dout, int cs_out, int co_off,
                         uint16_t *__restrict__ din, int cs_in, int ci_off)
{
    const int Ho = 2 * H, Wo = 2 * W;
    size_t total = (size_t)B * H * W * C;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        int c = t % C;
        size_t p = t / C;
        int w = p % W, h = (p / W) % H, n = p / ((size_t)W * H);
        float s = 0.f;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                s += bf2f(dout[((size_t)((n * Ho + 2 * h + a) * Wo + 2 * w + b)) * cs_out + co_off + c]);
        din[p * cs_in + ci_off + c] = f2bf(s);
    }
}

// ------------------------------------------------------------- head and loss
// The 1x1 head: 32 channels -> 3 outputs (FP32).  A pixel's 32 bf16 channels
// are four 16-byte loads (channel stride a multiple of 8).
constexpr int HC = 32, HO = 3;

__device__ __forceinline__ void load32(const uint16_t *p, float (&x)[HC])
{
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint4 v = *(const uint4 *)(p + 8 * q);
        const uint16_t *e = (const uint16_t *)&v;
#pragma unroll
        for (int k = 0; k < 8; k++) x[8 * q + k] = bf2f(e[k]);
    }
}

// y[m, o] = bh[o] + sum_c Wh[o][c] x[m, c]
__global__ void __launch_bounds__(256) k_head_fwd(int M, const uint16_t *__restrict__ x, int cs,
                                                  const float *__restrict__ Wh, const float *__restrict__ bh,
                                                  float *__restrict__ y)
{
    __shared__ float w[HO * HC + HO];
    for (int t = threadIdx.x; t < HO * HC + HO; t += blockDim.x) w[t] = t < HO * HC ? Wh[t] : bh[t - HO * HC];
    __syncthreads();
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        float xv[HC];
        load32(x + (size_t)m * cs, xv);
#pragma unroll
        for (int o = 0; o < HO; o++) {
            float s = w[HO * HC + o];
#pragma unroll
            for (int c = 0; c < HC; c++) s += w[o * HC + c] * xv[c];
            y[(size_t)m * HO + o] = s;
        }
    }
}

// dy = 2 (y - t) / (M O), loss partials; dx[m,c] = sum_o Wh[o][c] dy; dWh,
// dbh and the loss summed per thread in registers, then per workgroup, then
// one atomic per value and workgroup.
__global__ void __launch_bounds__(256) k_head_bwd(int M, const uint16_t *__restrict__ x, int cs,
                                                  const float *__restrict__ Wh, const float *__restrict__ y,
                                                  const float *__restrict__ tgt, float scale,
                                                  uint16_t *__restrict__ dx, float *__restrict__ dWh,
                                                  float *__restrict__ dbh, float *__restrict__ loss)
{
    constexpr int NV = HO * HC + HO + 1;   // dWh, dbh, loss
    __shared__ float w[HO * HC];
    __shared__ float red[4][NV];
    for (int t = threadIdx.x; t < HO * HC; t += blockDim.x) w[t] = Wh[t];
    __syncthreads();
    float aw[HO * HC], ab[HO], ls = 0.f;
#pragma unroll
    for (int k = 0; k < HO * HC; k++) aw[k] = 0.f;
#pragma unroll
    for (int o = 0; o < HO; o++) ab[o] = 0.f;
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        float dy[HO];
#pragma unroll
        for (int o = 0; o < HO; o++) {
            const float r = y[(size_t)m * HO + o] - tgt[(size_t)m * HO + o];
            ls += r * r;
            dy[o] = scale * r;
            ab[o] += dy[o];
        }
        float xv[HC];
        load32(x + (size_t)m * cs, xv);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint4 v;
            uint16_t *e = (uint16_t *)&v;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int c = 8 * q + k;
                float g = 0.f;
#pragma unroll
                for (int o = 0; o < HO; o++) {
                    g += w[o * HC + c] * dy[o];
                    aw[o * HC + c] += dy[o] * xv[c];
                }
                e[k] = f2bf(g);
            }
            *(uint4 *)(dx + (size_t)m * cs + 8 * q) = v;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        float v = k < HO * HC ? aw[k] : k < HO * HC + HO ? ab[k - HO * HC] : ls;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NV; k += blockDim.x) {
        const float v = red[0][k] + red[1][k] + red[2][k] + red[3][k];
        if (k < HO * HC) atomicAdd(&dWh[k], v);
        else if (k < HO * HC + HO) atomicAdd(&dbh[k - HO * HC], v);
        else atomicAdd(loss, v);
    }
}

// ----------------------------------------------------------------- Adam etc.
// Adam (Kingma & Ba 2015) on FP32 master weights; refreshes the bf16 copies.
__global__ void k_adam(int n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float gi = g[i];
        float a = b1 * m1[i] + (1.f - b1) * gi;
        float b = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    }
}

// Wf (bf16, [co][tap][ci]) and Wd (bf16, [ci][tap][co]) from the FP32 master [co][tap][ci].
__global__ void k_weights_bf16(int Cout, int Cin, const float *__restrict__ w, uint16_t *__restrict__ Wf,
                               uint16_t *__restrict__ Wd)
{
    int n = Cout * 9 * Cin;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int ci = i % Cin, tap = (i / Cin) % 9, co = i / (9 * Cin);
        uint16_t v = f2bf(w[i]);
        Wf[i] = v;
        Wd[((size_t)ci * 9 + tap) * Cout + co] = v;
    }
}

__global__ void k_f32_to_bf16(size_t n, const float *__restrict__ x, uint16_t *__restrict__ y)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

__global__ void k_bf16_to_f32(size_t n, const uint16_t *__restrict__ x, float *__restrict__ y)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = bf2f(x[i]);
}
