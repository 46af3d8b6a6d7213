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

// nearest 2x upsampling: out (B, 2H, 2W) from in (B, H, W)
__global__ void k_up_fwd(int B, int H, int W, int C, const uint16_t *__restrict__ in, int cs_in, int ci_off,
                         uint16_t *__restrict__ out, int cs_out, int co_off)
{
    const int Ho = 2 * H, Wo = 2 * W;
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
// y[m, o] = bh[o] + sum_c Wh[o][c] x[m, c]   (1x1 convolution, FP32 output)
__global__ void k_head_fwd(int M, int C, int O, const uint16_t *__restrict__ x, int cs, const float *__restrict__ Wh,
                           const float *__restrict__ bh, float *__restrict__ y)
{
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x)
        for (int o = 0; o < O; o++) {
            float s = bh[o];
            for (int c = 0; c < C; c++) s += Wh[o * C + c] * bf2f(x[(size_t)m * cs + c]);
            y[(size_t)m * O + o] = s;
        }
}

// dy = 2 (y - t) / (M O), loss partials; dx[m,c] = sum_o Wh[o][c] dy; dWh, dbh (atomics)
__global__ void k_head_bwd(int M, int C, int O, const uint16_t *__restrict__ x, int cs, const float *__restrict__ Wh,
                           const float *__restrict__ y, const float *__restrict__ tgt, float scale,
                           uint16_t *__restrict__ dx, float *__restrict__ dWh, float *__restrict__ dbh,
                           float *__restrict__ loss)
{
    __shared__ float sW[8 * 64], sb[8];
    for (int t = threadIdx.x; t < O * C + O; t += blockDim.x) {
        if (t < O * C) sW[t] = 0.f;
        else sb[t - O * C] = 0.f;
    }
    __syncthreads();
    float ls = 0.f;
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        float dy[8];
        for (int o = 0; o < O; o++) {
            float r = y[(size_t)m * O + o] - tgt[(size_t)m * O + o];
            ls += r * r;
            dy[o] = scale * r;
            atomicAdd(&sb[o], dy[o]);
        }
        for (int c = 0; c < C; c++) {
            float xv = bf2f(x[(size_t)m * cs + c]);
            float g = 0.f;
            for (int o = 0; o < O; o++) {
                g += Wh[o * C + c] * dy[o];
                atomicAdd(&sW[o * C + c], dy[o] * xv);
            }
            dx[(size_t)m * cs + c] = f2bf(g);
        }
    }
    atomicAdd(loss, ls);
    __syncthreads();
    for (int t = threadIdx.x; t < O * C + O; t += blockDim.x) {
        if (t < O * C) atomicAdd(&dWh[t], sW[t]);
        else atomicAdd(&dbh[t - O * C], sb[t - O * C]);
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
