// gp_nn.hip -- learned warm start / preconditioner for the elliptic solves.
//
// No reference counterpart: the reference solves with AMGCL only
// (EMField.mm:866, 1199, 1237, 1275).  The network maps a normalised
// right-hand side b/rms(b) of the SPD system S x = b (the solver's unknown
// grid, zero-padded to a multiple of 8 per axis) to the direction of x.
//
// Network: 4-level convolutional encoder-decoder, 3x3 kernels, channels
// c0, 2c0, 4c0, 4c0 (c0 = 32), BF16 weights and activations, FP32
// accumulation on MFMA (v_mfma_f32_16x16x32_bf16), NHWC activations.
//   a0 = relu(conv_in(r))            1 -> c0       level 0   (VALU)
//   a1 = relu(W1 pool(a0))           c0 -> 2c0     level 1   (MFMA)
//   a2 = relu(W2 pool(a1))           2c0 -> 4c0    level 2
//   a3 = relu(W3 pool(a2))           4c0 -> 4c0    level 3
//   g2 = relu(W4 a3);  u2 = up(g2) + a2            level 3 -> 2
//   g1 = relu(W5 u2);  u1 = up(g1) + a1            level 2 -> 1
//   g0 = relu(W6 u1);  u0 = up(g0) + a0            level 1 -> 0
//   y  = conv_out(u0)                 c0 -> 1      level 0   (VALU)
// pool is a 2x2 mean, up a nearest-neighbour copy.  Padding of every 3x3
// convolution follows the problem: wrap on a periodic axis, zero otherwise.
//
// Training: targets t = x / rms(x); loss (default) = the relative energy
// error (y - t)' S1 (y - t) / t' S1 t per sample (S1 the grid-unit operator),
// mean over the batch, or the mean square (y - t)^2; the
// backward pass has its own kernels (MFMA for the data and weight gradients
// of W1..W6), the weight-gradient partials are reduced in a fixed order, and
// Adam updates FP32 master weights from which the BF16 copies (and the
// tap-flipped transposes used by the data gradient) are regenerated.
#include "gp_internal.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16s;   // BF16 storage

namespace {

constexpr int NL = 8;            // layers: in, W1..W6, out
constexpr int NB = 256;          // threads of the elementwise kernels

// round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ bf16s f2b(float f) { return __builtin_bit_cast(bf16s, (__bf16)f); }
__device__ __forceinline__ float b2f(bf16s h) { return __uint_as_float((uint32_t)h << 16); }
// the 8 BF16 values of a 16-byte word, in memory (channel) order
__device__ __forceinline__ void unpack8(const uint4 &v, float *f)
{
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        f[2 * k] = __uint_as_float(w[k] << 16);
        f[2 * k + 1] = __uint_as_float(w[k] & 0xFFFF0000u);
    }
}
__device__ __forceinline__ uint4 pack8(const float *f)
{
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; k++) w[k] = (uint32_t)f2b(f[2 * k]) | ((uint32_t)f2b(f[2 * k + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}

struct Geo {
    int B, H, W;      // batch, rows, columns
    int perx, pery;   // periodic axes (wrap padding)
};
// source pixel of (y + dy, x + dx): -1 outside a non-periodic axis
__device__ __forceinline__ int src_pix(const Geo &g, int b, int y, int x)
{
    if (y < 0 || y >= g.H) {
        if (!g.pery) return -1;
        y = ((y % g.H) + g.H) % g.H;
    }
    if (x < 0 || x >= g.W) {
        if (!g.perx) return -1;
        x = ((x % g.W) + g.W) % g.W;
    }
    return (b * g.H + y) * g.W + x;
}

// ---------------------------------------------------------------------------
// 3x3 convolution on MFMA: out[p][co] = sum_t sum_ci in[p + off_t][ci] w[co][t][ci]
// One workgroup (4 waves) per 8x8 output tile of one image; wave w owns rows
// 2w, 2w+1 (16 pixels) and all CO channels (CO/16 accumulator tiles).  The
// 10x10 input halo and one tap's weights at a time are staged in LDS with a
// 16-byte pad per row, which keeps the 16-byte operand reads of a 16-lane
// group on distinct banks.
// MFMA 16x16x32 operand layout: lane l holds A[l%16][8(l/16) .. +7] and
// B[8(l/16) .. +7][l%16]; the result lane l holds D[4(l/16) + r][l%16].
// Epilogue flags: bias, relu, and mask (multiply by [mask[p][co] > 0]).
// ---------------------------------------------------------------------------
template <int CI, int CO>
__global__ __launch_bounds__(256) void k_conv(const bf16s *in, const bf16s *w, const float *bias, const bf16s *mask,
                                              bf16s *out, Geo g, int relu)
{
    constexpr int PP = CI + 8;
    __shared__ __attribute__((aligned(16))) bf16s sIn[100 * PP];
    __shared__ __attribute__((aligned(16))) bf16s sW[CO * PP];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int tx0 = blockIdx.x * 8, ty0 = blockIdx.y * 8, b = blockIdx.z;
    // stage the halo tile, 8 channels (16 bytes) per item
    constexpr int CH = CI / 8;
    for (int e = tid; e < 100 * CH; e += 256) {
        const int pix = e / CH, c8 = e % CH;
        const int hy = pix / 10, hx = pix % 10;
        const int sp = src_pix(g, b, ty0 + hy - 1, tx0 + hx - 1);
        uint4 v = make_uint4(0, 0, 0, 0);
        if (sp >= 0) v = *(const uint4 *)(in + (int64_t)sp * CI + c8 * 8);
        *(uint4 *)(sIn + pix * PP + c8 * 8) = v;
    }
    f32x4 acc[CO / 16];
#pragma unroll
    for (int n = 0; n < CO / 16; n++) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int i = lane & 15, kq = lane >> 4;
    const int ry = 2 * wv + (i >> 3), rx = i & 7;
    // the next tap's weights are loaded into registers while this tap's
    // MFMAs run (their global latency is off the critical path)
    constexpr int WPT = (CO * CH + 255) / 256;
    u32x4 wr[WPT];
#pragma unroll
    for (int k = 0; k < WPT; k++) {
        const int e = tid + 256 * k;
        if (e < CO * CH) wr[k] = *(const u32x4 *)(w + ((int64_t)(e / CH) * 9) * CI + (e % CH) * 8);
    }
    for (int t = 0; t < 9; t++) {
        __syncthreads();   // previous tap's weights consumed (and the halo staged)
#pragma unroll
        for (int k = 0; k < WPT; k++) {
            const int e = tid + 256 * k;
            if (e < CO * CH) *(u32x4 *)(sW + (e / CH) * PP + (e % CH) * 8) = wr[k];
        }
        __syncthreads();
        if (t < 8) {
#pragma unroll
            for (int k = 0; k < WPT; k++) {
                const int e = tid + 256 * k;
                if (e < CO * CH) wr[k] = *(const u32x4 *)(w + ((int64_t)(e / CH) * 9 + t + 1) * CI + (e % CH) * 8);
            }
        }
        const int dy = t / 3 - 1, dx = t % 3 - 1;
        const bf16s *arow = sIn + ((ry + 1 + dy) * 10 + (rx + 1 + dx)) * PP + kq * 8;
#pragma unroll
        for (int kc = 0; kc < CI / 32; kc++) {
            const bf16x8 a = __builtin_bit_cast(bf16x8, *(const uint4 *)(arow + kc * 32));
#pragma unroll
            for (int n = 0; n < CO / 16; n++) {
                const bf16x8 bb = __builtin_bit_cast(bf16x8, *(const uint4 *)(sW + (n * 16 + i) * PP + kc * 32 + kq * 8));
                acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[n], 0, 0, 0);
            }
        }
    }
    // epilogue: lane holds pixels 4 kq + r of the wave's 16, channel n*16 + i
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int pi = 4 * kq + r;
        const int oy = ty0 + 2 * wv + (pi >> 3), ox = tx0 + (pi & 7);
        if (oy >= g.H || ox >= g.W) continue;
        const int64_t p = ((int64_t)b * g.H + oy) * g.W + ox;
#pragma unroll
        for (int n = 0; n < CO / 16; n++) {
            const int co = n * 16 + i;
            float v = acc[n][r];
            if (bias) v += bias[co];
            if (relu) v = fmaxf(v, 0.0f);
            if (mask && !(b2f(mask[p * CO + co]) > 0.0f)) v = 0.0f;
            out[p * CO + co] = f2b(v);
        }
    }
}

// ---------------------------------------------------------------------------
// weight gradient on MFMA, one tap per workgroup row of the grid:
// part[chunk][t][co][ci] = sum_{p in chunk} d[p][co] x[p + off_t][ci]
// Each 32-pixel step stages d and the shifted x transposed (pixel-contiguous)
// in LDS; the MFMA reduction dimension is the pixel.  CO/16 x CI/16 output
// tiles are split over the 4 waves.
// ---------------------------------------------------------------------------
template <int CI, int CO>
__global__ __launch_bounds__(256) void k_dw(const bf16s *d, const bf16s *x, float *part, Geo g, int64_t npix,
                                            int64_t chunk)
{
    constexpr int KP = 40;   // 32 pixels + 8 pad (80-byte rows)
    __shared__ __attribute__((aligned(16))) bf16s sD[CO * KP];
    __shared__ __attribute__((aligned(16))) bf16s sX[CI * KP];
    constexpr int NT = (CO / 16) * (CI / 16), TPW = NT / 4;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int t = blockIdx.x, ck = blockIdx.y;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const int i = lane & 15, kq = lane >> 4;
    f32x4 acc[TPW];
#pragma unroll
    for (int n = 0; n < TPW; n++) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int64_t p0 = ck * chunk, p1 = std::min(npix, p0 + chunk);
    const int HW = g.H * g.W;
    for (int64_t q0 = p0; q0 < p1; q0 += 32) {
        __syncthreads();
        for (int e = tid; e < 32 * (CO / 8 + CI / 8); e += 256) {
            const int px = e & 31, item = e >> 5;
            const int64_t p = q0 + px;
            uint4 v = make_uint4(0, 0, 0, 0);
            bf16s *dst;
            int stride = KP;
            if (item < CO / 8) {
                if (p < p1) v = *(const uint4 *)(d + p * CO + item * 8);
                dst = sD + item * 8 * KP + px;
            } else {
                const int c8 = item - CO / 8;
                if (p < p1) {
                    const int b = (int)(p / HW), rem = (int)(p % HW);
                    const int sp = src_pix(g, b, rem / g.W + dy, rem % g.W + dx);
                    if (sp >= 0) v = *(const uint4 *)(x + (int64_t)sp * CI + c8 * 8);
                }
                dst = sX + c8 * 8 * KP + px;
            }
            const bf16s *h = (const bf16s *)&v;
#pragma unroll
            for (int c = 0; c < 8; c++) dst[c * stride] = h[c];
        }
        __syncthreads();
#pragma unroll
        for (int n = 0; n < TPW; n++) {
            const int tile = wv * TPW + n;
            const int ta = tile / (CI / 16), tb = tile % (CI / 16);
            const bf16x8 a = __builtin_bit_cast(bf16x8, *(const uint4 *)(sD + (ta * 16 + i) * KP + kq * 8));
            const bf16x8 bb = __builtin_bit_cast(bf16x8, *(const uint4 *)(sX + (tb * 16 + i) * KP + kq * 8));
            acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[n], 0, 0, 0);
        }
    }
    float *dst = part + ((int64_t)ck * 9 + t) * CO * CI;
#pragma unroll
    for (int n = 0; n < TPW; n++) {
        const int tile = wv * TPW + n;
        const int ta = tile / (CI / 16), tb = tile % (CI / 16);
#pragma unroll
        for (int r = 0; r < 4; r++) dst[(ta * 16 + 4 * kq + r) * CI + tb * 16 + i] = acc[n][r];
    }
}
// grad[co][t][ci] = sum_chunk part[chunk][t][co][ci] (fixed order)
__global__ void k_dw_reduce(const float *part, int nchunk, int CO, int CI, float *grad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = 9 * CO * CI;
    if (e >= n) return;
    const int t = e / (CO * CI), rem = e % (CO * CI), co = rem / CI, ci = rem % CI;
    float s = 0.0f;
    for (int k = 0; k < nchunk; k++) s += part[(int64_t)k * n + e];
    grad[((int64_t)co * 9 + t) * CI + ci] = s;
}
// bias gradient partials: part[blk][c] = sum over the block's pixels of d[p][c].
// Each thread reads 8 channels (16 bytes) of every (256 / (C / 8))-th pixel;
// the pixel lanes are added in lane order.
__global__ __launch_bounds__(NB) void k_colsum(const bf16s *d, int64_t npix, int C, int64_t chunk, float *part)
{
    __shared__ float sh[NB * 8];
    const int C8 = C / 8, PL = NB / C8;
    const int cg = threadIdx.x % C8, pl = threadIdx.x / C8;
    const int64_t p0 = blockIdx.x * chunk, p1 = std::min(npix, p0 + chunk);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (pl < PL)
        for (int64_t p = p0 + pl; p < p1; p += PL) {
            float v[8];
            unpack8(*(const uint4 *)(d + p * C + 8 * cg), v);
#pragma unroll
            for (int k = 0; k < 8; k++) s[k] += v[k];
        }
    if (pl < PL)
#pragma unroll
        for (int k = 0; k < 8; k++) sh[pl * C + 8 * cg + k] = s[k];
    __syncthreads();
    for (int c = threadIdx.x; c < C; c += NB) {
        float t = 0.0f;
        for (int l = 0; l < PL; l++) t += sh[l * C + c];
        part[(int64_t)blockIdx.x * C + c] = t;
    }
}
__global__ void k_colsum_f(const float *d, int64_t npix, int64_t chunk, float *part)
{
    __shared__ float sh[NB];
    const int64_t p0 = blockIdx.x * chunk, p1 = std::min(npix, p0 + chunk);
    float s = 0.0f;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += NB) s += d[p];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = NB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
// out[e] = sum over k < nblk of part[k][e]: a block takes 32 outputs with 8
// lanes each; lane l sums k = l, l + 8, ... (four loads in flight, added in
// k order) and the lanes add in lane order
__global__ __launch_bounds__(NB) void k_part_reduce(const float *part, int nblk, int n, float *out)
{
    __shared__ float sh[8][33];
    const int ex = threadIdx.x & 31, kl = threadIdx.x >> 5;
    const int e = blockIdx.x * 32 + ex;
    float s = 0.0f;
    if (e < n) {
        int k = kl;
        for (; k + 24 < nblk; k += 32) {
            const float a0 = part[(int64_t)k * n + e], a1 = part[(int64_t)(k + 8) * n + e];
            const float a2 = part[(int64_t)(k + 16) * n + e], a3 = part[(int64_t)(k + 24) * n + e];
            s += a0; s += a1; s += a2; s += a3;
        }
        for (; k < nblk; k += 8) s += part[(int64_t)k * n + e];
    }
    sh[kl][ex] = s;
    __syncthreads();
    if (kl == 0 && e < n) {
        float t = 0.0f;
        for (int l = 0; l < 8; l++) t += sh[l][ex];
        out[e] = t;
    }
}
// out[0] = sum of part[0 .. nblk): strided thread sums, then a fixed tree
__global__ __launch_bounds__(NB) void k_part_sum1(const float *part, int nblk, float *out)
{
    __shared__ float sh[NB];
    float s = 0.0f;
    for (int k = threadIdx.x; k < nblk; k += NB) s += part[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = NB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// ---------------------------------------------------------------------------
// input and output convolutions (one channel side, VALU)
// ---------------------------------------------------------------------------
// a0[p][c] = relu(sum_t w[c][t] r[p + off_t] + b[c]), C0 channels
template <int C0>
__global__ void k_conv_in(const float *r, const float *w, const float *bias, bf16s *out, Geo g)
{
    const int64_t p = blockIdx.x * (int64_t)NB + threadIdx.x;
    const int64_t np = (int64_t)g.B * g.H * g.W;
    if (p >= np) return;
    const int HW = g.H * g.W;
    const int b = (int)(p / HW), rem = (int)(p % HW), y = rem / g.W, x = rem % g.W;
    float v[9];
#pragma unroll
    for (int t = 0; t < 9; t++) {
        const int sp = src_pix(g, b, y + t / 3 - 1, x + t % 3 - 1);
        v[t] = sp >= 0 ? r[sp] : 0.0f;
    }
    for (int c = 0; c < C0; c += 2) {
        float s0 = bias[c], s1 = bias[c + 1];
#pragma unroll
        for (int t = 0; t < 9; t++) {
            s0 = __builtin_fmaf(w[c * 9 + t], v[t], s0);
            s1 = __builtin_fmaf(w[(c + 1) * 9 + t], v[t], s1);
        }
        const uint32_t pk = (uint32_t)f2b(fmaxf(s0, 0.0f)) | ((uint32_t)f2b(fmaxf(s1, 0.0f)) << 16);
        *(uint32_t *)(out + p * C0 + c) = pk;
    }
}
// y[p] = sum_t sum_c w[t][c] u[p + off_t][c] + b
template <int C0>
__global__ void k_conv_out(const bf16s *u, const float *w, const float *bias, float *y, Geo g)
{
    const int64_t p = blockIdx.x * (int64_t)NB + threadIdx.x;
    const int64_t np = (int64_t)g.B * g.H * g.W;
    if (p >= np) return;
    const int HW = g.H * g.W;
    const int b = (int)(p / HW), rem = (int)(p % HW), yy = rem / g.W, x = rem % g.W;
    float s = bias[0];
    for (int t = 0; t < 9; t++) {
        const int sp = src_pix(g, b, yy + t / 3 - 1, x + t % 3 - 1);
        if (sp < 0) continue;
        const bf16s *up = u + (int64_t)sp * C0;
#pragma unroll
        for (int c = 0; c < C0; c++) s = __builtin_fmaf(w[t * C0 + c], b2f(up[c]), s);
    }
    y[p] = s;
}
// data gradient of conv_out: du[p][c] = sum_t w[t][c] dy[p - off_t]
template <int C0>
__global__ void k_conv_out_dx(const float *dy, const float *w, bf16s *du, Geo g)
{
    const int64_t p = blockIdx.x * (int64_t)NB + threadIdx.x;
    const int64_t np = (int64_t)g.B * g.H * g.W;
    if (p >= np) return;
    const int HW = g.H * g.W;
    const int b = (int)(p / HW), rem = (int)(p % HW), yy = rem / g.W, x = rem % g.W;
    float dv[9];
#pragma unroll
    for (int t = 0; t < 9; t++) {
        const int sp = src_pix(g, b, yy - (t / 3 - 1), x - (t % 3 - 1));
        dv[t] = sp >= 0 ? dy[sp] : 0.0f;
    }
    for (int c = 0; c < C0; c++) {
        float s = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; t++) s = __builtin_fmaf(w[t * C0 + c], dv[t], s);
        du[p * C0 + c] = f2b(s);
    }
}
// Weight-gradient partials of the two one-channel convolutions at level 0:
//   part[blk][idx(c, t)] = sum over the block's pixels q of A[q][c] v[src(q + sg off_t)]
//   conv_in:  A = d (gradient of the pre-activation), v = r (input), sg = +1, idx = 9 c + t
//   conv_out: A = u (input activation), v = dy,                     sg = -1, idx = C0 t + c
// (sum_p dy[p] u[src(p + off)][c] = sum_q u[q][c] dy[src(q - off)]: both run
// over the pixel pairs inside the image, or wrap on a periodic axis.)
// Tiles of 8 x 32 pixels, taken by the blocks in a grid-stride loop: A of
// the tile and the 10 x 34 window of v are staged in LDS, and thread (c, t)
// accumulates its product over the tile's pixels in row order.
template <int C0>
__global__ __launch_bounds__(9 * C0) void k_dw1(const bf16s *A, const float *v, Geo g, int sg, int ct, float *part)
{
    constexpr int TH = 8, TW = 32, NT1 = 9 * C0, VW = TW + 2;
    __shared__ __attribute__((aligned(16))) bf16s sA[TH * TW * C0];
    __shared__ float sV[(TH + 2) * VW];
    const int tid = threadIdx.x, c = tid / 9, t = tid % 9;
    const int dy = sg * (t / 3 - 1), dx = sg * (t % 3 - 1);
    const int tyn = (g.H + TH - 1) / TH, txn = (g.W + TW - 1) / TW;
    const int64_t ntile = (int64_t)g.B * tyn * txn;
    float s = 0.0f;
    for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int b = (int)(tile / (tyn * txn)), rem = (int)(tile % (tyn * txn));
        const int ty0 = (rem / txn) * TH, tx0 = (rem % txn) * TW;
        __syncthreads();   // the previous tile is consumed
        for (int e = tid; e < TH * TW * (C0 / 8); e += NT1) {
            const int pix = e / (C0 / 8), c8 = e % (C0 / 8);
            const int yy = ty0 + pix / TW, xx = tx0 + pix % TW;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (yy < g.H && xx < g.W) val = *(const uint4 *)(A + (((int64_t)b * g.H + yy) * g.W + xx) * C0 + 8 * c8);
            *(uint4 *)(sA + pix * C0 + 8 * c8) = val;
        }
        for (int e = tid; e < (TH + 2) * VW; e += NT1) {
            const int sp = src_pix(g, b, ty0 + e / VW - 1, tx0 + e % VW - 1);
            sV[e] = sp >= 0 ? v[sp] : 0.0f;
        }
        __syncthreads();
        const float *vrow = sV + (1 + dy) * VW + 1 + dx;
        for (int yy = 0; yy < TH; yy++) {
#pragma unroll 8
            for (int xx = 0; xx < TW; xx++)
                s = __builtin_fmaf(b2f(sA[(yy * TW + xx) * C0 + c]), vrow[yy * VW + xx], s);
        }
    }
    part[(int64_t)blockIdx.x * NT1 + (ct ? t * C0 + c : 9 * c + t)] = s;
}

// ---------------------------------------------------------------------------
// Level 0 of the inference forward, fused: the C0-channel activations a0 and
// u0 (64 B per pixel each) are never written.  Both kernels evaluate every
// value exactly as the unfused chain does (conv_in's FMA order, the BF16
// rounding of a0, of the pooled p1 and of u0 = up(g0) + a0, conv_out's tap
// and channel order), so the output is the same to the bit.
// ---------------------------------------------------------------------------
// p1 = pool(relu(conv_in(r))): one thread per pooled pixel, all C0 channels
template <int C0>
__global__ void k_in_pool(const float *r, const float *w, const float *bias, bf16s *p1, Geo g)
{
    const int H2 = g.H / 2, W2 = g.W / 2;
    const int64_t e = blockIdx.x * (int64_t)NB + threadIdx.x;
    if (e >= (int64_t)g.B * H2 * W2) return;
    const int x2 = (int)(e % W2), y2 = (int)((e / W2) % H2), b = (int)(e / ((int64_t)W2 * H2));
    float v[4][4];   // r around the 2 x 2 block, rows 2 y2 - 1 .. 2 y2 + 2
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int sp = src_pix(g, b, 2 * y2 - 1 + a, 2 * x2 - 1 + c);
            v[a][c] = sp >= 0 ? r[sp] : 0.0f;
        }
    for (int c8 = 0; c8 < C0 / 8; c8++) {
        float o[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int c = 8 * c8 + k;
            float q[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {   // (0,0) (0,1) (1,0) (1,1)
                const int oy = j >> 1, ox = j & 1;
                float s0 = bias[c];
#pragma unroll
                for (int t = 0; t < 9; t++) s0 = __builtin_fmaf(w[c * 9 + t], v[oy + t / 3][ox + t % 3], s0);
                q[j] = b2f(f2b(fmaxf(s0, 0.0f)));
            }
            o[k] = 0.25f * ((q[0] + q[1]) + (q[2] + q[3]));
        }
        *(uint4 *)(p1 + e * C0 + 8 * c8) = pack8(o);
    }
}
// y = conv_out(up(g0) + relu(conv_in(r))) on 14 x 14 output tiles: the r
// window (18 x 18) and u0 on the 16 x 16 window (channel-major, odd pitch;
// one window position per thread) are built in LDS, then each of 196
// threads sums its pixel's 9 taps x C0 channels
template <int C0>
__global__ __launch_bounds__(256) void k_up_out(const float *r, const bf16s *g0, const float *win, const float *bin,
                                                const float *wout, const float *bout, float *y, Geo g)
{
    constexpr int T = 14, UW = T + 2, RW = T + 4, UP = UW * UW + 1;
    __shared__ float sr[RW * RW];
    __shared__ float su[C0 * UP];
    const int tid = threadIdx.x;
    const int tx0 = blockIdx.x * T, ty0 = blockIdx.y * T, b = blockIdx.z;
    for (int e = tid; e < RW * RW; e += 256) {
        const int sp = src_pix(g, b, ty0 - 2 + e / RW, tx0 - 2 + e % RW);
        sr[e] = sp >= 0 ? r[sp] : 0.0f;
    }
    __syncthreads();
    // one window position per thread: its 3 x 3 r values once, then every
    // channel (the weights are uniform across the block)
    for (int pix = tid; pix < UW * UW; pix += 256) {
        const int hy = pix / UW, hx = pix % UW;
        // the pixel this window position reads (wrapped on a periodic axis)
        const int sp = src_pix(g, b, ty0 - 1 + hy, tx0 - 1 + hx);
        if (sp < 0) {
            for (int c = 0; c < C0; c++) su[c * UP + pix] = 0.0f;
            continue;
        }
        const int rem = sp % (g.H * g.W), wy = rem / g.W, wx = rem % g.W;
        float v[9];
#pragma unroll
        for (int t = 0; t < 9; t++) v[t] = sr[(hy + t / 3) * RW + hx + t % 3];
        const bf16s *gp = g0 + (((int64_t)b * (g.H / 2) + wy / 2) * (g.W / 2) + wx / 2) * C0;
        for (int c8 = 0; c8 < C0 / 8; c8++) {
            float gv[8];
            unpack8(*(const uint4 *)(gp + 8 * c8), gv);
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int c = 8 * c8 + k;
                float s0 = bin[c];
#pragma unroll
                for (int t = 0; t < 9; t++) s0 = __builtin_fmaf(win[c * 9 + t], v[t], s0);
                su[c * UP + pix] = b2f(f2b(gv[k] + b2f(f2b(fmaxf(s0, 0.0f)))));
            }
        }
    }
    __syncthreads();
    const int ly = tid / T, lx = tid % T, oy = ty0 + ly, ox = tx0 + lx;
    if (tid >= T * T || oy >= g.H || ox >= g.W) return;
    float s = bout[0];
    for (int t = 0; t < 9; t++) {
        if (src_pix(g, b, oy + t / 3 - 1, ox + t % 3 - 1) < 0) continue;
        const float *up = su + (ly + t / 3) * UW + lx + t % 3;
#pragma unroll
        for (int c = 0; c < C0; c++) s = __builtin_fmaf(wout[t * C0 + c], up[c * UP], s);
    }
    y[((int64_t)b * g.H + oy) * g.W + ox] = s;
}

// ---------------------------------------------------------------------------
// pooling, upsampling and the fused backward combinations (NHWC bf16)
// ---------------------------------------------------------------------------
// One thread per 8 channels (16 bytes) of a pixel in the four kernels below;
// every channel is combined exactly as in a scalar loop.
// out (H/2 x W/2) = mean of 2x2 blocks of in (H x W); g = geometry of in
__global__ void k_pool(const bf16s *in, bf16s *out, Geo g, int C)
{
    const int64_t e = blockIdx.x * (int64_t)NB + threadIdx.x;
    const int C8 = C / 8, H2 = g.H / 2, W2 = g.W / 2;
    const int64_t n = (int64_t)g.B * H2 * W2 * C8;
    if (e >= n) return;
    const int c8 = (int)(e % C8);
    const int64_t q = e / C8;
    const int x = (int)(q % W2), y = (int)((q / W2) % H2), b = (int)(q / ((int64_t)W2 * H2));
    const int64_t s = (((int64_t)b * g.H + 2 * y) * g.W + 2 * x) * C + 8 * c8, row = (int64_t)g.W * C;
    float v0[8], v1[8], v2[8], v3[8], o[8];
    unpack8(*(const uint4 *)(in + s), v0);
    unpack8(*(const uint4 *)(in + s + C), v1);
    unpack8(*(const uint4 *)(in + s + row), v2);
    unpack8(*(const uint4 *)(in + s + row + C), v3);
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = 0.25f * ((v0[k] + v1[k]) + (v2[k] + v3[k]));
    *(uint4 *)(out + 8 * e) = pack8(o);
}
// out (H x W) = up(gin (H/2 x W/2)) + skip; g = geometry of out
__global__ void k_up_add(const bf16s *gin, const bf16s *skip, bf16s *out, Geo g, int C)
{
    const int64_t e = blockIdx.x * (int64_t)NB + threadIdx.x;
    const int C8 = C / 8;
    const int64_t n = (int64_t)g.B * g.H * g.W * C8;
    if (e >= n) return;
    const int c8 = (int)(e % C8);
    const int64_t q = e / C8;
    const int x = (int)(q % g.W), y = (int)((q / g.W) % g.H), b = (int)(q / ((int64_t)g.W * g.H));
    const int64_t s = (((int64_t)b * (g.H / 2) + y / 2) * (g.W / 2) + x / 2) * C + 8 * c8;
    float vg[8], vs[8], o[8];
    unpack8(*(const uint4 *)(gin + s), vg);
    unpack8(*(const uint4 *)(skip + 8 * e), vs);
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = vg[k] + vs[k];
    *(uint4 *)(out + 8 * e) = pack8(o);
}
// decoder backward: dpre (H/2) = (sum of the 2x2 block of du (H)) * [act > 0]; g = geometry of du
__global__ void k_sum2x2_mask(const bf16s *du, const bf16s *act, bf16s *dpre, Geo g, int C)
{
    const int64_t e = blockIdx.x * (int64_t)NB + threadIdx.x;
    const int C8 = C / 8, H2 = g.H / 2, W2 = g.W / 2;
    const int64_t n = (int64_t)g.B * H2 * W2 * C8;
    if (e >= n) return;
    const int c8 = (int)(e % C8);
    const int64_t q = e / C8;
    const int x = (int)(q % W2), y = (int)((q / W2) % H2), b = (int)(q / ((int64_t)W2 * H2));
    const int64_t s = (((int64_t)b * g.H + 2 * y) * g.W + 2 * x) * C + 8 * c8, row = (int64_t)g.W * C;
    float v0[8], v1[8], v2[8], v3[8], va[8], o[8];
    unpack8(*(const uint4 *)(du + s), v0);
    unpack8(*(const uint4 *)(du + s + C), v1);
    unpack8(*(const uint4 *)(du + s + row), v2);
    unpack8(*(const uint4 *)(du + s + row + C), v3);
    unpack8(*(const uint4 *)(act + 8 * e), va);
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = va[k] > 0.0f ? (v0[k] + v1[k]) + (v2[k] + v3[k]) : 0.0f;
    *(uint4 *)(dpre + 8 * e) = pack8(o);
}
// encoder backward: dpre (H) = (dskip + 0.25 up(dpool (H/2))) * [act > 0]; g = geometry of dskip
__global__ void k_skip_pool_mask(const bf16s *dskip, const bf16s *dpool, const bf16s *act, bf16s *dpre, Geo g, int C)
{
    const int64_t e = blockIdx.x * (int64_t)NB + threadIdx.x;
    const int C8 = C / 8;
    const int64_t n = (int64_t)g.B * g.H * g.W * C8;
    if (e >= n) return;
    const int c8 = (int)(e % C8);
    const int64_t q = e / C8;
    const int x = (int)(q % g.W), y = (int)((q / g.W) % g.H), b = (int)(q / ((int64_t)g.W * g.H));
    const int64_t s = (((int64_t)b * (g.H / 2) + y / 2) * (g.W / 2) + x / 2) * C + 8 * c8;
    float vs[8], vp[8], va[8], o[8];
    unpack8(*(const uint4 *)(dskip + 8 * e), vs);
    unpack8(*(const uint4 *)(dpool + s), vp);
    unpack8(*(const uint4 *)(act + 8 * e), va);
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = va[k] > 0.0f ? vs[k] + 0.25f * vp[k] : 0.0f;
    *(uint4 *)(dpre + 8 * e) = pack8(o);
}

// loss partials: dy = 2 (y - t) / n_total; part[blk] = sum (y - t)^2
__global__ void k_loss(const float *y, const float *t, float *dy, int64_t n, float inv_n, float *part)
{
    __shared__ float sh[NB];
    const int64_t e = blockIdx.x * (int64_t)NB + threadIdx.x;
    float s = 0.0f;
    if (e < n) {
        const float d = y[e] - t[e];
        dy[e] = 2.0f * d * inv_n;
        s = d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = NB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// grid-unit Poisson operator S1 on the unknown grid (mx x my) inside the
// padded image: cx = 1, cy = cyr; wrap on a periodic axis, zero otherwise
struct EGeo {
    int B, H, W, mx, my, perx, pery;
    float cyr;
};
__device__ __forceinline__ float s1_at(const float *d, const EGeo &e, int x, int y)
{
    const float c = d[x + y * e.W];
    const float l = x > 0 ? d[x - 1 + y * e.W] : (e.perx ? d[e.mx - 1 + y * e.W] : 0.0f);
    const float r = x < e.mx - 1 ? d[x + 1 + y * e.W] : (e.perx ? d[y * e.W] : 0.0f);
    const float dn = y > 0 ? d[x + (y - 1) * e.W] : (e.pery ? d[x + (e.my - 1) * e.W] : 0.0f);
    const float up = y < e.my - 1 ? d[x + (y + 1) * e.W] : (e.pery ? d[x] : 0.0f);
    return ((2.0f * c - l) - r) + e.cyr * ((2.0f * c - dn) - up);
}
// d = (y - t) on the unknown region, 0 in the padding
__global__ void k_diff_mask(const float *y, const float *t, float *d, EGeo e)
{
    const int64_t k = blockIdx.x * (int64_t)NB + threadIdx.x;
    if (k >= (int64_t)e.B * e.H * e.W) return;
    const int x = (int)(k % e.W), yy = (int)((k / e.W) % e.H);
    d[k] = (x < e.mx && yy < e.my) ? y[k] - t[k] : 0.0f;
}
// energy loss: dy = 2 w_b S1 d, part[blk] = sum w_b d S1 d, w_b per image
__global__ void k_energy_grad(const float *d, float *dy, EGeo e, const float *wb, float *part)
{
    __shared__ float sh[NB];
    const int64_t k = blockIdx.x * (int64_t)NB + threadIdx.x;
    float acc = 0.0f;
    if (k < (int64_t)e.B * e.H * e.W) {
        const int x = (int)(k % e.W), yy = (int)((k / e.W) % e.H), b = (int)(k / ((int64_t)e.W * e.H));
        float g = 0.0f;
        if (x < e.mx && yy < e.my) {
            const float sd = s1_at(d + (int64_t)b * e.H * e.W, e, x, yy);
            g = 2.0f * sd * wb[b];
            acc = d[k] * sd * wb[b];
        }
        dy[k] = g;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = NB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ void k_adam(float *p, const float *gr, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const int64_t e = blockIdx.x * (int64_t)NB + threadIdx.x;
    if (e >= n) return;
    const float gg = gr[e];
    const float mm = b1 * m[e] + (1.0f - b1) * gg;
    const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}
// BF16 copies: wb[co][t][ci] = w, wt[ci][8 - t][co] = w (the data gradient's kernel)
__global__ void k_wprep(const float *w, bf16s *wb, bf16s *wt, int CO, int CI)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= CO * 9 * CI) return;
    const int co = e / (9 * CI), rem = e % (9 * CI), t = rem / CI, ci = rem % CI;
    const bf16s h = f2b(w[e]);
    wb[e] = h;
    wt[((int64_t)ci * 9 + (8 - t)) * CO + co] = h;
}

// per-sample normalisation of a right-hand side: out = (in - mean) / rms of
// the unknown grid (mx x my) placed in the padded H x W image (zeros outside);
// Statistics of an unknown-grid vector v (mx x my) in two passes over
// NSB blocks, partials added in a fixed order:
//   stats[0] = mean (0 unless the problem is singular)
//   stats[1] = rms after the mean removal (kind 0), or sqrt(v' S1 v) (kind 1;
//              S1 is blind to the constant, so no mean removal is needed)
constexpr int NSB = 256;
__device__ double block_sum(double s, double *sh)
{
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = NB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    return sh[0];
}
// part[blk] = sum of v over the block's contiguous chunk
__global__ __launch_bounds__(NB) void k_vsum(const float *v, int64_t n, double *part)
{
    __shared__ double sh[NB];
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x, p0 = blockIdx.x * chunk;
    const int64_t p1 = std::min(n, p0 + chunk);
    double s = 0.0;
    for (int64_t k = p0 + threadIdx.x; k < p1; k += NB) s += v[k];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// part[blk] = sum over the chunk of (v - stats[0])^2 (kind 0) or v (S1 v) (kind 1)
__global__ __launch_bounds__(NB) void k_vsq(const float *v, EGeo e, const double *stats, int kind, double *part)
{
    __shared__ double sh[NB];
    const int64_t n = (int64_t)e.mx * e.my;
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x, p0 = blockIdx.x * chunk;
    const int64_t p1 = std::min(n, p0 + chunk);
    const double mean = stats[0];
    EGeo u = e;
    u.W = e.mx;
    double q = 0.0;
    for (int64_t k = p0 + threadIdx.x; k < p1; k += NB) {
        if (kind == 0) {
            const double d = v[k] - mean;
            q += d * d;
        } else {
            const int x = (int)(k % e.mx), y = (int)(k / e.mx);
            q += (double)v[k] * (double)s1_at(v, u, x, y);
        }
    }
    q = block_sum(q, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = q;
}
// stats[slot] from np partials: op 0 the mean (singular ? s / n : 0),
// op 1 sqrt(s / n), op 2 sqrt(max(s, 0))
__global__ __launch_bounds__(NB) void k_vfin(const double *part, int np, int64_t n, int op, int singular, double *stats,
                                             int slot)
{
    __shared__ double sh[NB];
    double s = 0.0;
    for (int k = threadIdx.x; k < np; k += NB) s += part[k];
    s = block_sum(s, sh);
    if (threadIdx.x == 0)
        stats[slot] = op == 0 ? (singular ? s / (double)n : 0.0) : op == 1 ? sqrt(s / (double)n) : sqrt(fmax(s, 0.0));
}
// img (H x W, zero padded) = (v - stats[0]) / stats[1] on the mx x my grid
__global__ void k_embed(const float *in, int mx, int my, const double *stats, float *img, int H, int W)
{
    const int64_t e = blockIdx.x * (int64_t)NB + threadIdx.x;
    if (e >= (int64_t)H * W) return;
    const int x = (int)(e % W), y = (int)(e / W);
    float v = 0.0f;
    if (x < mx && y < my) {
        const double r = stats[1] > 0.0 ? stats[1] : 1.0;
        v = (float)((in[x + (int64_t)y * mx] - stats[0]) / r);
    }
    img[e] = v;
}
__global__ void k_extract(const float *img, int H, int W, int mx, int my, float *out)
{
    const int64_t e = blockIdx.x * (int64_t)NB + threadIdx.x;
    if (e >= (int64_t)mx * my) return;
    const int x = (int)(e % mx), y = (int)(e / mx);
    out[e] = img[x + (int64_t)y * W];
    (void)H;
}
__global__ void k_d2f(const double *a, float *b, int64_t n)
{
    const int64_t e = blockIdx.x * (int64_t)NB + threadIdx.x;
    if (e < n) b[e] = (float)a[e];
}

inline int g1(int64_t n) { return (int)((n + NB - 1) / NB); }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct gp_nn {
    int c0 = 32;
    int mx = 0, my = 0, H = 0, W = 0;     // unknown grid and padded image
    int perx = 0, pery = 0, singular = 0;
    float cyr = 1.0f;                     // (1/dy^2) / (1/dx^2) of the unknown grid
    // layers: 0 in, 1..6 W1..W6, 7 out
    int ci[NL], co[NL], lvl[NL];
    int64_t woff[NL], boff[NL], nparam = 0;   // offsets into the FP32 vectors
    int64_t bfoff[NL], nbf = 0;               // offsets into the BF16 weight buffers
    float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
    bf16s *WB = nullptr, *WT = nullptr;
    int64_t adam_t = 0;
    // activations for a batch of up to bcap images of up to hw_cap pixels
    // (sized for the largest geometry met, so that switching between the
    // Poisson and Darwin hierarchies reallocates nothing)
    int bcap = 0;
    int64_t hw_cap = 0, m_cap = 0;
    float *x_in = nullptr, *y_out = nullptr, *tgt = nullptr, *dy = nullptr, *dtmp = nullptr;
    bf16s *a0 = nullptr, *p1 = nullptr, *a1 = nullptr, *p2 = nullptr, *a2 = nullptr, *p3 = nullptr, *a3 = nullptr;
    bf16s *g2 = nullptr, *u2 = nullptr, *g1 = nullptr, *u1 = nullptr, *g0 = nullptr, *u0 = nullptr;
    bf16s *d0a = nullptr, *d0b = nullptr, *d1a = nullptr, *d1b = nullptr, *d1c = nullptr, *d2a = nullptr, *d2b = nullptr,
          *d2c = nullptr, *d3a = nullptr, *d3b = nullptr;
    float *part = nullptr;
    int64_t part_n = 0;
    float *lpart = nullptr, *hloss = nullptr;
    double *stats = nullptr;
    // training set on the device: nsamp normalised (input, target) images
    std::vector<float *> ds_in, ds_tgt;
    std::vector<double> ds_E;      // target energy t' S1 t (t normalised to unit rms)
    int loss_kind = 1;             // training loss: 0 mean square, 1 relative energy norm
    int smooth = 2;                // red-black sweeps after the network in the preconditioner
    int fcg_cap = 0;               // flexible-CG iteration cap (0: the last plain MG-PCG count)
    float *bw = nullptr;           // per-image loss weights of the current batch
    std::vector<float> hbw;
    float *tmpf = nullptr;   // mx*my FP32 staging
    // time-loop increments, per stream (0 Poisson, 1..3 Darwin chi, Ax, Ay):
    // the previous solve's b and x; inc_b, inc_x hold the differences
    double *prev_b[4] = {}, *prev_x[4] = {}, *inc_b = nullptr, *inc_x = nullptr;
    int64_t prev_n[4] = {}, inc_n = 0;
    int have_prev[4] = {};
    int train_sys = 0;   // the hierarchy of the training set: 0 Poisson, 1 Darwin
};

namespace {

void nn_free_batch(gp_nn *n)
{
    bf16s *bs[] = {n->a0, n->p1, n->a1, n->p2, n->a2, n->p3, n->a3, n->g2, n->u2, n->g1, n->u1, n->g0, n->u0,
                   n->d0a, n->d0b, n->d1a, n->d1b, n->d1c, n->d2a, n->d2b, n->d2c, n->d3a, n->d3b};
    for (bf16s *p : bs) if (p) (void)hipFree(p);
    float *fs[] = {n->x_in, n->y_out, n->tgt, n->dy, n->dtmp, n->part};
    for (float *p : fs) if (p) (void)hipFree(p);
    n->dtmp = nullptr;
    n->a0 = n->p1 = n->a1 = n->p2 = n->a2 = n->p3 = n->a3 = n->g2 = n->u2 = n->g1 = n->u1 = n->g0 = n->u0 = nullptr;
    n->d0a = n->d0b = n->d1a = n->d1b = n->d1c = n->d2a = n->d2b = n->d2c = n->d3a = n->d3b = nullptr;
    n->x_in = n->y_out = n->tgt = n->dy = n->part = nullptr;
    n->bcap = 0;
    n->hw_cap = 0;
}

int nn_alloc_batch(gp_nn *n, int B)
{
    if (B <= n->bcap && (int64_t)n->H * n->W <= n->hw_cap) return GP_OK;
    nn_free_batch(n);
    const int64_t P0 = (int64_t)B * n->H * n->W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    const int c0 = n->c0, c1 = 2 * c0, c2 = 4 * c0, c3 = 4 * c0;
    auto al = [&](bf16s **p, int64_t cnt) { return hipMalloc((void **)p, cnt * sizeof(bf16s)); };
    GP_HIP(hipMalloc((void **)&n->x_in, P0 * 4));
    GP_HIP(hipMalloc((void **)&n->y_out, P0 * 4));
    GP_HIP(hipMalloc((void **)&n->tgt, P0 * 4));
    GP_HIP(hipMalloc((void **)&n->dy, P0 * 4));
    GP_HIP(hipMalloc((void **)&n->dtmp, P0 * 4));
    if (n->bw) (void)hipFree(n->bw);
    GP_HIP(hipMalloc((void **)&n->bw, (B + 16) * 4));
    GP_HIP(al(&n->a0, P0 * c0)); GP_HIP(al(&n->u0, P0 * c0));
    GP_HIP(al(&n->p1, P1 * c0)); GP_HIP(al(&n->a1, P1 * c1)); GP_HIP(al(&n->g0, P1 * c0)); GP_HIP(al(&n->u1, P1 * c1));
    GP_HIP(al(&n->p2, P2 * c1)); GP_HIP(al(&n->a2, P2 * c2)); GP_HIP(al(&n->g1, P2 * c1)); GP_HIP(al(&n->u2, P2 * c2));
    GP_HIP(al(&n->p3, P3 * c2)); GP_HIP(al(&n->a3, P3 * c3)); GP_HIP(al(&n->g2, P3 * c3));
    // gradient scratch per level (largest channel count at the level)
    GP_HIP(al(&n->d0a, P0 * c0)); GP_HIP(al(&n->d0b, P0 * c0));
    GP_HIP(al(&n->d1a, P1 * c1)); GP_HIP(al(&n->d1b, P1 * c1)); GP_HIP(al(&n->d1c, P1 * c1));
    GP_HIP(al(&n->d2a, P2 * c2)); GP_HIP(al(&n->d2b, P2 * c2)); GP_HIP(al(&n->d2c, P2 * c2));
    GP_HIP(al(&n->d3a, P3 * c3)); GP_HIP(al(&n->d3b, P3 * c3));
    // weight-gradient partials: chunks x 9 x CO x CI for the largest layer
    const int64_t nchunk = 64;
    n->part_n = nchunk * 9 * (int64_t)c3 * c3 + 4096 * 9 * c0;
    GP_HIP(hipMalloc((void **)&n->part, n->part_n * 4));
    if (n->lpart) { (void)hipFree(n->lpart); n->lpart = nullptr; }
    GP_HIP(hipMalloc((void **)&n->lpart, ((P0 + NB - 1) / NB + 16) * 4));
    n->bcap = B;
    n->hw_cap = (int64_t)n->H * n->W;
    return GP_OK;
}

Geo geo(const gp_nn *n, int B, int lvl)
{
    Geo g;
    g.B = B; g.H = n->H >> lvl; g.W = n->W >> lvl;
    // a periodic axis keeps wrap padding only while the level is the exact
    // coarsening of the periodic grid (no zero padding was added)
    g.perx = n->perx && (n->W == n->mx);
    g.pery = n->pery && (n->H == n->my);
    return g;
}

template <int CI, int CO>
int conv_t(hipStream_t st, const bf16s *in, const bf16s *w, const float *bias, const bf16s *mask, bf16s *out, Geo g,
           int relu)
{
    dim3 grid((g.W + 7) / 8, (g.H + 7) / 8, g.B);
    k_conv<CI, CO><<<grid, 256, 0, st>>>(in, w, bias, mask, out, g, relu);
    GP_HIP(hipGetLastError());
    return GP_OK;
}
int conv(hipStream_t st, int CI, int CO, const bf16s *in, const bf16s *w, const float *bias, const bf16s *mask, bf16s *out,
         Geo g, int relu)
{
#define GP_CONV_CASE(a, b) if (CI == a && CO == b) return conv_t<a, b>(st, in, w, bias, mask, out, g, relu);
    GP_CONV_CASE(32, 64) GP_CONV_CASE(64, 128) GP_CONV_CASE(128, 128) GP_CONV_CASE(128, 64) GP_CONV_CASE(64, 32)
#undef GP_CONV_CASE
    return gp_fail(GP_ERR_UNSUPPORTED, "convolution shape " + std::to_string(CI) + " -> " + std::to_string(CO) + " not instantiated");
}
template <int CI, int CO>
int dw_t(hipStream_t st, const bf16s *d, const bf16s *x, float *part, int64_t part_n, float *grad, Geo g)
{
    const int64_t npix = (int64_t)g.B * g.H * g.W;
    int nchunk = (int)std::min<int64_t>(64, (npix + 1023) / 1024);
    if ((int64_t)nchunk * 9 * CO * CI > part_n) return gp_fail(GP_ERR_STATE, "weight-gradient scratch too small");
    int64_t chunk = (npix + nchunk - 1) / nchunk;
    chunk = (chunk + 31) / 32 * 32;
    nchunk = (int)((npix + chunk - 1) / chunk);
    k_dw<CI, CO><<<dim3(9, nchunk), 256, 0, st>>>(d, x, part, g, npix, chunk);
    k_dw_reduce<<<gp_cdiv(9 * CO * CI, NB), NB, 0, st>>>(part, nchunk, CO, CI, grad);
    GP_HIP(hipGetLastError());
    return GP_OK;
}
int dw(hipStream_t st, int CI, int CO, const bf16s *d, const bf16s *x, float *part, int64_t part_n, float *grad, Geo g)
{
#define GP_DW_CASE(a, b) if (CI == a && CO == b) return dw_t<a, b>(st, d, x, part, part_n, grad, g);
    GP_DW_CASE(32, 64) GP_DW_CASE(64, 128) GP_DW_CASE(128, 128) GP_DW_CASE(128, 64) GP_DW_CASE(64, 32)
#undef GP_DW_CASE
    return gp_fail(GP_ERR_UNSUPPORTED, "weight gradient shape " + std::to_string(CI) + " -> " + std::to_string(CO) + " not instantiated");
}
// out[0 .. n) = column sums of the nblk x n partials (fixed order)
void part_reduce(hipStream_t st, const float *part, int nblk, int n, float *out)
{
    if (n == 1) k_part_sum1<<<1, NB, 0, st>>>(part, nblk, out);
    else k_part_reduce<<<gp_cdiv(n, 32), NB, 0, st>>>(part, nblk, n, out);
}
int bias_grad(hipStream_t st, const bf16s *d, int64_t npix, int C, float *part, float *out)
{
    if (C % 8 || NB % (C / 8)) return gp_fail(GP_ERR_UNSUPPORTED, "bias gradient of " + std::to_string(C) + " channels");
    const int nblk = (int)std::min<int64_t>(1024, (npix + 255) / 256);
    const int64_t chunk = (npix + nblk - 1) / nblk;
    k_colsum<<<nblk, NB, 0, st>>>(d, npix, C, chunk, part);
    part_reduce(st, part, nblk, C, out);
    GP_HIP(hipGetLastError());
    return GP_OK;
}
// weight gradient of a one-channel convolution at level 0 (k_dw1) into grad
int dw1(hipStream_t st, const gp_nn *n, const bf16s *A, const float *v, const Geo &g, int sg, int ct, float *grad)
{
    const int64_t ntile = (int64_t)g.B * ((g.H + 7) / 8) * ((g.W + 31) / 32);
    const int nblk = (int)std::min<int64_t>(2048, ntile);
    if ((int64_t)nblk * 9 * 32 > n->part_n) return gp_fail(GP_ERR_STATE, "weight-gradient scratch too small");
    k_dw1<32><<<nblk, 9 * 32, 0, st>>>(A, v, g, sg, ct, n->part);
    part_reduce(st, n->part, nblk, 9 * 32, grad);
    GP_HIP(hipGetLastError());
    return GP_OK;
}
// launch shapes of the elementwise kernels (one thread per 8 channels)
void pool(hipStream_t st, const bf16s *in, bf16s *out, const Geo &g, int C)
{
    k_pool<<<g1((int64_t)g.B * (g.H / 2) * (g.W / 2) * (C / 8)), NB, 0, st>>>(in, out, g, C);
}
void up_add(hipStream_t st, const bf16s *gin, const bf16s *skip, bf16s *out, const Geo &g, int C)
{
    k_up_add<<<g1((int64_t)g.B * g.H * g.W * (C / 8)), NB, 0, st>>>(gin, skip, out, g, C);
}
void sum2x2_mask(hipStream_t st, const bf16s *du, const bf16s *act, bf16s *dpre, const Geo &g, int C)
{
    k_sum2x2_mask<<<g1((int64_t)g.B * (g.H / 2) * (g.W / 2) * (C / 8)), NB, 0, st>>>(du, act, dpre, g, C);
}
void skip_pool_mask(hipStream_t st, const bf16s *dskip, const bf16s *dpool, const bf16s *act, bf16s *dpre,
                    const Geo &g, int C)
{
    k_skip_pool_mask<<<g1((int64_t)g.B * g.H * g.W * (C / 8)), NB, 0, st>>>(dskip, dpool, act, dpre, g, C);
}

void nn_layout(gp_nn *n)
{
    const int c0 = n->c0;
    const int cis[NL] = {1, c0, 2 * c0, 4 * c0, 4 * c0, 4 * c0, 2 * c0, c0};
    const int cos_[NL] = {c0, 2 * c0, 4 * c0, 4 * c0, 4 * c0, 2 * c0, c0, 1};
    const int lv[NL] = {0, 1, 2, 3, 3, 2, 1, 0};
    int64_t o = 0, ob = 0;
    for (int l = 0; l < NL; l++) {
        n->ci[l] = cis[l]; n->co[l] = cos_[l]; n->lvl[l] = lv[l];
        n->woff[l] = o; o += (int64_t)cos_[l] * 9 * cis[l];
        n->boff[l] = o; o += cos_[l];
        n->bfoff[l] = ob; ob += (l == 0 || l == NL - 1) ? 0 : (int64_t)cos_[l] * 9 * cis[l];
    }
    n->nparam = o;
    n->nbf = ob;
}

int nn_prep_weights(hipStream_t st, gp_nn *n)
{
    for (int l = 1; l < NL - 1; l++) {
        const int e = n->co[l] * 9 * n->ci[l];
        k_wprep<<<gp_cdiv(e, NB), NB, 0, st>>>(n->P + n->woff[l], n->WB + n->bfoff[l], n->WT + n->bfoff[l], n->co[l],
                                              n->ci[l]);
    }
    GP_HIP(hipGetLastError());
    return GP_OK;
}

// forward of B images already in n->x_in; output n->y_out (padded images)
int nn_forward(hipStream_t st, gp_nn *n, int B, bool infer = false)
{
    const int c0 = n->c0, c1 = 2 * c0, c2 = 4 * c0;
    const Geo G0 = geo(n, B, 0), G1 = geo(n, B, 1), G2 = geo(n, B, 2), G3 = geo(n, B, 3);
    const int64_t P0 = (int64_t)B * n->H * n->W;
    float *P = n->P;
    auto WBl = [&](int l) { return n->WB + n->bfoff[l]; };
    int rc;
    if (infer) {
        k_in_pool<32><<<g1(P0 / 4), NB, 0, st>>>(n->x_in, P + n->woff[0], P + n->boff[0], n->p1, G0);
    } else {
        k_conv_in<32><<<g1(P0), NB, 0, st>>>(n->x_in, P + n->woff[0], P + n->boff[0], n->a0, G0);
        pool(st, n->a0, n->p1, G0, c0);
    }
    if ((rc = conv(st, c0, c1, n->p1, WBl(1), P + n->boff[1], nullptr, n->a1, G1, 1))) return rc;
    pool(st, n->a1, n->p2, G1, c1);
    if ((rc = conv(st, c1, c2, n->p2, WBl(2), P + n->boff[2], nullptr, n->a2, G2, 1))) return rc;
    pool(st, n->a2, n->p3, G2, c2);
    if ((rc = conv(st, c2, c2, n->p3, WBl(3), P + n->boff[3], nullptr, n->a3, G3, 1))) return rc;
    if ((rc = conv(st, c2, c2, n->a3, WBl(4), P + n->boff[4], nullptr, n->g2, G3, 1))) return rc;
    up_add(st, n->g2, n->a2, n->u2, G2, c2);
    if ((rc = conv(st, c2, c1, n->u2, WBl(5), P + n->boff[5], nullptr, n->g1, G2, 1))) return rc;
    up_add(st, n->g1, n->a1, n->u1, G1, c1);
    if ((rc = conv(st, c1, c0, n->u1, WBl(6), P + n->boff[6], nullptr, n->g0, G1, 1))) return rc;
    if (infer) {
        k_up_out<32><<<dim3((G0.W + 13) / 14, (G0.H + 13) / 14, B), 256, 0, st>>>(
            n->x_in, n->g0, P + n->woff[0], P + n->boff[0], P + n->woff[7], P + n->boff[7], n->y_out, G0);
    } else {
        up_add(st, n->g0, n->a0, n->u0, G0, c0);
        k_conv_out<32><<<g1(P0), NB, 0, st>>>(n->u0, P + n->woff[7], P + n->boff[7], n->y_out, G0);
    }
    GP_HIP(hipGetLastError());
    return GP_OK;
}

// backward from n->dy (FP32, dL/dy); fills n->G
int nn_backward(hipStream_t st, gp_nn *n, int B)
{
    const int c0 = n->c0, c1 = 2 * c0, c2 = 4 * c0;
    const Geo G0 = geo(n, B, 0), G1 = geo(n, B, 1), G2 = geo(n, B, 2), G3 = geo(n, B, 3);
    const int64_t P0 = (int64_t)B * n->H * n->W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    float *Gr = n->G;
    auto WTl = [&](int l) { return n->WT + n->bfoff[l]; };
    int rc;
    // conv_out
    {
        const int nblk = (int)std::min<int64_t>(2048, (P0 + 255) / 256);
        const int64_t chunk = (P0 + nblk - 1) / nblk;
        if ((rc = dw1(st, n, n->u0, n->dy, G0, -1, 1, Gr + n->woff[7]))) return rc;
        k_colsum_f<<<nblk, NB, 0, st>>>(n->dy, P0, chunk, n->part);
        part_reduce(st, n->part, nblk, 1, Gr + n->boff[7]);
        k_conv_out_dx<32><<<g1(P0), NB, 0, st>>>(n->dy, n->P + n->woff[7], n->d0a, G0);   // d_u0
    }
    // u0 = up(g0) + a0 -> dpre6 at level 1
    sum2x2_mask(st, n->d0a, n->g0, n->d1a, G0, c0);
    if ((rc = dw(st, c1, c0, n->d1a, n->u1, n->part, n->part_n, Gr + n->woff[6], G1))) return rc;
    if ((rc = bias_grad(st, n->d1a, P1, c0, n->part, Gr + n->boff[6]))) return rc;
    if ((rc = conv(st, c0, c1, n->d1a, WTl(6), nullptr, nullptr, n->d1b, G1, 0))) return rc;   // d_u1
    // u1 = up(g1) + a1 -> dpre5 at level 2
    sum2x2_mask(st, n->d1b, n->g1, n->d2a, G1, c1);
    if ((rc = dw(st, c2, c1, n->d2a, n->u2, n->part, n->part_n, Gr + n->woff[5], G2))) return rc;
    if ((rc = bias_grad(st, n->d2a, P2, c1, n->part, Gr + n->boff[5]))) return rc;
    if ((rc = conv(st, c1, c2, n->d2a, WTl(5), nullptr, nullptr, n->d2b, G2, 0))) return rc;   // d_u2
    // u2 = up(g2) + a2 -> dpre4 at level 3
    sum2x2_mask(st, n->d2b, n->g2, n->d3a, G2, c2);
    if ((rc = dw(st, c2, c2, n->d3a, n->a3, n->part, n->part_n, Gr + n->woff[4], G3))) return rc;
    if ((rc = bias_grad(st, n->d3a, P3, c2, n->part, Gr + n->boff[4]))) return rc;
    // d_a3 masked by a3 > 0 -> dpre3
    if ((rc = conv(st, c2, c2, n->d3a, WTl(4), nullptr, n->a3, n->d3b, G3, 0))) return rc;
    if ((rc = dw(st, c2, c2, n->d3b, n->p3, n->part, n->part_n, Gr + n->woff[3], G3))) return rc;
    if ((rc = bias_grad(st, n->d3b, P3, c2, n->part, Gr + n->boff[3]))) return rc;
    if ((rc = conv(st, c2, c2, n->d3b, WTl(3), nullptr, nullptr, n->d3a, G3, 0))) return rc;   // d_p3
    // a2: skip (d_u2) + pool path -> dpre2
    skip_pool_mask(st, n->d2b, n->d3a, n->a2, n->d2c, G2, c2);
    if ((rc = dw(st, c1, c2, n->d2c, n->p2, n->part, n->part_n, Gr + n->woff[2], G2))) return rc;
    if ((rc = bias_grad(st, n->d2c, P2, c2, n->part, Gr + n->boff[2]))) return rc;
    if ((rc = conv(st, c2, c1, n->d2c, WTl(2), nullptr, nullptr, n->d2a, G2, 0))) return rc;   // d_p2
    // a1: skip (d_u1) + pool path -> dpre1
    skip_pool_mask(st, n->d1b, n->d2a, n->a1, n->d1c, G1, c1);
    if ((rc = dw(st, c0, c1, n->d1c, n->p1, n->part, n->part_n, Gr + n->woff[1], G1))) return rc;
    if ((rc = bias_grad(st, n->d1c, P1, c1, n->part, Gr + n->boff[1]))) return rc;
    if ((rc = conv(st, c1, c0, n->d1c, WTl(1), nullptr, nullptr, n->d1a, G1, 0))) return rc;   // d_p1
    // a0: skip (d_u0) + pool path -> dpre0
    skip_pool_mask(st, n->d0a, n->d1a, n->a0, n->d0b, G0, c0);
    if ((rc = dw1(st, n, n->d0b, n->x_in, G0, 1, 0, Gr + n->woff[0]))) return rc;
    if ((rc = bias_grad(st, n->d0b, P0, c0, n->part, Gr + n->boff[0]))) return rc;
    GP_HIP(hipGetLastError());
    return GP_OK;
}

EGeo egeo(const gp_nn *n, int B)
{
    EGeo e;
    e.B = B; e.H = n->H; e.W = n->W; e.mx = n->mx; e.my = n->my;
    e.perx = n->perx; e.pery = n->pery; e.cyr = n->cyr;
    return e;
}

// loss of the current batch (targets in n->tgt); fills n->dy.  kind 0: mean
// square over the images; kind 1: energy norm in S1, mean over the batch
int nn_loss(hipStream_t st, gp_nn *n, int B, int kind, double *loss)
{
    const int64_t P0 = (int64_t)B * n->H * n->W;
    const int nblk = g1(P0);
    if (kind == 0) {
        k_loss<<<nblk, NB, 0, st>>>(n->y_out, n->tgt, n->dy, P0, 1.0f / (float)P0, n->lpart);
    } else {
        const EGeo e = egeo(n, B);
        k_diff_mask<<<nblk, NB, 0, st>>>(n->y_out, n->tgt, n->dtmp, e);
        if ((int)n->hbw.size() < B) n->hbw.assign(B, 1.0f / (float)B);
        GP_HIP(hipMemcpyAsync(n->bw, n->hbw.data(), B * 4, hipMemcpyHostToDevice, st));
        k_energy_grad<<<nblk, NB, 0, st>>>(n->dtmp, n->dy, e, n->bw, n->lpart);
    }
    float *sum = n->lpart + nblk;
    part_reduce(st, n->lpart, nblk, 1, sum);
    GP_HIP(hipGetLastError());
    GP_HIP(hipMemcpyAsync(n->hloss, sum, 4, hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    *loss = (double)n->hloss[0] / (kind == 0 ? (double)P0 : 1.0);
    return GP_OK;
}

uint64_t sm64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// stats[0..1] of v (k_vsum .. k_vfin); the partials use n->stats[64 ..]
void vstats(hipStream_t st, gp_nn *n, const float *v, int kind, double *stats)
{
    const int64_t m = (int64_t)n->mx * n->my;
    double *part = n->stats + 64;
    k_vsum<<<NSB, NB, 0, st>>>(v, m, part);
    k_vfin<<<1, NB, 0, st>>>(part, NSB, m, 0, n->singular, stats, 0);
    k_vsq<<<NSB, NB, 0, st>>>(v, egeo(n, 1), stats, kind, part);
    k_vfin<<<1, NB, 0, st>>>(part, NSB, m, kind == 0 ? 1 : 2, n->singular, stats, 1);
}

// normalise one unknown-grid vector (device FP32, mx*my) into image slot
// img: by its rms (kind 0, network inputs) or by its S1 energy norm (kind 1,
// training targets); the mean is removed first when the problem is singular
int nn_embed(hipStream_t st, gp_nn *n, const float *v, float *img, double *stats_dev, int kind = 0)
{
    vstats(st, n, v, kind, stats_dev);
    k_embed<<<g1((int64_t)n->H * n->W), NB, 0, st>>>(v, n->mx, n->my, stats_dev, img, n->H, n->W);
    GP_HIP(hipGetLastError());
    return GP_OK;
}

// take the unknown grid of solver s (level 0), padded to a multiple of 8
int nn_set_grid(gp_nn *n, const gp_solver &s)
{
    const int mx = s.lv[0].mx, my = s.lv[0].my;
    const int W = (mx + 7) / 8 * 8, H = (my + 7) / 8 * 8;
    n->perx = s.lv[0].perx == GP_AX_PER;
    n->pery = s.lv[0].pery == GP_AX_PER;
    n->singular = s.singular;
    n->cyr = (float)(s.lv[0].hy2i / s.lv[0].hx2i);
    if (mx == n->mx && my == n->my && n->tmpf) return GP_OK;
    n->mx = mx; n->my = my; n->W = W; n->H = H;
    if ((int64_t)H * W > n->hw_cap) nn_free_batch(n);   // the next nn_alloc_batch sizes them anew
    if ((int64_t)mx * my > n->m_cap) {
        if (n->tmpf) (void)hipFree(n->tmpf);
        n->tmpf = nullptr;
        n->m_cap = 0;
        GP_HIP(hipMalloc((void **)&n->tmpf, (int64_t)mx * my * 4));
        n->m_cap = (int64_t)mx * my;
    }
    return GP_OK;
}

}  // namespace

int gp_nn_create(gp_ctx *c, int base_ch, uint64_t seed)
{
    if (base_ch != 32) return gp_fail(GP_ERR_UNSUPPORTED, "base_channels must be 32 (kernels are instantiated for 32/64/128)");
    // the training set survives a re-initialisation of the weights
    std::vector<float *> keep_in, keep_tgt;
    std::vector<double> keep_E;
    if (c->nn) {
        keep_in.swap(c->nn->ds_in);
        keep_tgt.swap(c->nn->ds_tgt);
        keep_E.swap(c->nn->ds_E);
    }
    gp_nn_destroy(c);
    gp_nn *n = new gp_nn();
    c->nn = n;
    n->ds_in.swap(keep_in);
    n->ds_tgt.swap(keep_tgt);
    n->ds_E.swap(keep_E);
    n->c0 = base_ch;
    int rc0 = nn_set_grid(n, c->ps);
    if (rc0) return rc0;
    nn_layout(n);
    GP_HIP(hipMalloc((void **)&n->P, n->nparam * 4));
    GP_HIP(hipMalloc((void **)&n->G, n->nparam * 4));
    GP_HIP(hipMalloc((void **)&n->M, n->nparam * 4));
    GP_HIP(hipMalloc((void **)&n->V, n->nparam * 4));
    GP_HIP(hipMemset(n->M, 0, n->nparam * 4));
    GP_HIP(hipMemset(n->V, 0, n->nparam * 4));
    GP_HIP(hipMalloc((void **)&n->WB, std::max<int64_t>(n->nbf, 1) * 2));
    GP_HIP(hipMalloc((void **)&n->WT, std::max<int64_t>(n->nbf, 1) * 2));
    GP_HIP(hipMalloc((void **)&n->stats, (64 + NSB) * sizeof(double)));   // slots, then partials
    GP_HIP(hipHostMalloc((void **)&n->hloss, 64, hipHostMallocDefault));
    // He-uniform initialisation from a splitmix64 stream, biases zero
    std::vector<float> h(n->nparam, 0.0f);
    uint64_t st = seed;
    for (int l = 0; l < NL; l++) {
        const double fan_in = 9.0 * n->ci[l];
        const double lim = std::sqrt(6.0 / fan_in) * (l == NL - 1 ? 0.5 : 1.0);
        for (int64_t k = 0; k < (int64_t)n->co[l] * 9 * n->ci[l]; k++) {
            const double u = (double)(sm64(st) >> 11) * 0x1p-53;
            h[n->woff[l] + k] = (float)((2.0 * u - 1.0) * lim);
        }
    }
    GP_HIP(hipMemcpy(n->P, h.data(), n->nparam * 4, hipMemcpyHostToDevice));
    int rc = nn_prep_weights(c->stream, n);
    if (rc) return rc;
    GP_HIP(hipStreamSynchronize(c->stream));
    return GP_OK;
}

void gp_nn_destroy(gp_ctx *c)
{
    gp_nn *n = c->nn;
    if (!n) return;
    nn_free_batch(n);
    void *ps[] = {n->P, n->G, n->M, n->V, n->WB, n->WT, n->stats, n->lpart, n->tmpf, n->bw, n->inc_b, n->inc_x,
                  n->prev_b[0], n->prev_b[1], n->prev_b[2], n->prev_b[3],
                  n->prev_x[0], n->prev_x[1], n->prev_x[2], n->prev_x[3]};
    for (void *p : ps) if (p) (void)hipFree(p);
    for (float *p : n->ds_in) (void)hipFree(p);
    for (float *p : n->ds_tgt) (void)hipFree(p);
    if (n->hloss) (void)hipHostFree(n->hloss);
    delete n;
    c->nn = nullptr;
}

// add one (b, x) pair (device FP64 unknown-grid vectors) to the training set
int gp_nn_add_pair_dev(gp_ctx *c, const double *b, const double *x)
{
    gp_nn *n = c->nn;
    hipStream_t st = c->stream;
    int rc0 = nn_set_grid(n, n->train_sys ? c->ds : c->ps);
    if (rc0) return rc0;
    const int64_t m = (int64_t)n->mx * n->my, HW = (int64_t)n->H * n->W;
    float *in = nullptr, *tg = nullptr;
    GP_HIP(hipMalloc((void **)&in, HW * 4));
    GP_HIP(hipMalloc((void **)&tg, HW * 4));
    k_d2f<<<g1(m), NB, 0, st>>>(b, n->tmpf, m);
    int rc = nn_embed(st, n, n->tmpf, in, n->stats);
    if (rc) return rc;
    k_d2f<<<g1(m), NB, 0, st>>>(x, n->tmpf, m);
    if ((rc = nn_embed(st, n, n->tmpf, tg, n->stats + 2, 0))) return rc;
    vstats(st, n, n->tmpf, 1, n->stats + 4);
    GP_HIP(hipGetLastError());
    double hs[6];
    GP_HIP(hipMemcpyAsync(hs, n->stats, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    const double E = hs[3] > 0.0 ? (hs[5] / hs[3]) * (hs[5] / hs[3]) : 1.0;
    n->ds_in.push_back(in);
    n->ds_tgt.push_back(tg);
    n->ds_E.push_back(E > 0.0 ? E : 1.0);
    return GP_OK;
}

namespace {
__global__ void k_sub_keep(const double *a, double *prev, double *d, int64_t n, int diff)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = a[i];
        if (diff) d[i] = v - prev[i];
        prev[i] = v;
    }
}
}  // namespace

// the time-loop increment (b_n - b_{n-1}, x_n - x_{n-1}) of stream k's last
// two solves on the device (b, x: the current ones, m unknowns); the first
// call of a stream only records b_n, x_n.  *added = 1 when a pair was appended.
int gp_nn_add_increment_dev(gp_ctx *c, int k, const double *b, const double *x, int64_t m, int *added)
{
    gp_nn *n = c->nn;
    hipStream_t st = c->stream;
    *added = 0;
    if (k < 0 || k > 3) return gp_fail(GP_ERR_ARG, "increment stream must be 0..3");
    if (n->inc_n < m) {
        if (n->inc_b) GP_HIP(hipFree(n->inc_b));
        if (n->inc_x) GP_HIP(hipFree(n->inc_x));
        n->inc_b = n->inc_x = nullptr;
        n->inc_n = 0;
        GP_HIP(hipMalloc((void **)&n->inc_b, m * sizeof(double)));
        GP_HIP(hipMalloc((void **)&n->inc_x, m * sizeof(double)));
        n->inc_n = m;
    }
    if (n->prev_n[k] != m) {
        if (n->prev_b[k]) GP_HIP(hipFree(n->prev_b[k]));
        if (n->prev_x[k]) GP_HIP(hipFree(n->prev_x[k]));
        n->prev_b[k] = n->prev_x[k] = nullptr;
        n->prev_n[k] = 0;
        n->have_prev[k] = 0;
        GP_HIP(hipMalloc((void **)&n->prev_b[k], m * sizeof(double)));
        GP_HIP(hipMalloc((void **)&n->prev_x[k], m * sizeof(double)));
        n->prev_n[k] = m;
    }
    const int g = (int)std::min<int64_t>(gp_cdiv(m, 256), 4096);
    k_sub_keep<<<g, 256, 0, st>>>(b, n->prev_b[k], n->inc_b, m, n->have_prev[k]);
    k_sub_keep<<<g, 256, 0, st>>>(x, n->prev_x[k], n->inc_x, m, n->have_prev[k]);
    GP_HIP(hipGetLastError());
    if (!n->have_prev[k]) {
        n->have_prev[k] = 1;
        return GP_OK;
    }
    int rc = gp_nn_add_pair_dev(c, n->inc_b, n->inc_x);
    if (rc) return rc;
    *added = 1;
    return GP_OK;
}

// the hierarchy the training pairs belong to (0 Poisson, 1 Darwin); a change
// drops the pairs held for the other one (the weights are kept)
int gp_nn_set_system_impl(gp_ctx *c, int sys)
{
    gp_nn *n = c->nn;
    if (!n) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    if (sys < 0 || sys > 1) return gp_fail(GP_ERR_ARG, "training system must be 0 (Poisson) or 1 (Darwin)");
    if (sys == 1 && c->cfg.em_model != 2) return gp_fail(GP_ERR_STATE, "Darwin training needs em_model = 2");
    if (sys == n->train_sys) return GP_OK;
    GP_HIP(hipStreamSynchronize(c->stream));
    for (float *p : n->ds_in) GP_HIP(hipFree(p));
    for (float *p : n->ds_tgt) GP_HIP(hipFree(p));
    n->ds_in.clear();
    n->ds_tgt.clear();
    n->ds_E.clear();
    n->train_sys = sys;
    return GP_OK;
}

int gp_nn_train_impl(gp_ctx *c, const double *rho, const double *phi, int nsamp, int epochs, int batch, float lr,
                     float *loss_out)
{
    gp_nn *n = c->nn;
    if (!n) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    hipStream_t st = c->stream;
    int rc0 = nn_set_grid(n, n->train_sys ? c->ds : c->ps);
    if (rc0) return rc0;
    const int64_t m = (int64_t)n->mx * n->my, HW = (int64_t)n->H * n->W;
    if (nsamp > 0) {
        if (!rho || !phi) return gp_fail(GP_ERR_ARG, "null training arrays");
        double *db = nullptr, *dx = nullptr;
        GP_HIP(hipMalloc((void **)&db, m * 8));
        GP_HIP(hipMalloc((void **)&dx, m * 8));
        for (int k = 0; k < nsamp; k++) {
            GP_HIP(hipMemcpyAsync(db, rho + k * m, m * 8, hipMemcpyHostToDevice, st));
            GP_HIP(hipMemcpyAsync(dx, phi + k * m, m * 8, hipMemcpyHostToDevice, st));
            int rc = gp_nn_add_pair_dev(c, db, dx);
            if (rc) { (void)hipFree(db); (void)hipFree(dx); return rc; }
        }
        GP_HIP(hipFree(db));
        GP_HIP(hipFree(dx));
    }
    const int ns = (int)n->ds_in.size();
    if (ns == 0) return gp_fail(GP_ERR_STATE, "no training pairs");
    if (batch < 1 || epochs < 0) return gp_fail(GP_ERR_ARG, "batch >= 1 and epochs >= 0");
    batch = std::min(batch, ns);
    int rc = nn_alloc_batch(n, batch);
    if (rc) return rc;
    std::vector<int> order(ns);
    for (int k = 0; k < ns; k++) order[k] = k;
    uint64_t rs = 0x5eed0000ull + (uint64_t)n->adam_t;
    double last = 0.0;
    for (int ep = 0; ep < epochs; ep++) {
        for (int k = ns - 1; k > 0; k--) std::swap(order[k], order[(int)(sm64(rs) % (uint64_t)(k + 1))]);
        double acc = 0.0;
        int nbat = 0;
        for (int s0 = 0; s0 + batch <= ns; s0 += batch) {
            n->hbw.assign(batch, 0.0f);
            for (int j = 0; j < batch; j++) {
                GP_HIP(hipMemcpyAsync(n->x_in + j * HW, n->ds_in[order[s0 + j]], HW * 4, hipMemcpyDeviceToDevice, st));
                GP_HIP(hipMemcpyAsync(n->tgt + j * HW, n->ds_tgt[order[s0 + j]], HW * 4, hipMemcpyDeviceToDevice, st));
                n->hbw[j] = (float)(1.0 / ((double)batch * n->ds_E[order[s0 + j]]));
            }
            if ((rc = nn_forward(st, n, batch))) return rc;
            double L = 0.0;
            if ((rc = nn_loss(st, n, batch, n->loss_kind, &L))) return rc;
            if ((rc = nn_backward(st, n, batch))) return rc;
            n->adam_t++;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.0f / (1.0f - std::pow(b1, (float)n->adam_t));
            const float c2 = 1.0f / (1.0f - std::pow(b2, (float)n->adam_t));
            k_adam<<<g1(n->nparam), NB, 0, st>>>(n->P, n->G, n->M, n->V, n->nparam, lr, b1, b2, 1e-8f, c1, c2);
            GP_HIP(hipGetLastError());
            if ((rc = nn_prep_weights(st, n))) return rc;
            acc += L;
            nbat++;
        }
        last = nbat ? acc / nbat : 0.0;
    }
    GP_HIP(hipStreamSynchronize(st));
    if (loss_out) *loss_out = (float)last;
    return GP_OK;
}

// one forward+backward on caller-provided padded images (host FP32, B x H x
// W inputs and targets): loss and the flat FP32 gradient (no update)
int gp_nn_grad_impl(gp_ctx *c, const float *in, const float *tgt, int B, int kind, float *loss, float *grad)
{
    gp_nn *n = c->nn;
    if (!n) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    hipStream_t st = c->stream;
    int rc = nn_alloc_batch(n, B);
    if (rc) return rc;
    const int64_t HW = (int64_t)n->H * n->W;
    GP_HIP(hipMemcpyAsync(n->x_in, in, B * HW * 4, hipMemcpyHostToDevice, st));
    GP_HIP(hipMemcpyAsync(n->tgt, tgt, B * HW * 4, hipMemcpyHostToDevice, st));
    if ((rc = nn_forward(st, n, B))) return rc;
    double L = 0.0;
    n->hbw.assign(B, 1.0f / (float)B);
    if ((rc = nn_loss(st, n, B, kind, &L))) return rc;
    if ((rc = nn_backward(st, n, B))) return rc;
    if (grad) GP_HIP(hipMemcpyAsync(grad, n->G, n->nparam * 4, hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    if (loss) *loss = (float)L;
    return GP_OK;
}

// forward of caller-provided padded images (host FP32): the outputs y
int gp_nn_forward_impl(gp_ctx *c, const float *in, int B, float *out)
{
    gp_nn *n = c->nn;
    if (!n) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    hipStream_t st = c->stream;
    int rc = nn_alloc_batch(n, B);
    if (rc) return rc;
    const int64_t HW = (int64_t)n->H * n->W;
    GP_HIP(hipMemcpyAsync(n->x_in, in, B * HW * 4, hipMemcpyHostToDevice, st));
    if ((rc = nn_forward(st, n, B, true))) return rc;
    GP_HIP(hipMemcpyAsync(out, n->y_out, B * HW * 4, hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    return GP_OK;
}

int gp_nn_params_impl(gp_ctx *c, float *get, const float *set, int64_t *count)
{
    gp_nn *n = c->nn;
    if (!n) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    if (count) *count = n->nparam;
    if (set) {
        GP_HIP(hipMemcpy(n->P, set, n->nparam * 4, hipMemcpyHostToDevice));
        int rc = nn_prep_weights(c->stream, n);
        if (rc) return rc;
        GP_HIP(hipStreamSynchronize(c->stream));
    }
    if (get) GP_HIP(hipMemcpy(get, n->P, n->nparam * 4, hipMemcpyDeviceToHost));
    return GP_OK;
}

// z = the network's estimate of the direction of S^-1 r for a device FP32
// unknown-grid vector r (input normalised to zero mean when singular and
// unit rms); the scale of z is left to the caller (CG is invariant to it)
int gp_nn_apply(gp_ctx *c, gp_solver *s, const float *r, float *z)
{
    gp_nn *n = c->nn;
    hipStream_t st = c->stream;
    int rc = nn_set_grid(n, *s);
    if (rc) return rc;
    if ((rc = nn_alloc_batch(n, 1))) return rc;
    if ((rc = nn_embed(st, n, r, n->x_in, n->stats))) return rc;
    if ((rc = nn_forward(st, n, 1, true))) return rc;
    k_extract<<<g1((int64_t)n->mx * n->my), NB, 0, st>>>(n->y_out, n->H, n->W, n->mx, n->my, z);
    GP_HIP(hipGetLastError());
    return GP_OK;
}

int gp_nn_infer_impl(gp_ctx *c, const double *rhs, double *out)
{
    gp_nn *n = c->nn;
    if (!n) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    hipStream_t st = c->stream;
    const int64_t m = (int64_t)n->mx * n->my;
    std::vector<float> h(m);
    for (int64_t k = 0; k < m; k++) h[k] = (float)rhs[k];
    float *r = nullptr, *z = nullptr;
    // every exit below frees both buffers
    struct Bufs {
        float **a, **b;
        ~Bufs() { if (*a) (void)hipFree(*a); if (*b) (void)hipFree(*b); }
    } guard{&r, &z};
    GP_HIP(hipMalloc((void **)&r, m * 4));
    GP_HIP(hipMalloc((void **)&z, m * 4));
    GP_HIP(hipMemcpy(r, h.data(), m * 4, hipMemcpyHostToDevice));
    int rc = gp_nn_apply(c, &c->ps, r, z);
    if (!rc) {
        GP_HIP(hipStreamSynchronize(st));
        GP_HIP(hipMemcpy(h.data(), z, m * 4, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < m; k++) out[k] = h[k];
    }
    return rc;
}

// file: "GPNN1\0" magic, int32 c0, mx, my, int64 nparam, FP32 params
int gp_nn_save_impl(gp_ctx *c, const char *path)
{
    gp_nn *n = c->nn;
    if (!n) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    std::vector<float> h(n->nparam);
    GP_HIP(hipMemcpy(h.data(), n->P, n->nparam * 4, hipMemcpyDeviceToHost));
    FILE *f = std::fopen(path, "wb");
    if (!f) return gp_fail(GP_ERR_ARG, std::string("cannot open ") + path);
    const char magic[6] = {'G', 'P', 'N', 'N', '1', 0};
    const int32_t hd[3] = {n->c0, n->mx, n->my};
    const int64_t np = n->nparam;
    bool ok = std::fwrite(magic, 1, 6, f) == 6 && std::fwrite(hd, 4, 3, f) == 3 && std::fwrite(&np, 8, 1, f) == 1 &&
              std::fwrite(h.data(), 4, np, f) == (size_t)np;
    std::fclose(f);
    return ok ? GP_OK : gp_fail(GP_ERR_ARG, std::string("write failed: ") + path);
}

int gp_nn_load_impl(gp_ctx *c, const char *path)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return gp_fail(GP_ERR_ARG, std::string("cannot open ") + path);
    char magic[6];
    int32_t hd[3];
    int64_t np = 0;
    bool ok = std::fread(magic, 1, 6, f) == 6 && std::memcmp(magic, "GPNN1", 6) == 0 && std::fread(hd, 4, 3, f) == 3 &&
              std::fread(&np, 8, 1, f) == 1;
    if (!ok) { std::fclose(f); return gp_fail(GP_ERR_ARG, std::string("not a network file: ") + path); }
    int rc = gp_nn_create(c, hd[0], 0);
    if (rc) { std::fclose(f); return rc; }
    gp_nn *n = c->nn;
    if (np != n->nparam) { std::fclose(f); return gp_fail(GP_ERR_ARG, std::string("parameter count mismatch in ") + path); }
    std::vector<float> h(np);
    ok = std::fread(h.data(), 4, np, f) == (size_t)np;
    std::fclose(f);
    if (!ok) return gp_fail(GP_ERR_ARG, std::string("short file: ") + path);
    return gp_nn_params_impl(c, nullptr, h.data(), nullptr);
}

// warm-start hook of the FP64 refinement: e = alpha * NN(r), alpha minimising
// the S-norm error along NN(r) (computed by gp_mg_scale)
int gp_nn_precorrect_cb(void *ctx, float *e, const float *r)
{
    gp_nn_bind *b = (gp_nn_bind *)ctx;
    int rc = gp_nn_apply(b->c, b->s, r, e);
    if (rc) return rc;
    return gp_mg_scale(b->c->stream, *b->s, e, r);
}
// preconditioner hook of the flexible CG: z = NN(r), scaled to its S-norm
// optimal multiple, then smoothed by red-black Gauss-Seidel sweeps on S z = r
// (the network carries the smooth part of S^-1 r, the sweeps the rough part)
int gp_nn_precond_cb(void *ctx, float *z, const float *r)
{
    gp_nn_bind *b = (gp_nn_bind *)ctx;
    int rc = gp_nn_apply(b->c, b->s, r, z);
    if (rc) return rc;
    if ((rc = gp_mg_scale(b->c->stream, *b->s, z, r))) return rc;
    if (b->c->nn->smooth > 0) return gp_mg_smooth(b->c->stream, *b->s, z, r, b->c->nn->smooth);
    return GP_OK;
}

int gp_nn_config_impl(gp_ctx *c, int loss_kind, int smooth, int fcg_cap)
{
    gp_nn *n = c->nn;
    if (!n) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    if (loss_kind < 0 || loss_kind > 1 || smooth < 0 || smooth > 64 || fcg_cap < 0)
        return gp_fail(GP_ERR_ARG, "bad network setting");
    n->loss_kind = loss_kind;
    n->smooth = smooth;
    n->fcg_cap = fcg_cap;
    return GP_OK;
}
int gp_nn_fcg_cap(const gp_ctx *c) { return c->nn ? c->nn->fcg_cap : 0; }
int gp_nn_train_system(const gp_ctx *c) { return c->nn ? c->nn->train_sys : 0; }
