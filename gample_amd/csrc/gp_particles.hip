// gp_particles.hip -- particle kernels for gfx950.
//
//  k_push<ORDER, UNIFORM_B, DEP, PUSH>
//      One workgroup per tile of GP_TILE x GP_TILE cells.  The workgroup
//      stages the field stencil window of its tile (Ex, Ey, and Bz unless B
//      is uniform) in LDS, streams its particles through registers with
//      coalesced SoA loads, gathers (B-spline order 1/3/5), applies the
//      Boris rotation, moves, applies the particle BCs, and deposits the new
//      position's charge (and current) into an int64 fixed-point window in
//      LDS.  The window is written whole to the tile's slot of a staging
//      area, and k_stage_reduce adds, for every grid node, the windows of the
//      tiles that cover it to the global int64 accumulators (plain stores and
//      one owner per node: no global atomics for binned particles).
//      Integer addition is associative, so the deposit is bitwise the same
//      for any particle order and any workgroup schedule.  Particles whose
//      cell is outside the tile window (drifted since the last sort, or
//      appended after it) take the same arithmetic through global memory.
//      Reference: updateParticles + integrateChargeDensity,
//      Particle.mm:65-278 and :281-389.
//
//  Compaction, radix sort, loading, injection and moments follow.
#include "gp_internal.h"
#include <algorithm>
#include <cmath>
#include <mutex>

// a species with more than this fraction of its particles changing cell per
// push is sorted without bank-class interleaving, and its hot push adds its
// charge in 32-bit halves
#ifndef GP_CLS_MOVED
#define GP_CLS_MOVED 0.01
#endif

namespace {

constexpr int WIN = GP_TILE + 2 * GP_DRIFT;   // cell window of a tile
// threads of a push workgroup (one workgroup per tile): 512 for the 1st-
// and 3rd-order kernels, register-budgeted for PUSH_WAVES waves per SIMD (80
// VGPRs; three workgroups with charge-only windows fit a CU's LDS); the
// 5th-order kernel needs more registers and keeps 256 threads and the
// compiler's budget, as do the current-deposit kernels, whose windows limit
// a CU to fewer workgroups anyway
#ifndef PUSH_BLOCK
#define PUSH_BLOCK 512
#endif
#ifndef PUSH_WAVES
#define PUSH_WAVES 6
#endif
constexpr int push_block(int order) { return order <= 3 ? PUSH_BLOCK : 256; }
// the 3rd-order charge-deposit push with uniform B (the hot kernel) stages
// its E windows as column coefficients (coef3): 16-byte rows in LDS, and no
// per-particle y weights (72 KB: 2 workgroups, 4 waves per SIMD)
constexpr bool push_cw(int order, bool unib, int dep, bool push) { return order == 3 && unib && dep == 1 && push; }
#define PUSH_OCC __attribute__((amdgpu_waves_per_eu(push_cw(ORDER, UNIB, DEP, PUSH) ? 4 : (ORDER <= 3 && DEP <= 1 ? PUSH_WAVES : 1))))

// ---------------------------------------------------------------------------
// B-spline weights.  Horner forms of the reference polynomials
// (Particle.mm:97-131) in h, mirrored in t = 1-h:
//   order 3: w0 = t^3/6, w1 = 2/3 + h^2 (h/2 - 1), w2 = same in t, w3 = h^3/6
//   order 5: 120 w1 = 26 - 50h + 20h^2 + 20h^3 - 20h^4 + 5h^5 (w4 in t)
//            120 w2 = 66 - 60h^2 + 30h^4 - 10h^5 (w3 in t)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float q5a(float s)
{
    return __builtin_fmaf(s, __builtin_fmaf(s, __builtin_fmaf(s, __builtin_fmaf(s,
           __builtin_fmaf(5.0f, s, -20.0f), 20.0f), 20.0f), -50.0f), 26.0f);
}
__device__ __forceinline__ float q5b(float s)
{
    const float s2 = s * s;
    return __builtin_fmaf(s2, __builtin_fmaf(s2, __builtin_fmaf(-10.0f, s, 30.0f), -60.0f), 66.0f);
}
template <int ORDER>
__device__ __forceinline__ void wts(float h, float *w)
{
    if constexpr (ORDER == 1) {
        w[0] = 1.0f - h;
        w[1] = h;
    } else if constexpr (ORDER == 3) {
        const float c6 = 1.0f / 6.0f, c23 = 2.0f / 3.0f;
        const float t = 1.0f - h, h2 = h * h, t2 = t * t;
        w[0] = (t2 * t) * c6;
        w[1] = __builtin_fmaf(h2, __builtin_fmaf(0.5f, h, -1.0f), c23);
        w[2] = __builtin_fmaf(t2, __builtin_fmaf(0.5f, t, -1.0f), c23);
        w[3] = (h2 * h) * c6;
    } else {
        const float c120 = 1.0f / 120.0f;
        const float t = 1.0f - h, h2 = h * h, t2 = t * t;
        w[0] = ((t2 * t2) * t) * c120;
        w[1] = q5a(h) * c120;
        w[2] = q5b(h) * c120;
        w[3] = q5b(t) * c120;
        w[4] = q5a(t) * c120;
        w[5] = ((h2 * h2) * h) * c120;
    }
}

// factorised gather  F = sum_i sx[i] * (sum_j sy[j] * b[i + j*pitch])
template <int NW, typename P>
__device__ __forceinline__ float gather(P b, int pitch, const float *sx, const float *sy)
{
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < NW; i++) {
        float r = sy[0] * b[i];
#pragma unroll
        for (int j = 1; j < NW; j++) r = __builtin_fmaf(sy[j], b[i + j * pitch], r);
        acc = (i == 0) ? sx[0] * r : __builtin_fmaf(sx[i], r, acc);
    }
    return acc;
}

// row r of a window of pitch W as an offset, for 0 <= r < 2^24 (the 24-bit
// multiply issues at full rate, a 32-bit one at a quarter)
template <int W>
__device__ __forceinline__ int rowW(int r) { return (int)__umul24((unsigned)r, (unsigned)W); }

// the same gather from an LDS window: the window address is passed through an
// empty asm statement, so the compiler keeps it in one register and reaches
// the taps (at most 3 + 3 W dwords away) through the instructions' offset
// fields, where it would otherwise add the array's LDS base for every pair
template <int NW, int PITCH>
__device__ __forceinline__ float gather_lds(const float *b, const float *sx, const float *sy)
{
    typedef const __attribute__((address_space(3))) float lds_f;
    lds_f *p = (lds_f *)b;
    asm volatile("" : "+v"(p));
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < NW; i++) {
        float r = sy[0] * p[i];
#pragma unroll
        for (int j = 1; j < NW; j++) r = __builtin_fmaf(sy[j], p[i + j * PITCH], r);
        acc = (i == 0) ? sx[0] * r : __builtin_fmaf(sx[i], r, acc);
    }
    return acc;
}

// Cubic gather along y as a polynomial in the fraction h.  The four taps
// f0..f3 of a stencil column weighted by the order-3 weights w0..w3(h) are
// d0 + h (d1 + h (d2 + h d3)) with
//   d0 = (f0 + 4 f1 + f2) / 6,  d1 = (f2 - f0) / 2,
//   d2 = (f0 + f2) / 2 - f1,    d3 = (f3 - f0 + 3 (f1 - f2)) / 6
// (the coefficients of h^k in w0..w3, Particle.mm:97-131).  The gather is
// then sum_i sx[i] r_i over the columns: two weight sets per gathered field
// instead of four, and three fmas per column instead of four.  coef3 is the
// one definition used by every path (the staged coefficient windows, the tap
// windows and global memory), so all give the same bits.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f4v coef3(float f0, float f1, float f2, float f3)
{
    const float a02 = f0 + f2;
    f4v d;
    d[0] = __builtin_fmaf(f1, 2.0f / 3.0f, a02 * (1.0f / 6.0f));
    d[1] = (f2 - f0) * 0.5f;
    d[2] = __builtin_fmaf(a02, 0.5f, -f1);
    d[3] = __builtin_fmaf(f1 - f2, 3.0f, f3 - f0) * (1.0f / 6.0f);
    return d;
}
__device__ __forceinline__ float horner3(const f4v d, float h)
{
    return __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(d[3], h, d[2]), h, d[1]), h, d[0]);
}
// from a coefficient window: q -> the 4 consecutive columns of the stencil
__device__ __forceinline__ float gather_cw(const f4v *q, float h, const float *sx)
{
    float acc = sx[0] * horner3(q[0], h);
#pragma unroll
    for (int i = 1; i < 4; i++) acc = __builtin_fmaf(sx[i], horner3(q[i], h), acc);
    return acc;
}
// from taps at b[i + j pitch] (LDS window or global field)
template <typename P>
__device__ __forceinline__ float gather_c3(P b, int pitch, float h, const float *sx)
{
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float r = horner3(coef3(b[i], b[i + pitch], b[i + 2 * pitch], b[i + 3 * pitch]), h);
        acc = (i == 0) ? sx[0] * r : __builtin_fmaf(sx[i], r, acc);
    }
    return acc;
}

// exact round-to-nearest-even of c * 2^40 as int64 (|c| < 2^22)
__device__ __forceinline__ long long fx40(float c)
{
    const float r = __builtin_rintf(c * 0x1p40f);
    const float a = fabsf(r);
    const float hi = truncf(a * 0x1p-32f);
    const float lo = __builtin_fmaf(hi, -4294967296.0f, a);
    const unsigned long long u = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
    return r < 0.0f ? -(long long)u : (long long)u;
}
// per-axis weight s in [0, 1] as an unsigned with 23 fractional bits,
// rint(s * 2^23).  One fma: s * 2^23 + 2^23 lies in [2^23, 2^24], where the
// float spacing is 1, so the single rounding of the fma is the round-to-
// nearest-even of s * 2^23, and the low 24 bits of the sum's encoding are that
// integer (2^24 itself encodes as 0x4B800000, whose low bits give 2^23).  The
// 24-bit value lets the compiler use the full-rate 24-bit multiplies
// (v_mul_u32_u24 / v_mul_hi_u32_u24) for the 46-bit product.
__device__ __forceinline__ unsigned q23(float s)
{
    return __float_as_uint(__builtin_fmaf(s, 8388608.0f, 8388608.0f)) & 0xFFFFFFu;
}
// rint(s * 2^16) the same way: s * 2^16 + 2^23 lies in [2^23, 2^23 + 2^16],
// so the low 23 bits of its encoding are the rounded value.  The order-3
// charge uses it: its largest weight is 2/3, so q <= 43691 and a product of
// two fits 31 bits (one 24-bit multiply), and the windows add the products
// unshifted; they are scaled by 2^14 to the 46-bit fixed point at the flush.
__device__ __forceinline__ unsigned q16(float s)
{
    return __float_as_uint(__builtin_fmaf(s, 65536.0f, 8388608.0f)) & 0x7FFFFFu;
}

// periodic wrap x = fmod(x + (ng-ngb), ng) + ngb (Particle.mm:216-221); the
// conditional subtraction is the exact remainder for 0 <= a < 2 ng.
__device__ __forceinline__ float wrap(float x, int ng, int ngb)
{
    const float L = (float)ng;
    float a = x + (float)(ng - ngb);
    // (for 0 <= a < 2L the one subtraction is exact, so it is fmod's result)
    if (a >= 0.0f && a < 2.0f * L) a = a >= L ? a - L : a;
    else a = fmodf(a, L);
    return a + (float)ngb;
}

struct PushArgs {
    float *x, *y, *vx, *vy, *vz;
    int64_t n, n_binned, pfirst;
    const int32_t *tile_off;
    int ntx, ntiles, chunk, use_tiles;
    const float *Ex, *Ey, *Bz;
    float bz0;
    float constE, constB, constX, constY;
    int ngx, ngy, ngb, nx, ny;
    int bc0, bc1, bc2, bc3, boris;
    // charge windows by position on periodic axes: a particle whose cell
    // lies one period away from the window deposits into the window at its
    // image, and the windows are folded onto the interior nodes as the
    // reference's ghost fold would (EMField.mm:749-850); rho deposit only
    int imgx, imgy;
    // E gathered at periodic images on the edge tiles: only while E is exactly
    // periodic (gp_ctx::e_periodic) and B uniform
    int gimgx, gimgy;
    int moved_mask;   // tiles b with (b & moved_mask) == 0 count cell changes
    unsigned long long *rho, *jx, *jy;
    unsigned long long *stage;   // per tile: W*W rho window, then jx and jy (DEP 2)
    uint8_t *stage_on;           // per tile: 1 when its window was written
    float inv_vnorm;
    uint32_t *dead;
    unsigned long long *counters;   // [0] dead, [1..4] walls, [5] escaped, [6] drifted, [7] moved
    double *ke_part;
    double ke_coef;
};

// BCK: 0 any particle boundary codes (read from the arguments), 1 periodic
// on all four sides (the wrap only: no absorbed or reflected particles)
// DC: 1 adds the hot kernel's charge in 32-bit halves (a species whose lanes
// address random cells: half the bytes per atomic, see the deposit)
template <int ORDER, bool UNIB, int DEP, bool PUSH, int BCK, int DC = 0>
__global__ __launch_bounds__(push_block(ORDER)) PUSH_OCC void k_push(PushArgs a)
{
    constexpr int PB = push_block(ORDER);
    constexpr int NGB = ORDER / 2, NW = ORDER + 1;
    // node window edge: relative to X0 = cx0 - NGB the stencils reach from 0
    // (rho: i2 - NGB with i2 = cx0) to WIN + NW - 1 (Ex, jx: i1 + 1 - NGB +
    // NW - 1 with i1 = i2 = cx0 + WIN - 1; likewise Ey, jy in y)
    constexpr int W = WIN + NW;
    // CW: E staged as column coefficients (push_cw); CROWS cell rows of them
    constexpr bool CW = push_cw(ORDER, UNIB, DEP, PUSH);
    // NOVZ: vz is not read (it never changes); the host adds the sum of vz^2
    constexpr bool NOVZ = CW && BCK == 1;
    constexpr int CROWS = WIN + 1;
    constexpr int NF = (PUSH && !CW) ? W * W : 1;
    constexpr int NC = CW ? CROWS * W : 1;
    constexpr int NB = (PUSH && !UNIB) ? W * W : 1;
    constexpr int NR = DEP >= 1 ? W * W : 1;
    constexpr int NJ = DEP >= 2 ? W * W : 1;
    constexpr int QSH = ORDER == 3 ? 14 : 0;   // window charge scaled by 2^QSH at the flush
    static_assert(!CW || NR * 2 >= 2 * W * W, "CW stages the E taps in the charge window");
    __shared__ f4v sCx[NC], sCy[NC];
    __shared__ float sEx[NF], sEy[NF];
    __shared__ float sBz[NB];
    __shared__ unsigned long long sRho[NR], sJx[NJ], sJy[NJ];
    __shared__ double sKE[PB / 64];

    const int tid = threadIdx.x, b = blockIdx.x;
    int64_t p0, p1;
    int tiled = 0, tx0 = 0, ty0 = 0;
    if (a.use_tiles && b < a.ntiles) {
        p0 = a.tile_off[b];
        p1 = a.tile_off[b + 1];
        if (p1 > a.n) p1 = a.n;
        tiled = 1;
        tx0 = a.ngb + (b % a.ntx) * GP_TILE;
        ty0 = a.ngb + (b / a.ntx) * GP_TILE;
    } else {
        const int64_t eb = a.use_tiles ? (b - a.ntiles) : b;
        p0 = a.pfirst + eb * a.chunk;
        p1 = p0 + a.chunk;
        if (p1 > a.n) p1 = a.n;
    }
    const int cx0 = tx0 - GP_DRIFT, cy0 = ty0 - GP_DRIFT;   // first window cell
    // images are needed only by the tiles on the periodic edges
    const bool img_x = a.imgx && tiled && (b % a.ntx == 0 || b % a.ntx == a.ntx - 1);
    const bool img_y = a.imgy && tiled && (b < a.ntx || b >= a.ntiles - a.ntx);
    // gather images: the E window of an edge tile holds, beyond the array, the
    // values one period away (a.gimgx/y only while the fields are periodic)
    const bool gimg_x = a.gimgx && tiled && (b % a.ntx == 0 || b % a.ntx == a.ntx - 1);
    const bool gimg_y = a.gimgy && tiled && (b < a.ntx || b >= a.ntiles - a.ntx);
    const bool count_moved = (b & a.moved_mask) == 0;
    const int X0 = cx0 - NGB, Y0 = cy0 - NGB;                // first window node
    const int pex = a.nx + 2, pey = a.nx + 1;
    // field node (gx, gy) of an array with ex x ey nodes and row pitch ex for
    // a window: wrapped by one period on gather-imaged axes, 0 outside
    auto fnode = [&](const float *F, int gx, int gy, int ex, int ey) -> float {
        if (gimg_x) gx = gx < 0 ? gx + a.ngx : (gx >= ex ? gx - a.ngx : gx);
        if (gimg_y) gy = gy < 0 ? gy + a.ngy : (gy >= ey ? gy - a.ngy : gy);
        return (gx >= 0 && gx < ex && gy >= 0 && gy < ey) ? F[gx + gy * ex] : 0.0f;
    };

    if (tiled && p0 < p1) {
        if constexpr (CW) {
            // E taps into the (not yet used) charge window, then each column's
            // cubic coefficients over cell rows r .. r+3
            float *tEx = (float *)sRho, *tEy = tEx + W * W;
            for (int e = tid; e < W * W; e += PB) {
                const int gx = X0 + e % W, gy = Y0 + e / W;
                tEx[e] = fnode(a.Ex, gx, gy, a.nx + 2, a.ny + 1);
                tEy[e] = fnode(a.Ey, gx, gy, a.nx + 1, a.ny + 2);
            }
            __syncthreads();
            for (int e = tid; e < CROWS * W; e += PB) {
                sCx[e] = coef3(tEx[e], tEx[e + W], tEx[e + 2 * W], tEx[e + 3 * W]);
                sCy[e] = coef3(tEy[e], tEy[e + W], tEy[e + 2 * W], tEy[e + 3 * W]);
            }
            __syncthreads();
        }
        for (int e = tid; e < W * W; e += PB) {
            const int gx = X0 + e % W, gy = Y0 + e / W;
            if constexpr (PUSH && !CW) {
                sEx[e] = fnode(a.Ex, gx, gy, a.nx + 2, a.ny + 1);
                sEy[e] = fnode(a.Ey, gx, gy, a.nx + 1, a.ny + 2);
            }
            if constexpr (PUSH && !UNIB) sBz[e] = fnode(a.Bz, gx, gy, a.nx + 2, a.ny + 2);
            if constexpr (DEP >= 1) sRho[e] = 0ull;
            if constexpr (DEP >= 2) { sJx[e] = 0ull; sJy[e] = 0ull; }
        }
    }
    __syncthreads();

    const float lox = (float)a.ngb, hix = (float)(a.ngb + a.ngx);
    const float loy = (float)a.ngb, hiy = (float)(a.ngb + a.ngy);
    double ke = 0.0;   // sum of |v|^2 (times ke_coef after the loop)
    int drift = 0;   // live particles gathered outside their tile window
    // uniform B: btz = constB Bz (scaled by fN1 at Boris order 4/6, with
    // btbt from the unscaled btz), s = 2 btz / (1 + btbt)
    float ub_btz = 0.0f, ub_ssz = 0.0f, ub_fn1 = 1.0f;
    if constexpr (UNIB) {
        float btz = a.constB * a.bz0;
        const float btbt = btz * btz;
        if (a.boris == 4 || a.boris == 6) {
            ub_fn1 = (a.boris == 4) ? (1.0f + btbt / 3.0f) : (1.0f + btbt / 3.0f + 2.0f * btbt * btbt / 15.0f);
            btz = ub_fn1 * btz;
        }
        ub_btz = btz;
        ub_ssz = (2.0f * btz) / (1.0f + btbt);
    }
    int moved = 0;   // live particles whose cell changed in this push

    // The segment is walked in sub-segments of at most 2^28 particles, so that
    // a particle's address is a uniform base plus a 32-bit byte offset (the
    // scalar-base + vector-offset form of the global loads and stores).
    constexpr int64_t SUB = int64_t(1) << 28;
    for (int64_t s0 = p0; s0 < p1; s0 += SUB) {
    const uint32_t cnt = (uint32_t)(p1 - s0 < SUB ? p1 - s0 : SUB);
    float *const bx = a.x + s0, *const by = a.y + s0, *const bvx = a.vx + s0, *const bvy = a.vy + s0;
    const float *const bvz = a.vz + s0;
    auto ld = [](const float *b, uint32_t o) { return *(const float *)((const char *)b + o); };
    auto st = [](float *b, uint32_t o, float v) { *(float *)((char *)b + o) = v; };
    // one particle ahead: the next particle's loads are in flight while the
    // current one is processed
    float nx_ = 0.f, ny_ = 0.f, nvx = 0.f, nvy = 0.f, nvz = 0.f;
    if ((uint32_t)tid < cnt) {
        const uint32_t o = 4u * tid;
        nx_ = ld(bx, o); ny_ = ld(by, o); nvx = ld(bvx, o); nvy = ld(bvy, o);
        if constexpr (!NOVZ) nvz = ld(bvz, o);
    }
    for (uint32_t k = tid; k < cnt; k += PB) {
        const int64_t p = s0 + k;
        const uint32_t ob = 4u * k;
        float x = nx_, y = ny_;
        float vx = nvx, vy = nvy;
        const float vz = nvz;
        if (k + PB < cnt) {
            const uint32_t o = ob + 4u * PB;
            nx_ = ld(bx, o); ny_ = ld(by, o); nvx = ld(bvx, o); nvy = ld(bvy, o);
        if constexpr (!NOVZ) nvz = ld(bvz, o);
        }
        // live particles lie in [ngb, ngb+ng); anything else (absorbed
        // marker, escaped, NaN) is left untouched and not deposited
        if (!(x >= lox && x < hix && y >= loy && y < hiy)) {
            if (x >= 0.0f) atomicAdd(&a.counters[5], 1ull);
            continue;
        }
        int flag = 0;
        int oi = 0, oj = 0;   // cell before the push
        if constexpr (PUSH) {
            const float xh = x - 0.5f, yh = y - 0.5f;
            const float fi1 = floorf(xh), fj1 = floorf(y), fi2 = floorf(x), fj2 = floorf(yh);
            int i1 = (int)fi1, j1 = (int)fj1, i2 = (int)fi2, j2 = (int)fj2;
            oi = i2; oj = j1;
            // window cell; on an edge tile whose E window is staged with the
            // periodic images, a cell one period away is gathered at its image
            // (E is periodic there, so the taps are the same values)
            bool inw = tiled && (unsigned)(i2 - cx0) < (unsigned)WIN && (unsigned)(j1 - cy0) < (unsigned)WIN;
            if (!inw && (gimg_x || gimg_y)) {
                const int gi = i2 - cx0, gj = j1 - cy0;
                const int shx = gimg_x ? (gi < 0 ? a.ngx : (gi >= WIN ? -a.ngx : 0)) : 0;
                const int shy = gimg_y ? (gj < 0 ? a.ngy : (gj >= WIN ? -a.ngy : 0)) : 0;
                inw = (unsigned)(gi + shx) < (unsigned)WIN && (unsigned)(gj + shy) < (unsigned)WIN;
                if (inw) { i1 += shx; i2 += shx; j1 += shy; j2 += shy; }
            }
            if constexpr (DEP == 0) drift += inw ? 0 : 1;
            float sx0[NW], sx1[NW];
            wts<ORDER>(xh - fi1, sx0);
            wts<ORDER>(x - fi2, sx1);
            float Epx, Epy, Bpz;
            if constexpr (ORDER == 3) {
                // cubic gather as column polynomials in the y fraction
                const float hy0 = y - fj1, hy1 = yh - fj2;
                if (inw) {
                    // (window rows are non-negative here: 24-bit multiplies)
                    if constexpr (CW) {
                        Epx = gather_cw(sCx + (i1 + 1 - cx0) + rowW<W>(j1 - cy0), hy0, sx0);
                        Epy = gather_cw(sCy + (i2 - cx0) + rowW<W>(j2 + 1 - cy0), hy1, sx1);
                    } else {
                        Epx = gather_c3(sEx + (i1 + 1 - NGB - X0) + rowW<W>(j1 - NGB - Y0), W, hy0, sx0);
                        Epy = gather_c3(sEy + (i2 - NGB - X0) + rowW<W>(j2 + 1 - NGB - Y0), W, hy1, sx1);
                    }
                    if constexpr (UNIB) Bpz = a.bz0;
                    else Bpz = gather_c3(sBz + (i1 + 1 - NGB - X0) + rowW<W>(j2 + 1 - NGB - Y0), W, hy1, sx0);
                } else {
                    Epx = gather_c3(a.Ex + (i1 + 1 - NGB) + (int64_t)(j1 - NGB) * pex, pex, hy0, sx0);
                    Epy = gather_c3(a.Ey + (i2 - NGB) + (int64_t)(j2 + 1 - NGB) * pey, pey, hy1, sx1);
                    if constexpr (UNIB) Bpz = a.bz0;
                    else Bpz = gather_c3(a.Bz + (i1 + 1 - NGB) + (int64_t)(j2 + 1 - NGB) * pex, pex, hy1, sx0);
                }
            } else {
                float sy0[NW], sy1[NW];
                wts<ORDER>(y - fj1, sy0);
                wts<ORDER>(yh - fj2, sy1);
                if (inw) {
                    Epx = gather_lds<NW, W>(sEx + (i1 + 1 - NGB - X0) + rowW<W>(j1 - NGB - Y0), sx0, sy0);
                    Epy = gather_lds<NW, W>(sEy + (i2 - NGB - X0) + rowW<W>(j2 + 1 - NGB - Y0), sx1, sy1);
                    if constexpr (UNIB) Bpz = a.bz0;
                    else Bpz = gather_lds<NW, W>(sBz + (i1 + 1 - NGB - X0) + rowW<W>(j2 + 1 - NGB - Y0), sx0, sy1);
                } else {
                    Epx = gather<NW>(a.Ex + (i1 + 1 - NGB) + (int64_t)(j1 - NGB) * pex, pex, sx0, sy0);
                    Epy = gather<NW>(a.Ey + (i2 - NGB) + (int64_t)(j2 + 1 - NGB) * pey, pey, sx1, sy1);
                    if constexpr (UNIB) Bpz = a.bz0;
                    else Bpz = gather<NW>(a.Bz + (i1 + 1 - NGB) + (int64_t)(j2 + 1 - NGB) * pex, pex, sx0, sy1);
                }
            }
            // Boris (Particle.mm:156-209) for the gathered components
            // Ex, Ey, Bz; Ez = Bx = By = 0 exactly as in the reference.
            // With a uniform B the rotation factors are the kernel-wide
            // constants computed before the loop (the same operations).
            const float ex0 = a.constE * Epx, ey0 = a.constE * Epy;
            float etx = ex0, ety = ey0, btz, ssz;
            if constexpr (UNIB) {
                btz = ub_btz;
                ssz = ub_ssz;
                if (a.boris == 4 || a.boris == 6) { etx = ub_fn1 * etx; ety = ub_fn1 * ety; }
            } else {
                btz = a.constB * Bpz;
                const float btbt = btz * btz;
                if (a.boris == 4 || a.boris == 6) {
                    const float fN1 = (a.boris == 4) ? (1.0f + btbt / 3.0f)
                                                     : (1.0f + btbt / 3.0f + 2.0f * btbt * btbt / 15.0f);
                    etx = fN1 * etx; ety = fN1 * ety; btz = fN1 * btz;
                }
                ssz = (2.0f * btz) / (1.0f + btbt);
            }
            const float umx = vx + etx, umy = vy + ety;
            const float v0x = __builtin_fmaf(umy, btz, umx);
            const float v0y = __builtin_fmaf(-umx, btz, umy);
            const float upx = __builtin_fmaf(v0y, ssz, umx);
            const float upy = __builtin_fmaf(-v0x, ssz, umy);
            vx = upx + ex0;
            vy = upy + ey0;
            x = __builtin_fmaf(vx, a.constX, x);
            y = __builtin_fmaf(vy, a.constY, y);
            // particle BCs (Particle.mm:215-269)
            if constexpr (BCK == 1) {
                x = wrap(x, a.ngx, a.ngb);
                y = wrap(y, a.ngy, a.ngb);
            } else {
                if (a.bc0 == GP_PBC_PERIODIC) x = wrap(x, a.ngx, a.ngb);
                if (a.bc2 == GP_PBC_PERIODIC) y = wrap(y, a.ngy, a.ngb);
                if (a.bc0 == GP_PBC_DELETE && x < lox) flag = 1;
                if (a.bc1 == GP_PBC_DELETE && x >= hix) flag = 2;
                if (a.bc2 == GP_PBC_DELETE && y < loy) flag = 3;
                if (a.bc3 == GP_PBC_DELETE && y >= hiy) flag = 4;
                if (a.bc0 == GP_PBC_REFLECT && x < lox) { x = lox + lox - x; vx = -vx; }
                if (a.bc1 == GP_PBC_REFLECT && x >= hix) { x = hix + hix - x; vx = -vx; }
                if (a.bc2 == GP_PBC_REFLECT && y < loy) { y = loy + loy - y; vy = -vy; }
                if (a.bc3 == GP_PBC_REFLECT && y >= hiy) { y = hiy + hiy - y; vy = -vy; }
            }
            // a wrapped or reflected position can round onto the upper wall
            if (!flag) {
                if (x >= hix) x = __int_as_float(__float_as_int(hix) - 1);
                if (y >= hiy) y = __int_as_float(__float_as_int(hiy) - 1);
            }
            st(bx, ob, flag ? -(float)flag : x);
            st(by, ob, y);
            st(bvx, ob, vx);
            st(bvy, ob, vy);
        }
        if (flag) {
            atomicAdd(&a.counters[flag], 1ull);
            const unsigned long long slot = atomicAdd(&a.counters[0], 1ull);
            a.dead[slot] = (uint32_t)p;
            continue;
        }
        if constexpr (PUSH) {
            const float vv = NOVZ ? vx * vx + vy * vy : vx * vx + vy * vy + vz * vz;
            ke += (double)vv;
        }
        if constexpr (DEP >= 1) {
            // charge (Particle.mm:357-361): node (i2+i-ngb, j1+j-ngb), weights
            // of x - floor(x) and y - floor(y)
            if (!(x >= lox && x < hix && y >= loy && y < hiy)) {
                atomicAdd(&a.counters[5], 1ull);
                continue;
            }
            const float fi2 = floorf(x), fj1 = floorf(y);
            const int i2 = (int)fi2, j1 = (int)fj1;
            // window cell (the stencil's first node is window node (wi, wj));
            // on an edge tile, a cell outside the window may have its image
            // one period away inside it (the rare branch)
            int wi = i2 - cx0, wj = j1 - cy0;
            bool inw = tiled && (unsigned)wi < (unsigned)WIN && (unsigned)wj < (unsigned)WIN;
            if (!inw && (img_x || img_y)) {
                if (img_x) wi += wi < 0 ? a.ngx : (wi >= WIN ? -a.ngx : 0);
                if (img_y) wj += wj < 0 ? a.ngy : (wj >= WIN ? -a.ngy : 0);
                inw = (unsigned)wi < (unsigned)WIN && (unsigned)wj < (unsigned)WIN;
            }
            if constexpr (PUSH) {
                if (count_moved) moved += (i2 != oi || j1 != oj) ? 1 : 0;
                drift += inw ? 0 : 1;
            }
            float sx[NW], sy[NW];
            wts<ORDER>(x - fi2, sx);
            wts<ORDER>(y - fj1, sy);
            // node charge qx[i] qy[j] 2^QSH in the 46-bit fixed point
            unsigned qx[NW], qy[NW];
#pragma unroll
            for (int k = 0; k < NW; k++) {
                if constexpr (ORDER == 3) { qx[k] = q16(sx[k]); qy[k] = q16(sy[k]); }
                else { qx[k] = q23(sx[k]); qy[k] = q23(sy[k]); }
            }
            auto qq = [&](int i, int j) -> unsigned long long {
                if constexpr (ORDER == 3) return (unsigned long long)(unsigned)__umul24(qx[i], qy[j]);
                else return (unsigned long long)qx[i] * qy[j];
            };
            if (inw) {
                if constexpr (ORDER == 3 && DC == 1) {
                    // the window as 32-bit low words and, W*W further, high
                    // words: a lane whose addition wraps its low word adds 1
                    // to the high word (the returned old value shows the wrap)
                    unsigned *lo = (unsigned *)sRho + wi + rowW<W>(wj), *hi = lo + W * W;
                    unsigned pp[NW * NW], ov[NW * NW];
#pragma unroll
                    for (int j = 0; j < NW; j++)
#pragma unroll
                        for (int i = 0; i < NW; i++) {
                            pp[i + NW * j] = __umul24(qx[i], qy[j]);
                            ov[i + NW * j] = atomicAdd(lo + i + j * W, pp[i + NW * j]);
                        }
#pragma unroll
                    for (int j = 0; j < NW; j++)
#pragma unroll
                        for (int i = 0; i < NW; i++)
                            if (ov[i + NW * j] + pp[i + NW * j] < ov[i + NW * j]) atomicAdd(hi + i + j * W, 1u);
                } else {
                    unsigned long long *r = sRho + wi + rowW<W>(wj);
#pragma unroll
                    for (int j = 0; j < NW; j++)
#pragma unroll
                        for (int i = 0; i < NW; i++) atomicAdd(r + i + j * W, qq(i, j));
                }
            } else {
                unsigned long long *r = a.rho + (i2 - NGB) + (int64_t)(j1 - NGB) * pey;
#pragma unroll
                for (int j = 0; j < NW; j++)
#pragma unroll
                    for (int i = 0; i < NW; i++) atomicAdd(r + i + (int64_t)j * pey, qq(i, j) << QSH);
            }
            if constexpr (DEP >= 2) {
                // current (Particle.mm:363-373)
                const float xh = x - 0.5f, yh = y - 0.5f;
                const float fi1 = floorf(xh), fj2 = floorf(yh);
                const int i1 = (int)fi1, j2 = (int)fj2;
                float sx0[NW], sy1[NW];
                wts<ORDER>(xh - fi1, sx0);
                wts<ORDER>(yh - fj2, sy1);
                const float vnx = vx * a.inv_vnorm, vny = vy * a.inv_vnorm;
                unsigned long long *jxb, *jyb;
                int pjx, pjy;
                if (inw) {
                    jxb = sJx + (i1 + 1 - NGB - X0) + rowW<W>(j1 - NGB - Y0); pjx = W;
                    jyb = sJy + (i2 - NGB - X0) + rowW<W>(j2 + 1 - NGB - Y0); pjy = W;
                } else {
                    jxb = a.jx + (i1 + 1 - NGB) + (int64_t)(j1 - NGB) * pex; pjx = pex;
                    jyb = a.jy + (i2 - NGB) + (int64_t)(j2 + 1 - NGB) * pey; pjy = pey;
                }
#pragma unroll
                for (int j = 0; j < NW; j++)
#pragma unroll
                    for (int i = 0; i < NW; i++) {
                        atomicAdd(jxb + i + (int64_t)j * pjx, (unsigned long long)fx40((sx0[i] * sy[j]) * vnx));
                        atomicAdd(jyb + i + (int64_t)j * pjy, (unsigned long long)fx40((sx[i] * sy1[j]) * vny));
                    }
            }
        }
    }
    }

    if constexpr (PUSH) {
        // deterministic workgroup sum of the kinetic energy
        ke *= a.ke_coef;
        for (int off = 32; off > 0; off >>= 1) ke += __shfl_down(ke, off, 64);
        if ((tid & 63) == 0) sKE[tid >> 6] = ke;
        for (int off = 32; off > 0; off >>= 1) drift += __shfl_down(drift, off, 64);
        if ((tid & 63) == 0 && drift) atomicAdd(&a.counters[6], (unsigned long long)drift);
        for (int off = 32; off > 0; off >>= 1) moved += __shfl_down(moved, off, 64);
        if ((tid & 63) == 0 && moved) atomicAdd(&a.counters[7], (unsigned long long)moved);
    }
    __syncthreads();
    if constexpr (PUSH) {
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < PB / 64; w++) s += sKE[w];
            a.ke_part[b] = s;
        }
    }
    if constexpr (DEP >= 1) {
        if (tiled) {
            const bool on = p0 < p1;
            if (tid == 0) a.stage_on[b] = on ? 1 : 0;
            if (on) {
                unsigned long long *st = a.stage + (int64_t)b * (W * W) * (DEP >= 2 ? 3 : 1);
#pragma unroll 1
                for (int e = tid; e < W * W; e += PB) {
                    if constexpr (ORDER == 3 && DC == 1) {
                        const unsigned *lo = (const unsigned *)sRho;
                        st[e] = (((unsigned long long)lo[W * W + e] << 32) | lo[e]) << QSH;
                    } else {
                        st[e] = sRho[e] << QSH;
                    }
                    if constexpr (DEP >= 2) { st[W * W + e] = sJx[e]; st[2 * W * W + e] = sJy[e]; }
                }
            }
        }
    }
}

// HBM streaming rate of the hot push's access pattern, for its roofline:
// read x, y, vx, vy and write four arrays (32 B per particle), one 512-thread
// workgroup per contiguous segment like a tile, into the species' scratch
// arrays (the live particles are not modified).
__global__ __launch_bounds__(512) void k_stream_probe(gp_soa in, gp_soa out, int64_t n, int64_t per)
{
    const int64_t p0 = (int64_t)blockIdx.x * per, p1 = p0 + per < n ? p0 + per : n;
    for (int64_t i = p0 + threadIdx.x; i < p1; i += 512) {
        const float x = in.x[i], y = in.y[i], vx = in.vx[i], vy = in.vy[i];
        out.x[i] = x + vy;
        out.y[i] = y + vx;
        out.vx[i] = vx + x;
        out.vy[i] = vy + y;
    }
}

// floor(a / 16) for any sign
__device__ __forceinline__ int fdiv16(int a) { return a >= 0 ? a / GP_TILE : -((-a + GP_TILE - 1) / GP_TILE); }

// Sum of the staged tile windows into the accumulators: node (gx, gy) of
// component comp (0 rho, 1 jx, 2 jy; extents ext_x x ext_y, row pitch ext_x)
// lies in the window of tile (tx, ty) when X0(tx) <= gx < X0(tx) + W, with
// X0(tx) = ngb + GP_TILE tx - GP_DRIFT - NGB (the push kernel's window origin).
template <int ORDER>
__global__ __launch_bounds__(GP_BLOCK) void k_stage_reduce(const unsigned long long *stage, const uint8_t *on,
                                                           int ncomp, int ntx, int nty, int ngb, int ngx, int ngy,
                                                           int nx, int ny, int imgx, int imgy,
                                                           unsigned long long *rho, unsigned long long *jx,
                                                           unsigned long long *jy)
{
    constexpr int NGB = ORDER / 2, NW = ORDER + 1, W = WIN + NW;
    const int comp = blockIdx.z;
    const int ext_x = nx + 1 + (comp == 1 ? 1 : 0), ext_y = ny + 1 + (comp == 2 ? 1 : 0);
    const int gx = blockIdx.x * GP_BLOCK + threadIdx.x, gy = blockIdx.y;
    if (gx >= ext_x || gy >= ext_y) return;
    unsigned long long *acc = comp == 0 ? rho : (comp == 1 ? jx : jy);
    const int c0 = ngb - GP_DRIFT - NGB;
    // window positions that land on this node: the node itself, or on an
    // imaged (periodic, rho) axis its interior positions one period either
    // side; a ghost node of an imaged axis receives nothing from the windows
    int px[3], py[3], npx = 0, npy = 0;
    if (imgx && comp == 0) {
        if (gx >= ngb && gx < ngb + ngx) { px[0] = gx - ngx; px[1] = gx; px[2] = gx + ngx; npx = 3; }
    } else { px[0] = gx; npx = 1; }
    if (imgy && comp == 0) {
        if (gy >= ngb && gy < ngb + ngy) { py[0] = gy - ngy; py[1] = gy; py[2] = gy + ngy; npy = 3; }
    } else { py[0] = gy; npy = 1; }
    // at most MT windows cover a node along an axis; the loads of all of
    // them are issued together (a window of an empty tile holds stale data
    // and is read, then masked out)
    constexpr int MT = (W + GP_TILE - 1) / GP_TILE;
    unsigned long long sum = 0;
    // (an image position beyond every window, the usual case away from the
    // periodic edges, is skipped before its loads are set up)
    for (int b = 0; b < npy; b++) {
        const int y = py[b];
        const int tyl = max(fdiv16(y - c0 - W) + 1, 0), tyh = min(fdiv16(y - c0), nty - 1);
        if (tyl > tyh) continue;
        for (int a = 0; a < npx; a++) {
            const int x = px[a];
            const int txl = max(fdiv16(x - c0 - W) + 1, 0), txh = min(fdiv16(x - c0), ntx - 1);
            if (txl > txh) continue;
            unsigned long long v[MT * MT];
            bool use[MT * MT];
#pragma unroll
            for (int qy = 0; qy < MT; qy++)
#pragma unroll
                for (int qx = 0; qx < MT; qx++) {
                    const int ty = tyl + qy, tx = txl + qx, q = qx + qy * MT;
                    use[q] = ty <= tyh && tx <= txh;
                    v[q] = 0;
                    if (use[q]) {
                        const int t = ty * ntx + tx;
                        const int lx = x - c0 - GP_TILE * tx, ly = y - c0 - GP_TILE * ty;
                        use[q] = on[t] != 0;
                        v[q] = stage[((int64_t)t * ncomp + comp) * (W * W) + lx + ly * W];
                    }
                }
#pragma unroll
            for (int q = 0; q < MT * MT; q++) sum += use[q] ? v[q] : 0ull;
        }
    }
    if (sum) acc[gx + (int64_t)gy * ext_x] += sum;
}

// extra (may be null): coef * *extra is added after the partials
__global__ void k_sum_partials(const double *part, int n, double *out, const double *extra = nullptr,
                               double coef = 0.0)
{
    // single workgroup, fixed order: thread t sums t, t+256, ...; then tree
    __shared__ double s[GP_BLOCK];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += GP_BLOCK) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = GP_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = extra ? s[0] + coef * *extra : s[0];
}

// kinetic energy of the live particles without a push
__global__ __launch_bounds__(GP_BLOCK) void k_ke(const float *x, const float *vx, const float *vy, const float *vz,
                                                int64_t n, double coef, double *part)
{
    __shared__ double s[GP_BLOCK / 64];
    double ke = 0.0;
    for (int64_t p = blockIdx.x * (int64_t)GP_BLOCK + threadIdx.x; p < n; p += (int64_t)gridDim.x * GP_BLOCK) {
        if (!(x[p] >= 0.0f)) continue;
        const float a = vx[p], b = vy[p], c = vz[p];
        const float vv = a * a + b * b + c * c;
        ke += coef * (double)vv;
    }
    for (int off = 32; off > 0; off >>= 1) ke += __shfl_down(ke, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = ke;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < GP_BLOCK / 64; w++) t += s[w];
        part[blockIdx.x] = t;
    }
}

// per-block sums of vz^2 over the live particles of [p0, p1) (the push's
// live test), in a fixed order
__global__ __launch_bounds__(GP_BLOCK) void k_vz2(const float *x, const float *y, const float *vz, int64_t p0,
                                                 int64_t p1, float lox, float hix, float loy, float hiy, double *part)
{
    __shared__ double s[GP_BLOCK / 64];
    double acc = 0.0;
    for (int64_t p = p0 + blockIdx.x * (int64_t)GP_BLOCK + threadIdx.x; p < p1; p += (int64_t)gridDim.x * GP_BLOCK) {
        const float px = x[p], py = y[p], c = vz[p];
        if (px >= lox && px < hix && py >= loy && py < hiy) acc += (double)(c * c);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < GP_BLOCK / 64; w++) t += s[w];
        part[blockIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// exclusive scan of uint32 (3-phase, recursive over block sums)
// ---------------------------------------------------------------------------
constexpr int SC_ITEMS = 8, SC_TILE = GP_BLOCK * SC_ITEMS;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *tot)
{
    __shared__ uint32_t ws[GP_BLOCK / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t inc = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) ws[wv] = inc;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (int w = 0; w < GP_BLOCK / 64; w++) {
        if (w < wv) base += ws[w];
        all += ws[w];
    }
    __syncthreads();
    *tot = all;
    return base + inc - v;
}

__global__ __launch_bounds__(GP_BLOCK) void k_scan_block(uint32_t *d, int64_t n, uint32_t *sums)
{
    const int64_t base = (int64_t)blockIdx.x * SC_TILE + (int64_t)threadIdx.x * SC_ITEMS;
    uint32_t v[SC_ITEMS], s = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        v[k] = (base + k < n) ? d[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t off = block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        if (base + k < n) d[base + k] = off;
        off += v[k];
    }
    if (threadIdx.x == 0 && sums) sums[blockIdx.x] = tot;
}
__global__ void k_scan_add(uint32_t *d, int64_t n, const uint32_t *sums)
{
    const int64_t i = (int64_t)blockIdx.x * SC_TILE + threadIdx.x;
    const uint32_t s = sums[blockIdx.x];
    for (int k = 0; k < SC_ITEMS; k++) {
        const int64_t j = i + (int64_t)k * GP_BLOCK;
        if (j < n) d[j] += s;
    }
}

// block sums of a scan of n values: one uint32 per block, recursively
size_t scan_scratch_bytes(int64_t n)
{
    const int64_t nb = n > 0 ? (n + SC_TILE - 1) / SC_TILE : 0;
    if (nb <= 1) return 0;
    return gp_align256(sizeof(uint32_t) * (size_t)nb) + scan_scratch_bytes(nb);
}

// Exclusive scan in place.  The block sums come from the arena when one is
// given (and are released to it on return), else from the stream allocator.
int scan_u32(hipStream_t s, uint32_t *d, int64_t n, gp_arena *ar)
{
    if (n <= 0) return GP_OK;
    const int nb = gp_cdiv(n, SC_TILE);
    if (nb == 1) {
        k_scan_block<<<1, GP_BLOCK, 0, s>>>(d, n, nullptr);
        GP_HIP(hipGetLastError());
        return GP_OK;
    }
    uint32_t *sums = nullptr;
    const size_t mark = ar ? ar->used : 0;
    if (ar) {
        sums = (uint32_t *)ar->take(sizeof(uint32_t) * nb);
        if (!sums) return gp_fail(GP_ERR_CAPACITY, "scan: scratch arena exhausted");
    } else {
        GP_HIP(hipMallocAsync((void **)&sums, sizeof(uint32_t) * nb, s));
    }
    k_scan_block<<<nb, GP_BLOCK, 0, s>>>(d, n, sums);
    GP_HIP(hipGetLastError());
    int rc = scan_u32(s, sums, nb, ar);
    if (rc) return rc;
    k_scan_add<<<nb, GP_BLOCK, 0, s>>>(d, n, sums);
    GP_HIP(hipGetLastError());
    if (ar) ar->used = mark;
    else GP_HIP(hipFreeAsync(sums, s));
    return GP_OK;
}

// ---------------------------------------------------------------------------
// stable LSD radix sort of (uint32 key, uint32 value), DIG-bit digits (8 or 9)
// ---------------------------------------------------------------------------
constexpr int RS_TILE = 4096;

template <int DIG>
__global__ __launch_bounds__(GP_BLOCK) void k_rs_hist(const uint32_t *keys, int64_t n, int shift,
                                                      uint32_t *hist, int nblk)
{
    constexpr int NB = 1 << DIG;
    __shared__ uint32_t h[NB];
    for (int d = threadIdx.x; d < NB; d += GP_BLOCK) h[d] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int k = threadIdx.x; k < RS_TILE; k += GP_BLOCK) {
        const int64_t i = base + k;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (NB - 1u)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < NB; d += GP_BLOCK) hist[(int64_t)d * nblk + blockIdx.x] = h[d];
}

// Each chunk of GP_BLOCK keys is ranked per digit with wavefront ballots (the
// lanes of equal digit, in lane order), then waves in order: stable.
template <int DIG>
__global__ __launch_bounds__(GP_BLOCK) void k_rs_scatter(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                                                         uint32_t *vout, int64_t n, int shift, int init_vals,
                                                         const uint32_t *offs, int nblk)
{
    constexpr int NB = 1 << DIG;
    __shared__ uint32_t run[NB];
    __shared__ uint32_t wc[GP_BLOCK / 64][NB];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int d = tid; d < NB; d += GP_BLOCK) run[d] = offs[(int64_t)d * nblk + blockIdx.x];
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int c = 0; c < RS_TILE / GP_BLOCK; c++) {
        for (int d = tid; d < NB; d += GP_BLOCK)
            for (int w = 0; w < GP_BLOCK / 64; w++) wc[w][d] = 0;
        __syncthreads();
        const int64_t i = base + (int64_t)c * GP_BLOCK + tid;
        const bool valid = i < n;
        const uint32_t k = valid ? kin[i] : 0u;
        const uint32_t d = (k >> shift) & (NB - 1u);
        const uint32_t v = valid ? (init_vals ? (uint32_t)i : vin[i]) : 0u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < DIG; bit++) {
            const bool on = (d >> bit) & 1u;
            const unsigned long long bb = __ballot(valid && on);
            peers &= on ? bb : ~bb;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wc[wv][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = run[d] + rank;
            for (int w = 0; w < wv; w++) off += wc[w][d];
            kout[off] = k;
            vout[off] = v;
        }
        __syncthreads();
        for (int e = tid; e < NB; e += GP_BLOCK) {
            uint32_t add = 0;
            for (int w = 0; w < GP_BLOCK / 64; w++) add += wc[w][e];
            run[e] += add;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// sort keys, bucket offsets, interleaved placement
//
// The particles of a tile are stored class-interleaved: class
// b = (cx + W cy) mod 64 of the particle's cell (cx, cy) within the tile,
// W the row pitch of the push kernel's LDS windows, is the LDS bank group of
// the cell's stencil base.  Position q of the tile holds the r-th particle
// (in the previous order) of a class, rank-major: all classes' particle 0 in
// class order, then all particle 1, ...  A wavefront of freshly sorted
// particles therefore addresses 64 different bank groups.
// ---------------------------------------------------------------------------
// the cell (ci, cj) of a position, clamped to the grid (the sort's binning)
__device__ __forceinline__ void bin_cell(float x, float y, int ngb, int ngx, int ngy, int &ci, int &cj)
{
    ci = (int)floorf(x) - ngb;
    cj = (int)floorf(y) - ngb;
    ci = ci < 0 ? 0 : (ci >= ngx ? ngx - 1 : ci);
    cj = cj < 0 ? 0 : (cj >= ngy ? ngy - 1 : cj);
}
__global__ void k_tile_key(const float *x, const float *y, int64_t n, int ngb, int ngx, int ngy, int ntx, int W,
                           int cls, uint32_t *key)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ci, cj;
    bin_cell(x[i], y[i], ngb, ngx, ngy, ci, cj);
    const uint32_t t = (uint32_t)((cj / GP_TILE) * ntx + ci / GP_TILE);
    const uint32_t b = (uint32_t)((ci % GP_TILE) + W * (cj % GP_TILE)) & 63u;
    key[i] = cls ? t * 64u + b : t;
}
__global__ void k_tile_off(const uint32_t *key, int64_t n, int ntiles, int32_t *off)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = (int)key[i];
    const int kp = i > 0 ? (int)key[i - 1] : -1;
    for (int t = kp + 1; t <= k; t++) off[t] = (int32_t)i;
    if (i == n - 1)
        for (int t = k + 1; t <= ntiles; t++) off[t] = (int32_t)n;
}
// first sorted position of every key value 0..nb-1, and off[nb] = n
__global__ void k_permute(gp_soa in, gp_soa out, const uint32_t *perm, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = perm[i];
    out.x[i] = in.x[s]; out.y[i] = in.y[s];
    out.vx[i] = in.vx[s]; out.vy[i] = in.vy[s]; out.vz[i] = in.vz[s];
}
__global__ void k_bucket_off(const uint32_t *key, int64_t n, int64_t nb, uint32_t *off)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t k = key[i];
    const int64_t kp = i > 0 ? (int64_t)key[i - 1] : -1;
    for (int64_t t = kp + 1; t <= k; t++) off[t] = (uint32_t)i;
    if (i == n - 1)
        for (int64_t t = k + 1; t <= nb; t++) off[t] = (uint32_t)n;
}
// One workgroup per tile; lane = class.  Rank r of the tile's classes goes
// to positions M(r) .. M(r) + popcount(classes with more than r particles),
// M(r) the particles of ranks below r: the popcounts of a chunk of ranks are
// ballots, M is their scan.  Only the permutation is interleaved here (perm
// holds the old index of each (tile, class)-sorted position; dst[d]
// receives the old index of new position d): a lane walks one bucket of
// perm.  The particles follow in one gather (k_permute).
__global__ __launch_bounds__(GP_BLOCK) void k_interleave(const uint32_t *__restrict__ boff,
                                                         const uint32_t *__restrict__ perm,
                                                         uint32_t *__restrict__ dst, int32_t *tile_off, int ntiles)
{
    __shared__ uint32_t sM[SC_TILE];
    const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t b0 = boff[t * 64 + lane], nbk = boff[t * 64 + lane + 1] - b0;
    const uint32_t t0 = boff[t * 64];
    if (tid == 0) tile_off[t] = (int32_t)t0;
    if (t == ntiles - 1 && tid == 0) tile_off[ntiles] = (int32_t)boff[ntiles * 64];
    uint32_t R = nbk;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) R = max(R, (uint32_t)__shfl_xor((int)R, o, 64));
    const unsigned long long lt = (1ull << lane) - 1ull;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < R; c0 += SC_TILE) {
        for (int k = wv; k < SC_TILE; k += GP_BLOCK / 64) {
            uint32_t pc = 0;
            if (c0 + k < R) pc = (uint32_t)__popcll(__ballot(c0 + k < nbk));
            if (lane == 0) sM[k] = pc;
        }
        __syncthreads();
        uint32_t v[SC_ITEMS], sum = 0;
#pragma unroll
        for (int q = 0; q < SC_ITEMS; q++) { v[q] = sM[tid * SC_ITEMS + q]; sum += v[q]; }
        uint32_t tot;
        uint32_t off = block_excl_scan(sum, &tot);
#pragma unroll
        for (int q = 0; q < SC_ITEMS; q++) { sM[tid * SC_ITEMS + q] = carry + off; off += v[q]; }
        __syncthreads();
        const uint32_t kend = min((uint32_t)SC_TILE, R - c0);
        for (uint32_t k = wv; k < kend; k += GP_BLOCK / 64) {
            const uint32_t r = c0 + k;
            const bool on = r < nbk;
            const unsigned long long m = __ballot(on);
            if (on) dst[t0 + sM[k] + (uint32_t)__popcll(m & lt)] = perm[b0 + r];
        }
        carry += tot;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// neighbour re-bin: the sort of a binned species whose particles mostly
// stayed in, or moved next to, the tile whose segment holds them.
//
// A particle at index i of segment s (tile_off[s] <= i < tile_off[s+1],
// i < n_binned) that is now in tile t arrives from the source offset
// e = s - t (tile coordinates, periodic wrap on both axes); it is "near" when
// both components of e lie in -1..1, "far" otherwise, and every particle of
// the unbinned tail (i >= n_binned) is far.  Tile t then holds, in order:
// its near arrivals grouped by e in row-major order from (-1, -1) to (+1, +1)
// (group k = 3 (ey + 1) + ex + 1), each group in the old order, then its far
// arrivals in the old order.  Without far particles or moves across the
// grid's edge this is the stable sort by tile.
//
// Passes: k_rb_count (moves per segment and code, far arrivals per tile),
// k_rb_total + scans (first position of every tile), k_rb_place (first
// position of every (segment, code) group), k_rb_scatter (near particles
// moved to their places; far ones listed as (tile, index) in index order),
// a stable radix sort of the far list by tile and k_rb_far, then
// k_rb_offsets (the new tile_off).  Particle data moves once; the count pass
// reads x and y.  Needs at least 3 tiles per axis (the nine source offsets
// of a tile are then distinct tiles).
// ---------------------------------------------------------------------------
constexpr int RB_CODES = 10;   // move codes 0..8 = 3 (dy + 1) + dx + 1 with d = t - s; RB_FAR
constexpr int RB_FAR = 9;
constexpr int64_t RB_CHUNK = 32768;   // tail particles per workgroup

struct RebinArgs {
    gp_soa a, b;                 // input order, output order
    const int32_t *tile_off;     // segments of the input (ntiles + 1)
    int64_t n, n_binned;
    int ngb, ngx, ngy, ntx, nty, ntiles;
    uint32_t *cnt;      // ntiles * RB_CODES: particles per (segment, code); group positions after k_rb_place
    uint32_t *start;    // ntiles + 1: arrivals per tile, scanned: first position of each tile
    uint32_t *fart;     // ntiles + 1: far arrivals per tile, scanned: first entry in the sorted far list
    uint32_t *fbase;    // ntiles: first position of each tile's far arrivals
    uint32_t *segfar;   // nseg + 1: far particles per workgroup, scanned: first entry in the far list
    uint32_t *fkey, *fval;       // far list: tile, old index
};

// workgroup b's particles [lo, hi) and its segment (-1 for the tail)
__device__ __forceinline__ void rb_range(const RebinArgs &a, int b, int64_t &lo, int64_t &hi, int &s)
{
    if (b < a.ntiles) {
        lo = min((int64_t)a.tile_off[b], a.n_binned);
        hi = min((int64_t)a.tile_off[b + 1], a.n_binned);
        s = b;
    } else {
        lo = a.n_binned + (int64_t)(b - a.ntiles) * RB_CHUNK;
        hi = min(lo + RB_CHUNK, a.n);
        s = -1;
    }
    lo = max(lo, (int64_t)0);
    hi = max(hi, lo);
}
// tile (tx, ty) of a particle and its move code from source tile (sx, sy)
__device__ __forceinline__ int rb_code(const RebinArgs &a, float x, float y, int sx, int sy, int &t)
{
    int ci, cj;
    bin_cell(x, y, a.ngb, a.ngx, a.ngy, ci, cj);
    const int tx = ci / GP_TILE, ty = cj / GP_TILE;
    t = ty * a.ntx + tx;
    if (sx < 0) return RB_FAR;
    int dx = tx - sx, dy = ty - sy;
    dx = dx == a.ntx - 1 ? -1 : (dx == 1 - a.ntx ? 1 : dx);
    dy = dy == a.nty - 1 ? -1 : (dy == 1 - a.nty ? 1 : dy);
    return (dx < -1 || dx > 1 || dy < -1 || dy > 1) ? RB_FAR : 3 * (dy + 1) + dx + 1;
}
__device__ __forceinline__ int rb_wrap(int v, int m) { return v < 0 ? v + m : (v >= m ? v - m : v); }

__global__ void k_rb_zero(RebinArgs a, int nseg)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k <= a.ntiles) a.fart[k] = 0u;
    if (k == 0) a.segfar[nseg] = 0u;
}

__global__ __launch_bounds__(GP_BLOCK) void k_rb_count(RebinArgs a)
{
    __shared__ uint32_t sc[GP_BLOCK / 64][RB_CODES];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int64_t lo, hi;
    int s;
    rb_range(a, blockIdx.x, lo, hi, s);
    const int sx = s >= 0 ? s % a.ntx : -1, sy = s >= 0 ? s / a.ntx : -1;
    uint32_t c[RB_CODES];
#pragma unroll
    for (int k = 0; k < RB_CODES; k++) c[k] = 0u;
    for (int64_t base = lo; base < hi; base += GP_BLOCK) {
        const int64_t i = base + tid;
        const bool valid = i < hi;
        int code = -1, t = 0;
        if (valid) {
            code = rb_code(a, a.a.x[i], a.a.y[i], sx, sy, t);
            if (code == RB_FAR) atomicAdd(&a.fart[t], 1u);
        }
#pragma unroll
        for (int k = 0; k < RB_CODES; k++) c[k] += (uint32_t)__popcll(__ballot(code == k));
    }
    if (lane == 0)
        for (int k = 0; k < RB_CODES; k++) sc[wv][k] = c[k];
    __syncthreads();
    if (tid < RB_CODES) {
        uint32_t v = 0;
        for (int w = 0; w < GP_BLOCK / 64; w++) v += sc[w][tid];
        if (tid == RB_FAR) a.segfar[blockIdx.x] = v;
        else if (s >= 0) a.cnt[(int64_t)s * RB_CODES + tid] = v;
    }
}

// arrivals per tile: near groups from the nine source tiles, then far
__global__ void k_rb_total(RebinArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > a.ntiles) return;
    uint32_t sum = 0;
    if (t < a.ntiles) {
        const int tx = t % a.ntx, ty = t / a.ntx;
        for (int k = 0; k < 9; k++) {
            const int s = rb_wrap(ty + k / 3 - 1, a.nty) * a.ntx + rb_wrap(tx + k % 3 - 1, a.ntx);
            sum += a.cnt[(int64_t)s * RB_CODES + (8 - k)];
        }
        sum += a.fart[t];
    }
    a.start[t] = sum;
}

// group k of tile t comes from source s = t + e_k with move code 8 - k; its
// count is replaced by its first position (each entry belongs to one tile)
__global__ void k_rb_place(RebinArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntiles) return;
    const int tx = t % a.ntx, ty = t / a.ntx;
    uint32_t p = a.start[t];
    for (int k = 0; k < 9; k++) {
        const int s = rb_wrap(ty + k / 3 - 1, a.nty) * a.ntx + rb_wrap(tx + k % 3 - 1, a.ntx);
        uint32_t *e = a.cnt + (int64_t)s * RB_CODES + (8 - k);
        const uint32_t v = *e;
        *e = p;
        p += v;
    }
    a.fbase[t] = p;
}

// Each chunk of GP_BLOCK particles is ranked per code with wavefront ballots
// (lane order), then waves in order: the old order within every group.
__global__ __launch_bounds__(GP_BLOCK) void k_rb_scatter(RebinArgs a)
{
    constexpr int NW = GP_BLOCK / 64;
    __shared__ uint32_t run[RB_CODES];
    __shared__ uint32_t wc[2][NW][RB_CODES];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int64_t lo, hi;
    int s;
    rb_range(a, blockIdx.x, lo, hi, s);
    const int sx = s >= 0 ? s % a.ntx : -1, sy = s >= 0 ? s / a.ntx : -1;
    if (tid < RB_CODES)
        run[tid] = tid == RB_FAR ? a.segfar[blockIdx.x] : (s >= 0 ? a.cnt[(int64_t)s * RB_CODES + tid] : 0u);
    const unsigned long long lt = (1ull << lane) - 1ull;
    int par = 0;
    for (int64_t base = lo; base < hi; base += GP_BLOCK, par ^= 1) {
        const int64_t i = base + tid;
        const bool valid = i < hi;
        float x = 0.0f, y = 0.0f, vx = 0.0f, vy = 0.0f, vz = 0.0f;
        int code = 0, t = 0;
        if (valid) {
            x = a.a.x[i]; y = a.a.y[i]; vx = a.a.vx[i]; vy = a.a.vy[i]; vz = a.a.vz[i];
            code = rb_code(a, x, y, sx, sy, t);
        }
        const unsigned long long vm = __ballot(valid);
        unsigned long long bm[4];
#pragma unroll
        for (int bit = 0; bit < 4; bit++) bm[bit] = __ballot(valid && ((code >> bit) & 1));
        unsigned long long peers = vm;
#pragma unroll
        for (int bit = 0; bit < 4; bit++) peers &= ((code >> bit) & 1) ? bm[bit] : ~bm[bit];
        const int rank = __popcll(peers & lt);
        if (lane < RB_CODES) {
            unsigned long long m = vm;
#pragma unroll
            for (int bit = 0; bit < 4; bit++) m &= ((lane >> bit) & 1) ? bm[bit] : ~bm[bit];
            wc[par][wv][lane] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (valid) {
            uint32_t off = run[code] + (uint32_t)rank;
            for (int w = 0; w < wv; w++) off += wc[par][w][code];
            if (off >= (uint64_t)a.n) {
                // unreachable with disjoint segments; keeps a corrupt
                // tile_off from writing out of bounds
            } else if (code == RB_FAR) {
                a.fkey[off] = (uint32_t)t;
                a.fval[off] = (uint32_t)i;
            } else {
                a.b.x[off] = x; a.b.y[off] = y; a.b.vx[off] = vx; a.b.vy[off] = vy; a.b.vz[off] = vz;
            }
        }
        __syncthreads();
        if (tid < RB_CODES) {
            uint32_t add = 0;
            for (int w = 0; w < NW; w++) add += wc[par][w][tid];
            run[tid] += add;
        }
    }
}

// far particles, sorted by tile (stable: old order within a tile)
__global__ void k_rb_far(RebinArgs a, const uint32_t *ks, const uint32_t *vs, int64_t nf)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= nf) return;
    const uint32_t t = ks[j], i = vs[j];
    if (t >= (uint32_t)a.ntiles || i >= (uint64_t)a.n) return;
    const uint32_t off = a.fbase[t] + ((uint32_t)j - a.fart[t]);
    if (off >= (uint64_t)a.n) return;
    a.b.x[off] = a.a.x[i]; a.b.y[off] = a.a.y[i];
    a.b.vx[off] = a.a.vx[i]; a.b.vy[off] = a.a.vy[i]; a.b.vz[off] = a.a.vz[i];
}

__global__ void k_rb_offsets(RebinArgs a, int32_t *tile_off)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= a.ntiles) tile_off[t] = (int32_t)a.start[t];
}

// ---------------------------------------------------------------------------
// compaction: holes (absorbed, index < n_new) ascending are filled by the
// live particles of the tail [n_new, n) taken from the end, which is the
// permutation of the reference's loop (Particle.mm:906-968).
// ---------------------------------------------------------------------------
// ascending sort of n <= SMALL_SORT distinct keys in one workgroup: a
// bitonic network in LDS, padded to a power of two with the largest key (the
// hundreds to about ten thousand particles a bounded step absorbs, where the
// radix sort's nine launches cost more than the sort)
constexpr int SMALL_SORT = 16384;   // 64 KB of LDS
__global__ __launch_bounds__(1024) void k_sort_small(const uint32_t *in, uint32_t *out, int n)
{
    __shared__ uint32_t s[SMALL_SORT];
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = threadIdx.x; i < m; i += 1024) s[i] = i < n ? in[i] : 0xffffffffu;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += 1024) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t x = s[i], y = s[l];
                    if ((x > y) == ((i & k) == 0)) {
                        s[i] = y;
                        s[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < n; i += 1024) out[i] = s[i];
}
__global__ void k_lower_bound(const uint32_t *a, int64_t n, uint32_t v, uint32_t *out)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    *out = (uint32_t)lo;
}
__global__ void k_tail_flags(const float *x, int64_t n_new, int64_t len, uint32_t *flag)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < len) flag[k] = (x[n_new + k] >= 0.0f) ? 1u : 0u;
}
__global__ void k_donors(const float *x, int64_t n_new, int64_t len, const uint32_t *pos, uint32_t *donor)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < len && x[n_new + k] >= 0.0f) donor[pos[k]] = (uint32_t)(n_new + k);
}
__global__ void k_fill(gp_soa s, const uint32_t *holes, const uint32_t *donor, const uint32_t *nh)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t h = *nh;
    if (k >= h) return;
    const uint32_t d = donor[h - 1 - k], t = holes[k];
    s.x[t] = s.x[d]; s.y[t] = s.y[d]; s.vx[t] = s.vx[d]; s.vy[t] = s.vy[d]; s.vz[t] = s.vz[d];
}

// ---------------------------------------------------------------------------
// seeded generation (Collision.mm:63-174, seeding :280-294)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t smix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
struct Pcg {
    uint64_t s, inc;
    __device__ __forceinline__ uint32_t next()
    {
        const uint64_t old = s;
        s = old * 6364136223846793005ULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    // rand01 of Collision.mm:28-30, float(u) 2^-32 (round to nearest even),
    // with u = 0 mapped to 2^-33 (the draws feed log) and a conversion that
    // rounds to 2^32 mapped to the largest float below 1
    __device__ __forceinline__ float u01()
    {
        const uint32_t u = next();
        const float r = (float)u * 0x1p-32f;
        return u == 0u ? 0x1p-33f : (r < 1.0f ? r : 0x1.fffffep-1f);
    }
};
__device__ __forceinline__ Pcg seed_stream(uint64_t master, int64_t id)
{
    const uint64_t g = 0x9E3779B97F4A7C15ULL;
    Pcg r;
    r.s = smix(master + (uint64_t)(2 * id + 1) * g);
    r.inc = smix(master + (uint64_t)(2 * id + 2) * g) | 1ULL;
    return r;
}

// seeded loads recorded per context whose draws a stream state recomputed
// from its seed must replay
constexpr int GP_RNG_RECS = 8;
struct LoadArgs {
    gp_gen g;
    float dx, dy, ngb, xdom, ydom;
    uint64_t seed;
    int64_t first_id, n, ntot;
    gp_soa e, i;
    int64_t ne, ni;
    int with_ion;
    // materialised stream states of ids < rng_n (read, then written back)
    uint64_t *rs, *ri;
    int64_t rng_n;
    // earlier loads on the same seed: ids in [rec_first, rec_first + rec_n)
    // have already drawn rec_draws numbers
    int nrec;
    int64_t rec_first[GP_RNG_RECS], rec_n[GP_RNG_RECS];
    int rec_draws[GP_RNG_RECS];
};

// numbers one generation call draws per thread (Collision.mm:63-174):
// position 2 (3 for the Gaussian profiles), electron velocity 4, ion
// velocity 4 when the call also writes the ion
int gen_draws(const gp_gen *g, bool with_ion)
{
    const int t = g->gen_type;
    return 2 + ((t == 5 || t == 6) ? 1 : 0) + 4 + ((with_ion && t != 3) ? 4 : 0);
}

// the stream of thread id: its materialised state, or the state seeded from
// the master seed and advanced past the draws of the recorded loads
__device__ __forceinline__ Pcg stream_of(const LoadArgs &a, int64_t id)
{
    if (id < a.rng_n) {
        Pcg r;
        r.s = a.rs[id];
        r.inc = a.ri[id];
        return r;
    }
    Pcg r = seed_stream(a.seed, id);
    for (int q = 0; q < a.nrec; q++)
        if (id >= a.rec_first[q] && id < a.rec_first[q] + a.rec_n[q])
            for (int d = 0; d < a.rec_draws[q]; d++) (void)r.next();
    return r;
}

// materialise the stream states of ids [lo, hi)
__global__ void k_rng_init(LoadArgs a, int64_t lo, int64_t hi, uint64_t *rs, uint64_t *ri)
{
    const int64_t id = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= hi) return;
    const Pcg r = stream_of(a, id);
    rs[id] = r.s;
    ri[id] = r.inc;
}

__global__ void k_load(LoadArgs a)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const float PIf = 3.14159265358979f;
    const int64_t id = a.first_id + k;
    Pcg r = stream_of(a, id);
    const gp_gen &g = a.g;
    const float rx = r.u01(), ry = r.u01();
    float x, y;
    switch (g.gen_type) {
    default:
    case 0: case 3:
        x = g.xmin + rx * (g.xmax - g.xmin);
        y = g.ymin + ry * (g.ymax - g.ymin);
        break;
    case 1:
        x = (g.xmin + g.xmax) / 2.0f + (g.xmax - g.xmin) / PIf * asinf(2.0f * rx - 1.0f);
        y = g.ymin + ry * (g.ymax - g.ymin);
        break;
    case 2:
        x = g.xmin + rx * (g.xmax - g.xmin);
        y = (g.ymin + g.ymax) / 2.0f + (g.ymax - g.ymin) / PIf * asinf(2.0f * ry - 1.0f);
        break;
    case 4: {
        int ax = (int)roundf(sqrtf((float)a.ntot * (g.xmax - g.xmin) / (g.ymax - g.ymin)));
        if (ax < 1) ax = 1;
        int ay = (int)(a.ntot / ax);
        if (ay < 1) ay = 1;
        const float dq = (g.xmax - g.xmin) / (float)ax, eq = (g.ymax - g.ymin) / (float)ay;
        x = g.xmin + dq * (0.5f + (float)(id % ax));
        y = g.ymin + eq * (0.5f + (float)(id / ax));
        break; }
    case 5: {
        const float rz = r.u01();
        x = (g.xmin + g.xmax) / 2.0f + (g.xmax - g.xmin) * sqrtf(-logf(rx)) * cosf(2.0f * PIf * rz);
        y = g.ymin + ry * (g.ymax - g.ymin);
        break; }
    case 6: {
        const float rz = r.u01();
        x = g.xmin + rx * (g.xmax - g.xmin);
        y = (g.ymin + g.ymax) / 2.0f + (g.ymax - g.ymin) * sqrtf(-logf(rz)) * cosf(2.0f * PIf * ry);
        break; }
    case 7: {
        const double L = (double)g.xmax - (double)g.xmin;
        const double kk = 2.0 * 3.141592653589793 * (double)g.pert_mode / L, d = g.pert_delta;
        const double tgt = (double)rx * L;
        double s = tgt;
        for (int it = 0; it < 8; it++) {
            const double f = s + d / kk * sin(kk * s) - tgt, fp = 1.0 + d * cos(kk * s);
            s -= f / fp;
        }
        x = (float)((double)g.xmin + s);
        y = g.ymin + ry * (g.ymax - g.ymin);
        break; }
    }
    float rv1 = r.u01(), rv2 = r.u01();
    float am = g.ele_vth * sqrtf(-logf(rv1));
    const float vx = g.ele_u[0] + am * cosf(2.0f * PIf * rv2);
    const float vy = g.ele_u[1] + am * sinf(2.0f * PIf * rv2);
    rv1 = r.u01(); rv2 = r.u01();
    const float vz = g.ele_u[2] + g.ele_vth * sqrtf(-logf(rv1)) * cosf(2.0f * PIf * rv2);
    // a position inside the domain or on its upper face (cm) that lands on
    // the upper face in grid units is moved just below it (the push takes
    // [ngb, ngb+ng))
    {
        const float hx = a.ngb + a.xdom, hy = a.ngb + a.ydom;
        const float xc = x, yc = y;
        x = x / a.dx + a.ngb;
        y = y / a.dy + a.ngb;
        if (xc <= a.xdom * a.dx && x >= hx) x = __int_as_float(__float_as_int(hx) - 1);
        if (yc <= a.ydom * a.dy && y >= hy) y = __int_as_float(__float_as_int(hy) - 1);
    }
    const int64_t je = a.ne + k;
    a.e.x[je] = x; a.e.y[je] = y; a.e.vx[je] = vx; a.e.vy[je] = vy; a.e.vz[je] = vz;
    if (a.with_ion && g.gen_type != 3) {
        rv1 = r.u01(); rv2 = r.u01();
        am = g.ion_vth * sqrtf(-logf(rv1));
        const int64_t ji = a.ni + k;
        a.i.vx[ji] = g.ion_u[0] + am * cosf(2.0f * PIf * rv2);
        a.i.vy[ji] = g.ion_u[1] + am * sinf(2.0f * PIf * rv2);
        rv1 = r.u01(); rv2 = r.u01();
        a.i.vz[ji] = g.ion_u[2] + g.ion_vth * sqrtf(-logf(rv1)) * cosf(2.0f * PIf * rv2);
        a.i.x[ji] = x; a.i.y[ji] = y;
    }
    if (id < a.rng_n) a.rs[id] = r.s;   // the stream continues in the next call
}

// half-Maxwellian flux through the x-min wall: vx Rayleigh (flux-weighted),
// vy, vz Gaussian; the particle entered during the step, x = wall + vx*dt*r
struct InjArgs {
    gp_soa s;
    int64_t base, n;
    uint64_t seed, stream;
    float vth, u0, u1, u2;
    float dtdx, lo, ngx, ngy;
};
__global__ void k_inject(InjArgs a)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const float PIf = 3.14159265358979f;
    Pcg r = seed_stream(a.seed, (int64_t)(a.stream + (uint64_t)k));
    const float r1 = r.u01(), r2 = r.u01(), r3 = r.u01(), r4 = r.u01(), r5 = r.u01();
    const float vx = a.u0 + a.vth * sqrtf(-2.0f * logf(r1));
    const float am = a.vth * sqrtf(-2.0f * logf(r2));
    const float vy = a.u1 + am * cosf(2.0f * PIf * r3);
    const float vz = a.u2 + am * sinf(2.0f * PIf * r3);
    float x = a.lo + vx * a.dtdx * r4;
    const float hi = a.lo + a.ngx;
    if (x >= hi) x = __int_as_float(__float_as_int(hi) - 1);   // a particle faster than L/dt stays inside
    // r5 = 1 - 2^-24 can round onto the upper face (as a loaded position can)
    float y = a.lo + r5 * a.ngy;
    const float hy = a.lo + a.ngy;
    if (y >= hy) y = __int_as_float(__float_as_int(hy) - 1);
    const int64_t j = a.base + k;
    a.s.x[j] = x; a.s.y[j] = y; a.s.vx[j] = vx; a.s.vy[j] = vy; a.s.vz[j] = vz;
}

// moments (Moment.mm:33-135): node weights of int(x + 1e-20)
struct MomArgs {
    const float *x, *y, *vx, *vy, *vz;
    int64_t n;
    int ngb, ngx, ngy, nx;
    float inv_vnorm;
    unsigned long long *acc;   // 10 grids
    int64_t ng;
    unsigned long long *escaped;
};
template <int ORDER>
__global__ void k_moments(MomArgs a)
{
    constexpr int NW = ORDER + 1, NGB = ORDER / 2;
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const float x = a.x[p], y = a.y[p];
    if (!(x >= (float)a.ngb && x < (float)(a.ngb + a.ngx) && y >= (float)a.ngb && y < (float)(a.ngb + a.ngy))) {
        if (x >= 0.0f) atomicAdd(a.escaped, 1ull);
        return;
    }
    const int i1 = (int)(x + 1e-20f), j1 = (int)(y + 1e-20f);
    float sx[NW], sy[NW];
    wts<ORDER>(x - (float)i1, sx);
    wts<ORDER>(y - (float)j1, sy);
    const float v[3] = {a.vx[p] * a.inv_vnorm, a.vy[p] * a.inv_vnorm, a.vz[p] * a.inv_vnorm};
    const int pit = a.nx + 1;
    for (int j = 0; j < NW; j++)
        for (int i = 0; i < NW; i++) {
            const int64_t node = (i1 + i - NGB) + (int64_t)(j1 + j - NGB) * pit;
            // the density takes the same FP32 node weight as the velocity
            // sums, so that n v / n at a node of few particles is their mean
            const float s = sx[i] * sy[j];
            atomicAdd(&a.acc[node], (unsigned long long)fx40(s));
            atomicAdd(&a.acc[1 * a.ng + node], (unsigned long long)fx40(s * v[0]));
            atomicAdd(&a.acc[2 * a.ng + node], (unsigned long long)fx40(s * v[1]));
            atomicAdd(&a.acc[3 * a.ng + node], (unsigned long long)fx40(s * v[2]));
            atomicAdd(&a.acc[4 * a.ng + node], (unsigned long long)fx40((s * v[0]) * v[0]));
            atomicAdd(&a.acc[5 * a.ng + node], (unsigned long long)fx40((s * v[0]) * v[1]));
            atomicAdd(&a.acc[6 * a.ng + node], (unsigned long long)fx40((s * v[0]) * v[2]));
            atomicAdd(&a.acc[7 * a.ng + node], (unsigned long long)fx40((s * v[1]) * v[1]));
            atomicAdd(&a.acc[8 * a.ng + node], (unsigned long long)fx40((s * v[1]) * v[2]));
            atomicAdd(&a.acc[9 * a.ng + node], (unsigned long long)fx40((s * v[2]) * v[2]));
        }
}
__global__ void k_mom_scale(const unsigned long long *acc, int64_t ng, double sn, double sv, double svv, double *out)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= 10 * ng) return;
    const int c = (int)(k / ng);
    const double f = c == 0 ? sn : (c <= 3 ? sv : svv);
    out[k] = (double)(long long)acc[k] * f;
}

// ---------------------------------------------------------------------------
template <int ORDER, bool UNIB, int DEP, bool PUSH>
void launch_push_t(int grid, hipStream_t s, const PushArgs &a, bool dc)
{
    const bool all_periodic = a.bc0 == GP_PBC_PERIODIC && a.bc1 == GP_PBC_PERIODIC && a.bc2 == GP_PBC_PERIODIC &&
                              a.bc3 == GP_PBC_PERIODIC;
    if constexpr (push_cw(ORDER, UNIB, DEP, PUSH)) {
        if (dc) {
            if (all_periodic) k_push<ORDER, UNIB, DEP, PUSH, 1, 1><<<grid, push_block(ORDER), 0, s>>>(a);
            else k_push<ORDER, UNIB, DEP, PUSH, 0, 1><<<grid, push_block(ORDER), 0, s>>>(a);
            return;
        }
    }
    if (PUSH && all_periodic) k_push<ORDER, UNIB, DEP, PUSH, 1><<<grid, push_block(ORDER), 0, s>>>(a);
    else k_push<ORDER, UNIB, DEP, PUSH, 0><<<grid, push_block(ORDER), 0, s>>>(a);
}
template <int ORDER>
int launch_push_o(int grid, hipStream_t s, const PushArgs &a, bool unib, int dep, bool push, bool dc)
{
    if (push) {
        if (unib) {
            if (dep == 0) launch_push_t<ORDER, true, 0, true>(grid, s, a, dc);
            else if (dep == 1) launch_push_t<ORDER, true, 1, true>(grid, s, a, dc);
            else launch_push_t<ORDER, true, 2, true>(grid, s, a, dc);
        } else {
            if (dep == 0) launch_push_t<ORDER, false, 0, true>(grid, s, a, dc);
            else if (dep == 1) launch_push_t<ORDER, false, 1, true>(grid, s, a, dc);
            else launch_push_t<ORDER, false, 2, true>(grid, s, a, dc);
        }
    } else {
        if (dep == 1) launch_push_t<ORDER, true, 1, false>(grid, s, a, dc);
        else launch_push_t<ORDER, true, 2, false>(grid, s, a, dc);
    }
    GP_HIP(hipGetLastError());
    return GP_OK;
}

// staging area of the tile windows (allocated on first use, kept by the context)
int ensure_stage(gp_ctx *c, int dep)
{
    const int W = WIN + c->order + 1;
    const size_t bytes = (size_t)c->ntiles * W * W * (dep >= 2 ? 3 : 1) * sizeof(unsigned long long);
    if (c->stage && c->stage_bytes >= bytes) return GP_OK;
    if (c->stage) GP_HIP(hipFree(c->stage));
    if (c->stage_on) GP_HIP(hipFree(c->stage_on));
    c->stage = nullptr; c->stage_on = nullptr; c->stage_bytes = 0;
    GP_HIP(hipMalloc((void **)&c->stage, bytes));
    GP_HIP(hipMalloc((void **)&c->stage_on, (size_t)c->ntiles));
    c->stage_bytes = bytes;
    return GP_OK;
}

int run_particle_kernel(gp_ctx *c, int sid, double dt, int dep, bool push, int64_t first, int64_t count)
{
    gp_species &S = c->sp[sid];
    PushArgs a{};
    a.x = S.a.x; a.y = S.a.y; a.vx = S.a.vx; a.vy = S.a.vy; a.vz = S.a.vz;
    const int64_t end = first + count;
    a.n = end;
    const bool whole = (first == 0 && end == S.n);
    a.use_tiles = whole && S.n_binned > 0 ? 1 : 0;
    a.n_binned = a.use_tiles ? std::min(S.n_binned, S.n) : 0;
    a.pfirst = a.use_tiles ? a.n_binned : first;
    a.tile_off = S.tile_off;
    a.ntx = c->ntx; a.ntiles = c->ntiles;
    a.Ex = c->Ex; a.Ey = c->Ey; a.Bz = c->Bz; a.bz0 = c->bz0;
    a.constE = (float)(0.5 * (double)S.qf * dt / (double)S.mf);             // Particle.mm:618
    a.constB = (float)(0.5 * (double)S.qf * dt / (double)S.mf / 2.99792458e10);  // Particle.mm:619
    a.constX = (float)(dt / (double)c->dxf);                                 // Particle.mm:620
    a.constY = (float)(dt / (double)c->dyf);
    a.ngx = c->ngx; a.ngy = c->ngy; a.ngb = c->ngb; a.nx = c->nx; a.ny = c->ny;
    a.moved_mask = c->ntiles >= 256 ? GP_MOVED_SAMPLE - 1 : 0;
    a.bc0 = c->cfg.pbc[0]; a.bc1 = c->cfg.pbc[1]; a.bc2 = c->cfg.pbc[2]; a.bc3 = c->cfg.pbc[3];
    a.boris = c->cfg.boris_order;
    // (a period of at least 2 tiles keeps every imaged stencil node within
    // one period of its interior node)
    a.imgx = dep <= 1 && c->cfg.pbc[0] == GP_PBC_PERIODIC && c->cfg.fbc[0] == GP_FBC_PERIODIC && c->ngx >= 2 * GP_TILE;
    a.imgy = dep <= 1 && c->cfg.pbc[2] == GP_PBC_PERIODIC && c->cfg.fbc[2] == GP_FBC_PERIODIC && c->ngy >= 2 * GP_TILE;
    a.gimgx = push && c->e_periodic && c->uniform_b && c->cfg.pbc[0] == GP_PBC_PERIODIC && c->cfg.fbc[0] == GP_FBC_PERIODIC &&
              c->ngx >= 2 * GP_TILE;
    a.gimgy = push && c->e_periodic && c->uniform_b && c->cfg.pbc[2] == GP_PBC_PERIODIC && c->cfg.fbc[2] == GP_FBC_PERIODIC &&
              c->ngy >= 2 * GP_TILE;
    a.rho = (unsigned long long *)S.rho_acc;
    a.jx = (unsigned long long *)S.jx_acc;
    a.jy = (unsigned long long *)S.jy_acc;
    if (a.use_tiles && dep >= 1) {
        int rc = ensure_stage(c, dep);
        if (rc) return rc;
        a.stage = (unsigned long long *)c->stage;
        a.stage_on = c->stage_on;
    }
    a.inv_vnorm = (float)S.inv_vnorm;
    a.dead = S.dead;
    a.counters = S.counters;
    a.ke_part = S.ke_part;
    a.ke_coef = 0.5 * (double)S.mf * (double)S.wf;
    const int64_t rest = end - a.pfirst;
    // particles outside the tile ranges go in chunks of up to 16 blocks; a
    // launch of few of them (an injection's deposit) spreads them over up to
    // 512 workgroups instead of leaving each of a few to loop over its chunk
    const int pb = push_block(c->order);
    a.chunk = pb * (int)std::min<int64_t>(16, std::max<int64_t>(1, gp_cdiv(rest, (int64_t)pb * 512)));
    const int extra = rest > 0 ? gp_cdiv(rest, a.chunk) : 0;
    const int grid = (a.use_tiles ? c->ntiles : 0) + extra;
    if (grid == 0) return GP_OK;
    if (grid > S.ke_nblk) return gp_fail(GP_ERR_STATE, "kinetic-energy partial buffer too small");
    // a species whose particles change cell too often to keep the bank-class
    // interleave of its sort (its lanes address random cells) adds its charge
    // in 32-bit halves; both forms give the same integers
    const bool dc = S.last_moved >= 0 && (double)S.last_moved >= GP_CLS_MOVED * (double)S.n;
    // the kernel that skips vz (k_push's NOVZ: order 3, uniform B, charge,
    // push, particles periodic on all sides) needs the sum of vz^2 over its
    // range, kept until the particles or the range change
    const bool novz = push_cw(c->order, c->uniform_b, dep, push) && a.bc0 == GP_PBC_PERIODIC &&
                      a.bc1 == GP_PBC_PERIODIC && a.bc2 == GP_PBC_PERIODIC && a.bc3 == GP_PBC_PERIODIC;
    if (push) S.ke_pending_n = 0;   // this push supersedes the last one's energy
    if (novz && (S.vz2_gen != S.vz_gen || S.vz2_first != first || S.vz2_end != end)) {
        const int gz = std::min(1024, std::max(1, gp_cdiv(end - first, GP_BLOCK)));
        k_vz2<<<gz, GP_BLOCK, 0, c->stream>>>(S.a.x, S.a.y, S.a.vz, first, end, (float)c->ngb, (float)(c->ngb + c->ngx),
                                              (float)c->ngb, (float)(c->ngb + c->ngy), S.ke_part);
        k_sum_partials<<<1, GP_BLOCK, 0, c->stream>>>(S.ke_part, gz, S.ke_out + 1);
        GP_HIP(hipGetLastError());
        S.vz2_gen = S.vz_gen; S.vz2_first = first; S.vz2_end = end;
    }
    const bool tk = push && c->timing;
    if (tk) { int rc = gp_time_mark(c, GP_TIME_PUSHK, -1); if (rc) return rc; }
    int rc;
    switch (c->order) {
    case 1: rc = launch_push_o<1>(grid, c->stream, a, c->uniform_b, dep, push, dc); break;
    case 3: rc = launch_push_o<3>(grid, c->stream, a, c->uniform_b, dep, push, dc); break;
    default: rc = launch_push_o<5>(grid, c->stream, a, c->uniform_b, dep, push, dc); break;
    }
    if (rc) return rc;
    if (tk) { rc = gp_time_mark(c, GP_TIME_PUSHK, -1); if (rc) return rc; }
    if (a.use_tiles && dep >= 1) {
        const int nc = dep >= 2 ? 3 : 1;
        const dim3 g(gp_cdiv(c->nx + 2, GP_BLOCK), c->ny + 2, nc);
        switch (c->order) {
        case 1: k_stage_reduce<1><<<g, GP_BLOCK, 0, c->stream>>>(a.stage, a.stage_on, nc, c->ntx, c->nty, c->ngb, c->ngx, c->ngy, c->nx, c->ny, a.imgx, a.imgy, a.rho, a.jx, a.jy); break;
        case 3: k_stage_reduce<3><<<g, GP_BLOCK, 0, c->stream>>>(a.stage, a.stage_on, nc, c->ntx, c->nty, c->ngb, c->ngx, c->ngy, c->nx, c->ny, a.imgx, a.imgy, a.rho, a.jx, a.jy); break;
        default: k_stage_reduce<5><<<g, GP_BLOCK, 0, c->stream>>>(a.stage, a.stage_on, nc, c->ntx, c->nty, c->ngb, c->ngx, c->ngy, c->nx, c->ny, a.imgx, a.imgy, a.rho, a.jx, a.jy); break;
        }
        GP_HIP(hipGetLastError());
    }
    if (push) {
        // summed only if the energy is read (gp_reduce, the step diagnostics)
        S.ke_pending_n = grid;
        S.ke_pending_novz = novz ? 1 : 0;
        S.ke_pending_coef = a.ke_coef;
    }
    return GP_OK;
}

// the last push's kinetic energy into ke_out when it has not been summed yet
int gp_ke_settle(gp_ctx *c, gp_species &S)
{
    if (S.ke_pending_n <= 0) return GP_OK;
    k_sum_partials<<<1, GP_BLOCK, 0, c->stream>>>(S.ke_part, S.ke_pending_n, S.ke_out,
                                                 S.ke_pending_novz ? S.ke_out + 1 : nullptr, S.ke_pending_coef);
    GP_HIP(hipGetLastError());
    S.ke_pending_n = 0;
    return GP_OK;
}

}  // namespace

// ===========================================================================
// host entry points used by gp_api.hip
// ===========================================================================
// digit width of a sort of `bits`-bit keys: the fewest passes of at most 9
// bits, 8-bit digits when they suffice
static int rs_digit(int bits, int *passes)
{
    if (bits < 1) bits = 1;
    *passes = (bits + 8) / 9;
    return bits <= 8 * *passes ? 8 : 9;
}

// scratch of a radix sort of n keys: the digit histograms and their scan
size_t radix_scratch_bytes(int64_t n, int bits)
{
    if (n <= 0) return 0;
    int passes;
    const int64_t hl = ((int64_t)1 << rs_digit(bits, &passes)) * gp_cdiv(n, RS_TILE);
    return gp_align256(sizeof(uint32_t) * (size_t)hl) + scan_scratch_bytes(hl);
}

int gp_radix_sort_u32(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp,
                      int64_t n, int bits, int init_vals, uint32_t **keys_out, uint32_t **vals_out, gp_arena *ar)
{
    uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
    if (n <= 0) { *keys_out = ki; *vals_out = vi; return GP_OK; }
    const int nblk = gp_cdiv(n, RS_TILE);
    int passes;
    const int dig = rs_digit(bits, &passes);
    uint32_t *hist = nullptr;
    const size_t hbytes = sizeof(uint32_t) * ((size_t)1 << dig) * (size_t)nblk;
    const size_t mark = ar ? ar->used : 0;
    if (ar) {
        hist = (uint32_t *)ar->take(hbytes);
        if (!hist) return gp_fail(GP_ERR_CAPACITY, "radix sort: scratch arena exhausted");
    } else {
        GP_HIP(hipMallocAsync((void **)&hist, hbytes, s));
    }
    for (int ps = 0; ps < passes; ps++) {
        const int shift = dig * ps;
        if (dig == 8) k_rs_hist<8><<<nblk, GP_BLOCK, 0, s>>>(ki, n, shift, hist, nblk);
        else k_rs_hist<9><<<nblk, GP_BLOCK, 0, s>>>(ki, n, shift, hist, nblk);
        GP_HIP(hipGetLastError());
        int rc = scan_u32(s, hist, ((int64_t)1 << dig) * (int64_t)nblk, ar);
        if (rc) return rc;
        if (dig == 8)
            k_rs_scatter<8><<<nblk, GP_BLOCK, 0, s>>>(ki, vi, ko, vo, n, shift, (ps == 0) ? init_vals : 0, hist, nblk);
        else
            k_rs_scatter<9><<<nblk, GP_BLOCK, 0, s>>>(ki, vi, ko, vo, n, shift, (ps == 0) ? init_vals : 0, hist, nblk);
        GP_HIP(hipGetLastError());
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    if (ar) ar->used = mark;
    else GP_HIP(hipFreeAsync(hist, s));
    *keys_out = ki;
    *vals_out = vi;
    return GP_OK;
}

int gp_launch_push(gp_ctx *c, int sid, double dt, int deposit)
{
    gp_species &S = c->sp[sid];
    const int dep = deposit ? (c->cfg.cur_dens ? 2 : 1) : 0;
    const int slot = GP_TIME_SPECIES + 3 * sid;
    if (c->timing && S.n > 0) { int rc = gp_time_mark(c, GP_TIME_PUSH, slot); if (rc) return rc; }
    int rc = run_particle_kernel(c, sid, dt, dep, true, 0, S.n);
    if (rc) return rc;
    if (c->timing && S.n > 0) { rc = gp_time_mark(c, GP_TIME_PUSH, slot); if (rc) return rc; }
    return GP_OK;
}

int gp_launch_deposit(gp_ctx *c, int sid, int64_t first, int64_t count)
{
    const int dep = c->cfg.cur_dens ? 2 : 1;
    return run_particle_kernel(c, sid, 0.0, dep, false, first, count);
}

int gp_launch_stream_probe(gp_ctx *c, int sid, int reps, double *ms)
{
    gp_species &S = c->sp[sid];
    const int64_t n = S.n;
    *ms = 0.0;
    if (n == 0) return GP_OK;
    const int64_t per = 32768;
    const int64_t nb = gp_cdiv(n, per);
    if (nb > INT32_MAX) return gp_fail(GP_ERR_ARG, "too many particles for the probe grid");
    hipEvent_t e0, e1;
    GP_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return gp_fail(GP_ERR_HIP, "hipEventCreate"); }
    hipError_t err = hipSuccess;
    k_stream_probe<<<(unsigned)nb, 512, 0, c->stream>>>(S.a, S.b, n, per);   // warm-up
    err = hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps && err == hipSuccess; r++) {
        k_stream_probe<<<(unsigned)nb, 512, 0, c->stream>>>(S.a, S.b, n, per);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipEventRecord(e1, c->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float t = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != hipSuccess) return gp_fail(GP_ERR_HIP, std::string("stream probe: ") + hipGetErrorString(err));
    *ms = (double)t / reps;
    return GP_OK;
}

// key bits of the absorbed-index sort of a species of n particles
static int dead_bits(int64_t n)
{
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < n) bits++;
    return bits;
}

// compaction of one species from its counters h (read from the device by
// the caller): no host synchronisation; ke (may be null) is filled by an
// asynchronous copy, valid once the stream has been synchronised
int gp_launch_reduce_from(gp_ctx *c, int sid, const unsigned long long *h, int64_t counts[4], double *ke, gp_arena *ar)
{
    gp_species &S = c->sp[sid];
    hipStream_t s = c->stream;
    if (h[5]) return gp_fail(GP_ERR_STATE, "particles outside the domain with no boundary condition: " + std::to_string(h[5]));
    const int64_t nd = (int64_t)h[0];
    if (ar && nd > 0) {
        // the arena must hold the larger of the two phases' scratch (the
        // sort's, none for a sort in LDS, is released before the scan of the
        // tail flags)
        const size_t need = std::max(nd <= SMALL_SORT ? 0 : radix_scratch_bytes(nd, dead_bits(S.n)),
                                     scan_scratch_bytes(nd));
        if (need > ar->cap - ar->used)
            return gp_fail(GP_ERR_CAPACITY, "gp_reduce_noalloc: " + std::to_string(nd) + " absorbed particles need " +
                                                std::to_string(need) + " B of scratch, more than the " +
                                                std::to_string(ar->cap - ar->used) + " B it has");
    }
    for (int w = 0; w < 4; w++) counts[w] = (int64_t)h[1 + w];
    const int slot = GP_TIME_SPECIES + 3 * sid + 1;
    if (c->timing) { int rc = gp_time_mark(c, GP_TIME_REDUCE, slot); if (rc) return rc; }
    if (ke) {
        int rc = gp_ke_settle(c, S);
        if (rc) return rc;
        GP_HIP(hipMemcpyAsync(ke, S.ke_out, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    if (nd > 0) {
        if (nd > S.n) return gp_fail(GP_ERR_STATE, "absorbed count exceeds particle count");
        const int64_t n_new = S.n - nd;
        // sort the absorbed indices ascending
        uint32_t *ks = nullptr, *vs = nullptr;
        int rc = GP_OK;
        if (nd <= SMALL_SORT) {
            k_sort_small<<<1, 1024, 0, s>>>(S.dead, S.dead_tmp, (int)nd);
            GP_HIP(hipGetLastError());
            ks = S.dead_tmp;
        } else {
            rc = gp_radix_sort_u32(s, S.dead, S.val, S.dead_tmp, S.val2, nd, dead_bits(S.n), 1, &ks, &vs, ar);
            if (rc) return rc;
        }
        uint32_t *nh = S.key2;         // device scalar: number of holes
        uint32_t *flags = S.key;       // tail flags -> positions (nd)
        uint32_t *donor = S.key2 + 1;  // donors (<= nd)
        k_lower_bound<<<1, 1, 0, s>>>(ks, nd, (uint32_t)n_new, nh);
        GP_HIP(hipGetLastError());
        const int g = gp_cdiv(nd, GP_BLOCK);
        k_tail_flags<<<g, GP_BLOCK, 0, s>>>(S.a.x, n_new, nd, flags);
        GP_HIP(hipGetLastError());
        rc = scan_u32(s, flags, nd, ar);
        if (rc) return rc;
        k_donors<<<g, GP_BLOCK, 0, s>>>(S.a.x, n_new, nd, flags, donor);
        GP_HIP(hipGetLastError());
        k_fill<<<g, GP_BLOCK, 0, s>>>(S.a, ks, donor, nh);
        GP_HIP(hipGetLastError());
        S.n = n_new;
        S.vz_gen++;
        if (S.n_binned > S.n) S.n_binned = S.n;
    }
    GP_HIP(hipMemsetAsync(S.counters, 0, sizeof(unsigned long long) * 8, s));
    if (c->timing) { int rc = gp_time_mark(c, GP_TIME_REDUCE, slot); if (rc) return rc; }
    return GP_OK;
}

int gp_launch_reduce(gp_ctx *c, int sid, int64_t counts[4], double *ke, gp_arena *ar)
{
    gp_species &S = c->sp[sid];
    hipStream_t s = c->stream;
    unsigned long long *h = c->h_counters + 8 * sid;
    GP_HIP(hipMemcpyAsync(h, S.counters, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    const int rc = gp_launch_reduce_from(c, sid, h, counts, ke, ar);
    if (rc) return rc;
    GP_HIP(hipStreamSynchronize(s));
    return GP_OK;
}

// the neighbour re-bin of a binned species (see k_rb_count); the caller
// checked that it applies
static int rebin_sort(gp_ctx *c, gp_species &S)
{
    hipStream_t s = c->stream;
    const int64_t n = S.n, nt = c->ntiles;
    const int64_t nseg = nt + gp_cdiv(n - S.n_binned, RB_CHUNK);
    const int64_t words = nt * RB_CODES + 3 * (nt + 1) + (nseg + 1);
    if (words > c->rb_words) {
        if (c->rb_buf) GP_HIP(hipFree(c->rb_buf));
        c->rb_buf = nullptr;
        c->rb_words = 0;
        const int64_t want = nt * RB_CODES + 3 * (nt + 1) + nt + gp_cdiv(S.cap, RB_CHUNK) + 1;
        GP_HIP(hipMalloc((void **)&c->rb_buf, sizeof(uint32_t) * std::max(words, want)));
        c->rb_words = std::max(words, want);
    }
    RebinArgs a{};
    a.a = S.a; a.b = S.b;
    a.tile_off = S.tile_off;
    a.n = n; a.n_binned = S.n_binned;
    a.ngb = c->ngb; a.ngx = c->ngx; a.ngy = c->ngy; a.ntx = c->ntx; a.nty = c->nty; a.ntiles = c->ntiles;
    uint32_t *w = c->rb_buf;
    a.cnt = w; w += nt * RB_CODES;
    a.start = w; w += nt + 1;
    a.fart = w; w += nt + 1;
    a.fbase = w; w += nt + 1;
    a.segfar = w;
    a.fkey = S.key; a.fval = S.val;
    const int gt = (int)gp_cdiv(nt + 1, GP_BLOCK);
    k_rb_zero<<<gt, GP_BLOCK, 0, s>>>(a, (int)nseg);
    k_rb_count<<<(int)nseg, GP_BLOCK, 0, s>>>(a);
    k_rb_total<<<gt, GP_BLOCK, 0, s>>>(a);
    GP_HIP(hipGetLastError());
    int rc = scan_u32(s, a.start, nt + 1, nullptr);
    if (!rc) rc = scan_u32(s, a.fart, nt + 1, nullptr);
    if (!rc) rc = scan_u32(s, a.segfar, nseg + 1, nullptr);
    if (rc) return rc;
    k_rb_place<<<gt, GP_BLOCK, 0, s>>>(a);
    k_rb_scatter<<<(int)nseg, GP_BLOCK, 0, s>>>(a);
    GP_HIP(hipGetLastError());
    uint32_t nf = 0;
    GP_HIP(hipMemcpyAsync(&nf, a.segfar + nseg, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    if (nf > (uint64_t)n) return gp_fail(GP_ERR_STATE, "re-bin: far count exceeds the particle count");
    if (nf > 0) {
        uint32_t *ks = nullptr, *vs = nullptr;
        rc = gp_radix_sort_u32(s, S.key, S.val, S.key2, S.val2, nf, c->tile_bits, 0, &ks, &vs);
        if (rc) return rc;
        k_rb_far<<<(int)gp_cdiv(nf, GP_BLOCK), GP_BLOCK, 0, s>>>(a, ks, vs, nf);
    }
    k_rb_offsets<<<gt, GP_BLOCK, 0, s>>>(a, S.tile_off);
    GP_HIP(hipGetLastError());
    return GP_OK;
}

int gp_launch_sort(gp_ctx *c, int sid)
{
    gp_species &S = c->sp[sid];
    hipStream_t s = c->stream;
    const int64_t n = S.n;
    if (n == 0) {
        GP_HIP(hipMemsetAsync(S.tile_off, 0, sizeof(int32_t) * (c->ntiles + 1), s));
        S.n_binned = 0;
        return GP_OK;
    }
    const int slot = GP_TIME_SPECIES + 3 * sid + 2;
    if (c->timing) { int rc = gp_time_mark(c, GP_TIME_SORT, slot); if (rc) return rc; }
    S.nsorts++;
    const int g = gp_cdiv(n, GP_BLOCK);
    const int W = WIN + c->order + 1;
    // interleave the bank classes unless the species' particles change cell
    // too often for the order to last (more than GP_CLS_MOVED of them in the
    // last push)
    const bool cls = S.last_moved < 0 || (double)S.last_moved < GP_CLS_MOVED * (double)n;
    // a binned species without the interleave is re-binned by its moves to
    // neighbouring tiles when most of it is binned
    if (!cls && c->ntx >= 3 && c->nty >= 3 && S.n_binned > 0 && 8 * (n - S.n_binned) <= n) {
        int rc = rebin_sort(c, S);
        if (rc) return rc;
        std::swap(S.a, S.b);
        S.n_binned = n;
        if (c->timing) { rc = gp_time_mark(c, GP_TIME_SORT, slot); if (rc) return rc; }
        return GP_OK;
    }
    k_tile_key<<<g, GP_BLOCK, 0, s>>>(S.a.x, S.a.y, n, c->ngb, c->ngx, c->ngy, c->ntx, W, cls ? 1 : 0, S.key);
    GP_HIP(hipGetLastError());
    uint32_t *ks = nullptr, *vs = nullptr;
    int rc = gp_radix_sort_u32(s, S.key, S.val, S.key2, S.val2, n, c->tile_bits + (cls ? 6 : 0), 1, &ks, &vs);
    if (rc) return rc;
    if (cls) {
        const int64_t nb = (int64_t)c->ntiles * 64;
        if (!c->boff) GP_HIP(hipMalloc((void **)&c->boff, sizeof(uint32_t) * (nb + 1)));
        k_bucket_off<<<g, GP_BLOCK, 0, s>>>(ks, n, nb, c->boff);
        GP_HIP(hipGetLastError());
        uint32_t *dst = vs == S.val ? S.val2 : S.val;   // the value buffer the sort left free
        k_interleave<<<c->ntiles, GP_BLOCK, 0, s>>>(c->boff, vs, dst, S.tile_off, c->ntiles);
        GP_HIP(hipGetLastError());
        k_permute<<<g, GP_BLOCK, 0, s>>>(S.a, S.b, dst, n);
        GP_HIP(hipGetLastError());
    } else {
        k_permute<<<g, GP_BLOCK, 0, s>>>(S.a, S.b, vs, n);
        GP_HIP(hipGetLastError());
        k_tile_off<<<g, GP_BLOCK, 0, s>>>(ks, n, c->ntiles, S.tile_off);
        GP_HIP(hipGetLastError());
    }
    std::swap(S.a, S.b);
    S.n_binned = n;
    if (c->timing) { int rc = gp_time_mark(c, GP_TIME_SORT, slot); if (rc) return rc; }
    return GP_OK;
}

static void load_args(gp_ctx *c, int se, int si, const gp_gen *g, uint64_t seed, int64_t first_id, int64_t n,
                      LoadArgs &a)
{
    gp_species &E = c->sp[se];
    a = LoadArgs{};
    a.g = *g;
    a.dx = c->dxf; a.dy = c->dyf; a.ngb = (float)c->ngb;
    a.xdom = (float)c->ngx; a.ydom = (float)c->ngy;
    a.seed = seed; a.first_id = first_id; a.n = n; a.ntot = first_id + n;
    a.e = E.a; a.ne = E.n;
    a.with_ion = si >= 0;
    if (si >= 0) { a.i = c->sp[si].a; a.ni = c->sp[si].n; }
}

// the context's generation streams for LoadArgs a: its materialised states
// and the recorded loads on its seed
static int bind_streams(gp_ctx *c, LoadArgs &a)
{
    a.rs = c->rng_s; a.ri = c->rng_inc; a.rng_n = c->rng_n;
    if ((int)c->rng_loads.size() > GP_RNG_RECS)
        return gp_fail(GP_ERR_STATE, "more than 8 seeded loads on one context's generation streams");
    a.nrec = (int)c->rng_loads.size();
    for (int q = 0; q < a.nrec; q++) {
        a.rec_first[q] = c->rng_loads[q].first;
        a.rec_n[q] = c->rng_loads[q].n;
        a.rec_draws[q] = c->rng_loads[q].draws;
    }
    return GP_OK;
}

int gp_launch_load(gp_ctx *c, int se, int si, const gp_gen *g, uint64_t seed, int64_t first_id, int64_t n)
{
    if (n <= 0) return GP_OK;
    gp_species &E = c->sp[se];
    if (E.n + n > E.cap || (si >= 0 && c->sp[si].n + n > c->sp[si].cap))
        return gp_fail(GP_ERR_CAPACITY, "pNum > pNumMax in particle generation");
    LoadArgs a;
    load_args(c, se, si, g, seed, first_id, n, a);
    // the first seeded load fixes the context's generation streams; a load on
    // another seed draws from independent streams and is not recorded
    if (!c->rng_seeded) { c->rng_seed = seed; c->rng_seeded = 1; }
    const bool ours = seed == c->rng_seed;
    if (ours) {
        int rc = bind_streams(c, a);
        if (rc) return rc;
    }
    k_load<<<gp_cdiv(n, GP_BLOCK), GP_BLOCK, 0, c->stream>>>(a);
    GP_HIP(hipGetLastError());
    if (ours) c->rng_loads.push_back({first_id, n, gen_draws(g, si >= 0)});
    E.n += n;
    E.vz_gen++;
    if (si >= 0 && g->gen_type != 3) { c->sp[si].n += n; c->sp[si].vz_gen++; }
    return GP_OK;
}

int gp_launch_generate(gp_ctx *c, int se, int si, const gp_gen *g, int64_t addn, int deposit)
{
    if (addn <= 0) return GP_OK;
    gp_species &E = c->sp[se];
    const bool ion = si >= 0 && g->gen_type != 3;
    if (E.n + addn > E.cap || (ion && c->sp[si].n + addn > c->sp[si].cap))
        return gp_fail(GP_ERR_CAPACITY, "pNum + addn > pNumMax in particle generation");
    hipStream_t s = c->stream;
    if (!c->rng_seeded) { c->rng_seed = 0; c->rng_seeded = 1; }   // the reference's master seed
    LoadArgs a;
    load_args(c, se, ion ? si : -1, g, c->rng_seed, 0, addn, a);
    int rc = bind_streams(c, a);
    if (rc) return rc;
    if (addn > c->rng_n) {
        // materialise the states of ids [rng_n, addn)
        if (addn > c->rng_cap) {
            const int64_t cap = std::max(addn, 2 * c->rng_cap);
            uint64_t *ns = nullptr, *ni = nullptr;
            hipError_t e = hipMalloc((void **)&ns, sizeof(uint64_t) * cap);
            if (e == hipSuccess) e = hipMalloc((void **)&ni, sizeof(uint64_t) * cap);
            if (e == hipSuccess && c->rng_n > 0) {
                e = hipMemcpyAsync(ns, c->rng_s, sizeof(uint64_t) * c->rng_n, hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(ni, c->rng_inc, sizeof(uint64_t) * c->rng_n, hipMemcpyDeviceToDevice, s);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) {   // the old states stay in place
                if (ns) (void)hipFree(ns);
                if (ni) (void)hipFree(ni);
                return gp_fail(GP_ERR_NOMEM, std::string("generation streams: ") + hipGetErrorString(e));
            }
            if (c->rng_s) GP_HIP(hipFree(c->rng_s));
            if (c->rng_inc) GP_HIP(hipFree(c->rng_inc));
            c->rng_s = ns; c->rng_inc = ni; c->rng_cap = cap;
        }
        const int64_t lo = c->rng_n;
        a.rs = c->rng_s; a.ri = c->rng_inc; a.rng_n = lo;
        k_rng_init<<<gp_cdiv(addn - lo, GP_BLOCK), GP_BLOCK, 0, s>>>(a, lo, addn, c->rng_s, c->rng_inc);
        GP_HIP(hipGetLastError());
        c->rng_n = addn;
    }
    a.rs = c->rng_s; a.ri = c->rng_inc; a.rng_n = c->rng_n;
    a.nrec = 0;   // every id of the call is materialised
    k_load<<<gp_cdiv(addn, GP_BLOCK), GP_BLOCK, 0, s>>>(a);
    GP_HIP(hipGetLastError());
    const int64_t e0 = E.n, i0 = ion ? c->sp[si].n : 0;
    E.n += addn;
    E.vz_gen++;
    if (ion) { c->sp[si].n += addn; c->sp[si].vz_gen++; }
    if (deposit) {
        rc = gp_launch_deposit(c, se, e0, addn);
        if (rc) return rc;
        if (ion && (rc = gp_launch_deposit(c, si, i0, addn))) return rc;
    }
    return GP_OK;
}

int gp_launch_inject_wall(gp_ctx *c, int sid, int64_t n, float vth, const float u[3], double dt,
                          uint64_t seed, uint64_t stream, int deposit)
{
    if (n <= 0) return GP_OK;
    gp_species &S = c->sp[sid];
    if (S.n + n > S.cap) return gp_fail(GP_ERR_CAPACITY, "pNum > pNumMax in wall injection");
    InjArgs a{};
    a.s = S.a; a.base = S.n; a.n = n; a.seed = seed; a.stream = stream;
    a.vth = vth; a.u0 = u[0]; a.u1 = u[1]; a.u2 = u[2];
    a.dtdx = (float)(dt / (double)c->dxf);
    a.lo = (float)c->ngb;
    a.ngx = (float)c->ngx;
    a.ngy = (float)c->ngy;
    k_inject<<<gp_cdiv(n, GP_BLOCK), GP_BLOCK, 0, c->stream>>>(a);
    GP_HIP(hipGetLastError());
    const int64_t first = S.n;
    S.n += n;
    S.vz_gen++;
    if (deposit) return gp_launch_deposit(c, sid, first, n);
    return GP_OK;
}

int gp_launch_moments(gp_ctx *c, int sid, double *host10)
{
    gp_species &S = c->sp[sid];
    hipStream_t s = c->stream;
    const int64_t ng = (int64_t)(c->nx + 1) * (c->ny + 1);
    unsigned long long *acc = nullptr;
    double *out = nullptr;
    GP_HIP(hipMallocAsync((void **)&acc, sizeof(unsigned long long) * 10 * ng, s));
    GP_HIP(hipMallocAsync((void **)&out, sizeof(double) * 10 * ng, s));
    GP_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * 10 * ng, s));
    MomArgs a{};
    a.x = S.a.x; a.y = S.a.y; a.vx = S.a.vx; a.vy = S.a.vy; a.vz = S.a.vz;
    a.n = S.n; a.ngb = c->ngb; a.ngx = c->ngx; a.ngy = c->ngy; a.nx = c->nx;
    a.inv_vnorm = (float)S.inv_vnorm; a.acc = acc; a.ng = ng; a.escaped = S.counters + 5;
    if (S.n > 0) {
        const int g = gp_cdiv(S.n, GP_BLOCK);
        if (c->order == 1) k_moments<1><<<g, GP_BLOCK, 0, s>>>(a);
        else if (c->order == 3) k_moments<3><<<g, GP_BLOCK, 0, s>>>(a);
        else k_moments<5><<<g, GP_BLOCK, 0, s>>>(a);
        GP_HIP(hipGetLastError());
    }
    // Moment.mm:242 scale = w/(dx*dy) on n; velocity sums are unscaled
    const double sn = ldexp(1.0, -GP_CUR_FRAC) * ((double)S.wf / ((double)c->dxf * (double)c->dyf));
    const double sv = ldexp(1.0, -GP_CUR_FRAC) * S.vnorm;
    const double svv = ldexp(1.0, -GP_CUR_FRAC) * S.vnorm * S.vnorm;
    k_mom_scale<<<gp_cdiv(10 * ng, GP_BLOCK), GP_BLOCK, 0, s>>>(acc, ng, sn, sv, svv, out);
    GP_HIP(hipGetLastError());
    GP_HIP(hipMemcpyAsync(host10, out, sizeof(double) * 10 * ng, hipMemcpyDeviceToHost, s));
    GP_HIP(hipFreeAsync(acc, s));
    GP_HIP(hipFreeAsync(out, s));
    GP_HIP(hipStreamSynchronize(s));
    return GP_OK;
}

int gp_launch_ke(gp_ctx *c, int sid)
{
    gp_species &S = c->sp[sid];
    const int g = std::min(1024, std::max(1, gp_cdiv(S.n, GP_BLOCK)));
    S.ke_pending_n = 0;   // recomputed from the particles here
    k_ke<<<g, GP_BLOCK, 0, c->stream>>>(S.a.x, S.a.vx, S.a.vy, S.a.vz, S.n, 0.5 * (double)S.mf * (double)S.wf, S.ke_part);
    GP_HIP(hipGetLastError());
    k_sum_partials<<<1, GP_BLOCK, 0, c->stream>>>(S.ke_part, g, S.ke_out);
    GP_HIP(hipGetLastError());
    return GP_OK;
}

// ---------------------------------------------------------------------------
// Static scratch of gp_reduce_noalloc: the radix histograms and scan block
// sums of the compaction, for callers that cannot allocate device memory.
// Each device holds its own copy for the life of the process; host code
// reaches it only through the address hipGetSymbolAddress returned for that
// device (the array's name is a host-side stand-in).  Every call on a device
// shares the array: calls that overlap on one device race.
// ---------------------------------------------------------------------------
#define GP_NOALLOC_SCRATCH (12u << 20)
static __device__ __align__(256) unsigned char scratch[GP_NOALLOC_SCRATCH];

namespace {
constexpr int GP_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[GP_MAX_DEVICES] = {};
}  // namespace

// resolve the device's copy of the scratch array, with that device current
int gp_noalloc_setup(int device)
{
    if (device < 0 || device >= GP_MAX_DEVICES) return gp_fail(GP_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return GP_OK;
    int cur = -1;
    GP_HIP(hipGetDevice(&cur));
    if (cur != device) GP_HIP(hipSetDevice(device));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !ptr)
        return gp_fail(GP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return GP_OK;
}

int gp_noalloc_arena(int device, gp_arena *ar)
{
    if (device < 0 || device >= GP_MAX_DEVICES) return gp_fail(GP_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) return gp_fail(GP_ERR_STATE, "no reduce scratch on this device (gp_ctx_create sets it up)");
    ar->base = (char *)g_scratch_ptr[device];
    ar->cap = GP_NOALLOC_SCRATCH;
    ar->used = 0;
    return GP_OK;
}
