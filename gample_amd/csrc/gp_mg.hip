// gp_mg.hip -- geometric multigrid preconditioned CG with FP64 refinement.
//
// Replaces the AMGCL solves of EMField.mm:866, 1199, 1237, 1275.  The system
// is S x = b with S = D (-A): A is the reference's 5-point operator
// (EMField.mm:236-405 Poisson, 483-623 Darwin) and D halves the rows of
// Neumann boundary nodes (making the ghost-mirror rows symmetric).
//
// V-cycle (symmetric, so it is a valid CG preconditioner):
//   pre-smoothing  NSW red-black Gauss-Seidel sweeps (red, then black) from 0
//   restriction    R = P^T / 4 (full weighting) of the residual
//   coarse level   recursion; rediscretised operator (h doubles per level)
//   prolongation   bilinear P, added
//   post-smoothing NSW sweeps in the adjoint order (black, then red)
// Each grid level costs two launches: k_mg_down (the pre-sweeps, residual
// and restriction on a 32x32 tile -- 64x64 on a level with more than 512
// such tiles -- whose (2 NSW + 1)-wide halo is recomputed
// locally) and k_mg_up (prolongation and the post-sweeps with a 2 NSW-wide
// halo).  Recomputing the halo reproduces global red-black Gauss-Seidel
// exactly.  All levels with at
// most 8192 unknowns run inside one workgroup in LDS (k_mg_bottom).
//
// CG in mixed precision: the iterate in FP64, the residual recursion, the
// V-cycle and the search directions in FP32 (dot products accumulated in
// FP64), with the recursive residual replaced by the true FP64 one every few
// iterations; the solve stops at ||b - S x|| / ||b|| <= tol, tested in FP64.
#include "gp_internal.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace {

constexpr int TB = 32;                 // tile of unknowns per workgroup
constexpr int HU = 2, WU = TB + 2 * HU;   // window of k_rb_sweep (one full sweep)
#ifndef MG_SWEEPS
#define MG_SWEEPS 1
#endif
constexpr int NSW = MG_SWEEPS;            // smoothing sweeps per side of the V-cycle
// smallest reduction asked of one FP32 flexible-CG solve inside the FP64
// refinement of the network solvers (its own rounding limits it)
#ifndef MG_INNER_FLOOR
#define MG_INNER_FLOOR 2e-6
#endif
// MG-PCG: the reduction of the residual between two replacements of the
// recursive residual by the true one (the FP32 recursion drifts by about
// 1e-7 of the residual it starts from), and the recursion's target as a
// fraction of tol (the true residual is tested after the recursion meets it)
#ifndef MG_REPLACE_DROP
#define MG_REPLACE_DROP 1e-5
#endif
#ifndef MG_TARGET
#define MG_TARGET 0.8
#endif
// relative difference of the true and the recursive residual norms at a
// replacement above which the search direction restarts
#ifndef MG_DRIFT
#define MG_DRIFT 0.1
#endif
#ifndef MG_BY
#define MG_BY 32
#endif
constexpr int BY = MG_BY;                 // rows of the (32, BY) workgroup of the level kernels
constexpr int NT = 32 * BY;
// unknowns of the largest LDS level: 8192 takes the coarsest levels that
// an odd interval count stops (124 x 50 for the 500 x 200 Dirichlet/periodic
// grids of the reference's thruster conditions) into the bottom kernel
constexpr int BOT_MAX = 8192;
constexpr int BOT_LEVELS = 12;
constexpr int DENSE_MAX = 64;          // coarsest level solved by a dense inverse
constexpr int RB = 256;                // workgroup of the reduction kernels

struct Lev {
    int mx, my;     // unknowns
    int tx, ty;     // axis types GP_AX_*
    int Mx, My;     // intervals
    float cx, cy;   // 1/h^2
    int rb;         // red-black colouring valid (no odd periodic axis)
};

// wrap for an index at most one period outside [0, m)
__device__ __forceinline__ int wrp(int g, int m) { return g < 0 ? g + m : (g >= m ? g - m : g); }
// wrap of a tile-window index: a window reaches up to TB + 2 HD past the
// tile, more than one period when the periodic extent is small
__device__ __forceinline__ int wwrap(int g, int m)
{
    const int w = wrp(g, m);
    return (w >= 0 && w < m) ? w : ((g % m) + m) % m;
}
// wrap of a window index on a level at least a window wide (the index is at
// most one period outside [0, m))
__device__ __forceinline__ int wnear(int g, int m) { return g < 0 ? g + m : (g >= m ? g - m : g); }
__device__ __forceinline__ float wgt(int t, int i, int m) { return (t == GP_AX_NEU && (i == 0 || i == m - 1)) ? 0.5f : 1.0f; }
__device__ __forceinline__ bool in_axis(int t, int g, int m) { return t == GP_AX_PER || (g >= 0 && g < m); }
// 1/w of the row scaling D (w = 1/2 on Neumann boundary nodes): 1, 2 or 4
__device__ __forceinline__ float iweight(const Lev &L, int gi, int gj)
{
    float a = (L.tx == GP_AX_NEU && (gi == 0 || gi == L.mx - 1)) ? 2.0f : 1.0f;
    float b = (L.ty == GP_AX_NEU && (gj == 0 || gj == L.my - 1)) ? 2.0f : 1.0f;
    return a * b;
}
// w = 1 / iweight (1, 1/2 or 1/4: exact, so x * w equals x / iweight)
__device__ __forceinline__ float rweight(const Lev &L, int gi, int gj)
{
    float a = (L.tx == GP_AX_NEU && (gi == 0 || gi == L.mx - 1)) ? 0.5f : 1.0f;
    float b = (L.ty == GP_AX_NEU && (gj == 0 || gj == L.my - 1)) ? 0.5f : 1.0f;
    return a * b;
}
// red/black colour; on a periodic axis the extent is even (checked when the
// hierarchy is built), so the parity of an unwrapped index is that of the
// wrapped one
__device__ __forceinline__ int colour(int gi, int gj) { return (gi + gj) & 1; }

// cx (uL + uR) + cy (uD + uU) at window point k = li + lj*W whose global
// (unwrapped) coordinates are gi, gj: Dirichlet walls are 0, Neumann
// boundaries mirror, periodic axes read the window's wrapped copy
__device__ __forceinline__ float nsum(const float *w, int W, const Lev &L, int k, int gi, int gj)
{
    float uL, uR, uD, uU;
    if (L.tx == GP_AX_PER || gi > 0) uL = w[k - 1]; else uL = (L.tx == GP_AX_NEU) ? w[k + 1] : 0.0f;
    if (L.tx == GP_AX_PER || gi < L.mx - 1) uR = w[k + 1]; else uR = (L.tx == GP_AX_NEU) ? w[k - 1] : 0.0f;
    if (L.ty == GP_AX_PER || gj > 0) uD = w[k - W]; else uD = (L.ty == GP_AX_NEU) ? w[k + W] : 0.0f;
    if (L.ty == GP_AX_PER || gj < L.my - 1) uU = w[k + W]; else uU = (L.ty == GP_AX_NEU) ? w[k - W] : 0.0f;
    return L.cx * (uL + uR) + L.cy * (uD + uU);
}

// fine unknown of fine node n (-1 for a Dirichlet wall or outside a Neumann axis)
__device__ __forceinline__ int f_unk(int t, int n, int M)
{
    if (t == GP_AX_DIR) return (n <= 0 || n >= M) ? -1 : n - 1;
    if (t == GP_AX_PER) return n;          // caller wraps
    return (n < 0 || n > M) ? -1 : n;
}
__device__ __forceinline__ int c_node(int t, int I) { return t == GP_AX_DIR ? I + 1 : I; }
__device__ __forceinline__ int c_unk(int t, int N, int M)
{
    if (t == GP_AX_DIR) return (N <= 0 || N >= M) ? -1 : N - 1;
    if (t == GP_AX_PER) return wrp(N, M);
    return N;
}
// bilinear prolongation of the coarse correction at fine unknown (i, j)
__device__ __forceinline__ float prolong(const float *uc, const Lev &F, const Lev &C, int i, int j)
{
    const int nx = F.tx == GP_AX_DIR ? i + 1 : i, ny = F.ty == GP_AX_DIR ? j + 1 : j;
    const int ax0 = nx >> 1, ax1 = (nx + 1) >> 1, ay0 = ny >> 1, ay1 = (ny + 1) >> 1;
    const int I0 = c_unk(C.tx, ax0, C.Mx), I1 = c_unk(C.tx, ax1, C.Mx);
    const int J0 = c_unk(C.ty, ay0, C.My), J1 = c_unk(C.ty, ay1, C.My);
    const float a00 = (I0 >= 0 && J0 >= 0) ? uc[I0 + J0 * C.mx] : 0.0f;
    const float a10 = (I1 >= 0 && J0 >= 0) ? uc[I1 + J0 * C.mx] : 0.0f;
    const float a01 = (I0 >= 0 && J1 >= 0) ? uc[I0 + J1 * C.mx] : 0.0f;
    const float a11 = (I1 >= 0 && J1 >= 0) ? uc[I1 + J1 * C.mx] : 0.0f;
    // even index: ax0 == ax1 and the two halves add to the coarse value
    return 0.25f * ((a00 + a10) + (a01 + a11));
}

// a level descriptor with its axis types fixed at compile time (the
// kernels are instantiated per pair of axis types, so every boundary test
// on an axis type folds away)
template <int TX, int TY>
__device__ __forceinline__ Lev fixed(Lev L)
{
    L.tx = TX;
    L.ty = TY;
    return L;
}

// device scalars; S_SC scales the true residual to unit RMS for the FP32
// recursion (S_ISC = 1 / S_SC), S_TRR is the true residual's r.r (S_TRR0 at
// the start of the solve)
enum { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_TMP, S_BB, S_MEAN, S_DONE, S_TGT, S_IT, S_MAXIT, S_RZ2, S_SC, S_ISC,
       S_TRR, S_TRR0, S_NS };
// every kernel of an iteration returns at once when the device-side
// convergence flag is set (the host enqueues iterations in batches)
__device__ __forceinline__ bool done(const double *scal) { return scal && scal[S_DONE] != 0.0; }

// the scalar update of a completed reduction (see k_final)
__device__ __forceinline__ void apply_op(double *scal, int slot, int op, double s)
{
    scal[slot] = s;
    if (op == 1) scal[S_ALPHA] = s != 0.0 ? scal[S_RZ] / s : 0.0;
}
// Sum of np reduction partials in a fixed order, computed by a whole block
// (threads 0..255 read, every thread passes the barriers and receives the
// result).  Every block of a grid gets the same bits, which lets the kernel
// that consumes a reduction finish it itself instead of a separate launch.
__device__ double sum_parts(const double *part, int np, double *sh, int tid)
{
    // threads 0..255: strided partial sums, then a fixed-order wave tree
    // (shuffles) and the four wave sums in order
    double a = 0.0;
    if (tid < 256)
        for (int k = tid; k < np; k += 256) a += part[k];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    if (tid < 256 && (tid & 63) == 0) sh[tid >> 6] = a;
    __syncthreads();
    const double s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return s;
}
// Residual test of PCG iteration it (1-based), made by the first kernel of
// the V-cycle that follows the r update: every block sums the r.r partials;
// block 0 records ||r||^2 and the count and raises the convergence flag at
// the target or the iteration limit.  Returns true when converged.
struct Gate {
    const double *rr_part;   // null: no test in this launch
    int np, it;
    double *scal;
};
__device__ bool gate_converged(const Gate &g, double *sh, int tid, bool block0)
{
    const double rr = sum_parts(g.rr_part, g.np, sh, tid);
    const bool conv = !(sqrt(rr) > g.scal[S_TGT]) || (double)g.it >= g.scal[S_MAXIT];
    if (block0 && tid == 0) {
        g.scal[S_RR] = rr;
        g.scal[S_IT] = (double)g.it;
        if (conv) g.scal[S_DONE] = 1.0;
    }
    return conv;
}

// ---------------------------------------------------------------------------
// grid level, down leg: u = NS symmetric-order red-black sweeps from 0
// (red, black, red, black, ...); fc = R (f - S u).  blockDim (32, 8).
// Every half-sweep is computed on a band one point narrower than the last,
// from a window with a (2 NS + 1)-wide halo, which reproduces the global
// sweeps exactly on the tile and its 1-wide rim (needed by the restriction).
// The window's global loads are all issued before any is consumed.
// ---------------------------------------------------------------------------
template <int NS, int TB, int TX, int TY>
__global__ __launch_bounds__(NT) void k_mg_down(const float *f, float *u, float *fc, Lev Lin, Lev Cin, const double *scal,
                                                 Gate gt)
{
    const Lev L = fixed<TX, TY>(Lin), C = fixed<TX, TY>(Cin);
    constexpr int HD = 2 * NS + 1, WD = TB + 2 * HD, NE = (WD * WD + NT - 1) / NT;
    __shared__ float su[WD * WD], sf[WD * WD], sr[(TB + 2) * (TB + 2)];
    __shared__ double shr[256];
    if (done(scal)) return;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + 32 * ty;
    if (gt.rr_part && gate_converged(gt, shr, tid, blockIdx.x == 0 && blockIdx.y == 0)) return;
    const int ti0 = blockIdx.x * TB, tj0 = blockIdx.y * TB;
    const int gi0 = ti0 - HD, gj0 = tj0 - HD;
    const float invd = 1.0f / (2.0f * (L.cx + L.cy)), d = 2.0f * (L.cx + L.cy);
    float fv[NE];
    // (a level at least a window wide needs no general wrap: the integer
    // remainders of wwrap would dominate this phase)
    auto load = [&](auto wr) {
#pragma unroll
        for (int k = 0; k < NE; k++) {
            const int e = tid + NT * k;
            const int li = e % WD, lj = e / WD;
            const int gi = gi0 + li, gj = gj0 + lj;
            const bool v = e < WD * WD && in_axis(L.tx, gi, L.mx) && in_axis(L.ty, gj, L.my);
            fv[k] = v ? f[wr(gi, L.mx) + wr(gj, L.my) * L.mx] : 0.0f;
        }
    };
    if (L.mx >= WD && L.my >= WD) load(wnear);
    else load(wwrap);
#pragma unroll
    for (int k = 0; k < NE; k++) {
        const int e = tid + NT * k;
        if (e < WD * WD) {
            const int li = e % WD, lj = e / WD;
            const int gi = gi0 + li, gj = gj0 + lj;
            sf[e] = fv[k];
            // red from zero (its neighbours are black = 0); black starts at 0
            su[e] = colour(gi, gj) == 0 ? fv[k] * iweight(L, gi, gj) * invd : 0.0f;
        }
    }
    __syncthreads();
    // half-sweeps 1 .. 2NS-1 (black, red, ...) on bands 1 .. 2NS-1; the
    // threads take the band's points of the colour in row order (every
    // band row holds hw of them: WD is even)
#pragma unroll
    for (int h = 1; h < 2 * NS; h++) {
        const int col = h & 1;   // 1 black, 0 red
        const int hw = (WD - 2 * h) / 2;
        for (int p = tid; p < hw * (WD - 2 * h); p += NT) {
            const int r = p / hw, lj = h + r;
            const int gj = gj0 + lj;
            const int li = h + ((col - gj - gi0 - h) & 1) + 2 * (p - r * hw);
            const int gi = gi0 + li, k = li + lj * WD;
            if (in_axis(L.ty, gj, L.my) && in_axis(L.tx, gi, L.mx))
                su[k] = (sf[k] * iweight(L, gi, gj) + nsum(su, WD, L, k, gi, gj)) * invd;
        }
        __syncthreads();
    }
    // residual on band 2NS (the tile and its rim)
    constexpr int RB0 = 2 * NS, WR = TB + 2;
    for (int p = tid; p < WR * WR; p += NT) {
        const int lj = RB0 + p / WR, li = RB0 + p % WR;
        const int gj = gj0 + lj, gi = gi0 + li, k = li + lj * WD;
        float r = 0.0f;
        if (in_axis(L.ty, gj, L.my) && in_axis(L.tx, gi, L.mx))
            r = sf[k] - (d * su[k] - nsum(su, WD, L, k, gi, gj)) * rweight(L, gi, gj);
        sr[(li - RB0) + (lj - RB0) * WR] = r;
    }
    __syncthreads();
    // pre-smoothed u of the tile
    for (int p = tid; p < TB * TB; p += NT) {
        const int li = p % TB, lj = p / TB;
        const int gi = ti0 + li, gj = tj0 + lj;
        if (gi < L.mx && gj < L.my) u[gi + gj * L.mx] = su[(HD + li) + (HD + lj) * WD];
    }
    // restriction for the coarse unknowns centred in this tile
    const int ci0 = ti0 > 0 ? ti0 / 2 - 1 : 0, cj0 = tj0 > 0 ? tj0 / 2 - 1 : 0;
    constexpr int CW = TB / 2 + 2;
    for (int p = tid; p < CW * CW; p += NT) {
        const int J = cj0 + p / CW, I = ci0 + p % CW;
        if (J >= C.my || I >= C.mx) continue;
        const int nyc = 2 * c_node(C.ty, J);
        const int fv2 = f_unk(L.ty, nyc, L.My);
        const int nxc = 2 * c_node(C.tx, I);
        const int fu = f_unk(L.tx, nxc, L.Mx);
        if (fv2 < tj0 || fv2 >= tj0 + TB || fu < ti0 || fu >= ti0 + TB) continue;
        {
            float acc = 0.0f;
#pragma unroll
            for (int b = -1; b <= 1; b++) {
                if (L.ty != GP_AX_PER && f_unk(L.ty, nyc + b, L.My) < 0) continue;
                const float wy = b == 0 ? 0.5f : 0.25f;
#pragma unroll
                for (int a = -1; a <= 1; a++) {
                    if (L.tx != GP_AX_PER && f_unk(L.tx, nxc + a, L.Mx) < 0) continue;
                    acc += (a == 0 ? 0.5f : 0.25f) * wy * sr[(fu + a - ti0 + 1) + (fv2 + b - tj0 + 1) * (TB + 2)];
                }
            }
            fc[I + J * C.mx] = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// grid level, up leg: u = upre + P uc, then NS sweeps in the adjoint order
// (black, red, black, red, ...) on bands 1 .. 2NS of a window with a 2NS-wide
// halo.  Optional partial of f.u (the CG's r.z at the finest level).
// blockDim (32, BY)
// ---------------------------------------------------------------------------
template <int NS, int TB, int TX, int TY>
__global__ __launch_bounds__(NT) void k_mg_up(const float *upre, const float *f, const float *uc, float *u, Lev Lin,
                                               Lev Cin, double *part, const double *scal)
{
    const Lev L = fixed<TX, TY>(Lin), C = fixed<TX, TY>(Cin);
    constexpr int HU = 2 * NS, WU = TB + 2 * HU, NE = (WU * WU + NT - 1) / NT;
    __shared__ float su[WU * WU], sf[WU * WU];
    __shared__ double red[NT / 64];
    if (done(scal)) return;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + 32 * ty;
    const int ti0 = blockIdx.x * TB, tj0 = blockIdx.y * TB;
    const int gi0 = ti0 - HU, gj0 = tj0 - HU;
    const float invd = 1.0f / (2.0f * (L.cx + L.cy));
    float uv[NE], fvv[NE];
    auto load = [&](auto wr) {
#pragma unroll
        for (int k = 0; k < NE; k++) {
            const int e = tid + NT * k;
            const int li = e % WU, lj = e / WU;
            const int gi = gi0 + li, gj = gj0 + lj;
            uv[k] = 0.0f;
            fvv[k] = 0.0f;
            if (e < WU * WU && in_axis(L.tx, gi, L.mx) && in_axis(L.ty, gj, L.my)) {
                const int wi = wr(gi, L.mx), wj = wr(gj, L.my), kk = wi + wj * L.mx;
                uv[k] = upre[kk] + prolong(uc, L, C, wi, wj);
                fvv[k] = f[kk];
            }
        }
    };
    if (L.mx >= WU && L.my >= WU) load(wnear);
    else load(wwrap);
#pragma unroll
    for (int k = 0; k < NE; k++) {
        const int e = tid + NT * k;
        if (e < WU * WU) { su[e] = uv[k]; sf[e] = fvv[k]; }
    }
    __syncthreads();
    // the band's points of the colour in row order (WU is even)
#pragma unroll
    for (int h = 1; h <= 2 * NS; h++) {
        const int col = h & 1;   // 1 black, 0 red
        const int hw = (WU - 2 * h) / 2;
        for (int p = tid; p < hw * (WU - 2 * h); p += NT) {
            const int r = p / hw, lj = h + r;
            const int gj = gj0 + lj;
            const int li = h + ((col - gj - gi0 - h) & 1) + 2 * (p - r * hw);
            const int gi = gi0 + li, k = li + lj * WU;
            if (in_axis(L.ty, gj, L.my) && in_axis(L.tx, gi, L.mx))
                su[k] = (sf[k] * iweight(L, gi, gj) + nsum(su, WU, L, k, gi, gj)) * invd;
        }
        __syncthreads();
    }
    double acc = 0.0;
    for (int lj = ty; lj < TB; lj += BY)
        for (int li = tx; li < TB; li += 32) {
            const int gj = tj0 + lj, gi = ti0 + li;
            if (gj < L.my && gi < L.mx) {
                const int k = (HU + li) + (HU + lj) * WU;
                u[gi + gj * L.mx] = su[k];
                acc += (double)sf[k] * (double)su[k];
            }
        }
    if (part) {
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((tid & 63) == 0) red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < NT / 64; w++) t += red[w];
            part[blockIdx.x + blockIdx.y * gridDim.x] = t;
        }
    }
}

// one red-black sweep on a grid level (coarsest fallback when a level cannot
// be coarsened below BOT_MAX): first colour on tile+1, second on the tile
__global__ __launch_bounds__(256) void k_rb_sweep(const float *uin, const float *f, float *u, Lev L, int first,
                                                  const double *scal)
{
    __shared__ float su[WU * WU], sf[WU * WU];
    if (done(scal)) return;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int ti0 = blockIdx.x * TB, tj0 = blockIdx.y * TB;
    const int gi0 = ti0 - HU, gj0 = tj0 - HU;
    const float invd = 1.0f / (2.0f * (L.cx + L.cy));
    for (int lj = ty; lj < WU; lj += 8) {
        const int gj = gj0 + lj;
        const bool vj = in_axis(L.ty, gj, L.my);
        const int wj = wwrap(gj, L.my);
        for (int li = tx; li < WU; li += 32) {
            const int gi = gi0 + li, e = li + lj * WU;
            const bool v = vj && in_axis(L.tx, gi, L.mx);
            const int k = v ? wwrap(gi, L.mx) + wj * L.mx : 0;
            su[e] = v ? uin[k] : 0.0f;
            sf[e] = v ? f[k] : 0.0f;
        }
    }
    __syncthreads();
    for (int lj = 1 + ty; lj < WU - 1; lj += 8) {
        const int gj = gj0 + lj;
        if (!in_axis(L.ty, gj, L.my)) continue;
        for (int li = 1 + tx; li < WU - 1; li += 32) {
            const int gi = gi0 + li, k = li + lj * WU;
            if (in_axis(L.tx, gi, L.mx) && colour(gi, gj) == first)
                su[k] = (sf[k] * iweight(L, gi, gj) + nsum(su, WU, L, k, gi, gj)) * invd;
        }
    }
    __syncthreads();
    for (int lj = HU + ty; lj < HU + TB; lj += 8) {
        const int gj = gj0 + lj;
        if (!in_axis(L.ty, gj, L.my)) continue;
        const int li = HU + tx, gi = gi0 + li, k = li + lj * WU;
        if (in_axis(L.tx, gi, L.mx) && colour(gi, gj) == 1 - first)
            su[k] = (sf[k] * iweight(L, gi, gj) + nsum(su, WU, L, k, gi, gj)) * invd;
    }
    __syncthreads();
    for (int lj = ty; lj < TB; lj += 8) {
        const int gj = tj0 + lj, gi = ti0 + tx;
        if (gj < L.my && gi < L.mx) u[gi + gj * L.mx] = su[(HU + tx) + (HU + lj) * WU];
    }
}

// ---------------------------------------------------------------------------
// bottom: the whole V-cycle of the small levels inside one workgroup
// (blockDim 32 x 32), every level resident in LDS
// ---------------------------------------------------------------------------
struct Bottom {
    Lev lv[BOT_LEVELS];
    int off[BOT_LEVELS];   // LDS offset of level l (u at off, f at off + n)
    int nl;
    int singular;
    int sweeps;            // symmetric sweep pairs on the coarsest level (no dense inverse)
    const float *dense;    // inverse of the coarsest operator, column-major (or null)
    // fast diagonalisation of the coarsest level (or null): eigenvector
    // matrices of the two 1D operators and the inverse eigenvalues
    const float *fvx, *fvy, *finv;
};

// Gauss-Seidel value at unknown (i, j) with the neighbour values given by
// v(i', j'): Dirichlet walls are 0, Neumann boundaries mirror, periodic
// axes wrap
template <class V>
__device__ __forceinline__ float gs_with(const float *f, const Lev &L, int i, int j, float invd, V v)
{
    float uL, uR, uD, uU;
    uL = i > 0 ? v(i - 1, j) : (L.tx == GP_AX_PER ? v(L.mx - 1, j) : (L.tx == GP_AX_NEU ? v(i + 1, j) : 0.0f));
    uR = i < L.mx - 1 ? v(i + 1, j) : (L.tx == GP_AX_PER ? v(0, j) : (L.tx == GP_AX_NEU ? v(i - 1, j) : 0.0f));
    uD = j > 0 ? v(i, j - 1) : (L.ty == GP_AX_PER ? v(i, L.my - 1) : (L.ty == GP_AX_NEU ? v(i, j + 1) : 0.0f));
    uU = j < L.my - 1 ? v(i, j + 1) : (L.ty == GP_AX_PER ? v(i, 0) : (L.ty == GP_AX_NEU ? v(i, j - 1) : 0.0f));
    return (f[i + j * L.mx] * iweight(L, i, j) + L.cx * (uL + uR) + L.cy * (uD + uU)) * invd;
}
__device__ __forceinline__ float gs_point(const float *u, const float *f, const Lev &L, int i, int j, float invd)
{
    return gs_with(f, L, i, j, invd, [&](int a, int b) { return u[a + b * L.mx]; });
}
// the first red half-sweep of a smoothing from u = 0 at red unknown (i, j):
// the neighbours are black and still zero, so gs_point gives (f w^-1 + 0 + 0) / d
__device__ __forceinline__ float red0_point(const float *f, const Lev &L, int i, int j, float invd)
{
    return (f[i + j * L.mx] * iweight(L, i, j) + L.cx * (0.0f + 0.0f) + L.cy * (0.0f + 0.0f)) * invd;
}
// r = w d (g - u) with g the Gauss-Seidel value, w = 1/iweight
__device__ __forceinline__ float resid_point(const float *u, const float *f, const Lev &L, int i, int j, float invd)
{
    const float g = gs_point(u, f, L, i, j, invd);
    return (g - u[i + j * L.mx]) * (2.0f * (L.cx + L.cy)) * rweight(L, i, j);
}
#define BOT_LOOP(L) \
    for (int j = threadIdx.y; j < (L).my; j += 32) \
        for (int i = threadIdx.x; i < (L).mx; i += 32)

// every level has a valid red-black colouring (gp_mg_build stops before an
// odd periodic extent)
__device__ void bot_sweep(float *u, const float *f, const Lev &L, int color)
{
    // the points of one colour only: column 2 tx + ((j + color) & 1) of row j
    const float invd = 1.0f / (2.0f * (L.cx + L.cy));
    for (int j = threadIdx.y; j < L.my; j += 32)
        for (int i = 2 * threadIdx.x + ((j + color) & 1); i < L.mx; i += 64)
            u[i + j * L.mx] = gs_point(u, f, L, i, j, invd);
    __syncthreads();
}
// the first red sweep of a smoothing from u = 0: a red point's neighbours
// are black and still zero, so gs_point gives (f w^-1 + 0 + 0) / d; the black
// points need no zeroing (a sweep never reads the point it updates)
__device__ void bot_sweep_red0(float *u, const float *f, const Lev &L)
{
    const float invd = 1.0f / (2.0f * (L.cx + L.cy));
    for (int j = threadIdx.y; j < L.my; j += 32)
        for (int i = 2 * threadIdx.x + (j & 1); i < L.mx; i += 64) u[i + j * L.mx] = red0_point(f, L, i, j, invd);
    __syncthreads();
}
__device__ void bot_project(float *a, const Lev &L, float *scratch)
{
    const int tid = threadIdx.x + 32 * threadIdx.y;
    float s = 0.0f;
    BOT_LOOP(L) s += a[i + j * L.mx];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((tid & 63) == 0) scratch[tid >> 6] = s;
    __syncthreads();
    float t = 0.0f;
    for (int w = 0; w < 16; w++) t += scratch[w];   // the same order in every thread
    const float m = t / (float)(L.mx * L.my);
    BOT_LOOP(L) a[i + j * L.mx] -= m;
    __syncthreads();
}

template <int TX, int TY>
__global__ __launch_bounds__(1024) void k_mg_bottom(const float *fin, float *uout, Bottom B, double *part,
                                                    const double *scal, Gate gt)
{
    extern __shared__ float lds[];
    __shared__ float scratch[1024];
    __shared__ float res[BOT_MAX];   // residual of the level being restricted
    __shared__ double dred[16];
    __shared__ double shr[256];
    if (done(scal)) return;
    const int tid = threadIdx.x + 32 * threadIdx.y;
    if (gt.rr_part && gate_converged(gt, shr, tid, true)) return;
    {
        const Lev L = fixed<TX, TY>(B.lv[0]);
        const int n = L.mx * L.my;
        float *f = lds + B.off[0] + n;
        float v[BOT_MAX / 1024];
#pragma unroll
        for (int q = 0; q < BOT_MAX / 1024; q++) {
            const int k = tid + 1024 * q;
            v[q] = k < n ? fin[k] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < BOT_MAX / 1024; q++) {
            const int k = tid + 1024 * q;
            if (k < n) f[k] = v[q];
        }
    }
    __syncthreads();
    for (int l = 0; l + 1 < B.nl; l++) {
        const Lev L = fixed<TX, TY>(B.lv[l]), C = fixed<TX, TY>(B.lv[l + 1]);
        const int n = L.mx * L.my, nc = C.mx * C.my;
        float *u = lds + B.off[l], *f = u + n;
        float *fc = lds + B.off[l + 1] + nc;
        const float invd = 1.0f / (2.0f * (L.cx + L.cy));
        if constexpr (NSW == 1) {
            // one red-black sweep from u = 0: the red values of the first
            // half-sweep are a function of f alone, so the black half-sweep
            // recomputes its neighbours' and comes first; the residual is then
            // evaluated at the red unknowns only, since the black half-sweep
            // has just made it exactly zero at the black ones (the restriction
            // skips them)
            auto red0 = [&](int a, int b) { return red0_point(f, L, a, b, invd); };
            for (int j = threadIdx.y; j < L.my; j += 32)
                for (int i = 2 * threadIdx.x + ((j + 1) & 1); i < L.mx; i += 64)
                    u[i + j * L.mx] = gs_with(f, L, i, j, invd, red0);
            __syncthreads();
            auto uv = [&](int a, int b) { return u[a + b * L.mx]; };
            for (int j = threadIdx.y; j < L.my; j += 32)
                for (int i = 2 * threadIdx.x + (j & 1); i < L.mx; i += 64)
                    res[i + j * L.mx] = (gs_with(f, L, i, j, invd, uv) - red0_point(f, L, i, j, invd)) *
                                        (2.0f * (L.cx + L.cy)) * rweight(L, i, j);
            __syncthreads();
            BOT_LOOP(C) {
                const int cxn = 2 * c_node(C.tx, i), cyn = 2 * c_node(C.ty, j);
                float acc = 0.0f;
                for (int b = -1; b <= 1; b++) {
                    int jj = L.ty == GP_AX_PER ? wrp(cyn + b, L.my) : f_unk(L.ty, cyn + b, L.My);
                    if (jj < 0) continue;
                    const float wy = b == 0 ? 0.5f : 0.25f;
                    for (int a = -1; a <= 1; a++) {
                        int ii = L.tx == GP_AX_PER ? wrp(cxn + a, L.mx) : f_unk(L.tx, cxn + a, L.Mx);
                        if (ii < 0 || ((ii + jj) & 1)) continue;
                        acc += (a == 0 ? 0.5f : 0.25f) * wy * res[ii + jj * L.mx];
                    }
                }
                fc[i + j * C.mx] = acc;
            }
            __syncthreads();
            continue;
        }
        // from u = 0: the first red sweep reads no u
        bot_sweep_red0(u, f, L);
        bot_sweep(u, f, L, 1);
        for (int sw = 1; sw < NSW; sw++) { bot_sweep(u, f, L, 0); bot_sweep(u, f, L, 1); }
        BOT_LOOP(L) res[i + j * L.mx] = resid_point(u, f, L, i, j, invd);
        __syncthreads();
        BOT_LOOP(C) {
            const int cxn = 2 * c_node(C.tx, i), cyn = 2 * c_node(C.ty, j);
            float acc = 0.0f;
            for (int b = -1; b <= 1; b++) {
                int jj = L.ty == GP_AX_PER ? wrp(cyn + b, L.my) : f_unk(L.ty, cyn + b, L.My);
                if (jj < 0) continue;
                const float wy = b == 0 ? 0.5f : 0.25f;
                for (int a = -1; a <= 1; a++) {
                    int ii = L.tx == GP_AX_PER ? wrp(cxn + a, L.mx) : f_unk(L.tx, cxn + a, L.Mx);
                    if (ii < 0) continue;
                    acc += (a == 0 ? 0.5f : 0.25f) * wy * res[ii + jj * L.mx];
                }
            }
            fc[i + j * C.mx] = acc;
        }
        __syncthreads();
    }
    {
        const Lev L = fixed<TX, TY>(B.lv[B.nl - 1]);
        const int n = L.mx * L.my;
        float *u = lds + B.off[B.nl - 1], *f = u + n;
        // (the dense inverse of a singular level has the projections folded
        // in; the fast diagonalisation drops the null mode)
        if (B.singular && !B.dense && !B.fvx) bot_project(f, L, scratch);
        if (B.dense) {
            // u = M f with the precomputed inverse (n <= DENSE_MAX): thread t
            // sums quarter t/256 of row t%256, the four quarters add in LDS;
            // 16 independent loads in flight per step (the matrix streams
            // from L2; one load per dependent fma would serialise them)
            float acc = 0.0f;
            const int row = tid & 255, qd = tid >> 8;
            const int j0 = (n * qd) / 4, j1 = (n * (qd + 1)) / 4;
            if (row < n) {
                int jj = j0;
                for (; jj + 16 <= j1; jj += 16) {
                    float mv[16];
#pragma unroll
                    for (int q = 0; q < 16; q++) mv[q] = B.dense[(jj + q) * n + row];
#pragma unroll
                    for (int q = 0; q < 16; q++) acc = __builtin_fmaf(mv[q], f[jj + q], acc);
                }
                for (; jj < j1; jj++) acc = __builtin_fmaf(B.dense[jj * n + row], f[jj], acc);
            }
            scratch[tid] = acc;
            __syncthreads();
            if (tid < n) u[tid] = (scratch[tid] + scratch[tid + 256]) + (scratch[tid + 512] + scratch[tid + 768]);
            __syncthreads();
        } else if (B.fvx) {
            // exact solve by fast diagonalisation: with the weighted operator
            // S = (Wx (x) Wy)(cx Kx (x) I + cy I (x) Ky) and K V = V Lambda,
            // V' W V = I per axis,  u = (Vx (x) Vy) D^-1 (Vx' (x) Vy') f
            // (res holds the intermediate transforms; a null mode has 1/0 -> 0)
            const int mx = L.mx, my = L.my;
            float *t = res;
            for (int k = tid; k < n; k += 1024) {        // along x: t = Vx' f
                const int ip = k % mx, j = k / mx;
                float acc = 0.0f;
                for (int i = 0; i < mx; i++) acc = __builtin_fmaf(B.fvx[i * mx + ip], f[i + j * mx], acc);
                t[k] = acc;
            }
            __syncthreads();
            for (int k = tid; k < n; k += 1024) {        // along y, scaled: u = D^-1 (t Vy)
                const int ip = k % mx, jp = k / mx;
                float acc = 0.0f;
                for (int j = 0; j < my; j++) acc = __builtin_fmaf(B.fvy[j * my + jp], t[ip + j * mx], acc);
                u[k] = acc * B.finv[k];
            }
            __syncthreads();
            for (int k = tid; k < n; k += 1024) {        // back along y: t = u Vy'
                const int ip = k % mx, j = k / mx;
                float acc = 0.0f;
                for (int jp = 0; jp < my; jp++) acc = __builtin_fmaf(B.fvy[j * my + jp], u[ip + jp * mx], acc);
                t[k] = acc;
            }
            __syncthreads();
            for (int k = tid; k < n; k += 1024) {        // back along x: u = Vx t
                const int i = k % mx, j = k / mx;
                float acc = 0.0f;
                for (int ip = 0; ip < mx; ip++) acc = __builtin_fmaf(B.fvx[i * mx + ip], t[ip + j * mx], acc);
                u[k] = acc;
            }
            __syncthreads();
        } else {
            bot_sweep_red0(u, f, L);
            bot_sweep(u, f, L, 1);
            for (int s = 1; s < B.sweeps; s++) { bot_sweep(u, f, L, 0); bot_sweep(u, f, L, 1); }
            for (int s = 0; s < B.sweeps; s++) { bot_sweep(u, f, L, 1); bot_sweep(u, f, L, 0); }
        }
        if (B.singular && !B.dense && !B.fvx) bot_project(u, L, scratch);
    }
    for (int l = B.nl - 2; l >= 0; l--) {
        const Lev L = fixed<TX, TY>(B.lv[l]), C = fixed<TX, TY>(B.lv[l + 1]);
        const int n = L.mx * L.my;
        float *u = lds + B.off[l], *f = u + n;
        const float *uc = lds + B.off[l + 1];
        if constexpr (NSW == 1) {
            // the black half-sweep reads only red values and the red one only
            // black values, so the coarse correction is added at the red
            // unknowns only (to their pre-smoothed values, recomputed from f)
            const float invd = 1.0f / (2.0f * (L.cx + L.cy));
            for (int j = threadIdx.y; j < L.my; j += 32)
                for (int i = 2 * threadIdx.x + (j & 1); i < L.mx; i += 64)
                    u[i + j * L.mx] = red0_point(f, L, i, j, invd) + prolong(uc, L, C, i, j);
            __syncthreads();
            bot_sweep(u, f, L, 1);
            bot_sweep(u, f, L, 0);
            continue;
        }
        BOT_LOOP(L) u[i + j * L.mx] += prolong(uc, L, C, i, j);
        __syncthreads();
        for (int sw = 0; sw < NSW; sw++) { bot_sweep(u, f, L, 1); bot_sweep(u, f, L, 0); }
    }
    {
        const Lev L = fixed<TX, TY>(B.lv[0]);
        const int n = L.mx * L.my;
        const float *u = lds + B.off[0], *f = u + n;
        double acc = 0.0;
        for (int k = tid; k < n; k += 1024) {
            uout[k] = u[k];
            acc += (double)f[k] * (double)u[k];
        }
        if (part) {
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
            if ((tid & 63) == 0) dred[tid >> 6] = acc;
            __syncthreads();
            if (tid == 0) {
                double t = 0.0;
                for (int w = 0; w < 16; w++) t += dred[w];
                part[0] = t;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// coarsest level of at most 64 x 64 unknowns solved exactly by fast
// diagonalisation on MFMA, in one workgroup of 16 waves:
//   u = Vy (Linv o (Vy' F Vx)) Vx'
// with F[j][i] = f(i, j) (row j of the unknown grid), K V = V Lambda per axis
// (V' W V = I) and Linv = 1 / (cx lambda_x + cy lambda_y) (0 for a null mode).
// Every factor is zero-padded to 64 x 64 in LDS with row pitch FDP, so that
// the strided operand reads of the products hit distinct banks; each wave
// computes one 16 x 16 tile of each of the four products with
// v_mfma_f32_16x16x4_f32 (A: lane l gives A[l % 16][k + l / 16], B: B[k +
// l / 16][l % 16]; D: register q of lane l holds D[4 (l / 16) + q][l % 16]).
// ---------------------------------------------------------------------------
constexpr int FD64 = 64, FDP = FD64 + 1;
constexpr int FD64_LDS = 6 * FD64 * FDP * (int)sizeof(float);   // dynamic LDS of k_mg_fd64
typedef float f32x4 __attribute__((ext_vector_type(4)));
struct Fd64 {
    const float *v;   // Vx, Vy, Linv: 3 x 64 x 64
    int mx, my;
};
// tile (ti, tj) of A B with A[i][k] = a(i, k), B[k][j] = b(k, j), K = 64
template <class FA, class FB>
__device__ __forceinline__ f32x4 fd_tile(FA a, FB b, int ti, int tj, int lane)
{
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int k0 = 0; k0 < FD64; k0 += 4)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a(ti * 16 + r, k0 + kq), b(k0 + kq, tj * 16 + r), acc, 0, 0, 0);
    return acc;
}
__global__ __launch_bounds__(1024) void k_mg_fd64(const float *fin, float *uout, Fd64 A, double *part,
                                                  const double *scal, Gate gt)
{
    extern __shared__ float lds[];   // F, Vx, Vy, Linv, T, U: 64 x FDP each
    __shared__ double shr[256];
    __shared__ double dred[16];
    if (done(scal)) return;
    const int tid = threadIdx.x;
    if (gt.rr_part && gate_converged(gt, shr, tid, true)) return;
    float *F = lds, *Vx = F + FD64 * FDP, *Vy = Vx + FD64 * FDP, *Li = Vy + FD64 * FDP;
    float *T = Li + FD64 * FDP, *U = T + FD64 * FDP;
    const int mx = A.mx, my = A.my;
    {
        // every load is issued before any is consumed (indices clamped into
        // the level; the padding is selected afterwards)
        float fv[4], xv[4], yv[4], lv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 1024 * q, i = e & 63, j = e >> 6;
            fv[q] = fin[std::min(i, mx - 1) + std::min(j, my - 1) * mx];
            xv[q] = A.v[e];
            yv[q] = A.v[4096 + e];
            lv[q] = A.v[8192 + e];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + 1024 * q, i = e & 63, j = e >> 6, k = j * FDP + i;
            F[k] = (i < mx && j < my) ? fv[q] : 0.0f;
            Vx[k] = xv[q];
            Vy[k] = yv[q];
            Li[k] = lv[q];
        }
    }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6, ti = wv >> 2, tj = wv & 3;
    const int r = lane & 15, kq = lane >> 4;
    {   // T = F Vx
        const f32x4 c = fd_tile([&](int i, int k) { return F[i * FDP + k]; },
                                [&](int k, int j) { return Vx[k * FDP + j]; }, ti, tj, lane);
#pragma unroll
        for (int q = 0; q < 4; q++) T[(ti * 16 + 4 * kq + q) * FDP + tj * 16 + r] = c[q];
    }
    __syncthreads();
    {   // U = Linv o (Vy' T)
        const f32x4 c = fd_tile([&](int i, int k) { return Vy[k * FDP + i]; },
                                [&](int k, int j) { return T[k * FDP + j]; }, ti, tj, lane);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = (ti * 16 + 4 * kq + q) * FDP + tj * 16 + r;
            U[k] = c[q] * Li[k];
        }
    }
    __syncthreads();
    {   // T = Vy U
        const f32x4 c = fd_tile([&](int i, int k) { return Vy[i * FDP + k]; },
                                [&](int k, int j) { return U[k * FDP + j]; }, ti, tj, lane);
#pragma unroll
        for (int q = 0; q < 4; q++) T[(ti * 16 + 4 * kq + q) * FDP + tj * 16 + r] = c[q];
    }
    __syncthreads();
    double acc = 0.0;
    {   // u = T Vx'
        const f32x4 c = fd_tile([&](int i, int k) { return T[i * FDP + k]; },
                                [&](int k, int j) { return Vx[j * FDP + k]; }, ti, tj, lane);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int j = ti * 16 + 4 * kq + q, i = tj * 16 + r;
            if (i < mx && j < my) {
                uout[i + j * mx] = c[q];
                acc += (double)F[j * FDP + i] * (double)c[q];
            }
        }
    }
    if (part) {   // this level is the finest: the partial of r.z
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if (lane == 0) dred[wv] = acc;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < 16; w++) t += dred[w];
            part[0] = t;
        }
    }
}

// ---------------------------------------------------------------------------
// CG vector kernels (FP32 data, FP64 accumulation)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < RB / 64; w++) s += sh[w];
    return s;
}
// final reduction of partials into scal[slot]; op 1 also sets
// alpha = rz / sum.  gate != 0: skip once the convergence flag is up.
__global__ void k_final(const double *part, int nb, double *scal, int slot, int op, int gate = 0)
{
    __shared__ double sh[RB];
    if (gate && done(scal)) return;
    double acc = 0.0;
    for (int k = threadIdx.x; k < nb; k += RB) acc += part[k];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) apply_op(scal, slot, op, sh[0]);
}
// the residual test alone (a hierarchy without grid levels above its
// coarsest, solved by sweeps)
__global__ void k_gate(Gate g)
{
    __shared__ double shr[256];
    if (done(g.scal)) return;
    (void)gate_converged(g, shr, threadIdx.x, true);
}
// p_new = z + beta p_old (beta from scal), q = S p_new, partial p_new.q
// Iteration it of the PCG: rz_it = r.z from the partials of the V-cycle,
// beta = rz_it / rz_(it-1) (0 on the first iteration, which ignores p_old);
// block 0 stores rz_it in slot cur (the previous value stays in slot prev,
// which the other blocks read).
// rr_true (may be null): the previous iteration replaced its recursive
// residual (partials rr_rec) by the true one (partials rr_true); when their
// norms differ by more than MG_DRIFT the old direction carries the
// recursion's drift and the search restarts from z (beta = 0).
template <int TX, int TY>
__global__ __launch_bounds__(RB) void k_pupdate_spmv(const float *z, const float *pold, float *pnew, float *q, Lev Lin,
                                                    double *scal, double *part, const double *rz_part, int nrz,
                                                    int first, int cur, int prev, const double *rr_rec = nullptr,
                                                    const double *rr_true = nullptr, int nrr = 0)
{
    const Lev L = fixed<TX, TY>(Lin);
    __shared__ double sh[RB / 64];
    __shared__ double shr[256];
    if (done(scal)) return;
    float beta = 0.0f;
    {
        const double rz = sum_parts(rz_part, nrz, shr, threadIdx.x);
        if (rr_true) {
            const double a = sqrt(sum_parts(rr_true, nrr, shr, threadIdx.x));
            const double b = sqrt(sum_parts(rr_rec, nrr, shr, threadIdx.x));
            if (!(fabs(a - b) <= MG_DRIFT * a)) first = 1;
        }
        if (!first) {
            const double ro = scal[prev];
            beta = ro != 0.0 ? (float)(rz / ro) : 0.0f;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) scal[cur] = rz;
    }
    const int n = L.mx * L.my;
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) {
        const int i = k % L.mx, j = k / L.mx;
        auto P = [&](int kk) { return __builtin_fmaf(beta, pold[kk], z[kk]); };
        const float c = P(k);
        float uL, uR, uD, uU;
        uL = i > 0 ? P(k - 1) : (L.tx == GP_AX_PER ? P(k + L.mx - 1) : (L.tx == GP_AX_NEU ? P(k + 1) : 0.0f));
        uR = i < L.mx - 1 ? P(k + 1) : (L.tx == GP_AX_PER ? P(k - (L.mx - 1)) : (L.tx == GP_AX_NEU ? P(k - 1) : 0.0f));
        uD = j > 0 ? P(k - L.mx) : (L.ty == GP_AX_PER ? P(k + (L.my - 1) * L.mx) : (L.ty == GP_AX_NEU ? P(k + L.mx) : 0.0f));
        uU = j < L.my - 1 ? P(k + L.mx) : (L.ty == GP_AX_PER ? P(k - (L.my - 1) * L.mx) : (L.ty == GP_AX_NEU ? P(k - L.mx) : 0.0f));
        const float w = wgt(L.tx, i, L.mx) * wgt(L.ty, j, L.my);
        const float v = w * (L.cx * (2.0f * c - uL - uR) + L.cy * (2.0f * c - uD - uU));
        pnew[k] = c;
        q[k] = v;
        acc += (double)c * (double)v;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// e += alpha p; r -= alpha q; partial r.r
// alpha = scal[S_ALPHA], or, with pq_part, rz (slot rzs) / (sum of the p.q
// partials), reduced by every block itself
__global__ __launch_bounds__(RB) void k_update_xr(float *x, float *r, const float *p, const float *q, int n,
                                                 const double *scal, double *part, const double *pq_part = nullptr,
                                                 int npq = 0, int rzs = S_RZ)
{
    __shared__ double sh[RB / 64];
    __shared__ double shr[256];
    if (done(scal)) return;
    float alpha;
    if (pq_part) {
        const double pq = sum_parts(pq_part, npq, shr, threadIdx.x);
        alpha = pq != 0.0 ? (float)(scal[rzs] / pq) : 0.0f;
    } else {
        alpha = (float)scal[S_ALPHA];
    }
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) {
        x[k] = __builtin_fmaf(alpha, p[k], x[k]);
        const float rv = __builtin_fmaf(-alpha, q[k], r[k]);
        r[k] = rv;
        acc += (double)rv * (double)rv;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ __launch_bounds__(RB) void k_dot(const float *a, const float *b, int n, double *part)
{
    __shared__ double sh[RB / 64];
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) acc += (double)a[k] * (double)b[k];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void k_set_scalar(double *scal, int slot, double v) { scal[slot] = v; }
__global__ void k_zero_f(float *d, int n)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) d[k] = 0.0f;
}
// q = S x, partial x.q
__global__ __launch_bounds__(RB) void k_spmv_dot(const float *x, float *q, Lev L, double *part)
{
    __shared__ double sh[RB / 64];
    const int n = L.mx * L.my;
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) {
        const int i = k % L.mx, j = k / L.mx;
        const float c = x[k];
        float uL, uR, uD, uU;
        uL = i > 0 ? x[k - 1] : (L.tx == GP_AX_PER ? x[k + L.mx - 1] : (L.tx == GP_AX_NEU ? x[k + 1] : 0.0f));
        uR = i < L.mx - 1 ? x[k + 1] : (L.tx == GP_AX_PER ? x[k - (L.mx - 1)] : (L.tx == GP_AX_NEU ? x[k - 1] : 0.0f));
        uD = j > 0 ? x[k - L.mx] : (L.ty == GP_AX_PER ? x[k + (L.my - 1) * L.mx] : (L.ty == GP_AX_NEU ? x[k + L.mx] : 0.0f));
        uU = j < L.my - 1 ? x[k + L.mx] : (L.ty == GP_AX_PER ? x[k - (L.my - 1) * L.mx] : (L.ty == GP_AX_NEU ? x[k - L.mx] : 0.0f));
        const float w = wgt(L.tx, i, L.mx) * wgt(L.ty, j, L.my);
        const float v = w * (L.cx * (2.0f * c - uL - uR) + L.cy * (2.0f * c - uD - uU));
        q[k] = v;
        acc += (double)c * (double)v;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// p = z + beta p (beta = scal[S_BETA])
__global__ void k_xpby(const float *z, float *p, int n, const double *scal)
{
    const float beta = (float)scal[S_BETA];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        p[k] = __builtin_fmaf(beta, p[k], z[k]);
}
// flexible-CG coefficients (Polak-Ribiere form, r_new - r_old = -alpha q):
// beta = -alpha (z.q) / rz_old, rz = rz_new; the sums are in S_TMP (z.r)
// and S_BB (z.q)
__global__ void k_fcg_beta(double *scal)
{
    scal[S_BETA] = scal[S_RZ] != 0.0 ? -scal[S_ALPHA] * scal[S_BB] / scal[S_RZ] : 0.0;
    scal[S_RZ] = scal[S_TMP];
}
// e *= (e.r) / (e.S e): S_TMP = e.r, S_PQ = e.S e
__global__ void k_scale_opt(float *e, int n, const double *scal)
{
    const double d = scal[S_PQ];
    const float a = d > 0.0 ? (float)(scal[S_TMP] / d) : 0.0f;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) e[k] *= a;
}

// FP64 kernels of the solve
// (S x)_k in FP64 at unknown k = i + j mx
__device__ __forceinline__ double apply64(const double *x, const Lev &L, double cx, double cy, int k)
{
    const int i = k % L.mx, j = k / L.mx;
    const double c = x[k];
    double xl, xr, yl, yr;
    if (i > 0) xl = x[k - 1]; else xl = L.tx == GP_AX_PER ? x[k + L.mx - 1] : (L.tx == GP_AX_NEU ? x[k + 1] : 0.0);
    if (i < L.mx - 1) xr = x[k + 1]; else xr = L.tx == GP_AX_PER ? x[k - (L.mx - 1)] : (L.tx == GP_AX_NEU ? x[k - 1] : 0.0);
    if (j > 0) yl = x[k - L.mx]; else yl = L.ty == GP_AX_PER ? x[k + (L.my - 1) * L.mx] : (L.ty == GP_AX_NEU ? x[k + L.mx] : 0.0);
    if (j < L.my - 1) yr = x[k + L.mx]; else yr = L.ty == GP_AX_PER ? x[k - (L.my - 1) * L.mx] : (L.ty == GP_AX_NEU ? x[k - L.mx] : 0.0);
    const double w = (double)wgt(L.tx, i, L.mx) * (double)wgt(L.ty, j, L.my);
    return w * (cx * (2.0 * c - xl - xr) + cy * (2.0 * c - yl - yr));
}
// r = b - S x (FP64), partials of r.r and b.b.
// bsum (may be null): the nbs partials of sum(b); b is first centred in
// place (b -= sum / n), every block finishing the sum itself.
// part_x (may be null): partials of sum(x)
__global__ __launch_bounds__(RB) void k_resid64(const double *x, double *b, double *r, Lev L, double cx, double cy,
                                               double *part_rr, double *part_bb, const double *bsum = nullptr,
                                               int nbs = 0, double *part_x = nullptr)
{
    __shared__ double sh[RB / 64];
    __shared__ double shr[256];
    const int n = L.mx * L.my;
    const double m = bsum ? sum_parts(bsum, nbs, shr, threadIdx.x) / (double)n : 0.0;
    double arr = 0.0, abb = 0.0, ax = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) {
        double bk = b[k];
        if (bsum) {
            bk -= m;
            b[k] = bk;
        }
        const double rv = bk - apply64(x, L, cx, cy, k);
        r[k] = rv;
        arr += rv * rv;
        abb += bk * bk;
        if (part_x) ax += x[k];
    }
    double s = block_sum(arr, sh);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = s;
    __syncthreads();
    s = block_sum(abb, sh);
    if (threadIdx.x == 0) part_bb[blockIdx.x] = s;
    if (part_x) {
        __syncthreads();
        s = block_sum(ax, sh);
        if (threadIdx.x == 0) part_x[blockIdx.x] = s;
    }
}
// start of an inner solve: rf = FP32 of scale * r, e = p = 0, and the block
// partials of rf.rf (FP64, one per block, grid-stride order)
__global__ __launch_bounds__(RB) void k_inner_start(const double *r, float *rf, float *e, float *p, int n,
                                                    double scale, double *part)
{
    __shared__ double sh[RB / 64];
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) {
        const float v = (float)(r[k] * scale);
        rf[k] = v;
        e[k] = 0.0f;
        p[k] = 0.0f;
        acc += (double)v * (double)v;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// rf = FP32 of sc r, p = 0 when given, partials of rf.rf.
// Start of the MG-PCG (rpcg) when prr is given: every block finishes the true
// residual's r.r and b.b from the nb partials of k_resid64 itself and takes
// the scale sc that brings r to unit RMS; block 0 records the solve's
// scalars, among them the target of the scaled FP32 recursion,
// ||r|| <= gf tol ||b|| sc, and raises the flag at once when the iterate
// already meets tol.  Otherwise sc = scal[S_SC].
__global__ __launch_bounds__(RB) void k_rscale(const double *r, float *rf, float *p, int n, double *scal,
                                               double *part, const double *prr = nullptr, const double *pbb = nullptr,
                                               int nb = 0, double tol = 0.0, double gf = 0.0, int maxit = 0)
{
    __shared__ double sh[RB / 64];
    __shared__ double shr[256];
    double sc;
    if (prr) {
        const double rr = sum_parts(prr, nb, shr, threadIdx.x);
        const double bb = sum_parts(pbb, nb, shr, threadIdx.x);
        const double rms = sqrt(rr / (double)n);
        sc = rms > 0.0 ? 1.0 / rms : 1.0;
        const bool conv = rr <= tol * tol * bb || bb == 0.0;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal[S_TRR] = rr;
            scal[S_TRR0] = rr;
            scal[S_BB] = bb;
            scal[S_SC] = sc;
            scal[S_ISC] = 1.0 / sc;
            scal[S_TGT] = gf * tol * sqrt(bb) * sc;
            scal[S_DONE] = conv ? 1.0 : 0.0;
            scal[S_IT] = 0.0;
            scal[S_MAXIT] = (double)maxit;
        }
        if (conv) return;
    } else {
        if (done(scal)) return;
        sc = scal[S_SC];
    }
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) {
        const float v = (float)(r[k] * sc);
        rf[k] = v;
        if (p) p[k] = 0.0f;
        acc += (double)v * (double)v;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// x += (alpha / sc) p in FP64, r -= alpha q in FP32, partials of r.r;
// alpha = rz (slot rzs) / (sum of the p.q partials), reduced by every block
__global__ __launch_bounds__(RB) void k_update_xr64(double *x, float *r, const float *p, const float *q, int n,
                                                    const double *scal, double *part, const double *pq_part, int npq,
                                                    int rzs)
{
    __shared__ double sh[RB / 64];
    __shared__ double shr[256];
    if (done(scal)) return;
    const double pq = sum_parts(pq_part, npq, shr, threadIdx.x);
    const double ad = pq != 0.0 ? scal[rzs] / pq : 0.0;
    const float alpha = (float)ad;
    const double ax = ad * scal[S_ISC];
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) {
        x[k] = __builtin_fma(ax, (double)p[k], x[k]);
        const float rv = __builtin_fmaf(-alpha, q[k], r[k]);
        r[k] = rv;
        acc += (double)rv * (double)rv;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// residual replacement: rf = FP32 of sc (b - S x) from the FP64 iterate,
// partials of rf.rf
__global__ __launch_bounds__(RB) void k_replace(const double *x, const double *b, float *rf, Lev L, double cx, double cy,
                                                const double *scal, double *part)
{
    __shared__ double sh[RB / 64];
    if (done(scal)) return;
    const double sc = scal[S_SC];
    const int n = L.mx * L.my;
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) {
        const float v = (float)((b[k] - apply64(x, L, cx, cy, k)) * sc);
        rf[k] = v;
        acc += (double)v * (double)v;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// the recursion met its target but the true residual did not: lower the
// target by f and clear the flag
__global__ void k_retarget(double *scal, double f)
{
    scal[S_TGT] *= f;
    scal[S_DONE] = 0.0;
}
// two reductions into two slots in one launch (the refinement's r.r, b.b),
// and a third when pc is given
__global__ void k_final2(const double *pa, const double *pb, int nb, double *scal, int sa, int sb,
                         const double *pc = nullptr, int scs = 0)
{
    __shared__ double sh[256];
    const double a = sum_parts(pa, nb, sh, threadIdx.x);
    const double b = sum_parts(pb, nb, sh, threadIdx.x);
    const double c = pc ? sum_parts(pc, nb, sh, threadIdx.x) : 0.0;
    if (threadIdx.x == 0) {
        scal[sa] = a;
        scal[sb] = b;
        if (pc) scal[scs] = c;
    }
}
__global__ void k_axpy64(double *x, const float *e, int n, double scale)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) x[k] += (double)e[k] * scale;
}
__global__ __launch_bounds__(RB) void k_sum64(const double *a, int n, double *part)
{
    __shared__ double sh[RB / 64];
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) acc += a[k];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void k_sub_mean64(double *a, int n, const double *scal, int slot)
{
    const double m = scal[slot] / (double)n;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) a[k] -= m;
}

Lev mklev(const gp_mg_level &l)
{
    Lev L;
    L.mx = l.mx; L.my = l.my;
    L.tx = l.perx; L.ty = l.pery;
    L.Mx = L.tx == GP_AX_DIR ? l.mx + 1 : (L.tx == GP_AX_NEU ? l.mx - 1 : l.mx);
    L.My = L.ty == GP_AX_DIR ? l.my + 1 : (L.ty == GP_AX_NEU ? l.my - 1 : l.my);
    L.cx = (float)l.hx2i; L.cy = (float)l.hy2i;
    L.rb = !((L.tx == GP_AX_PER && (l.mx & 1)) || (L.ty == GP_AX_PER && (l.my & 1)));
    return L;
}
inline dim3 tiles(const Lev &L, int tb = TB) { return dim3((L.mx + tb - 1) / tb, (L.my + tb - 1) / tb); }
// regions of the solver's reduction buffer s.red (doubles): partials of p.q,
// of r.r (both at most 1024 blocks), of r.z (one per fine-level tile, at
// most the 32 x 32 tiles of level 0), then sums of b or x, the true
// residual's r.r at a replacement, and b.b of a true-residual test (whose
// r.r goes to the p.q region: no iteration is under way)
constexpr int RED_PQ = 0, RED_RR = 1024, RED_RZ = 2048;
inline int RED_X(const gp_solver &s)
{
    const int64_t t = (int64_t)gp_cdiv(s.lv[0].mx, TB) * gp_cdiv(s.lv[0].my, TB);
    return RED_RZ + (int)std::max<int64_t>(t, 1024);
}
inline int RED_TR(const gp_solver &s) { return RED_X(s) + 1024; }
inline int RED_B(const gp_solver &s) { return RED_TR(s) + 1024; }
// tile edge of the level kernels: 64 where 32 x 32 tiles would need more
// than one round of workgroups (two 1024-thread groups per CU), so that a
// fine level runs in one round with a smaller share of recomputed halo
inline int level_tile(const Lev &L)
{
    const dim3 g = tiles(L);
    return (int64_t)g.x * g.y > 512 ? 2 * TB : TB;
}

// z = M r (s.lv[0].f aliases r, s.lv[0].u aliases z); part: partial r.z.
// scal (may be null) carries the convergence flag that makes every kernel a
// no-op once the CG has converged.
// gate (may be null): the residual test of the PCG iteration, made by the
// V-cycle's first kernel
// launch CALL(TX, TY) for the runtime axis types (GP_AX_PER 0, DIR 1, NEU 2)
#define MG_AXES(tx, ty, CALL)                                            \
    do {                                                                 \
        switch ((tx) * 3 + (ty)) {                                       \
        case 0: CALL(0, 0); break;                                       \
        case 1: CALL(0, 1); break;                                       \
        case 2: CALL(0, 2); break;                                       \
        case 3: CALL(1, 0); break;                                       \
        case 4: CALL(1, 1); break;                                       \
        case 5: CALL(1, 2); break;                                       \
        case 6: CALL(2, 0); break;                                       \
        case 7: CALL(2, 1); break;                                       \
        default: CALL(2, 2); break;                                      \
        }                                                                \
    } while (0)

int vcycle(hipStream_t st, gp_solver &s, double *part, int *nparts, const double *scal, const Gate *gate = nullptr)
{
    const Gate nog{nullptr, 0, 0, nullptr};
    bool gated = false;
    auto take = [&]() { const Gate g = (gate && !gated) ? *gate : nog; gated = true; return g; };
    const int nl = (int)s.lv.size();
    const int lb = s.bottom;   // first level handled by k_mg_bottom (nl if none)
    const dim3 blk(32, BY), blk8(32, 8);
    for (int l = 0; l < std::min(lb, nl - 1); l++) {
        const Lev L = mklev(s.lv[l]), C = mklev(s.lv[l + 1]);
        const Gate g = take();
        const int tb = level_tile(L);
        if (tb == TB) {
#define MG_DOWN(X, Y) k_mg_down<NSW, TB, X, Y><<<tiles(L, tb), blk, 0, st>>>(s.lv[l].f, s.lv[l].r, s.lv[l + 1].f, L, C, scal, g)
            MG_AXES(L.tx, L.ty, MG_DOWN);
#undef MG_DOWN
        } else {
#define MG_DOWN(X, Y) k_mg_down<NSW, 2 * TB, X, Y><<<tiles(L, tb), blk, 0, st>>>(s.lv[l].f, s.lv[l].r, s.lv[l + 1].f, L, C, scal, g)
            MG_AXES(L.tx, L.ty, MG_DOWN);
#undef MG_DOWN
        }
    }
    GP_HIP(hipGetLastError());
    *nparts = 0;
    if (lb < nl && s.fd64) {
        // the coarsest level (lb = nl - 1) solved exactly on MFMA
        const bool top = lb == 0;
        const Gate bg = take();
        const Fd64 A{s.fd64, s.lv[lb].mx, s.lv[lb].my};
        k_mg_fd64<<<1, 1024, FD64_LDS, st>>>(s.lv[lb].f, s.lv[lb].u, A, top ? part : nullptr, scal, bg);
        GP_HIP(hipGetLastError());
        if (top) *nparts = 1;
    } else if (lb < nl) {
        Bottom B{};
        int off = 0;
        B.nl = nl - lb;
        for (int l = lb; l < nl; l++) {
            B.lv[l - lb] = mklev(s.lv[l]);
            B.off[l - lb] = off;
            off += 2 * s.lv[l].mx * s.lv[l].my;
        }
        const Lev &Lc = B.lv[B.nl - 1];
        B.singular = s.singular;
        B.sweeps = std::max(2, std::max(Lc.mx, Lc.my) / 2);
        B.dense = s.dinv;
        B.fvx = s.fd_vx; B.fvy = s.fd_vy; B.finv = s.fd_inv;
        const bool top = lb == 0;
        double *bp = top ? part : nullptr;
        const Gate bg = take();
#define MG_BOT(X, Y) k_mg_bottom<X, Y><<<1, dim3(32, 32), off * sizeof(float), st>>>(s.lv[lb].f, s.lv[lb].u, B, bp, scal, bg)
        MG_AXES(B.lv[0].tx, B.lv[0].ty, MG_BOT);
#undef MG_BOT
        GP_HIP(hipGetLastError());
        if (top) *nparts = 1;
    } else {
        // the coarsest grid level could not be coarsened into the LDS range:
        // symmetric red-black sweeps with grid kernels, ping-pong buffers
        // (sweep k reads a, writes b; an even count ends in u)
        const Lev L = mklev(s.lv[nl - 1]);
        const int sweeps = std::min(400, std::max(4, 2 * std::max(L.mx, L.my)));
        float *a = s.lv[nl - 1].u, *b = s.lv[nl - 1].r;
        if (gate && !gated) {   // no grid level above: test the residual on its own
            k_gate<<<1, 256, 0, st>>>(*gate);
            gated = true;
        }
        k_zero_f<<<256, 256, 0, st>>>(a, L.mx * L.my);
        for (int k = 0; k < 2 * sweeps; k++) {
            k_rb_sweep<<<tiles(L), blk8, 0, st>>>(a, s.lv[nl - 1].f, b, L, k < sweeps ? 0 : 1, scal);
            std::swap(a, b);
        }
        GP_HIP(hipGetLastError());
        if (nl == 1) {
            const int n = L.mx * L.my;
            k_dot<<<std::min(1024, (n + RB - 1) / RB), RB, 0, st>>>(s.lv[0].f, s.lv[0].u, n, part);
            *nparts = std::min(1024, (n + RB - 1) / RB);
        }
    }
    for (int l = std::min(lb, nl - 1) - 1; l >= 0; l--) {
        const Lev L = mklev(s.lv[l]), C = mklev(s.lv[l + 1]);
        const int tb = level_tile(L);
        const dim3 g = tiles(L, tb);
        double *pp = l == 0 ? part : nullptr;
        if (tb == TB) {
#define MG_UP(X, Y) k_mg_up<NSW, TB, X, Y><<<g, blk, 0, st>>>(s.lv[l].r, s.lv[l].f, s.lv[l + 1].u, s.lv[l].u, L, C, pp, scal)
            MG_AXES(L.tx, L.ty, MG_UP);
#undef MG_UP
        } else {
#define MG_UP(X, Y) k_mg_up<NSW, 2 * TB, X, Y><<<g, blk, 0, st>>>(s.lv[l].r, s.lv[l].f, s.lv[l + 1].u, s.lv[l].u, L, C, pp, scal)
            MG_AXES(L.tx, L.ty, MG_UP);
#undef MG_UP
        }
        if (l == 0) *nparts = (int)(g.x * g.y);
    }
    GP_HIP(hipGetLastError());
    return GP_OK;
}

// MG-PCG on S x = b64 from the FP64 iterate s.x64, to ||b - S x|| <= tol ||b||.
// The iterate is accumulated in FP64 at every iteration; the residual
// recursion, the V-cycle and the search directions are FP32 (dot products
// in FP64), on the residual scaled to unit RMS.  Every s.rep_every
// iterations the recursive residual is replaced by the true one computed
// from the FP64 iterate, so the recursion's rounding (about eps32 cond(S)
// of the residual it starts from) never accumulates into the stopping test,
// and the CG keeps its search direction instead of restarting.  The
// interval is the number of iterations that reduced the residual by
// MG_REPLACE_DROP at the previous solve's rate.
// Iterations are enqueued in batches (the previous solve's count first); a
// device-side flag, raised by the residual test at the target or the
// iteration limit, turns the rest of a batch into no-ops.  The last
// iteration of a batch makes its residual test alone and leaves its V-cycle
// pending (a converged solve never needs it).  Each batch ends with the true
// FP64 residual, the one host synchronisation; when the recursion met its
// target and the true residual did not, the solve continues from the true
// residual with a lower target.
// center: b is centred (a singular system) by the first residual kernel.
// On a singular system the checks also leave sum(x) in scal[S_MEAN].
// iters: iterations made; checks: true-residual tests; rel: the last one.
int rpcg(hipStream_t st, gp_solver &s, double tol, int maxit, bool center, int *iters, int *checks, double *relo)
{
    const Lev L0 = mklev(s.lv[0]);
    const int n = s.n, nb = s.red_blocks;
    const double cx = s.lv[0].hx2i, cy = s.lv[0].hy2i;
    float *p = s.p, *p2 = s.q2;
    double *red_pq = s.red + RED_PQ, *red_rr = s.red + RED_RR, *red_rz = s.red + RED_RZ, *red_x = s.red + RED_X(s);
    double *red_tr = s.red + RED_TR(s), *red_b = s.red + RED_B(s);
    if (center) k_sum64<<<nb, RB, 0, st>>>(s.b64, n, red_x);
    k_resid64<<<nb, RB, 0, st>>>(s.x64, s.b64, s.r64, L0, cx, cy, red_pq, red_b, center ? red_x : nullptr, nb);
    k_rscale<<<nb, RB, 0, st>>>(s.r64, s.r, p, n, s.scal, red_rr, red_pq, red_b, nb, tol, MG_TARGET, maxit);
    GP_HIP(hipGetLastError());
    int np = 0;
    int rc = vcycle(st, s, red_rz, &np, s.scal);            // z = M r, partials of r.z
    if (rc) return rc;
    int it = 0, first = 1, nchk = 0, status = GP_OK;
    bool pending = false;   // the V-cycle of iteration it is still to be made
    bool replaced = false;  // iteration it replaced its recursive residual
    double rel = 0.0;
    // each kernel finishes the reduction it consumes (r.z in the p update,
    // p.q in the x/r update, r.r in the residual test), so an iteration is
    // one launch per kernel
    for (int batch = std::max(1, s.last_it);; batch = 1) {
        for (int b = 0; b < batch && it < maxit; b++) {
            it++;
            const int cur = (it & 1) ? S_RZ : S_RZ2, prev = (it & 1) ? S_RZ2 : S_RZ;
#define MG_PUP(X, Y) k_pupdate_spmv<X, Y><<<nb, RB, 0, st>>>(s.z, p, p2, s.q, L0, s.scal, red_pq, red_rz, np, first, \
                                                             cur, prev, red_rr, replaced ? red_tr : nullptr, nb)
            MG_AXES(L0.tx, L0.ty, MG_PUP);
#undef MG_PUP
            std::swap(p, p2);
            first = 0;
            k_update_xr64<<<nb, RB, 0, st>>>(s.x64, s.r, p, s.q, n, s.scal, red_rr, red_pq, nb, cur);
            replaced = it % s.rep_every == 0;
            if (replaced) k_replace<<<nb, RB, 0, st>>>(s.x64, s.b64, s.r, L0, cx, cy, s.scal, red_tr);
            const Gate g{replaced ? red_tr : red_rr, nb, it, s.scal};
            if (b + 1 < batch && it < maxit) {
                rc = vcycle(st, s, red_rz, &np, s.scal, &g);
                if (rc) return rc;
            } else {
                k_gate<<<1, 256, 0, st>>>(g);
                pending = true;
            }
        }
        k_resid64<<<nb, RB, 0, st>>>(s.x64, s.b64, s.r64, L0, cx, cy, red_pq, red_b, nullptr, 0,
                                     s.singular ? red_x : nullptr);
        k_final2<<<1, 256, 0, st>>>(red_pq, red_b, nb, s.scal, S_TRR, S_BB, s.singular ? red_x : nullptr, S_MEAN);
        GP_HIP(hipGetLastError());
        GP_HIP(hipMemcpyAsync(s.hscal, s.scal, S_NS * sizeof(double), hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        nchk++;
        const double rr = s.hscal[S_TRR], bb = s.hscal[S_BB];
        rel = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
        if (!(rel == rel)) return gp_fail(GP_ERR_NONFINITE, "non-finite residual in the elliptic solve");
        if (rel <= tol || bb == 0.0) break;
        if (it >= maxit) { status = GP_NOT_CONVERGED; break; }
        if (s.hscal[S_DONE] != 0.0) {
            // the recursion met its target at iteration S_IT and the rest of
            // the batch did nothing: continue from that iteration (the r.z
            // slots alternate with its parity, and p is the buffer its p
            // update wrote); z = M r of the true residual (k_resid64 just
            // wrote it) takes the place of its skipped V-cycle.  The
            // recursion had drifted from the true residual, so the search
            // direction restarts from z
            const int made = (int)s.hscal[S_IT];
            if ((it - made) & 1) std::swap(p, p2);
            it = made;
            first = 1;
            replaced = false;
            k_retarget<<<1, 1, 0, st>>>(s.scal, 0.5 * tol / rel);
            k_rscale<<<nb, RB, 0, st>>>(s.r64, s.r, nullptr, n, s.scal, red_rr);
            GP_HIP(hipGetLastError());
            const Gate g{red_rr, nb, it, s.scal};
            rc = vcycle(st, s, red_rz, &np, s.scal, &g);
            if (rc) return rc;
        } else if (pending) {
            rc = vcycle(st, s, red_rz, &np, s.scal);
            if (rc) return rc;
        }
        pending = false;
    }
    const int made = s.hscal[S_DONE] != 0.0 ? (int)s.hscal[S_IT] : it;
    *iters = made;
    *checks = nchk;
    *relo = rel;
    // the next solve's first batch: this count, one more when this solve
    // needed a second batch; a solve stopped by its limit leaves it as it was
    if (status == GP_OK) s.last_it = std::max(1, made + (nchk > 1 ? 1 : 0));
    // the next solve's replacement interval from this solve's rate
    const double rr0 = s.hscal[S_TRR0];
    if (status == GP_OK && made > 0 && rr0 > 0.0 && s.hscal[S_TRR] > 0.0) {
        const double rate = std::pow(std::sqrt(s.hscal[S_TRR] / rr0), 1.0 / made);
        if (rate < 1.0) s.rep_every = std::min(16, std::max(1, (int)(std::log(MG_REPLACE_DROP) / std::log(rate))));
    }
    return status;
}

// FP32 flexible CG on S e = r with a caller-supplied (possibly nonlinear)
// preconditioner z = M(r); e in s.e, r in s.r (overwritten).  Stops at
// ||r|| <= inner_tol ||r0|| (converged = 1) or after maxit iterations.
int fcg32(hipStream_t st, gp_solver &s, double inner_tol, int maxit, int (*M)(void *, float *, const float *), void *mctx,
          int *iters, int *converged)
{
    const Lev L0 = mklev(s.lv[0]);
    const int n = s.n, nb = s.red_blocks;
    const int ge = std::min(1024, gp_cdiv(n, 256));
    double *red_a = s.red, *red_b = s.red + 1024;
    // k_update_xr skips its work once the convergence flag is up; the flag
    // belongs to rpcg's batched iterations and may still be set
    k_set_scalar<<<1, 1, 0, st>>>(s.scal, S_DONE, 0.0);
    k_zero_f<<<ge, 256, 0, st>>>(s.e, n);
    k_dot<<<nb, RB, 0, st>>>(s.r, s.r, n, red_a);
    k_final<<<1, RB, 0, st>>>(red_a, nb, s.scal, S_RR, 0);
    GP_HIP(hipGetLastError());
    GP_HIP(hipMemcpyAsync(s.hscal, s.scal, S_NS * sizeof(double), hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    const double target = inner_tol * std::sqrt(s.hscal[S_RR]);
    *iters = 0;
    *converged = s.hscal[S_RR] > 0.0 ? 0 : 1;
    if (*converged) return GP_OK;
    int rc = M(mctx, s.z, s.r);
    if (rc) return rc;
    GP_HIP(hipMemcpyAsync(s.p, s.z, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, st));
    k_dot<<<nb, RB, 0, st>>>(s.z, s.r, n, red_a);
    k_final<<<1, RB, 0, st>>>(red_a, nb, s.scal, S_RZ, 0);
    for (int it = 1; it <= maxit; it++) {
        k_spmv_dot<<<nb, RB, 0, st>>>(s.p, s.q, L0, red_a);
        k_final<<<1, RB, 0, st>>>(red_a, nb, s.scal, S_PQ, 1);                 // alpha = rz / pq
        k_update_xr<<<nb, RB, 0, st>>>(s.e, s.r, s.p, s.q, n, s.scal, red_b);   // e += a p, r -= a q
        k_final<<<1, RB, 0, st>>>(red_b, nb, s.scal, S_RR, 0);
        GP_HIP(hipGetLastError());
        GP_HIP(hipMemcpyAsync(s.hscal, s.scal, S_NS * sizeof(double), hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        *iters = it;
        const double rr = s.hscal[S_RR];
        if (!(rr == rr)) return gp_fail(GP_ERR_NONFINITE, "non-finite residual in the flexible CG");
        if (std::sqrt(rr) <= target) { *converged = 1; break; }
        if (it == maxit) break;
        if ((rc = M(mctx, s.z, s.r))) return rc;
        k_dot<<<nb, RB, 0, st>>>(s.z, s.r, n, red_a);
        k_final<<<1, RB, 0, st>>>(red_a, nb, s.scal, S_TMP, 0);
        k_dot<<<nb, RB, 0, st>>>(s.z, s.q, n, red_b);
        k_final<<<1, RB, 0, st>>>(red_b, nb, s.scal, S_BB, 0);
        k_fcg_beta<<<1, 1, 0, st>>>(s.scal);
        k_xpby<<<ge, 256, 0, st>>>(s.z, s.p, n, s.scal);
        GP_HIP(hipGetLastError());
    }
    return GP_OK;
}

}  // namespace

// nsweeps red-black Gauss-Seidel sweeps (odd: red then black; the count is
// rounded up to even so the result lands back in z) on S z = r, level 0
int gp_mg_smooth(hipStream_t st, gp_solver &s, float *z, const float *r, int nsweeps)
{
    const Lev L = mklev(s.lv[0]);
    float *a = z, *b = s.lv[0].r;
    const int ns = (nsweeps + 1) / 2 * 2;
    for (int k = 0; k < ns; k++) {
        k_rb_sweep<<<tiles(L), dim3(32, 8), 0, st>>>(a, r, b, L, k & 1, nullptr);
        std::swap(a, b);
    }
    GP_HIP(hipGetLastError());
    return GP_OK;
}

// e *= (e.r) / (e.S e): the S-norm optimal multiple of a correction e
int gp_mg_scale(hipStream_t st, gp_solver &s, float *e, const float *r)
{
    const Lev L0 = mklev(s.lv[0]);
    const int n = s.n, nb = s.red_blocks;
    // s.q2 is scratch here: the flexible CG keeps S p in s.q across this call
    k_spmv_dot<<<nb, RB, 0, st>>>(e, s.q2, L0, s.red);
    k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_PQ, 0);
    k_dot<<<nb, RB, 0, st>>>(e, r, n, s.red + 1024);
    k_final<<<1, RB, 0, st>>>(s.red + 1024, nb, s.scal, S_TMP, 0);
    k_scale_opt<<<std::min(1024, gp_cdiv(n, 256)), 256, 0, st>>>(e, n, s.scal);
    GP_HIP(hipGetLastError());
    return GP_OK;
}

// largest coarsest-level extent solved by fast diagonalisation
constexpr int FD_MAX = 256;

// eigenvectors of the symmetric matrix a (n x n, row-major; destroyed) by
// cyclic Jacobi rotations: a = q diag(lam) q', q row-major with the
// eigenvectors as columns
static void jacobi_eig(std::vector<double> &a, int n, std::vector<double> &q, std::vector<double> &lam)
{
    q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) q[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                const double v = a[(size_t)i * n + j] * a[(size_t)i * n + j];
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < n - 1; p++)
            for (int r = p + 1; r < n; r++) {
                const double apr = a[(size_t)p * n + r];
                if (std::fabs(apr) < 1e-300) continue;
                const double app = a[(size_t)p * n + p], arr = a[(size_t)r * n + r];
                const double th = 0.5 * (arr - app) / apr;
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; k++) {   // columns p, r
                    const double akp = a[(size_t)k * n + p], akr = a[(size_t)k * n + r];
                    a[(size_t)k * n + p] = c * akp - sn * akr;
                    a[(size_t)k * n + r] = sn * akp + c * akr;
                }
                for (int k = 0; k < n; k++) {   // rows p, r
                    const double apk = a[(size_t)p * n + k], ark = a[(size_t)r * n + k];
                    a[(size_t)p * n + k] = c * apk - sn * ark;
                    a[(size_t)r * n + k] = sn * apk + c * ark;
                }
                for (int k = 0; k < n; k++) {
                    const double qkp = q[(size_t)k * n + p], qkr = q[(size_t)k * n + r];
                    q[(size_t)k * n + p] = c * qkp - sn * qkr;
                    q[(size_t)k * n + r] = sn * qkp + c * qkr;
                }
            }
    }
    lam.resize(n);
    for (int i = 0; i < n; i++) lam[i] = a[(size_t)i * n + i];
}

// the 1D operator of an axis of m unknowns, K u_i = 2 u_i - u_(i-1) - u_(i+1)
// (zero beyond a Dirichlet wall, mirrored across a Neumann boundary, wrapped
// on a periodic axis), with row weights w (1/2 on Neumann boundary rows):
// W K is symmetric, and K V = V diag(lam) with V' W V = I, V row-major
static void axis_eig(int t, int m, std::vector<double> &v, std::vector<double> &lam)
{
    std::vector<double> k((size_t)m * m, 0.0), w(m, 1.0);
    for (int i = 0; i < m; i++) {
        k[(size_t)i * m + i] += 2.0;
        int l = i - 1, r = i + 1;
        if (t == GP_AX_PER) { l = (l + m) % m; r = r % m; }
        if (t == GP_AX_NEU) { if (l < 0) l = 1; if (r >= m) r = m - 2; }
        if (l >= 0 && l < m) k[(size_t)i * m + l] -= 1.0;
        if (r >= 0 && r < m) k[(size_t)i * m + r] -= 1.0;
        if (t == GP_AX_NEU && (i == 0 || i == m - 1)) w[i] = 0.5;
    }
    // symmetric form W^(1/2) K W^(-1/2) = W^(-1/2) (W K) W^(-1/2)
    std::vector<double> a((size_t)m * m), q;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) a[(size_t)i * m + j] = std::sqrt(w[i] / w[j]) * k[(size_t)i * m + j];
    jacobi_eig(a, m, q, lam);
    v.resize((size_t)m * m);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) v[(size_t)i * m + j] = q[(size_t)i * m + j] / std::sqrt(w[i]);
}

int gp_mg_build(gp_solver &s, int tx, int ty, int mx, int my, double cx, double cy)
{
    s.lv.clear();
    gp_mg_level l;
    l.mx = mx; l.my = my; l.perx = tx; l.pery = ty; l.hx2i = cx; l.hy2i = cy;
    s.lv.push_back(l);
    auto M = [](int t, int m) { return t == GP_AX_DIR ? m + 1 : (t == GP_AX_NEU ? m - 1 : m); };
    for (;;) {
        const gp_mg_level &f = s.lv.back();
        if ((int64_t)f.mx * f.my <= DENSE_MAX) break;
        if (f.mx <= FD64 && f.my <= FD64) break;   // solved exactly on MFMA
        const int Mx = M(tx, f.mx), My = M(ty, f.my);
        if ((Mx & 1) || (My & 1) || Mx < 4 || My < 4) break;
        if ((int)s.lv.size() >= 24) break;
        gp_mg_level c;
        const int Mxc = Mx / 2, Myc = My / 2;
        if ((tx == GP_AX_PER && (Mxc & 1)) || (ty == GP_AX_PER && (Myc & 1))) break;   // keep red-black valid
        c.mx = tx == GP_AX_DIR ? Mxc - 1 : (tx == GP_AX_NEU ? Mxc + 1 : Mxc);
        c.my = ty == GP_AX_DIR ? Myc - 1 : (ty == GP_AX_NEU ? Myc + 1 : Myc);
        c.perx = tx; c.pery = ty;
        c.hx2i = f.hx2i / 4.0; c.hy2i = f.hy2i / 4.0;
        s.lv.push_back(c);
    }
    // the first level small enough for the LDS bottom solver; a coarsest
    // level of at most 64 x 64 (and more than DENSE_MAX) unknowns is the
    // bottom alone, solved by k_mg_fd64
    const int nl = (int)s.lv.size();
    const gp_mg_level &cl = s.lv[nl - 1];
    const bool fd64 = cl.mx <= FD64 && cl.my <= FD64 && (int64_t)cl.mx * cl.my > DENSE_MAX;
    s.bottom = nl;
    for (int k = 0; k < nl; k++)
        if ((int64_t)s.lv[k].mx * s.lv[k].my <= BOT_MAX) { s.bottom = k; break; }
    if (fd64) s.bottom = nl - 1;
    if (s.bottom < nl && nl - s.bottom > BOT_LEVELS) return gp_fail(GP_ERR_UNSUPPORTED, "too many multigrid levels");
    if (fd64) {
        GP_HIP(hipFuncSetAttribute((const void *)k_mg_fd64, hipFuncAttributeMaxDynamicSharedMemorySize, FD64_LDS));
    } else if (s.bottom < nl) {
        // the bottom kernel's dynamic LDS (u and f of every level it holds)
        size_t bytes = 0;
        for (int k = s.bottom; k < nl; k++) bytes += 2 * sizeof(float) * (size_t)s.lv[k].mx * s.lv[k].my;
        if (bytes > 64 * 1024) {
#define MG_ATTR(X, Y) GP_HIP(hipFuncSetAttribute((const void *)k_mg_bottom<X, Y>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes))
            MG_AXES(tx, ty, MG_ATTR);
#undef MG_ATTR
        }
    }
    const int n = mx * my;
    s.n = n;
    if (!mklev(s.lv[0]).rb)
        return gp_fail(GP_ERR_UNSUPPORTED, "a periodic axis needs an even number of cells (red-black smoothing)");
    for (int k = 0; k < nl; k++) {
        const size_t m = (size_t)s.lv[k].mx * s.lv[k].my;
        if (k > 0) {
            GP_HIP(hipMalloc((void **)&s.lv[k].u, m * sizeof(float)));
            GP_HIP(hipMalloc((void **)&s.lv[k].f, m * sizeof(float)));
        }
        GP_HIP(hipMalloc((void **)&s.lv[k].r, m * sizeof(float)));
    }
    GP_HIP(hipMalloc((void **)&s.b64, n * sizeof(double)));
    GP_HIP(hipMalloc((void **)&s.x64, n * sizeof(double)));
    GP_HIP(hipMalloc((void **)&s.r64, n * sizeof(double)));
    GP_HIP(hipMemset(s.x64, 0, n * sizeof(double)));
    GP_HIP(hipMalloc((void **)&s.r, n * sizeof(float)));
    GP_HIP(hipMalloc((void **)&s.z, n * sizeof(float)));
    GP_HIP(hipMalloc((void **)&s.p, n * sizeof(float)));
    GP_HIP(hipMalloc((void **)&s.q, n * sizeof(float)));
    GP_HIP(hipMalloc((void **)&s.q2, n * sizeof(float)));
    GP_HIP(hipMalloc((void **)&s.e, n * sizeof(float)));
    s.lv[0].f = s.r;     // the finest level's rhs is the CG residual
    s.lv[0].u = s.z;     // and its output the preconditioned residual
    s.red_blocks = std::min(1024, gp_cdiv(n, RB));
    GP_HIP(hipMalloc((void **)&s.red, (RED_B(s) + 1024 + 64) * sizeof(double)));
    GP_HIP(hipMalloc((void **)&s.scal, S_NS * sizeof(double)));
    // the coarsest level, when it is small and inside the LDS bottom solver,
    // is solved exactly: the inverse of its operator (the one the smoother
    // uses: S = W (d I - neighbours), W the Neumann row weights), regularised
    // by (lambda/n) 1 1' on a singular level, which leaves zero-mean
    // right-hand sides their zero-mean solution
    {
        const gp_mg_level &c = s.lv[nl - 1];
        const int m = c.mx * c.my;
        if (s.bottom < nl && m <= DENSE_MAX) {
            std::vector<double> Sm((size_t)m * m, 0.0), Inv((size_t)m * m, 0.0);
            const Lev L = mklev(c);
            const double cxl = c.hx2i, cyl = c.hy2i, dg = 2.0 * (cxl + cyl);
            auto nbr = [&](int t, int g, int mm) { return t == GP_AX_PER ? (g + mm) % mm : (t == GP_AX_NEU ? (g < 0 ? 1 : (g >= mm ? mm - 2 : g)) : g); };
            for (int j = 0; j < c.my; j++)
                for (int i = 0; i < c.mx; i++) {
                    const int k = i + j * c.mx;
                    const double w = ((L.tx == GP_AX_NEU && (i == 0 || i == c.mx - 1)) ? 0.5 : 1.0) *
                                     ((L.ty == GP_AX_NEU && (j == 0 || j == c.my - 1)) ? 0.5 : 1.0);
                    Sm[(size_t)k * m + k] += w * dg;
                    const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
                    for (int q = 0; q < 4; q++) {
                        int ii = i + di[q], jj = j + dj[q];
                        if (di[q]) { if (L.tx == GP_AX_DIR && (ii < 0 || ii >= c.mx)) continue; ii = nbr(L.tx, ii, c.mx); }
                        if (dj[q]) { if (L.ty == GP_AX_DIR && (jj < 0 || jj >= c.my)) continue; jj = nbr(L.ty, jj, c.my); }
                        Sm[(size_t)k * m + ii + (size_t)jj * c.mx] -= w * (di[q] ? cxl : cyl);
                    }
                }
            const bool sing = (L.tx != GP_AX_DIR) && (L.ty != GP_AX_DIR);
            if (sing)
                for (size_t e = 0; e < (size_t)m * m; e++) Sm[e] += dg / m;
            // Gauss-Jordan with partial pivoting (FP64)
            for (int k = 0; k < m; k++) Inv[(size_t)k * m + k] = 1.0;
            for (int col = 0; col < m; col++) {
                int piv = col;
                for (int r = col + 1; r < m; r++)
                    if (std::fabs(Sm[(size_t)r * m + col]) > std::fabs(Sm[(size_t)piv * m + col])) piv = r;
                if (!(std::fabs(Sm[(size_t)piv * m + col]) > 0.0)) return gp_fail(GP_ERR_STATE, "singular coarsest operator");
                if (piv != col)
                    for (int q = 0; q < m; q++) {
                        std::swap(Sm[(size_t)piv * m + q], Sm[(size_t)col * m + q]);
                        std::swap(Inv[(size_t)piv * m + q], Inv[(size_t)col * m + q]);
                    }
                const double pv = 1.0 / Sm[(size_t)col * m + col];
                for (int q = 0; q < m; q++) { Sm[(size_t)col * m + q] *= pv; Inv[(size_t)col * m + q] *= pv; }
                for (int r = 0; r < m; r++) {
                    if (r == col) continue;
                    const double fct = Sm[(size_t)r * m + col];
                    if (fct == 0.0) continue;
                    for (int q = 0; q < m; q++) {
                        Sm[(size_t)r * m + q] -= fct * Sm[(size_t)col * m + q];
                        Inv[(size_t)r * m + q] -= fct * Inv[(size_t)col * m + q];
                    }
                }
            }
            if (sing) {
                // fold the projections of the singular level (zero-mean rhs in,
                // zero-mean solution out) into the matrix: Inv <- P Inv P with
                // P = I - 1 1' / m
                std::vector<double> rs(m, 0.0), cs(m, 0.0);
                double tot = 0.0;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) {
                        const double v = Inv[(size_t)i * m + j];
                        rs[i] += v; cs[j] += v; tot += v;
                    }
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++)
                        Inv[(size_t)i * m + j] += (tot / m - rs[i] - cs[j]) / m;
            }
            std::vector<float> cm((size_t)m * m);   // column-major: cm[j*m + i] = Inv[i][j]
            for (int i = 0; i < m; i++)
                for (int j = 0; j < m; j++) cm[(size_t)j * m + i] = (float)Inv[(size_t)i * m + j];
            GP_HIP(hipMalloc((void **)&s.dinv, (size_t)m * m * sizeof(float)));
            GP_HIP(hipMemcpy(s.dinv, cm.data(), (size_t)m * m * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    // a coarsest level in the bottom but too large for the dense inverse (an
    // odd interval count stopped the coarsening) is solved exactly by fast
    // diagonalisation instead of Gauss-Seidel sweeps
    {
        const gp_mg_level &c = s.lv[nl - 1];
        if (s.bottom < nl && !fd64 && !s.dinv && c.mx <= FD_MAX && c.my <= FD_MAX) {
            std::vector<double> vx, lx, vy, ly;
            axis_eig(tx, c.mx, vx, lx);
            axis_eig(ty, c.my, vy, ly);
            const double cxl = c.hx2i, cyl = c.hy2i;
            std::vector<float> fvx(vx.begin(), vx.end()), fvy(vy.begin(), vy.end());
            std::vector<float> finv((size_t)c.mx * c.my);
            for (int j = 0; j < c.my; j++)
                for (int i = 0; i < c.mx; i++) {
                    const double ev = cxl * lx[i] + cyl * ly[j];
                    finv[i + (size_t)j * c.mx] = std::fabs(ev) > 1e-9 * (cxl + cyl) ? (float)(1.0 / ev) : 0.0f;
                }
            GP_HIP(hipMalloc((void **)&s.fd_vx, fvx.size() * sizeof(float)));
            GP_HIP(hipMalloc((void **)&s.fd_vy, fvy.size() * sizeof(float)));
            GP_HIP(hipMalloc((void **)&s.fd_inv, finv.size() * sizeof(float)));
            GP_HIP(hipMemcpy(s.fd_vx, fvx.data(), fvx.size() * sizeof(float), hipMemcpyHostToDevice));
            GP_HIP(hipMemcpy(s.fd_vy, fvy.data(), fvy.size() * sizeof(float), hipMemcpyHostToDevice));
            GP_HIP(hipMemcpy(s.fd_inv, finv.data(), finv.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    }

    if (fd64) {
        // the MFMA fast diagonalisation: Vx, Vy and 1 / (cx lx + cy ly), each
        // zero-padded to 64 x 64 (a null mode has 0)
        std::vector<double> vx, lx, vy, ly;
        axis_eig(tx, cl.mx, vx, lx);
        axis_eig(ty, cl.my, vy, ly);
        std::vector<float> h(3 * FD64 * FD64, 0.0f);
        for (int i = 0; i < cl.mx; i++)
            for (int k = 0; k < cl.mx; k++) h[i * FD64 + k] = (float)vx[(size_t)i * cl.mx + k];
        for (int j = 0; j < cl.my; j++)
            for (int k = 0; k < cl.my; k++) h[FD64 * FD64 + j * FD64 + k] = (float)vy[(size_t)j * cl.my + k];
        for (int j = 0; j < cl.my; j++)
            for (int i = 0; i < cl.mx; i++) {
                const double ev = cl.hx2i * lx[i] + cl.hy2i * ly[j];
                h[2 * FD64 * FD64 + j * FD64 + i] = std::fabs(ev) > 1e-9 * (cl.hx2i + cl.hy2i) ? (float)(1.0 / ev) : 0.0f;
            }
        GP_HIP(hipMalloc((void **)&s.fd64, h.size() * sizeof(float)));
        GP_HIP(hipMemcpy(s.fd64, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }

    GP_HIP(hipHostMalloc((void **)&s.hscal, S_NS * sizeof(double), hipHostMallocDefault));
    return GP_OK;
}

void gp_mg_free(gp_solver &s)
{
    for (size_t k = 0; k < s.lv.size(); k++) {
        if (k > 0) { (void)hipFree(s.lv[k].u); (void)hipFree(s.lv[k].f); }
        (void)hipFree(s.lv[k].r);
    }
    s.lv.clear();
    void *ps[] = {s.b64, s.x64, s.r64, s.rhs_bc, s.r, s.z, s.p, s.q, s.q2, s.e, s.red, s.scal, s.dinv,
                  s.fd_vx, s.fd_vy, s.fd_inv, s.fd64};
    for (void *p : ps) if (p) (void)hipFree(p);
    if (s.hscal) (void)hipHostFree(s.hscal);
    s = gp_solver();
}

// Solve S x = b64 (already in s.b64) starting from s.x64 by MG-PCG (rpcg).
// Hooks (may be null): hk->warm supplies the first correction (a learned
// warm start) before the MG-PCG; hk->precond first runs FP64 refinement
// rounds whose corrections come from a flexible CG with that preconditioner,
// and hands over to the MG-PCG once its iterations exceed those of the last
// plain MG-PCG solve on this hierarchy.
int gp_solve_spd(hipStream_t st, gp_solver &s, double tol, int maxiter, int singular, gp_solve_stats *stats,
                 const gp_hooks *hk)
{
    s.singular = singular;
    const Lev L0 = mklev(s.lv[0]);
    const int n = s.n, nb = s.red_blocks;
    const int ge = std::min(1024, gp_cdiv(n, 256));
    const double cx = s.lv[0].hx2i, cy = s.lv[0].hy2i;
    auto t0 = std::chrono::steady_clock::now();
    // a singular (all-Neumann/periodic) system: b is projected to zero mean,
    // and so is the solution.  The FP32 corrections carry a constant
    // component that S does not see; left in x over many refinements it would
    // grow until its FP64 rounding (ulp of the constant times the stencil)
    // sets a floor on the residual, so from the network stage's third
    // refinement on every iterate is projected too
    auto sub_mean = [&](double *a) {
        k_sum64<<<nb, RB, 0, st>>>(a, n, s.red);
        k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_MEAN, 0);
        k_sub_mean64<<<ge, 256, 0, st>>>(a, n, s.scal, S_MEAN);
    };
    int total_it = 0, outer = 0, nn_it = 0, fell_back = 0;
    const bool warm = hk && hk->warm, flex = hk && hk->precond;
    // (without the network stage, the MG-PCG's first kernel centres b)
    if (singular && (warm || flex)) {
        sub_mean(s.b64);
        GP_HIP(hipGetLastError());
    }
    const int cap = (hk && hk->cap > 0) ? hk->cap : (s.mg_ref_it > 0 ? s.mg_ref_it : 100);
    double rel = 0.0;
    int status = GP_OK;
    bool finished = false;
    // a warm start alone: one network correction of the first residual, with
    // no host check before it (the embedding and the optimal scaling both map
    // a zero residual to a zero correction, and the MG-PCG that follows makes
    // the first convergence test); the residual is used unscaled
    bool warm_pending = warm;
    if (warm && !flex) {
        k_resid64<<<nb, RB, 0, st>>>(s.x64, s.b64, s.r64, L0, cx, cy, s.red, s.red + nb);
        k_inner_start<<<nb, RB, 0, st>>>(s.r64, s.r, s.e, s.p, n, 1.0, s.red + 1024);
        GP_HIP(hipGetLastError());
        int rc = hk->warm(hk->ctx, s.e, s.r);
        if (rc) return rc;
        k_axpy64<<<ge, 256, 0, st>>>(s.x64, s.e, n, 1.0);
        GP_HIP(hipGetLastError());
        warm_pending = false;
        outer++;
    }
    // the network stage: FP64 refinement rounds
    for (; warm_pending || (flex && !fell_back);) {
        k_resid64<<<nb, RB, 0, st>>>(s.x64, s.b64, s.r64, L0, cx, cy, s.red, s.red + nb);
        k_final2<<<1, 256, 0, st>>>(s.red, s.red + nb, nb, s.scal, S_RR, S_BB);
        GP_HIP(hipGetLastError());
        GP_HIP(hipMemcpyAsync(s.hscal, s.scal, S_NS * sizeof(double), hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        const double rr = s.hscal[S_RR], bb = s.hscal[S_BB];
        rel = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
        if (!(rel == rel)) return gp_fail(GP_ERR_NONFINITE, "non-finite residual in the FP64 refinement");
        if (rel <= tol || bb == 0.0) { finished = true; break; }
        if (total_it >= maxiter || outer >= 40) { status = GP_NOT_CONVERGED; finished = true; break; }
        const double rms = std::sqrt(rr / n);
        const double sc = rms > 0 ? 1.0 / rms : 1.0;
        // FP32 residual, e = p = 0
        k_inner_start<<<nb, RB, 0, st>>>(s.r64, s.r, s.e, s.p, n, sc, s.red + 1024);
        GP_HIP(hipGetLastError());
        if (warm_pending) {
            int rc = hk->warm(hk->ctx, s.e, s.r);
            if (rc) return rc;
            warm_pending = false;
        } else {
            // ask the FP32 solve for just the reduction still needed (FP32
            // cannot usefully go below ~1e-6 of its own starting residual)
            const double inner = std::max(MG_INNER_FLOOR, std::min(1e-2, 0.3 * tol / rel));
            int it = 0, conv = 0;
            const int lim = std::max(1, std::min(cap - nn_it, maxiter - total_it));
            int rc = fcg32(st, s, inner, lim, hk->precond, hk->ctx, &it, &conv);
            if (rc) return rc;
            nn_it += it;
            total_it += it;
            if (!conv && nn_it >= cap) fell_back = 1;
        }
        k_axpy64<<<ge, 256, 0, st>>>(s.x64, s.e, n, 1.0 / sc);
        if (singular && outer >= 2) sub_mean(s.x64);
        GP_HIP(hipGetLastError());
        outer++;
    }
    if (!finished) {
        int it = 0, chk = 0;
        const int rc = rpcg(st, s, tol, std::max(1, maxiter - total_it), singular && !warm && !flex, &it, &chk, &rel);
        if (rc < 0) return rc;
        status = rc;
        total_it += it;
        outer += chk;
        // the MG-PCG's last check left sum(x) in S_MEAN
        if (singular) k_sub_mean64<<<ge, 256, 0, st>>>(s.x64, n, s.scal, S_MEAN);
    } else if (singular) {
        sub_mean(s.x64);
    }
    GP_HIP(hipGetLastError());
    // no synchronisation here: the last residual check has waited for the
    // iterations, and the mean projection runs on ahead of whatever the
    // caller enqueues next (the seconds cover the solve up to that check)
    auto t1 = std::chrono::steady_clock::now();
    if (!warm && !flex && status == GP_OK) s.mg_ref_it = std::max(1, total_it);
    if (stats) {
        stats->iterations = total_it;
        stats->refinements = outer;
        stats->relres = rel;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        stats->used_nn = (warm || flex) ? 1 : 0;
        stats->fell_back = fell_back;
    }
    return status;
}

// a -= plain mean of a (n = the solver's unknown count)
int gp_mg_sub_mean(hipStream_t st, gp_solver &s, double *a)
{
    const int n = s.n, nb = s.red_blocks;
    k_sum64<<<nb, RB, 0, st>>>(a, n, s.red);
    k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_MEAN, 0);
    k_sub_mean64<<<std::min(1024, gp_cdiv(n, 256)), 256, 0, st>>>(a, n, s.scal, S_MEAN);
    GP_HIP(hipGetLastError());
    return GP_OK;
}

// apply one V-cycle to an FP32 vector (used by the learned preconditioner's
// flexible CG): out = M in
int gp_mg_apply(hipStream_t st, gp_solver &s, const float *in, float *out)
{
    const int n = s.n;
    GP_HIP(hipMemcpyAsync(s.r, in, n * sizeof(float), hipMemcpyDeviceToDevice, st));
    int np = 0;
    int rc = vcycle(st, s, s.red + 2048, &np, nullptr);
    if (rc) return rc;
    GP_HIP(hipMemcpyAsync(out, s.z, n * sizeof(float), hipMemcpyDeviceToDevice, st));
    return GP_OK;
}
