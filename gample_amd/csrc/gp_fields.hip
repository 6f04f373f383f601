// gp_fields.hip -- grid kernels and the GPU elliptic solver.
//
// Replaces EMField's CPU path (EMField.mm:745-1361): ghost-charge fold,
// Poisson rhs, the AMGCL CG solve, phi reconstruction, E = -grad phi, the
// energy diagnostics and the Darwin chain.  All of it stays on the device.
//
// Solver.  The SPD form S x = b is solved with S = D (-A): A is the
// reference's 5-point operator (EMField.mm:236-405, 483-623) and D halves the
// rows of Neumann boundary nodes (the reference's ghost-mirror rows carry
// 2*Cx, which makes A unsymmetric; D restores symmetry without changing the
// solution).  Outer loop in FP64: r = b - S x, stop at ||r||/||b|| <= tol;
// inner correction S e = r solved in FP32 by CG preconditioned with one
// symmetric geometric V-cycle (red-black Gauss-Seidel, full weighting
// R = P^T/4, bilinear P, rediscretised coarse operators), x += e.
// Periodic or all-Neumann problems are singular: b is projected to zero mean
// and the returned solution has zero mean.
#include "gp_internal.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace {

constexpr double PI_D = 3.141592653589793;
constexpr double C_LIGHT = 2.99792458e10;   // Constant.h:5

// axis types of an MG level
enum { AX_PER = 0, AX_DIR = 1, AX_NEU = 2 };

struct Lev {
    int mx, my;           // unknowns
    int tx, ty;           // axis types
    int Mx, My;           // intervals (node count - 1, or node count if periodic)
    float cx, cy;         // 1/h^2
    int rb;               // red-black safe (no odd periodic axis)
};

__device__ __forceinline__ float wgt(int t, int i, int m) { return (t == AX_NEU && (i == 0 || i == m - 1)) ? 0.5f : 1.0f; }

// neighbour value along an axis, index i +- 1 in unknown space
__device__ __forceinline__ float nb_lo(const float *u, int t, int i, int m, int stride, int base)
{
    if (i > 0) return u[base - stride];
    if (t == AX_PER) return u[base + (m - 1) * stride];
    if (t == AX_NEU) return u[base + stride];        // ghost mirror
    return 0.0f;                                     // Dirichlet wall
}
__device__ __forceinline__ float nb_hi(const float *u, int t, int i, int m, int stride, int base)
{
    if (i < m - 1) return u[base + stride];
    if (t == AX_PER) return u[base - (m - 1) * stride];
    if (t == AX_NEU) return u[base - stride];
    return 0.0f;
}

// S u at unknown (i,j)
__device__ __forceinline__ float applyS(const float *u, const Lev &L, int i, int j)
{
    const int k = i + j * L.mx;
    const float c = u[k];
    const float ax = L.cx * (2.0f * c - nb_lo(u, L.tx, i, L.mx, 1, k) - nb_hi(u, L.tx, i, L.mx, 1, k));
    const float ay = L.cy * (2.0f * c - nb_lo(u, L.ty, j, L.my, L.mx, k) - nb_hi(u, L.ty, j, L.my, L.mx, k));
    return wgt(L.tx, i, L.mx) * wgt(L.ty, j, L.my) * (ax + ay);
}

// red-black Gauss-Seidel half sweep (color = (i+j)&1); weighted Jacobi
// (omega 0.8, into out) when an odd periodic axis breaks the colouring
__global__ void k_rbgs(float *u, const float *f, Lev L, int color)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= L.mx || j >= L.my || ((i + j) & 1) != color) return;
    const int k = i + j * L.mx;
    const float w = wgt(L.tx, i, L.mx) * wgt(L.ty, j, L.my);
    const float s = L.cx * (nb_lo(u, L.tx, i, L.mx, 1, k) + nb_hi(u, L.tx, i, L.mx, 1, k)) +
                    L.cy * (nb_lo(u, L.ty, j, L.my, L.mx, k) + nb_hi(u, L.ty, j, L.my, L.mx, k));
    u[k] = (f[k] / w + s) / (2.0f * (L.cx + L.cy));
}
__global__ void k_jacobi(const float *u, const float *f, float *out, Lev L)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= L.mx || j >= L.my) return;
    const int k = i + j * L.mx;
    const float d = wgt(L.tx, i, L.mx) * wgt(L.ty, j, L.my) * 2.0f * (L.cx + L.cy);
    out[k] = u[k] + 0.8f * (f[k] - applyS(u, L, i, j)) / d;
}

// residual r = f - S u, restricted with R = P^T/4 into the coarse rhs
__global__ void k_resid(const float *u, const float *f, float *r, Lev L)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= L.mx || j >= L.my) return;
    r[i + j * L.mx] = f[i + j * L.mx] - applyS(u, L, i, j);
}
// fine unknown index of fine node n on an axis; -1 if n is a Dirichlet wall
__device__ __forceinline__ int f_unk(int t, int n, int M)
{
    if (t == AX_DIR) return (n <= 0 || n >= M) ? -1 : n - 1;
    if (t == AX_PER) return ((n % M) + M) % M;
    return n;   // Neumann: n in [0, M]
}
__device__ __forceinline__ int c_node(int t, int I) { return t == AX_DIR ? I + 1 : I; }

__global__ void k_restrict(const float *rf, Lev F, float *fc, Lev C)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= C.mx || J >= C.my) return;
    const int nx = 2 * c_node(C.tx, I), ny = 2 * c_node(C.ty, J);
    float acc = 0.0f;
    for (int b = -1; b <= 1; b++) {
        const int fy = ny + b;
        if (F.ty == AX_NEU && (fy < 0 || fy > F.My)) continue;
        const int jj = f_unk(F.ty, fy, F.My);
        if (jj < 0) continue;
        const float wy = b == 0 ? 0.5f : 0.25f;
        for (int a = -1; a <= 1; a++) {
            const int fx = nx + a;
            if (F.tx == AX_NEU && (fx < 0 || fx > F.Mx)) continue;
            const int ii = f_unk(F.tx, fx, F.Mx);
            if (ii < 0) continue;
            const float wx = a == 0 ? 0.5f : 0.25f;
            acc += wx * wy * rf[ii + jj * F.mx];
        }
    }
    fc[I + J * C.mx] = acc;
}
// coarse value at coarse node N (walls of a Dirichlet axis are 0)
__device__ __forceinline__ int c_unk(int t, int N, int M)
{
    if (t == AX_DIR) return (N <= 0 || N >= M) ? -1 : N - 1;
    if (t == AX_PER) return ((N % M) + M) % M;
    return N;
}
__global__ void k_prolong_add(float *uf, Lev F, const float *uc, Lev C)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= F.mx || j >= F.my) return;
    const int nx = F.tx == AX_DIR ? i + 1 : i, ny = F.ty == AX_DIR ? j + 1 : j;
    int ax[2], ay[2];
    float wx[2], wy[2];
    int cxn = 1, cyn = 1;
    if (nx & 1) { ax[0] = (nx - 1) / 2; ax[1] = (nx + 1) / 2; wx[0] = wx[1] = 0.5f; cxn = 2; }
    else { ax[0] = nx / 2; wx[0] = 1.0f; }
    if (ny & 1) { ay[0] = (ny - 1) / 2; ay[1] = (ny + 1) / 2; wy[0] = wy[1] = 0.5f; cyn = 2; }
    else { ay[0] = ny / 2; wy[0] = 1.0f; }
    float acc = 0.0f;
    for (int b = 0; b < cyn; b++) {
        const int J = c_unk(C.ty, ay[b], C.My);
        if (J < 0) continue;
        for (int a = 0; a < cxn; a++) {
            const int I = c_unk(C.tx, ax[a], C.Mx);
            if (I < 0) continue;
            acc += wx[a] * wy[b] * uc[I + J * C.mx];
        }
    }
    uf[i + j * F.mx] += acc;
}

// coarsest level: one workgroup, grid resident in LDS, fixed symmetric sweeps
constexpr int COARSE_MAX = 4096;
__global__ __launch_bounds__(1024) void k_coarse(float *u, const float *f, Lev L, int sweeps, int singular)
{
    __shared__ float su[COARSE_MAX], sf[COARSE_MAX], st[COARSE_MAX];
    __shared__ float red[1024];
    const int n = L.mx * L.my, tid = threadIdx.x;
    for (int k = tid; k < n; k += 1024) { sf[k] = f[k]; su[k] = 0.0f; }
    __syncthreads();
    if (singular) {
        float s = 0.0f;
        for (int k = tid; k < n; k += 1024) s += sf[k];
        red[tid] = s;
        __syncthreads();
        for (int w = 512; w > 0; w >>= 1) { if (tid < w) red[tid] += red[tid + w]; __syncthreads(); }
        const float m = red[0] / (float)n;
        __syncthreads();
        for (int k = tid; k < n; k += 1024) sf[k] -= m;
        __syncthreads();
    }
    for (int sw = 0; sw < 2 * sweeps; sw++) {
        // forward half: colours 0,1; backward half: 1,0 (symmetric)
        const bool back = sw >= sweeps;
        for (int cc = 0; cc < 2; cc++) {
            const int color = back ? 1 - cc : cc;
            if (L.rb) {
                for (int k = tid; k < n; k += 1024) {
                    const int i = k % L.mx, j = k / L.mx;
                    if (((i + j) & 1) != color) continue;
                    const float w = wgt(L.tx, i, L.mx) * wgt(L.ty, j, L.my);
                    const float s = L.cx * (nb_lo(su, L.tx, i, L.mx, 1, k) + nb_hi(su, L.tx, i, L.mx, 1, k)) +
                                    L.cy * (nb_lo(su, L.ty, j, L.my, L.mx, k) + nb_hi(su, L.ty, j, L.my, L.mx, k));
                    su[k] = (sf[k] / w + s) / (2.0f * (L.cx + L.cy));
                }
            } else if (cc == 0) {
                for (int k = tid; k < n; k += 1024) {
                    const int i = k % L.mx, j = k / L.mx;
                    const float d = wgt(L.tx, i, L.mx) * wgt(L.ty, j, L.my) * 2.0f * (L.cx + L.cy);
                    st[k] = su[k] + 0.8f * (sf[k] - applyS(su, L, i, j)) / d;
                }
                __syncthreads();
                for (int k = tid; k < n; k += 1024) su[k] = st[k];
            }
            __syncthreads();
        }
    }
    if (singular) {
        float s = 0.0f;
        for (int k = tid; k < n; k += 1024) s += su[k];
        red[tid] = s;
        __syncthreads();
        for (int w = 512; w > 0; w >>= 1) { if (tid < w) red[tid] += red[tid + w]; __syncthreads(); }
        const float m = red[0] / (float)n;
        __syncthreads();
        for (int k = tid; k < n; k += 1024) su[k] -= m;
        __syncthreads();
    }
    for (int k = tid; k < n; k += 1024) u[k] = su[k];
}

// ---------------- FP32 CG pieces (dot products accumulated in FP64) -------
constexpr int RB = 256;   // reduction workgroup size
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
// q = S p, partial (p.q)
__global__ __launch_bounds__(RB) void k_spmv_dot(const float *p, float *q, Lev L, double *part)
{
    __shared__ double sh[RB / 64];
    const int n = L.mx * L.my;
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) {
        const int i = k % L.mx, j = k / L.mx;
        const float v = applyS(p, L, i, j);
        q[k] = v;
        acc += (double)p[k] * (double)v;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// partial dot a.b
__global__ __launch_bounds__(RB) void k_dot(const float *a, const float *b, int n, double *part)
{
    __shared__ double sh[RB / 64];
    double acc = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) acc += (double)a[k] * (double)b[k];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// final reduction of partials into scal[slot]; op: 0 store, 1 alpha = scal[RZ]/sum, 2 beta = sum/scal[RZ]
enum { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_TMP, S_BB, S_MEAN, S_NS };
__global__ void k_final(const double *part, int nb, double *scal, int slot, int op)
{
    __shared__ double sh[RB];
    double acc = 0.0;
    for (int k = threadIdx.x; k < nb; k += RB) acc += part[k];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double s = sh[0];
        scal[slot] = s;
        if (op == 1) scal[S_ALPHA] = s != 0.0 ? scal[S_RZ] / s : 0.0;
        if (op == 2) { scal[S_BETA] = scal[S_RZ] != 0.0 ? s / scal[S_RZ] : 0.0; scal[S_RZ] = s; }
    }
}
// x += alpha p; r -= alpha q; partial r.r
__global__ __launch_bounds__(RB) void k_update_xr(float *x, float *r, const float *p, const float *q, int n,
                                                 const double *scal, double *part)
{
    __shared__ double sh[RB / 64];
    const float alpha = (float)scal[S_ALPHA];
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
__global__ void k_update_p(float *p, const float *z, int n, const double *scal)
{
    const float beta = (float)scal[S_BETA];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        p[k] = __builtin_fmaf(beta, p[k], z[k]);
}
__global__ void k_copyf(float *d, const float *s, int n)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) d[k] = s[k];
}
__global__ void k_add_scalar_f(float *d, int n, const double *scal, int slot, double mul)
{
    const float v = (float)(scal[slot] * mul);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) d[k] += v;
}

// ---------------- FP64 refinement pieces ----------------
// S x in FP64 at the fine level, r = b - Sx, partial r.r and b.b
__global__ __launch_bounds__(RB) void k_resid64(const double *x, const double *b, double *r, Lev L,
                                               double cx, double cy, double *part_rr, double *part_bb)
{
    __shared__ double sh[RB / 64];
    const int n = L.mx * L.my;
    double arr = 0.0, abb = 0.0;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += gridDim.x * RB) {
        const int i = k % L.mx, j = k / L.mx;
        const double c = x[k];
        double xl, xr, yl, yr;
        if (i > 0) xl = x[k - 1]; else xl = L.tx == AX_PER ? x[k + L.mx - 1] : (L.tx == AX_NEU ? x[k + 1] : 0.0);
        if (i < L.mx - 1) xr = x[k + 1]; else xr = L.tx == AX_PER ? x[k - (L.mx - 1)] : (L.tx == AX_NEU ? x[k - 1] : 0.0);
        if (j > 0) yl = x[k - L.mx]; else yl = L.ty == AX_PER ? x[k + (L.my - 1) * L.mx] : (L.ty == AX_NEU ? x[k + L.mx] : 0.0);
        if (j < L.my - 1) yr = x[k + L.mx]; else yr = L.ty == AX_PER ? x[k - (L.my - 1) * L.mx] : (L.ty == AX_NEU ? x[k - L.mx] : 0.0);
        const double w = (double)wgt(L.tx, i, L.mx) * (double)wgt(L.ty, j, L.my);
        const double sx = w * (cx * (2.0 * c - xl - xr) + cy * (2.0 * c - yl - yr));
        const double rv = b[k] - sx;
        r[k] = rv;
        arr += rv * rv;
        abb += b[k] * b[k];
    }
    double s = block_sum(arr, sh);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = s;
    __syncthreads();
    s = block_sum(abb, sh);
    if (threadIdx.x == 0) part_bb[blockIdx.x] = s;
}
__global__ void k_to_f32(const double *r, float *rf, int n, double scale)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) rf[k] = (float)(r[k] * scale);
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

// ---------------- charge / current / rhs / reconstruction ----------------
struct AccSet {
    const int64_t *a[GP_MAX_SPECIES];
    double f[GP_MAX_SPECIES];
    int ns;
};
__global__ void k_scale_sum(AccSet s, double *out, int64_t n)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    double v = 0.0;
    for (int q = 0; q < s.ns; q++) v += (double)s.a[q][k] * s.f[q];
    out[k] = v;
}
// fold (EMField.mm:749-850), one thread per row (x) or column (y)
__global__ void k_fold_x(double *rho, int ngb, int ngx, int nx, int ny, int bmin, int bmax, int periodic)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > ny) return;
    double *r = rho + (int64_t)j * (nx + 1);
    if (!periodic) {
        if (bmin == GP_FBC_DIRICHLET) for (int i = 0; i <= ngb; i++) r[ngb + 1] += r[i];
        else if (bmin == GP_FBC_NEUMANN) for (int i = 0; i < ngb; i++) r[ngb] += r[i];
        if (bmax == GP_FBC_DIRICHLET) for (int i = 0; i <= ngb; i++) r[ngb + ngx - 1] += r[ngb + ngx + i];
        else if (bmax == GP_FBC_NEUMANN) for (int i = 0; i < ngb; i++) r[ngb + ngx] += r[ngb + ngx + i];
    } else {
        for (int i = 0; i < ngb; i++) r[ngx + i] += r[i];
        for (int i = 0; i <= ngb; i++) r[ngb + i] += r[ngb + ngx + i];
    }
}
__global__ void k_fold_y(double *rho, int ngb, int ngy, int nx, int bmin, int bmax, int periodic)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nx) return;
    const int P = nx + 1;
    double *r = rho + i;
    if (!periodic) {
        if (bmin == GP_FBC_DIRICHLET) for (int j = 0; j <= ngb; j++) r[(ngb + 1) * P] += r[j * P];
        else if (bmin == GP_FBC_NEUMANN) for (int j = 0; j < ngb; j++) r[ngb * P] += r[j * P];
        if (bmax == GP_FBC_DIRICHLET) for (int j = 0; j <= ngb; j++) r[(ngb + ngy - 1) * P] += r[(ngb + ngy + j) * P];
        else if (bmax == GP_FBC_NEUMANN) for (int j = 0; j < ngb; j++) r[(ngb + ngy) * P] += r[(ngb + ngy + j) * P];
    } else {
        for (int j = 0; j < ngb; j++) r[(ngy + j) * P] += r[j * P];
        for (int j = 0; j <= ngb; j++) r[(ngb + j) * P] += r[(ngb + ngy + j) * P];
    }
}
// b = -(rhs_bc - 4 pi rho) (SPD form; EMField.mm:853-861)
__global__ void k_rhs(const double *rhs_bc, const double *rho, double *b, int mx, int my, int ikmin, int jkmin, int P)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= mx * my) return;
    const int i = k % mx, j = k / mx;
    b[k] = -(rhs_bc[k] - 4 * PI_D * rho[(i + ikmin) + (int64_t)(j + jkmin) * P]);
}
// phi (nx+3)(ny+3) from the unknowns (EMField.mm:868-1061; periodic x index
// (i-(ngb+1)+ngx)%ngx as the reference's y branch)
struct RecArgs {
    int ngb, ngx, ngy, nx, ny, nkx, nky;
    int b[4];
    double v[4];
};
__global__ void k_reconstruct(const double *sol, double *phi, RecArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > a.nx + 2 || j > a.ny + 2) return;
    const int ngb = a.ngb, K = a.nkx + 1, nkx = a.nkx, nky = a.nky;
    int isB = 0, isT = 0, isL = 0, isR = 0, jn, in;
    if (a.b[2] == GP_FBC_DIRICHLET) { isB = (j <= ngb + 1); jn = j - (ngb + 1) - 1; }
    else jn = (j - (ngb + 1) + a.ngy) % a.ngy;
    if (a.b[3] == GP_FBC_DIRICHLET) isT = (j >= ngb + 1 + a.ngy);
    if (a.b[0] == GP_FBC_DIRICHLET) { isL = (i <= ngb + 1); in = i - (ngb + 1) - 1; }
    else in = (i - (ngb + 1) + a.ngx) % a.ngx;
    if (a.b[1] == GP_FBC_DIRICHLET) isR = (i >= ngb + 1 + a.ngx);
    const int di = i - (ngb + 1), dj = j - (ngb + 1), diR = i - (ngb + 1 + a.ngx), djT = j - (ngb + 1 + a.ngy);
    const double pXmin = a.v[0], pXmax = a.v[1], pYmin = a.v[2], pYmax = a.v[3];
    double v = 0.0;
    if (!isL && !isR && !isB && !isT) v = sol[in + (int64_t)jn * K];
    else if (isL && !isR && !isB && !isT) v = pXmin + di * (sol[(int64_t)jn * K] - pXmin);
    else if (!isL && isR && !isB && !isT) v = pXmax + diR * (pXmax - sol[nkx + (int64_t)jn * K]);
    else if (!isL && !isR && isB && !isT) v = pYmin + dj * (sol[in] - pYmin);
    else if (!isL && !isR && !isB && isT) v = pYmax + djT * (pYmax - sol[in + (int64_t)nky * K]);
    else if (isL && isB) v = pXmin + di * (sol[1] - sol[0]) + dj * (sol[K] - sol[0]);
    else if (isR && isB) v = pXmax + diR * (sol[nkx] - sol[nkx - 1]) + dj * (sol[nkx + K] - sol[nkx]);
    else if (isL && isT) v = pXmin + di * (sol[1 + (int64_t)nky * K] - sol[(int64_t)nky * K]) +
                             djT * (sol[(int64_t)nky * K] - sol[(int64_t)(nky - 1) * K]);
    else if (isR && isT) v = pXmax + diR * (sol[nkx + (int64_t)nky * K] - sol[nkx - 1 + (int64_t)nky * K]) +
                             djT * (sol[nkx + (int64_t)nky * K] - sol[nkx + (int64_t)(nky - 1) * K]);
    phi[i + (int64_t)j * (a.nx + 3)] = v;
}
// E = -grad phi (EMField.mm:1080-1094)
__global__ void k_gradient(const double *phi, float *Ex, float *Ey, int nx, int ny, double dx, double dy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    const int P = nx + 3;
    if (i <= nx + 1 && j <= ny)
        Ex[i + j * (nx + 2)] = (float)(-(phi[i + 1 + (j + 1) * P] - phi[i + (j + 1) * P]) / dx);
    if (i <= nx && j <= ny + 1)
        Ey[i + j * (nx + 1)] = (float)(-(phi[i + 1 + (j + 1) * P] - phi[i + 1 + j * P]) / dy);
}
// energy partials: 0 sum E^2 (whole), 1 sum E.J, 2 CGS sum over unique points
struct EnArgs {
    int ngb, ngx, ngy, nx, ny, perx, pery;
};
__global__ __launch_bounds__(RB) void k_energy(const float *Ex, const float *Ey, const double *jx, const double *jy,
                                              EnArgs a, double *part, int nb)
{
    __shared__ double sh[RB / 64];
    double e = 0.0, jl = 0.0, u = 0.0;
    const int nEx = (a.nx + 2) * (a.ny + 1), nEy = (a.nx + 1) * (a.ny + 2);
    for (int k = blockIdx.x * RB + threadIdx.x; k < nEx + nEy; k += gridDim.x * RB) {
        if (k < nEx) {
            const double v = Ex[k];
            e += v * v;
            if (jx) jl += v * jx[k];
            const int i = k % (a.nx + 2), j = k / (a.nx + 2);
            if (i >= a.ngb + 1 && i <= a.ngb + a.ngx && j >= a.ngb && j <= a.ngb + a.ngy - (a.pery ? 1 : 0)) {
                const double wy = (!a.pery && (j == a.ngb || j == a.ngb + a.ngy)) ? 0.5 : 1.0;
                u += wy * v * v;
            }
        } else {
            const int q = k - nEx;
            const double v = Ey[q];
            e += v * v;
            if (jy) jl += v * jy[q];
            const int i = q % (a.nx + 1), j = q / (a.nx + 1);
            if (j >= a.ngb + 1 && j <= a.ngb + a.ngy && i >= a.ngb && i <= a.ngb + a.ngx - (a.perx ? 1 : 0)) {
                const double wx = (!a.perx && (i == a.ngb || i == a.ngb + a.ngx)) ? 0.5 : 1.0;
                u += wx * v * v;
            }
        }
    }
    double s = block_sum(e, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
    __syncthreads();
    s = block_sum(jl, sh);
    if (threadIdx.x == 0) part[nb + blockIdx.x] = s;
    __syncthreads();
    s = block_sum(u, sh);
    if (threadIdx.x == 0) part[2 * nb + blockIdx.x] = s;
}
__global__ void k_fill_f(float *d, int64_t n, float v)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) d[k] = v;
}

// Darwin rhs (EMField.mm:1147-1192, 1203-1275), result in SPD form
// b = -D rhs (rows of Neumann boundary nodes halved), plain zero mean after
struct DarArgs {
    int ngb, nx, nkx, nky, perx, pery;
    double dx, dy;
};
__global__ void k_darwin_chi_rhs(const double *jx, const double *jy, double *b, DarArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int mx = a.nkx + 1, my = a.nky + 1;
    if (k >= mx * my) return;
    const int i = k % mx, j = k / mx, ngb = a.ngb, nx = a.nx;
    double r = 0.0;
    if (!a.perx) {
        if (i == 0) r += -2 * jx[(i + ngb) + (j + ngb) * (nx + 2)] / a.dx;
        else if (i == a.nkx) r += 2 * jx[(i + ngb) + (j + ngb) * (nx + 2)] / a.dx;
    }
    if (!a.pery) {
        if (j == 0) r += -2 * jy[(i + ngb) + (j + ngb) * (nx + 1)] / a.dy;
        else if (j == a.nky) r += 2 * jy[(i + ngb) + (j + ngb) * (nx + 1)] / a.dy;
    }
    const double djxdx = (jx[(i + ngb + 1) + (j + ngb) * (nx + 2)] - jx[(i + ngb) + (j + ngb) * (nx + 2)]) / a.dx;
    const double djydy = (jy[(i + ngb) + (j + ngb + 1) * (nx + 1)] - jy[(i + ngb) + (j + ngb) * (nx + 1)]) / a.dy;
    r += djxdx + djydy;
    const double w = ((!a.perx && (i == 0 || i == a.nkx)) ? 0.5 : 1.0) * ((!a.pery && (j == 0 || j == a.nky)) ? 0.5 : 1.0);
    b[k] = -w * r;
}
__global__ void k_darwin_a_rhs(const double *jx, const double *jy, const double *chi, double *bx, double *by, DarArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int mx = a.nkx + 1, my = a.nky + 1;
    if (k >= mx * my) return;
    const int i = k % mx, j = k / mx, ngb = a.ngb, nx = a.nx;
    double rx = 0.0, ry = 0.0;
    if (i < a.nkx) rx = -4 * PI_D / C_LIGHT * (jx[(i + ngb + 1) + (j + ngb) * (nx + 2)] - (chi[k + 1] - chi[k]) / a.dx);
    if (j < a.nky) ry = -4 * PI_D / C_LIGHT * (jy[(i + ngb) + (j + ngb + 1) * (nx + 1)] - (chi[k + mx] - chi[k]) / a.dy);
    const double w = ((!a.perx && (i == 0 || i == a.nkx)) ? 0.5 : 1.0) * ((!a.pery && (j == 0 || j == a.nky)) ? 0.5 : 1.0);
    bx[k] = -w * rx;
    by[k] = -w * ry;
}
// delta fields (EMField.mm:1277-1297)
__global__ void k_darwin_fields(const double *ax, const double *ay, double *lax, double *lay, float *dEx, float *dEy,
                                float *dBz, DarArgs a, double cdt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    const int mx = a.nkx + 1;
    if (i < a.nkx && j < a.nky) {
        const int ngb = a.ngb, nx = a.nx;
        const int kpx = i + ngb + 1 + (j + ngb) * (nx + 2);
        const int kpy = i + ngb + (j + ngb + 2) * (nx + 1);
        const int kf = i + j * mx;
        dEx[kpx] = (float)(-(ax[kf] - lax[kf]) / cdt);
        dEy[kpy] = (float)(-(ay[kf] - lay[kf]) / cdt);
        const double dyAx = (ax[kf + mx] - ax[kf]) / a.dy;
        const double dxAy = (ay[kf + 1] - ay[kf]) / a.dx;
        dBz[kpx + (nx + 2)] = (float)(dxAy - dyAx);
    }
}
__global__ void k_copy64(double *d, const double *s, int64_t n)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) d[k] = s[k];
}

// ---------------------------------------------------------------------------
// host side of the solver
// ---------------------------------------------------------------------------
inline dim3 g2(int mx, int my) { return dim3((mx + 31) / 32, (my + 7) / 8); }
const dim3 b2(32, 8);

Lev mklev(const gp_mg_level &l)
{
    Lev L;
    L.mx = l.mx; L.my = l.my;
    L.tx = l.perx; L.ty = l.pery;   // perx/pery fields hold the axis types
    L.Mx = L.tx == AX_DIR ? l.mx + 1 : (L.tx == AX_NEU ? l.mx - 1 : l.mx);
    L.My = L.ty == AX_DIR ? l.my + 1 : (L.ty == AX_NEU ? l.my - 1 : l.my);
    L.cx = (float)l.hx2i; L.cy = (float)l.hy2i;
    L.rb = !((L.tx == AX_PER && (l.mx & 1)) || (L.ty == AX_PER && (l.my & 1)));
    return L;
}

int build_levels(gp_solver &s, int tx, int ty, int mx, int my, double cx, double cy)
{
    s.lv.clear();
    gp_mg_level l;
    l.mx = mx; l.my = my; l.perx = tx; l.pery = ty; l.hx2i = cx; l.hy2i = cy;
    s.lv.push_back(l);
    for (;;) {
        const gp_mg_level &f = s.lv.back();
        if ((int64_t)f.mx * f.my <= 1024) break;
        auto M = [](int t, int m) { return t == AX_DIR ? m + 1 : (t == AX_NEU ? m - 1 : m); };
        const int Mx = M(tx, f.mx), My = M(ty, f.my);
        if ((Mx & 1) || (My & 1) || Mx < 4 || My < 4) break;
        gp_mg_level c;
        const int Mxc = Mx / 2, Myc = My / 2;
        c.mx = tx == AX_DIR ? Mxc - 1 : (tx == AX_NEU ? Mxc + 1 : Mxc);
        c.my = ty == AX_DIR ? Myc - 1 : (ty == AX_NEU ? Myc + 1 : Myc);
        c.perx = tx; c.pery = ty;
        c.hx2i = f.hx2i / 4.0; c.hy2i = f.hy2i / 4.0;
        s.lv.push_back(c);
    }
    if ((int64_t)s.lv.back().mx * s.lv.back().my > COARSE_MAX)
        return gp_fail(GP_ERR_UNSUPPORTED, "grid cannot be coarsened below 4096 unknowns (need even interval counts)");
    for (auto &L : s.lv) {
        const size_t n = (size_t)L.mx * L.my;
        GP_HIP(hipMalloc((void **)&L.u, n * sizeof(float)));
        GP_HIP(hipMalloc((void **)&L.f, n * sizeof(float)));
        GP_HIP(hipMalloc((void **)&L.r, n * sizeof(float)));
    }
    const int n = mx * my;
    s.n = n;
    GP_HIP(hipMalloc((void **)&s.b64, n * sizeof(double)));
    GP_HIP(hipMalloc((void **)&s.x64, n * sizeof(double)));
    GP_HIP(hipMalloc((void **)&s.r64, n * sizeof(double)));
    GP_HIP(hipMemset(s.x64, 0, n * sizeof(double)));
    GP_HIP(hipMalloc((void **)&s.r, n * sizeof(float)));
    GP_HIP(hipMalloc((void **)&s.z, n * sizeof(float)));
    GP_HIP(hipMalloc((void **)&s.p, n * sizeof(float)));
    GP_HIP(hipMalloc((void **)&s.q, n * sizeof(float)));
    GP_HIP(hipMalloc((void **)&s.e, n * sizeof(float)));
    s.red_blocks = std::min(1024, gp_cdiv(n, RB));
    GP_HIP(hipMalloc((void **)&s.red, 2 * s.red_blocks * sizeof(double)));
    GP_HIP(hipMalloc((void **)&s.scal, S_NS * sizeof(double)));
    GP_HIP(hipHostMalloc((void **)&s.hscal, S_NS * sizeof(double), hipHostMallocDefault));
    return GP_OK;
}
void free_solver(gp_solver &s)
{
    for (auto &L : s.lv) { (void)hipFree(L.u); (void)hipFree(L.f); (void)hipFree(L.r); }
    s.lv.clear();
    void *ps[] = {s.b64, s.x64, s.r64, s.rhs_bc, s.r, s.z, s.p, s.q, s.e, s.red, s.scal};
    for (void *p : ps) if (p) (void)hipFree(p);
    if (s.hscal) (void)hipHostFree(s.hscal);
    s = gp_solver();
}

// z = V-cycle(r) on level l (u_l is the output)
int vcycle(hipStream_t st, gp_solver &s, int l, int singular)
{
    gp_mg_level &ml = s.lv[l];
    const Lev L = mklev(ml);
    const int n = L.mx * L.my;
    if (l == (int)s.lv.size() - 1) {
        const int sweeps = std::max(8, 2 * std::max(L.mx, L.my));
        k_coarse<<<1, 1024, 0, st>>>(ml.u, ml.f, L, sweeps, singular);
        GP_HIP(hipGetLastError());
        return GP_OK;
    }
    GP_HIP(hipMemsetAsync(ml.u, 0, n * sizeof(float), st));
    auto smooth = [&](bool forward) -> int {
        if (L.rb) {
            k_rbgs<<<g2(L.mx, L.my), b2, 0, st>>>(ml.u, ml.f, L, forward ? 0 : 1);
            k_rbgs<<<g2(L.mx, L.my), b2, 0, st>>>(ml.u, ml.f, L, forward ? 1 : 0);
        } else {
            k_jacobi<<<g2(L.mx, L.my), b2, 0, st>>>(ml.u, ml.f, ml.r, L);
            k_copyf<<<gp_cdiv(n, 256), 256, 0, st>>>(ml.u, ml.r, n);
        }
        GP_HIP(hipGetLastError());
        return GP_OK;
    };
    int rc = smooth(true);
    if (rc) return rc;
    gp_mg_level &mc = s.lv[l + 1];
    const Lev C = mklev(mc);
    k_resid<<<g2(L.mx, L.my), b2, 0, st>>>(ml.u, ml.f, ml.r, L);
    k_restrict<<<g2(C.mx, C.my), b2, 0, st>>>(ml.r, L, mc.f, C);
    GP_HIP(hipGetLastError());
    rc = vcycle(st, s, l + 1, singular);
    if (rc) return rc;
    k_prolong_add<<<g2(L.mx, L.my), b2, 0, st>>>(ml.u, L, mc.u, C);
    GP_HIP(hipGetLastError());
    return smooth(false);
}

// FP32 PCG on S e = r32 (s.lv[0].f holds nothing; uses s.r etc.)
// returns iterations; the relative reduction target is inner_tol
int pcg32(hipStream_t st, gp_solver &s, int singular, double inner_tol, int maxit, int *iters)
{
    const Lev L0 = mklev(s.lv[0]);
    const int n = s.n, nb = s.red_blocks;
    const int ge = std::min(1024, gp_cdiv(n, 256));
    GP_HIP(hipMemsetAsync(s.e, 0, n * sizeof(float), st));
    // z = M r
    k_copyf<<<ge, 256, 0, st>>>(s.lv[0].f, s.r, n);
    int rc = vcycle(st, s, 0, singular);
    if (rc) return rc;
    k_copyf<<<ge, 256, 0, st>>>(s.z, s.lv[0].u, n);
    k_copyf<<<ge, 256, 0, st>>>(s.p, s.z, n);
    k_dot<<<nb, RB, 0, st>>>(s.r, s.z, n, s.red);
    k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_RZ, 0);
    k_dot<<<nb, RB, 0, st>>>(s.r, s.r, n, s.red);
    k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_RR, 0);
    GP_HIP(hipGetLastError());
    GP_HIP(hipMemcpyAsync(s.hscal, s.scal, S_NS * sizeof(double), hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    const double r0 = std::sqrt(s.hscal[S_RR]);
    int it = 0;
    if (r0 == 0.0) { *iters = 0; return GP_OK; }
    while (it < maxit) {
        k_spmv_dot<<<nb, RB, 0, st>>>(s.p, s.q, L0, s.red);
        k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_PQ, 1);
        k_update_xr<<<nb, RB, 0, st>>>(s.e, s.r, s.p, s.q, n, s.scal, s.red);
        k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_RR, 0);
        GP_HIP(hipGetLastError());
        it++;
        GP_HIP(hipMemcpyAsync(s.hscal + S_RR, s.scal + S_RR, sizeof(double), hipMemcpyDeviceToHost, st));
        k_copyf<<<ge, 256, 0, st>>>(s.lv[0].f, s.r, n);
        rc = vcycle(st, s, 0, singular);
        if (rc) return rc;
        GP_HIP(hipStreamSynchronize(st));
        const double rr = std::sqrt(std::max(0.0, s.hscal[S_RR]));
        if (!(rr == rr)) return gp_fail(GP_ERR_NONFINITE, "non-finite residual in the Poisson solve");
        if (rr <= inner_tol * r0) break;
        k_copyf<<<ge, 256, 0, st>>>(s.z, s.lv[0].u, n);
        k_dot<<<nb, RB, 0, st>>>(s.r, s.z, n, s.red);
        k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_TMP, 2);
        k_update_p<<<ge, 256, 0, st>>>(s.p, s.z, n, s.scal);
        GP_HIP(hipGetLastError());
    }
    *iters = it;
    return GP_OK;
}

}  // namespace

// Solve S x = b64 (already in s.b64) starting from s.x64; FP64 refinement.
int gp_solve_spd(hipStream_t st, gp_solver &s, double tol, int maxiter, int singular, gp_solve_stats *stats,
                 int (*precorrect)(void *, float *, const float *), void *pc_ctx)
{
    const Lev L0 = mklev(s.lv[0]);
    const int n = s.n, nb = s.red_blocks;
    const int ge = std::min(1024, gp_cdiv(n, 256));
    const double cx = s.lv[0].hx2i, cy = s.lv[0].hy2i;
    auto t0 = std::chrono::steady_clock::now();
    if (singular) {
        k_sum64<<<nb, RB, 0, st>>>(s.b64, n, s.red);
        k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_MEAN, 0);
        k_sub_mean64<<<ge, 256, 0, st>>>(s.b64, n, s.scal, S_MEAN);
        GP_HIP(hipGetLastError());
    }
    int total_it = 0, outer = 0;
    double rel = 0.0;
    int status = GP_OK;
    for (;;) {
        k_resid64<<<nb, RB, 0, st>>>(s.x64, s.b64, s.r64, L0, cx, cy, s.red, s.red + nb);
        k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_RR, 0);
        k_final<<<1, RB, 0, st>>>(s.red + nb, nb, s.scal, S_BB, 0);
        GP_HIP(hipGetLastError());
        GP_HIP(hipMemcpyAsync(s.hscal, s.scal, S_NS * sizeof(double), hipMemcpyDeviceToHost, st));
        GP_HIP(hipStreamSynchronize(st));
        const double rr = s.hscal[S_RR], bb = s.hscal[S_BB];
        rel = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
        if (!(rel == rel)) return gp_fail(GP_ERR_NONFINITE, "non-finite residual in the FP64 refinement");
        if (rel <= tol || bb == 0.0) break;
        if (total_it >= maxiter || outer >= 20) { status = GP_NOT_CONVERGED; break; }
        // scale the residual to unit RMS for FP32
        const double rms = std::sqrt(rr / n);
        const double sc = rms > 0 ? 1.0 / rms : 1.0;
        k_to_f32<<<ge, 256, 0, st>>>(s.r64, s.r, n, sc);
        GP_HIP(hipGetLastError());
        int it = 0;
        if (precorrect) {
            // learned correction: e0 = N(r); the FP64 loop absorbs its error
            int rc = precorrect(pc_ctx, s.e, s.r);
            if (rc) return rc;
            k_axpy64<<<ge, 256, 0, st>>>(s.x64, s.e, n, 1.0 / sc);
            GP_HIP(hipGetLastError());
            precorrect = nullptr;
            outer++;
            continue;
        }
        int rc = pcg32(st, s, singular, 1e-5, std::max(1, maxiter - total_it), &it);
        if (rc) return rc;
        total_it += it;
        k_axpy64<<<ge, 256, 0, st>>>(s.x64, s.e, n, 1.0 / sc);
        GP_HIP(hipGetLastError());
        outer++;
    }
    if (singular) {
        k_sum64<<<nb, RB, 0, st>>>(s.x64, n, s.red);
        k_final<<<1, RB, 0, st>>>(s.red, nb, s.scal, S_MEAN, 0);
        k_sub_mean64<<<ge, 256, 0, st>>>(s.x64, n, s.scal, S_MEAN);
        GP_HIP(hipGetLastError());
    }
    GP_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->iterations = total_it;
        stats->refinements = outer;
        stats->relres = rel;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    return status;
}

int gp_fields_init(gp_ctx *c)
{
    const int nx = c->nx, ny = c->ny;
    const size_t nEx = (size_t)(nx + 2) * (ny + 1), nEy = (size_t)(nx + 1) * (ny + 2), nN = (size_t)(nx + 1) * (ny + 1);
    const size_t nBz = (size_t)(nx + 2) * (ny + 2), nPhi = (size_t)(nx + 3) * (ny + 3);
    GP_HIP(hipMalloc((void **)&c->Ex, nEx * 4));
    GP_HIP(hipMalloc((void **)&c->Ey, nEy * 4));
    GP_HIP(hipMalloc((void **)&c->Ez, nN * 4));
    GP_HIP(hipMalloc((void **)&c->Bx, nEy * 4));
    GP_HIP(hipMalloc((void **)&c->By, nEx * 4));
    GP_HIP(hipMalloc((void **)&c->Bz, nBz * 4));
    GP_HIP(hipMalloc((void **)&c->dEx, nEx * 4));
    GP_HIP(hipMalloc((void **)&c->dEy, nEy * 4));
    GP_HIP(hipMalloc((void **)&c->dBz, nBz * 4));
    GP_HIP(hipMalloc((void **)&c->rho, nN * 8));
    GP_HIP(hipMalloc((void **)&c->jx, nEx * 8));
    GP_HIP(hipMalloc((void **)&c->jy, nEy * 8));
    GP_HIP(hipMalloc((void **)&c->phi, nPhi * 8));
    GP_HIP(hipMemset(c->rho, 0, nN * 8));
    GP_HIP(hipMemset(c->jx, 0, nEx * 8));
    GP_HIP(hipMemset(c->jy, 0, nEy * 8));
    GP_HIP(hipMemset(c->phi, 0, nPhi * 8));
    GP_HIP(hipMemset(c->dEx, 0, nEx * 4));
    GP_HIP(hipMemset(c->dEy, 0, nEy * 4));
    GP_HIP(hipMemset(c->dBz, 0, nBz * 4));
    // uniform initial fields (EMField.mm:146-174)
    auto fill = [&](float *d, size_t n, double v) -> int {
        k_fill_f<<<gp_cdiv(n, 256), 256, 0, c->stream>>>(d, n, (float)v);
        GP_HIP(hipGetLastError());
        return GP_OK;
    };
    int rc = fill(c->Ex, nEx, c->cfg.e0[0]); if (rc) return rc;
    rc = fill(c->Ey, nEy, c->cfg.e0[1]); if (rc) return rc;
    rc = fill(c->Ez, nN, c->cfg.e0[2]); if (rc) return rc;
    rc = fill(c->Bx, nEy, c->cfg.b0[0]); if (rc) return rc;
    rc = fill(c->By, nEx, c->cfg.b0[1]); if (rc) return rc;
    rc = fill(c->Bz, nBz, c->cfg.b0[2]); if (rc) return rc;
    c->bz0 = (float)c->cfg.b0[2];
    c->uniform_b = 1;
    GP_HIP(hipMalloc((void **)&c->red, 3 * 1024 * sizeof(double)));
    GP_HIP(hipHostMalloc((void **)&c->hred, 8 * sizeof(double), hipHostMallocDefault));

    // Poisson unknown grid (EMField.mm:247-271)
    for (int w = 0; w < 4; w++)
        if (c->cfg.fbc[w] == GP_FBC_NEUMANN)
            return gp_fail(GP_ERR_UNSUPPORTED, "Neumann boundary in the Poisson solve (the reference rejects it too, EMField.mm:434-437)");
    int nkx = c->ngx, ikmin = c->ngb, nky = c->ngy, jkmin = c->ngb;
    if (c->cfg.fbc[0] == GP_FBC_DIRICHLET) { nkx--; ikmin++; }
    if (c->cfg.fbc[1] == GP_FBC_DIRICHLET) nkx--;
    if (c->cfg.fbc[0] == GP_FBC_PERIODIC) nkx--;
    if (c->cfg.fbc[2] == GP_FBC_DIRICHLET) { nky--; jkmin++; }
    if (c->cfg.fbc[3] == GP_FBC_DIRICHLET) nky--;
    if (c->cfg.fbc[2] == GP_FBC_PERIODIC) nky--;
    if ((c->cfg.fbc[0] == GP_FBC_PERIODIC) != (c->cfg.fbc[1] == GP_FBC_PERIODIC) ||
        (c->cfg.fbc[2] == GP_FBC_PERIODIC) != (c->cfg.fbc[3] == GP_FBC_PERIODIC))
        return gp_fail(GP_ERR_UNSUPPORTED, "a periodic field boundary must be periodic on both sides");
    gp_solver &s = c->ps;
    s.nkx = nkx; s.nky = nky; s.ikmin = ikmin; s.jkmin = jkmin;
    const int tx = c->cfg.fbc[0] == GP_FBC_PERIODIC ? AX_PER : AX_DIR;
    const int ty = c->cfg.fbc[2] == GP_FBC_PERIODIC ? AX_PER : AX_DIR;
    s.singular = (tx == AX_PER && ty == AX_PER);
    const double dx = c->dxf, dy = c->dyf;
    rc = build_levels(s, tx, ty, nkx + 1, nky + 1, 1.0 / dx / dx, 1.0 / dy / dy);
    if (rc) return rc;
    // rhs_bc (EMField.mm:288-395: Dirichlet values enter the rhs)
    {
        const int mx = nkx + 1, my = nky + 1;
        std::vector<double> rb((size_t)mx * my, 0.0);
        const double Cx = 1.0 / dx / dx, Cy = 1.0 / dy / dy;
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                double v = 0.0;
                if (i == 0 && c->cfg.fbc[0] == GP_FBC_DIRICHLET) v += -Cx * c->cfg.fbc_value[0];
                if (i == nkx && c->cfg.fbc[1] == GP_FBC_DIRICHLET) v += -Cx * c->cfg.fbc_value[1];
                if (j == 0 && c->cfg.fbc[2] == GP_FBC_DIRICHLET) v += -Cy * c->cfg.fbc_value[2];
                if (j == nky && c->cfg.fbc[3] == GP_FBC_DIRICHLET) v += -Cy * c->cfg.fbc_value[3];
                rb[(size_t)i + (size_t)j * mx] = v;
            }
        GP_HIP(hipMalloc((void **)&s.rhs_bc, rb.size() * sizeof(double)));
        GP_HIP(hipMemcpy(s.rhs_bc, rb.data(), rb.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    // Darwin: node grid, Neumann where not periodic (EMField.mm:483-499)
    if (c->cfg.em_model == 2) {
        gp_solver &d = c->ds;
        const int perx = c->cfg.fbc[0] == GP_FBC_PERIODIC, pery = c->cfg.fbc[2] == GP_FBC_PERIODIC;
        d.nkx = c->ngx - (perx ? 1 : 0);
        d.nky = c->ngy - (pery ? 1 : 0);
        d.singular = 1;
        rc = build_levels(d, perx ? AX_PER : AX_NEU, pery ? AX_PER : AX_NEU, d.nkx + 1, d.nky + 1,
                          1.0 / dx / dx, 1.0 / dy / dy);
        if (rc) return rc;
        const size_t m = (size_t)(d.nkx + 1) * (d.nky + 1);
        GP_HIP(hipMalloc((void **)&c->dsol_chi, m * 8));
        GP_HIP(hipMalloc((void **)&c->dsol_ax, m * 8));
        GP_HIP(hipMalloc((void **)&c->dsol_ay, m * 8));
        GP_HIP(hipMalloc((void **)&c->last_ax, m * 8));
        GP_HIP(hipMalloc((void **)&c->last_ay, m * 8));
        GP_HIP(hipMemset(c->dsol_chi, 0, m * 8));
        GP_HIP(hipMemset(c->dsol_ax, 0, m * 8));
        GP_HIP(hipMemset(c->dsol_ay, 0, m * 8));
        GP_HIP(hipMemset(c->last_ax, 0, m * 8));
        GP_HIP(hipMemset(c->last_ay, 0, m * 8));
    }
    GP_HIP(hipStreamSynchronize(c->stream));
    return GP_OK;
}

void gp_fields_free(gp_ctx *c)
{
    float *fs[] = {c->Ex, c->Ey, c->Ez, c->Bx, c->By, c->Bz, c->dEx, c->dEy, c->dBz};
    for (float *p : fs) if (p) (void)hipFree(p);
    double *ds[] = {c->rho, c->jx, c->jy, c->phi, c->dsol_chi, c->dsol_ax, c->dsol_ay, c->last_ax, c->last_ay, c->red};
    for (double *p : ds) if (p) (void)hipFree(p);
    if (c->hred) (void)hipHostFree(c->hred);
    free_solver(c->ps);
    free_solver(c->ds);
}

int gp_finalize_grids(gp_ctx *c)
{
    // scale and sum the species accumulators; clear them for the next deposit
    hipStream_t st = c->stream;
    const int64_t nN = (int64_t)(c->nx + 1) * (c->ny + 1);
    const int64_t nEx = (int64_t)(c->nx + 2) * (c->ny + 1), nEy = (int64_t)(c->nx + 1) * (c->ny + 2);
    AccSet a{}, jxs{}, jys{};
    a.ns = jxs.ns = jys.ns = (int)c->sp.size();
    for (size_t q = 0; q < c->sp.size(); q++) {
        const gp_species &S = c->sp[q];
        a.a[q] = S.rho_acc;
        a.f[q] = ldexp(1.0, -GP_RHO_FRAC) * (double)S.scale;
        jxs.a[q] = S.jx_acc; jys.a[q] = S.jy_acc;
        jxs.f[q] = jys.f[q] = ldexp(1.0, -GP_CUR_FRAC) * S.vnorm * (double)S.scale;
    }
    k_scale_sum<<<gp_cdiv(nN, 256), 256, 0, st>>>(a, c->rho, nN);
    GP_HIP(hipGetLastError());
    if (c->cfg.cur_dens) {
        k_scale_sum<<<gp_cdiv(nEx, 256), 256, 0, st>>>(jxs, c->jx, nEx);
        k_scale_sum<<<gp_cdiv(nEy, 256), 256, 0, st>>>(jys, c->jy, nEy);
        GP_HIP(hipGetLastError());
    }
    for (auto &S : c->sp) {
        GP_HIP(hipMemsetAsync(S.rho_acc, 0, nN * 8, st));
        if (c->cfg.cur_dens) {
            GP_HIP(hipMemsetAsync(S.jx_acc, 0, nEx * 8, st));
            GP_HIP(hipMemsetAsync(S.jy_acc, 0, nEy * 8, st));
        }
    }
    // fold order of EMField.mm:749-850: non-periodic x, non-periodic y,
    // periodic x, periodic y
    const int *b = c->cfg.fbc;
    const int perx = b[0] == GP_FBC_PERIODIC, pery = b[2] == GP_FBC_PERIODIC;
    if (!perx) k_fold_x<<<gp_cdiv(c->ny + 1, 64), 64, 0, st>>>(c->rho, c->ngb, c->ngx, c->nx, c->ny, b[0], b[1], 0);
    if (!pery) k_fold_y<<<gp_cdiv(c->nx + 1, 64), 64, 0, st>>>(c->rho, c->ngb, c->ngy, c->nx, b[2], b[3], 0);
    if (perx) k_fold_x<<<gp_cdiv(c->ny + 1, 64), 64, 0, st>>>(c->rho, c->ngb, c->ngx, c->nx, c->ny, b[0], b[1], 1);
    if (pery) k_fold_y<<<gp_cdiv(c->nx + 1, 64), 64, 0, st>>>(c->rho, c->ngb, c->ngy, c->nx, b[2], b[3], 1);
    GP_HIP(hipGetLastError());
    return GP_OK;
}

int gp_energy_impl(gp_ctx *c, double dt, double *we_ref, double *we_cgs, double *joule)
{
    hipStream_t st = c->stream;
    EnArgs a;
    a.ngb = c->ngb; a.ngx = c->ngx; a.ngy = c->ngy; a.nx = c->nx; a.ny = c->ny;
    a.perx = c->cfg.fbc[0] == GP_FBC_PERIODIC; a.pery = c->cfg.fbc[2] == GP_FBC_PERIODIC;
    const int nb = 256;
    k_energy<<<nb, RB, 0, st>>>(c->Ex, c->Ey, c->cfg.cur_dens ? c->jx : nullptr, c->cfg.cur_dens ? c->jy : nullptr, a,
                                c->red, nb);
    double *sc = c->red + 3 * nb;   // three scalars after the partials (red holds 3*1024)
    k_final<<<1, RB, 0, st>>>(c->red, nb, sc, 0, 0);
    k_final<<<1, RB, 0, st>>>(c->red + nb, nb, sc, 1, 0);
    k_final<<<1, RB, 0, st>>>(c->red + 2 * nb, nb, sc, 2, 0);
    GP_HIP(hipGetLastError());
    GP_HIP(hipMemcpyAsync(c->hred, sc, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    GP_HIP(hipStreamSynchronize(st));
    const double dxdy = (double)c->dxf * (double)c->dyf;
    if (we_ref) *we_ref = c->hred[0] * (dxdy / 2);
    if (joule) *joule = c->hred[1] * (dt * c->dxf * c->dyf);
    if (we_cgs) *we_cgs = c->hred[2] * dxdy / (8.0 * PI_D);
    return GP_OK;
}

int gp_nn_precorrect_cb(void *ctx, float *e, const float *r);

int gp_solve_fields_impl(gp_ctx *c, double dt, double tol, int maxiter, int solver, gp_solve_stats *st,
                         gp_diag *diag)
{
    hipStream_t sm = c->stream;
    int rc = gp_finalize_grids(c);
    if (rc) return rc;
    gp_solver &s = c->ps;
    const int mx = s.nkx + 1, my = s.nky + 1;
    k_rhs<<<gp_cdiv(mx * my, 256), 256, 0, sm>>>(s.rhs_bc, c->rho, s.b64, mx, my, s.ikmin, s.jkmin, c->nx + 1);
    GP_HIP(hipGetLastError());
    gp_solve_stats local{};
    int (*pc)(void *, float *, const float *) = nullptr;
    if (solver == 1) {
        if (!c->nn) return gp_fail(GP_ERR_STATE, "NN solver requested before gp_nn_init/gp_nn_load");
        pc = gp_nn_precorrect_cb;
        local.used_nn = 1;
    }
    const int status = gp_solve_spd(sm, s, tol, maxiter, s.singular, &local, pc, c);
    if (status < 0) return status;
    if (st) { int un = local.used_nn; *st = local; st->used_nn = un; }
    RecArgs ra;
    ra.ngb = c->ngb; ra.ngx = c->ngx; ra.ngy = c->ngy; ra.nx = c->nx; ra.ny = c->ny; ra.nkx = s.nkx; ra.nky = s.nky;
    for (int w = 0; w < 4; w++) { ra.b[w] = c->cfg.fbc[w]; ra.v[w] = c->cfg.fbc_value[w]; }
    k_reconstruct<<<g2(c->nx + 3, c->ny + 3), b2, 0, sm>>>(s.x64, c->phi, ra);
    k_gradient<<<g2(c->nx + 2, c->ny + 2), b2, 0, sm>>>(c->phi, c->Ex, c->Ey, c->nx, c->ny, (double)c->dxf, (double)c->dyf);
    GP_HIP(hipGetLastError());
    c->have_solution = 1;
    if (diag) {
        rc = gp_energy_impl(c, dt, &diag->we_ref, &diag->we_cgs, &diag->joule);
        if (rc) return rc;
    }
    return status;
}

int gp_solve_darwin_impl(gp_ctx *c, double dt, double tol, int maxiter, int solver, gp_solve_stats st[3])
{
    if (c->cfg.em_model != 2) return gp_fail(GP_ERR_STATE, "Darwin solve needs em_model = 2");
    if (!c->cfg.cur_dens) return gp_fail(GP_ERR_STATE, "Darwin solve needs the current deposit");
    hipStream_t sm = c->stream;
    gp_solver &d = c->ds;
    DarArgs a;
    a.ngb = c->ngb; a.nx = c->nx; a.nkx = d.nkx; a.nky = d.nky;
    a.perx = c->cfg.fbc[0] == GP_FBC_PERIODIC; a.pery = c->cfg.fbc[2] == GP_FBC_PERIODIC;
    a.dx = c->dxf; a.dy = c->dyf;
    const int m = (d.nkx + 1) * (d.nky + 1);
    int (*pc)(void *, float *, const float *) = nullptr;
    (void)solver;
    // chi
    k_darwin_chi_rhs<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->jx, c->jy, d.b64, a);
    GP_HIP(hipGetLastError());
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.x64, c->dsol_chi, m);
    int s0 = gp_solve_spd(sm, d, tol, maxiter, 1, &st[0], pc, c);
    if (s0 < 0) return s0;
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->dsol_chi, d.x64, m);
    // Ax, Ay rhs from chi; solve each
    double *bx = nullptr, *by = nullptr;
    GP_HIP(hipMallocAsync((void **)&bx, m * 8, sm));
    GP_HIP(hipMallocAsync((void **)&by, m * 8, sm));
    k_darwin_a_rhs<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->jx, c->jy, c->dsol_chi, bx, by, a);
    GP_HIP(hipGetLastError());
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.b64, bx, m);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.x64, c->dsol_ax, m);
    int s1 = gp_solve_spd(sm, d, tol, maxiter, 1, &st[1], pc, c);
    if (s1 < 0) return s1;
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->dsol_ax, d.x64, m);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.b64, by, m);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.x64, c->dsol_ay, m);
    int s2 = gp_solve_spd(sm, d, tol, maxiter, 1, &st[2], pc, c);
    if (s2 < 0) return s2;
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->dsol_ay, d.x64, m);
    GP_HIP(hipFreeAsync(bx, sm));
    GP_HIP(hipFreeAsync(by, sm));
    k_darwin_fields<<<g2(d.nkx + 1, d.nky + 1), b2, 0, sm>>>(c->dsol_ax, c->dsol_ay, c->last_ax, c->last_ay, c->dEx,
                                                             c->dEy, c->dBz, a, C_LIGHT * dt);
    GP_HIP(hipGetLastError());
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->last_ax, c->dsol_ax, m);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->last_ay, c->dsol_ay, m);
    GP_HIP(hipGetLastError());
    GP_HIP(hipStreamSynchronize(sm));
    return (s0 == GP_NOT_CONVERGED || s1 == GP_NOT_CONVERGED || s2 == GP_NOT_CONVERGED) ? GP_NOT_CONVERGED : GP_OK;
}
