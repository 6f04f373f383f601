// gp_fields.hip -- grid kernels and the GPU elliptic solver.
//
// Replaces EMField's CPU path (EMField.mm:745-1361): ghost-charge fold,
// Poisson rhs, the AMGCL CG solve, phi reconstruction, E = -grad phi, the
// energy diagnostics and the Darwin chain.  All of it stays on the device.
//
// The elliptic solver itself (MG-PCG with FP64 refinement) is gp_mg.hip;
// this unit assembles its right-hand sides (SPD form b = -D rhs) and turns
// its solutions into phi, E and the Darwin corrections.
#include "gp_internal.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace {

constexpr double PI_D = 3.141592653589793;
constexpr double C_LIGHT = 2.99792458e10;   // Constant.h:5

// ---------------- deterministic FP64 reductions ----------------
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
// sum of nb partials into out[slot] (one workgroup, fixed order)
__global__ void k_final(const double *part, int nb, double *out, int slot, int)
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
    if (threadIdx.x == 0) out[slot] = sh[0];
}
// ---------------- charge / current / rhs / reconstruction ----------------
struct AccSet {
    int64_t *a[GP_MAX_SPECIES];
    double f[GP_MAX_SPECIES];
    int ns;
};
// out = sum over species of acc * f, and the accumulators cleared for the
// next deposit
__global__ void k_scale_sum(AccSet s, double *out, int64_t n)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    double v = 0.0;
    for (int q = 0; q < s.ns; q++) {
        v += (double)s.a[q][k] * s.f[q];
        s.a[q][k] = 0;
    }
    out[k] = v;
}
// fold (EMField.mm:749-850), one thread per row (x) or column (y); applied
// to each species' int64 accumulator before the conversion to FP64, so the
// folded charge is exact and independent of where a deposit landed
// (blockIdx.y selects the species' accumulator)
template <typename T>
__global__ void k_fold_x(AccSet as, int ngb, int ngx, int nx, int ny, int bmin, int bmax, int periodic)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > ny) return;
    T *r = (T *)as.a[blockIdx.y] + (int64_t)j * (nx + 1);
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
template <typename T>
__global__ void k_fold_y(AccSet as, int ngb, int ngy, int nx, int bmin, int bmax, int periodic)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nx) return;
    const int P = nx + 1;
    T *r = (T *)as.a[blockIdx.y] + i;
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
// hollow-cathode ramp (EMField.mm:1063-1077): the mean of the solution over
// the nky+1 rows of unknown column col, then phi -= mean ((i - ngb - 2) dx) / pos
// with (i - ngb - 2) dx a float product, as the reference's float dx makes it
__global__ __launch_bounds__(256) void k_cathode_mean(const double *sol, int col, int K, int rows, double *out)
{
    __shared__ double sh[256];
    double m = 0.0;
    for (int j = threadIdx.x; j < rows; j += 256) m += sol[col + (int64_t)j * K];
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] / (double)rows;
}
__global__ void k_cathode_ramp(double *phi, int nx, int ny, int ngb, float dx, double pos, const double *mean)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > nx + 2 || j > ny + 2) return;
    const float xi = (float)(i - ngb - 2) * dx;
    phi[i + (int64_t)j * (nx + 3)] -= mean[0] * (double)xi / pos;
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

// Darwin rhs (EMField.mm:1147-1192, 1203-1275): the kernels write the
// reference's rhs; its plain mean is removed (EMField.mm:1184-1187) and
// k_darwin_weight then turns it into the SPD form b = -D rhs (rows of Neumann
// boundary nodes halved)
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
    b[k] = r;
}
__global__ void k_darwin_weight(double *b, DarArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int mx = a.nkx + 1, my = a.nky + 1;
    if (k >= mx * my) return;
    const int i = k % mx, j = k / mx;
    const double w = ((!a.perx && (i == 0 || i == a.nkx)) ? 0.5 : 1.0) * ((!a.pery && (j == 0 || j == a.nky)) ? 0.5 : 1.0);
    b[k] = -w * b[k];
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
    bx[k] = rx;
    by[k] = ry;
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

}  // namespace

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
    GP_HIP(hipMalloc((void **)&c->red, (3 * 1024 + 8) * sizeof(double)));   // + the cathode mean
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
    const int tx = c->cfg.fbc[0] == GP_FBC_PERIODIC ? GP_AX_PER : GP_AX_DIR;
    const int ty = c->cfg.fbc[2] == GP_FBC_PERIODIC ? GP_AX_PER : GP_AX_DIR;
    s.singular = (tx == GP_AX_PER && ty == GP_AX_PER);
    const double dx = c->dxf, dy = c->dyf;
    rc = gp_mg_build(s, tx, ty, nkx + 1, nky + 1, 1.0 / dx / dx, 1.0 / dy / dy);
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
        rc = gp_mg_build(d, perx ? GP_AX_PER : GP_AX_NEU, pery ? GP_AX_PER : GP_AX_NEU, d.nkx + 1, d.nky + 1,
                         1.0 / dx / dx, 1.0 / dy / dy);
        if (rc) return rc;
        const size_t m = (size_t)(d.nkx + 1) * (d.nky + 1);
        GP_HIP(hipMalloc((void **)&c->dsol_chi, m * 8));
        GP_HIP(hipMalloc((void **)&c->dsol_ax, m * 8));
        GP_HIP(hipMalloc((void **)&c->dsol_ay, m * 8));
        GP_HIP(hipMalloc((void **)&c->last_ax, m * 8));
        GP_HIP(hipMalloc((void **)&c->last_ay, m * 8));
        GP_HIP(hipMalloc((void **)&c->dar_bx, m * 8));
        GP_HIP(hipMalloc((void **)&c->dar_by, m * 8));
        GP_HIP(hipMalloc((void **)&c->dar_bchi, m * 8));
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
    double *ds[] = {c->rho, c->jx, c->jy, c->phi, c->dsol_chi, c->dsol_ax, c->dsol_ay, c->last_ax, c->last_ay,
                    c->dar_bx, c->dar_by, c->dar_bchi, c->red};
    for (double *p : ds) if (p) (void)hipFree(p);
    if (c->hred) (void)hipHostFree(c->hred);
    gp_mg_free(c->ps);
    gp_mg_free(c->ds);
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
    // fold order of EMField.mm:749-850: non-periodic x, non-periodic y,
    // periodic x, periodic y
    const int *b = c->cfg.fbc;
    const int perx = b[0] == GP_FBC_PERIODIC, pery = b[2] == GP_FBC_PERIODIC;
    const int ns = (int)c->sp.size();
    if (ns > 0) {   // every species in one launch per fold
        const dim3 gx(gp_cdiv(c->ny + 1, 64), ns), gy(gp_cdiv(c->nx + 1, 64), ns);
        if (!perx) k_fold_x<long long><<<gx, 64, 0, st>>>(a, c->ngb, c->ngx, c->nx, c->ny, b[0], b[1], 0);
        if (!pery) k_fold_y<long long><<<gy, 64, 0, st>>>(a, c->ngb, c->ngy, c->nx, b[2], b[3], 0);
        if (perx) k_fold_x<long long><<<gx, 64, 0, st>>>(a, c->ngb, c->ngx, c->nx, c->ny, b[0], b[1], 1);
        if (pery) k_fold_y<long long><<<gy, 64, 0, st>>>(a, c->ngb, c->ngy, c->nx, b[2], b[3], 1);
        GP_HIP(hipGetLastError());
    }
    // the sums also clear the accumulators for the next deposit
    k_scale_sum<<<gp_cdiv(nN, 256), 256, 0, st>>>(a, c->rho, nN);
    GP_HIP(hipGetLastError());
    if (c->cfg.cur_dens) {
        k_scale_sum<<<gp_cdiv(nEx, 256), 256, 0, st>>>(jxs, c->jx, nEx);
        k_scale_sum<<<gp_cdiv(nEy, 256), 256, 0, st>>>(jys, c->jy, nEy);
        GP_HIP(hipGetLastError());
    }
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

// solver: 0 MG-PCG; 1 network warm start + network flexible CG (falling
// back to MG-PCG); 2 network warm start + MG-PCG; 3 network flexible CG
// from the current iterate
int make_hooks(gp_ctx *c, int solver, gp_nn_bind *bind, gp_hooks *hk)
{
    if (solver < 0 || solver > 3) return gp_fail(GP_ERR_ARG, "solver must be 0..3");
    if (solver == 0) return GP_OK;
    if (!c->nn) return gp_fail(GP_ERR_STATE, "network solver requested before gp_nn_init/gp_nn_load");
    hk->ctx = bind;
    if (solver == 1 || solver == 2) hk->warm = gp_nn_precorrect_cb;
    if (solver == 1 || solver == 3) hk->precond = gp_nn_precond_cb;
    hk->cap = gp_nn_fcg_cap(c);
    return GP_OK;
}

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
    gp_hooks hk;
    gp_nn_bind bind{c, &s};
    int rc2 = make_hooks(c, solver, &bind, &hk);
    if (rc2) return rc2;
    const int status = gp_solve_spd(sm, s, tol, maxiter, s.singular, &local, solver ? &hk : nullptr);
    if (status < 0) return status;
    if (st) *st = local;
    RecArgs ra;
    ra.ngb = c->ngb; ra.ngx = c->ngx; ra.ngy = c->ngy; ra.nx = c->nx; ra.ny = c->ny; ra.nkx = s.nkx; ra.nky = s.nky;
    for (int w = 0; w < 4; w++) { ra.b[w] = c->cfg.fbc[w]; ra.v[w] = c->cfg.fbc_value[w]; }
    k_reconstruct<<<g2(c->nx + 3, c->ny + 3), b2, 0, sm>>>(s.x64, c->phi, ra);
    if (c->cathode_pos > 0.0) {
        const int col = (int)(c->cathode_pos / (double)c->dxf);
        if (col < 0 || col > s.nkx) return gp_fail(GP_ERR_ARG, "hollow-cathode position outside the potential unknowns");
        double *mean = c->red + 3 * 1024;
        k_cathode_mean<<<1, 256, 0, sm>>>(s.x64, col, s.nkx + 1, s.nky + 1, mean);
        k_cathode_ramp<<<g2(c->nx + 3, c->ny + 3), b2, 0, sm>>>(c->phi, c->nx, c->ny, c->ngb, c->dxf, c->cathode_pos, mean);
        GP_HIP(hipGetLastError());
        GP_HIP(hipMemcpyAsync(c->hred + 4, mean, sizeof(double), hipMemcpyDeviceToHost, sm));
        GP_HIP(hipStreamSynchronize(sm));
        c->cathode_mean = c->hred[4];
    }
    k_gradient<<<g2(c->nx + 2, c->ny + 2), b2, 0, sm>>>(c->phi, c->Ex, c->Ey, c->nx, c->ny, (double)c->dxf, (double)c->dyf);
    GP_HIP(hipGetLastError());
    c->have_solution = 1;
    c->e_periodic = 1;   // phi repeats its periodic nodes exactly (k_reconstruct)
    if (diag) {
        rc = gp_energy_impl(c, dt, &diag->we_ref, &diag->we_cgs, &diag->joule);
        if (rc) return rc;
        diag->mean_cathode = c->cathode_pos > 0.0 ? c->cathode_mean : 0.0;
    }
    return status;
}

// S x = b on the Poisson hierarchy for a caller-given b (host FP64), x in/out
// as the initial guess; the context's phi and E are not touched
int gp_solve_rhs_impl(gp_ctx *c, const double *b, double *x, double tol, int maxiter, int solver, gp_solve_stats *st)
{
    hipStream_t sm = c->stream;
    gp_solver &s = c->ps;
    const int64_t m = s.n;
    // keep the context's own solution (the warm start of its next step)
    double *keep = nullptr;
    // every exit frees it (an early error exit leaves the warm start as the
    // failed solve left it)
    struct Keep {
        double **p;
        hipStream_t s;
        ~Keep() { if (*p) (void)hipFreeAsync(*p, s); }
    } guard{&keep, sm};
    GP_HIP(hipMallocAsync((void **)&keep, m * 8, sm));
    GP_HIP(hipMemcpyAsync(keep, s.x64, m * 8, hipMemcpyDeviceToDevice, sm));
    GP_HIP(hipMemcpyAsync(s.b64, b, m * 8, hipMemcpyHostToDevice, sm));
    GP_HIP(hipMemcpyAsync(s.x64, x, m * 8, hipMemcpyHostToDevice, sm));
    gp_solve_stats local{};
    gp_hooks hk;
    gp_nn_bind bind{c, &s};
    int rc = make_hooks(c, solver, &bind, &hk);
    if (rc) return rc;
    const int status = gp_solve_spd(sm, s, tol, maxiter, s.singular, &local, solver ? &hk : nullptr);
    if (status < 0) return status;
    GP_HIP(hipMemcpyAsync(x, s.x64, m * 8, hipMemcpyDeviceToHost, sm));
    GP_HIP(hipMemcpyAsync(s.x64, keep, m * 8, hipMemcpyDeviceToDevice, sm));
    GP_HIP(hipStreamSynchronize(sm));
    if (st) *st = local;
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
    gp_hooks hk;
    gp_nn_bind bind{c, &d};
    int rc3 = make_hooks(c, solver, &bind, &hk);
    if (rc3) return rc3;
    const gp_hooks *hp = solver ? &hk : nullptr;
    // chi
    k_darwin_chi_rhs<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->jx, c->jy, d.b64, a);
    GP_HIP(hipGetLastError());
    int rc = gp_mg_sub_mean(sm, d, d.b64);
    if (rc) return rc;
    k_darwin_weight<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.b64, a);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.x64, c->dsol_chi, m);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->dar_bchi, d.b64, m);
    int s0 = gp_solve_spd(sm, d, tol, maxiter, 1, &st[0], hp);
    if (s0 < 0) return s0;
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->dsol_chi, d.x64, m);
    // Ax, Ay rhs from chi; solve each
    double *const bx = c->dar_bx, *const by = c->dar_by;
    k_darwin_a_rhs<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->jx, c->jy, c->dsol_chi, bx, by, a);
    GP_HIP(hipGetLastError());
    if ((rc = gp_mg_sub_mean(sm, d, bx)) || (rc = gp_mg_sub_mean(sm, d, by))) return rc;
    k_darwin_weight<<<gp_cdiv(m, 256), 256, 0, sm>>>(bx, a);
    k_darwin_weight<<<gp_cdiv(m, 256), 256, 0, sm>>>(by, a);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.b64, bx, m);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.x64, c->dsol_ax, m);
    int s1 = gp_solve_spd(sm, d, tol, maxiter, 1, &st[1], hp);
    if (s1 < 0) return s1;
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->dsol_ax, d.x64, m);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.b64, by, m);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(d.x64, c->dsol_ay, m);
    int s2 = gp_solve_spd(sm, d, tol, maxiter, 1, &st[2], hp);
    if (s2 < 0) return s2;
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->dsol_ay, d.x64, m);
    k_darwin_fields<<<g2(d.nkx + 1, d.nky + 1), b2, 0, sm>>>(c->dsol_ax, c->dsol_ay, c->last_ax, c->last_ay, c->dEx,
                                                             c->dEy, c->dBz, a, C_LIGHT * dt);
    GP_HIP(hipGetLastError());
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->last_ax, c->dsol_ax, m);
    k_copy64<<<gp_cdiv(m, 256), 256, 0, sm>>>(c->last_ay, c->dsol_ay, m);
    GP_HIP(hipGetLastError());
    GP_HIP(hipStreamSynchronize(sm));
    return (s0 == GP_NOT_CONVERGED || s1 == GP_NOT_CONVERGED || s2 == GP_NOT_CONVERGED) ? GP_NOT_CONVERGED : GP_OK;
}
