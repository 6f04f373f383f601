/*
 * gp_oracle.c -- CPU oracle for the gample PIC step (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library.  See gp_oracle.h for the list of reference functions that
 * each routine restates.  Compile with -ffp-contract=off: every product and
 * sum not written as fmaf() is rounded on its own, as in the HIP kernels.
 */
#include "gp_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.141592653589793
#define C_LIGHT 2.99792458e10          /* src/Constant.h:5 */

static inline int nthr(int n) { return n > 0 ? n : 1; }

/* ======================================================================
 * B-spline weights
 * ==================================================================== */

/* Literal reference form, src/Particle.mm:97-131: Metal float arithmetic
 * with pow(); `sc` is the distance argument exactly as written there. */
void gpo_weights_ref(int order, float h, float *w)
{
    float sc;
    if (order == 1) {
        w[0] = 1.0f - h;
        w[1] = h;
    } else if (order == 3) {
        sc = 1.0f + h; w[0] = 1.0f / 6.0f * powf(2.0f - sc, 3.0f);
        sc = h;        w[1] = 1.0f / 6.0f * (4.0f - 6.0f * powf(sc, 2.0f) + 3.0f * powf(sc, 3.0f));
        sc = 1.0f - h; w[2] = 1.0f / 6.0f * (4.0f - 6.0f * powf(sc, 2.0f) + 3.0f * powf(sc, 3.0f));
        sc = 2.0f - h; w[3] = 1.0f / 6.0f * powf(2.0f - sc, 3.0f);
    } else {
        sc = 2.0f + h; w[0] = 1.0f / 120.0f * powf(3.0f - sc, 5.0f);
        sc = 1.0f + h; w[1] = 1.0f / 120.0f * (51.0f + 75.0f * sc - 210.0f * powf(sc, 2.0f)
                              + 150.0f * powf(sc, 3.0f) - 45.0f * powf(sc, 4.0f) + 5.0f * powf(sc, 5.0f));
        sc = h;        w[2] = 1.0f / 60.0f * (33.0f - 30.0f * powf(sc, 2.0f) + 15.0f * powf(sc, 4.0f)
                              - 5.0f * powf(sc, 5.0f));
        sc = 1.0f - h; w[3] = 1.0f / 60.0f * (33.0f - 30.0f * powf(sc, 2.0f) + 15.0f * powf(sc, 4.0f)
                              - 5.0f * powf(sc, 5.0f));
        sc = 2.0f - h; w[4] = 1.0f / 120.0f * (51.0f + 75.0f * sc - 210.0f * powf(sc, 2.0f)
                              + 150.0f * powf(sc, 3.0f) - 45.0f * powf(sc, 4.0f) + 5.0f * powf(sc, 5.0f));
        sc = 3.0f - h; w[5] = 1.0f / 120.0f * powf(3.0f - sc, 5.0f);
    }
}

/* Same literal formulas evaluated in FP64 (used to state errors). */
void gpo_weights_f64(int order, double h, double *w)
{
    double sc;
    if (order == 1) {
        w[0] = 1.0 - h; w[1] = h;
    } else if (order == 3) {
        sc = 1.0 + h; w[0] = 1.0 / 6.0 * pow(2.0 - sc, 3);
        sc = h;       w[1] = 1.0 / 6.0 * (4.0 - 6.0 * pow(sc, 2) + 3.0 * pow(sc, 3));
        sc = 1.0 - h; w[2] = 1.0 / 6.0 * (4.0 - 6.0 * pow(sc, 2) + 3.0 * pow(sc, 3));
        sc = 2.0 - h; w[3] = 1.0 / 6.0 * pow(2.0 - sc, 3);
    } else {
        sc = 2.0 + h; w[0] = 1.0 / 120.0 * pow(3.0 - sc, 5);
        sc = 1.0 + h; w[1] = 1.0 / 120.0 * (51.0 + 75.0 * sc - 210.0 * pow(sc, 2) + 150.0 * pow(sc, 3)
                             - 45.0 * pow(sc, 4) + 5.0 * pow(sc, 5));
        sc = h;       w[2] = 1.0 / 60.0 * (33.0 - 30.0 * pow(sc, 2) + 15.0 * pow(sc, 4) - 5.0 * pow(sc, 5));
        sc = 1.0 - h; w[3] = 1.0 / 60.0 * (33.0 - 30.0 * pow(sc, 2) + 15.0 * pow(sc, 4) - 5.0 * pow(sc, 5));
        sc = 2.0 - h; w[4] = 1.0 / 120.0 * (51.0 + 75.0 * sc - 210.0 * pow(sc, 2) + 150.0 * pow(sc, 3)
                             - 45.0 * pow(sc, 4) + 5.0 * pow(sc, 5));
        sc = 3.0 - h; w[5] = 1.0 / 120.0 * pow(3.0 - sc, 5);
    }
}

/* Kernel op order.  The same polynomials rewritten in h (and t = 1-h for
 * the mirrored half) and evaluated by Horner with fmaf:
 *   order 3: w0 = t^3/6, w1 = 2/3 + h^2 (h/2 - 1), w2 = same in t, w3 = h^3/6
 *   order 5: 120 w1 = 26 - 50h + 20h^2 + 20h^3 - 20h^4 + 5h^5 (w4: same in t)
 *            120 w2 = 66 - 60h^2 + 30h^4 - 10h^5            (w3: same in t)
 *            120 w0 = t^5, 120 w5 = h^5. */
static inline float q5a(float s) { /* 26 - 50s + 20s^2 + 20s^3 - 20s^4 + 5s^5 */
    return fmaf(s, fmaf(s, fmaf(s, fmaf(s, fmaf(5.0f, s, -20.0f), 20.0f), 20.0f), -50.0f), 26.0f);
}
static inline float q5b(float s) { /* 66 - 60s^2 + 30s^4 - 10s^5 */
    float s2 = s * s;
    return fmaf(s2, fmaf(s2, fmaf(-10.0f, s, 30.0f), -60.0f), 66.0f);
}
static inline void wfast(int order, float h, float *w)
{
    if (order == 1) {
        w[0] = 1.0f - h;
        w[1] = h;
    } else if (order == 3) {
        const float c6 = 1.0f / 6.0f, c23 = 2.0f / 3.0f;
        float t = 1.0f - h;
        float h2 = h * h, t2 = t * t;
        w[0] = (t2 * t) * c6;
        w[1] = fmaf(h2, fmaf(0.5f, h, -1.0f), c23);
        w[2] = fmaf(t2, fmaf(0.5f, t, -1.0f), c23);
        w[3] = (h2 * h) * c6;
    } else {
        const float c120 = 1.0f / 120.0f;
        float t = 1.0f - h;
        float h2 = h * h, t2 = t * t;
        w[0] = ((t2 * t2) * t) * c120;
        w[1] = q5a(h) * c120;
        w[2] = q5b(h) * c120;
        w[3] = q5b(t) * c120;
        w[4] = q5a(t) * c120;
        w[5] = ((h2 * h2) * h) * c120;
    }
}
void gpo_weights_fast(int order, float h, float *w) { wfast(order, h, w); }

/* ======================================================================
 * Push
 * ==================================================================== */

static inline int row_ex(int nx) { return nx + 2; }   /* Ex, jx, Bz pitch */
static inline int row_ey(int nx) { return nx + 1; }   /* Ey, jy, rho pitch */

/* Periodic wrap as the reference writes it (src/Particle.mm:216-221):
 * x = fmod(x + (ngx-ngb), ngx) + ngb, FP32. */
static inline float wrap_ref(float x, int ng, int ngb)
{
    return fmodf(x + (float)(ng - ngb), (float)ng) + (float)ngb;
}

/* BCs after the position update, src/Particle.mm:215-269.  Returns piflag.
 * int(floor(p)) < k is written as p < k (identical for integer k, and
 * well defined for NaN). */
static inline int apply_bc(const gpo_grid *g, float *x, float *y, float *vx, float *vy)
{
    const int ngb = g->ngb, ngx = g->ngx, ngy = g->ngy;
    int flag = 0;
    if (g->pbc[0] == GPO_PBC_PERIODIC) *x = wrap_ref(*x, ngx, ngb);
    if (g->pbc[2] == GPO_PBC_PERIODIC) *y = wrap_ref(*y, ngy, ngb);
    if (g->pbc[0] == GPO_PBC_DELETE && *x < (float)ngb) flag = 1;
    if (g->pbc[1] == GPO_PBC_DELETE && *x >= (float)(ngb + ngx)) flag = 2;
    if (g->pbc[2] == GPO_PBC_DELETE && *y < (float)ngb) flag = 3;
    if (g->pbc[3] == GPO_PBC_DELETE && *y >= (float)(ngb + ngy)) flag = 4;
    if (g->pbc[0] == GPO_PBC_REFLECT && *x < (float)ngb) {
        *x = (float)ngb + (float)ngb - *x; *vx = -*vx;
    }
    if (g->pbc[1] == GPO_PBC_REFLECT && *x >= (float)(ngb + ngx)) {
        *x = (float)(ngb + ngx) + (float)(ngb + ngx) - *x; *vx = -*vx;
    }
    if (g->pbc[2] == GPO_PBC_REFLECT && *y < (float)ngb) {
        *y = (float)ngb + (float)ngb - *y; *vy = -*vy;
    }
    if (g->pbc[3] == GPO_PBC_REFLECT && *y >= (float)(ngb + ngy)) {
        *y = (float)(ngb + ngy) + (float)(ngb + ngy) - *y; *vy = -*vy;
    }
    return flag;
}

/* Literal restatement of updateParticles, src/Particle.mm:65-278. */
int gpo_push_ref(const gpo_grid *g, const float *Ex, const float *Ey, const float *Bz,
                 float constE, float constB, float constX, float constY,
                 int64_t n, float *x, float *y, float *vx, float *vy, float *vz,
                 int32_t *piflag, int nthreads)
{
    const int ngb = g->ngb, nx = g->ngx + 2 * g->ngb, order = g->order, bo = g->boris_order;
    (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
    for (int64_t id = 0; id < n; id++) {
        float px = x[id], py = y[id];
        float xh = px - 0.5f, yh = py - 0.5f;
        int i1 = (int)floorf(xh), j1 = (int)floorf(py);
        int i2 = (int)floorf(px), j2 = (int)floorf(yh);
        float hv[2][2];
        hv[0][0] = xh - (float)i1; hv[0][1] = py - (float)j1;
        hv[1][0] = px - (float)i2; hv[1][1] = yh - (float)j2;
        float sf[6][2][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                float w[6];
                gpo_weights_ref(order, hv[a][b], w);
                for (int k = 0; k <= order; k++) sf[k][a][b] = w[k];
            }
        float Epx = 0.0f, Epy = 0.0f, Epz = 0.0f, Bpx = 0.0f, Bpy = 0.0f, Bpz = 0.0f;
        for (int i = 0; i < order + 1; i++)
            for (int j = 0; j < order + 1; j++) {
                int ii = i1 + (i - ngb + 1), jj = j1 + (j - ngb);
                Epx += sf[i][0][0] * sf[j][0][1] * Ex[ii + jj * (nx + 2)];
                ii = i2 + (i - ngb); jj = j2 + (j - ngb + 1);
                Epy += sf[i][1][0] * sf[j][1][1] * Ey[ii + jj * (nx + 1)];
                ii = i1 + (i - ngb + 1); jj = j2 + (j - ngb + 1);
                Bpz += sf[i][0][0] * sf[j][1][1] * Bz[ii + jj * (nx + 2)];
            }
        float etx = constE * Epx, ety = constE * Epy, etz = constE * Epz;
        float btx = constB * Bpx, bty = constB * Bpy, btz = constB * Bpz;
        float btbt = btx * btx + bty * bty + btz * btz;
        if (bo != 2) {
            float fN1 = 1.0f, fN2 = 0.0f;
            if (bo == 4) { fN1 = (1.0f + btbt / 3.0f); fN2 = 1.0f / 3.0f; }
            else if (bo == 6) { fN1 = (1.0f + btbt / 3.0f + 2.0f * btbt * btbt / 15.0f);
                                fN2 = (1.0f / 3.0f + 2.0f * btbt / 15.0f); }
            float etbt = etx * btx + ety * bty + etz * btz;
            etx = fN1 * etx + fN2 * etbt * btx;
            ety = fN1 * ety + fN2 * etbt * bty;
            etz = fN1 * etz + fN2 * etbt * btz;
            btx = fN1 * btx; bty = fN1 * bty; btz = fN1 * btz;
        }
        float umx = vx[id] + etx, umy = vy[id] + ety, umz = vz[id] + etz;
        float v0x = umx + umy * btz - umz * bty;
        float v0y = umy + umz * btx - umx * btz;
        float v0z = umz + umx * bty - umy * btx;
        float ssx = 2 * btx / (1 + btbt), ssy = 2 * bty / (1 + btbt), ssz = 2 * btz / (1 + btbt);
        float upx = umx + v0y * ssz - v0z * ssy;
        float upy = umy + v0z * ssx - v0x * ssz;
        float upz = umz + v0x * ssy - v0y * ssx;
        float nvx = upx + constE * Epx, nvy = upy + constE * Epy, nvz = upz + constE * Epz;
        px = px + nvx * constX;
        py = py + nvy * constY;
        int flag = apply_bc(g, &px, &py, &nvx, &nvy);
        x[id] = px; y[id] = py; vx[id] = nvx; vy[id] = nvy; vz[id] = nvz;
        if (piflag && flag) piflag[id] = flag;
    }
    return 0;
}

/* factorised tap gather F = sum_i sx[i] * (sum_j sy[j] * b[i + j*pitch]),
 * each sum an fmaf chain started by a plain product */
static inline float gather_taps(const float *b, int pitch, int nw, const float *sx, const float *sy)
{
    float acc = 0.0f;
    for (int i = 0; i < nw; i++) {
        float r = sy[0] * b[i];
        for (int j = 1; j < nw; j++) r = fmaf(sy[j], b[i + j * pitch], r);
        acc = (i == 0) ? sx[0] * r : fmaf(sx[i], r, acc);
    }
    return acc;
}

/* Order 3: a stencil column's taps f0..f3 weighted by w0..w3(h) is the cubic
 * d0 + h (d1 + h (d2 + h d3)), with d_k the coefficients of h^k in the
 * weights (src/Particle.mm:97-131):
 *   d0 = (f0 + 4 f1 + f2)/6, d1 = (f2 - f0)/2, d2 = (f0 + f2)/2 - f1,
 *   d3 = (f3 - f0 + 3 (f1 - f2))/6,
 * each in the kernel's FP32 operation order (coef3, horner3, gather_c3 in
 * gample_amd/csrc/gp_particles.hip). */
static inline float gather_c3(const float *b, int pitch, float h, const float *sx)
{
    float acc = 0.0f;
    for (int i = 0; i < 4; i++) {
        const float f0 = b[i], f1 = b[i + pitch], f2 = b[i + 2 * pitch], f3 = b[i + 3 * pitch];
        const float a02 = f0 + f2;
        const float d0 = fmaf(f1, 2.0f / 3.0f, a02 * (1.0f / 6.0f));
        const float d1 = (f2 - f0) * 0.5f;
        const float d2 = fmaf(a02, 0.5f, -f1);
        const float d3 = fmaf(f1 - f2, 3.0f, f3 - f0) * (1.0f / 6.0f);
        const float r = fmaf(fmaf(fmaf(d3, h, d2), h, d1), h, d0);
        acc = (i == 0) ? sx[0] * r : fmaf(sx[i], r, acc);
    }
    return acc;
}

/* Kernel op order (the HIP push restates this line by line):
 *   - Horner weights (wfast),
 *   - gather at orders 1 and 5 factorised, F_p = sum_i sx[i] *
 *     (sum_j sy[j] * F[i][j]), each sum an fmaf chain started by a plain
 *     product (gather_taps); at order 3 the column polynomials in the y
 *     fraction (gather_c3),
 *   - Boris specialised to the only components the reference gathers
 *     (Ex, Ey, Bz; Ez = Bx = By = 0 exactly, src/Particle.mm:134-154), with
 *     the cross products written as fmaf,
 *   - x = fmaf(vx, constX, x),
 *   - absorbed particles marked x = -piflag. */
int gpo_push_fast(const gpo_grid *g, const float *Ex, const float *Ey, const float *Bz,
                  int uniform_bz, float bz0,
                  float constE, float constB, float constX, float constY,
                  int64_t n, float *x, float *y, float *vx, float *vy, float *vz,
                  int nthreads)
{
    const int ngb = g->ngb, nx = g->ngx + 2 * g->ngb, order = g->order, bo = g->boris_order;
    const int pex = row_ex(nx), pey = row_ey(nx);
    const int nw = order + 1;
#pragma omp parallel for schedule(static) num_threads(nthr(nthreads))
    for (int64_t id = 0; id < n; id++) {
        float px = x[id], py = y[id];
        if (!(px >= (float)ngb && px < (float)(ngb + g->ngx) && py >= (float)ngb && py < (float)(ngb + g->ngy)))
            continue;                    /* absorbed marker or outside: untouched */
        float xh = px - 0.5f, yh = py - 0.5f;
        float fi1 = floorf(xh), fj1 = floorf(py), fi2 = floorf(px), fj2 = floorf(yh);
        int i1 = (int)fi1, j1 = (int)fj1, i2 = (int)fi2, j2 = (int)fj2;
        float sx0[6], sy0[6], sx1[6], sy1[6];
        wfast(order, xh - fi1, sx0);
        wfast(order, py - fj1, sy0);
        wfast(order, px - fi2, sx1);
        wfast(order, yh - fj2, sy1);
        /* Ex at (i1+1-ngb+i, j1-ngb+j), Ey at (i2-ngb+i, j2+1-ngb+j),
         * Bz at (i1+1-ngb+i, j2+1-ngb+j); order 3 in the column-polynomial
         * form (gather_c3), orders 1 and 5 as weighted tap sums */
        float Epx, Epy, Bpz;
        if (order == 3) {
            Epx = gather_c3(Ex + (i1 + 1 - ngb) + (int64_t)(j1 - ngb) * pex, pex, py - fj1, sx0);
            Epy = gather_c3(Ey + (i2 - ngb) + (int64_t)(j2 + 1 - ngb) * pey, pey, yh - fj2, sx1);
            Bpz = uniform_bz ? bz0 : gather_c3(Bz + (i1 + 1 - ngb) + (int64_t)(j2 + 1 - ngb) * pex, pex, yh - fj2, sx0);
        } else {
            Epx = gather_taps(Ex + (i1 + 1 - ngb) + (int64_t)(j1 - ngb) * pex, pex, nw, sx0, sy0);
            Epy = gather_taps(Ey + (i2 - ngb) + (int64_t)(j2 + 1 - ngb) * pey, pey, nw, sx1, sy1);
            Bpz = uniform_bz ? bz0 : gather_taps(Bz + (i1 + 1 - ngb) + (int64_t)(j2 + 1 - ngb) * pex, pex, nw, sx0, sy1);
        }
        const float ex0 = constE * Epx, ey0 = constE * Epy;
        float etx = ex0, ety = ey0;
        float btz = constB * Bpz;
        const float btbt = btz * btz;
        if (bo == 4 || bo == 6) {
            float fN1 = (bo == 4) ? (1.0f + btbt / 3.0f)
                                  : (1.0f + btbt / 3.0f + 2.0f * btbt * btbt / 15.0f);
            etx = fN1 * etx; ety = fN1 * ety; btz = fN1 * btz;
        }
        const float umx = vx[id] + etx, umy = vy[id] + ety;
        const float v0x = fmaf(umy, btz, umx);
        const float v0y = fmaf(-umx, btz, umy);
        const float ssz = (2.0f * btz) / (1.0f + btbt);
        const float upx = fmaf(v0y, ssz, umx);
        const float upy = fmaf(-v0x, ssz, umy);
        float nvx = upx + ex0, nvy = upy + ey0;
        px = fmaf(nvx, constX, px);
        py = fmaf(nvy, constY, py);
        int flag = apply_bc(g, &px, &py, &nvx, &nvy);
        /* a wrapped or reflected position can round onto the upper wall;
         * the kernel keeps it inside the last cell */
        if (!flag) {
            const float hx = (float)(ngb + g->ngx), hy = (float)(ngb + g->ngy);
            if (px >= hx) px = nextafterf(hx, 0.0f);
            if (py >= hy) py = nextafterf(hy, 0.0f);
        }
        x[id] = flag ? -(float)flag : px;
        y[id] = py; vx[id] = nvx; vy[id] = nvy;   /* vz unchanged: Ez = Bx = By = 0 */
    }
    return 0;
}

/* ======================================================================
 * Deposit
 * ==================================================================== */

int64_t gpo_fx_from_float(float c, int frac_bits)
{
    return (int64_t)llrintf(ldexpf(c, frac_bits));
}

static inline uint32_t q23(float s) { return (uint32_t)rintf(s * 8388608.0f); } /* 2^23 */
static inline uint32_t q16(float s) { return (uint32_t)rintf(s * 65536.0f); }   /* 2^16 */
/* Charge weights: order 3 quantises each per-axis weight to 16 bits, so that
 * the product fits 32 bits (the largest order-3 weight is 2/3: q <= 43691,
 * qx qy < 2^31), and adds qx qy 2^14 in the 46-bit fixed point; orders 1 and
 * 5 quantise to 23 bits and add qx qy. */
static inline uint32_t qw(int order, float s) { return order == 3 ? q16(s) : q23(s); }
static inline int qshift(int order) { return order == 3 ? 14 : 0; }

int gpo_deposit_fx(const gpo_grid *g, int64_t n, const float *x, const float *y,
                   const float *vx, const float *vy, float inv_vnorm,
                   int64_t *rho_acc, int64_t *jx_acc, int64_t *jy_acc, int nthreads)
{
    const int ngb = g->ngb, nx = g->ngx + 2 * g->ngb, ny = g->ngy + 2 * g->ngb, order = g->order;
    const int nw = order + 1;
    const int64_t nrho = (int64_t)(nx + 1) * (ny + 1);
    const int64_t njx = (int64_t)(nx + 2) * (ny + 1), njy = (int64_t)(nx + 1) * (ny + 2);
    const int want_j = (jx_acc != NULL && jy_acc != NULL);
    const int T = nthr(nthreads);
    const int64_t per = nrho + (want_j ? njx + njy : 0);
    int64_t *priv = (int64_t *)calloc((size_t)T * per, sizeof(int64_t));
    if (!priv) return -1;
#pragma omp parallel num_threads(T)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int64_t *r = priv + (size_t)tid * per;
        int64_t *jxp = r + nrho, *jyp = jxp + njx;
#pragma omp for schedule(static)
        for (int64_t id = 0; id < n; id++) {
            float px = x[id], py = y[id];
            if (!(px >= 0.0f)) continue;
            float fi2 = floorf(px), fj1 = floorf(py);
            int i2 = (int)fi2, j1 = (int)fj1;
            float sx1[6], sy0[6];
            wfast(order, px - fi2, sx1);
            wfast(order, py - fj1, sy0);
            uint32_t qx[6], qy[6];
            for (int k = 0; k < nw; k++) { qx[k] = qw(order, sx1[k]); qy[k] = qw(order, sy0[k]); }
            const int sh = qshift(order);
            for (int i = 0; i < nw; i++)
                for (int j = 0; j < nw; j++)
                    r[(i2 + i - ngb) + (int64_t)(j1 + j - ngb) * (nx + 1)] +=
                        (int64_t)(((uint64_t)qx[i] * (uint64_t)qy[j]) << sh);
            if (want_j) {
                float xh = px - 0.5f, yh = py - 0.5f;
                float fi1 = floorf(xh), fj2 = floorf(yh);
                int i1 = (int)fi1, j2 = (int)fj2;
                float sx0[6], sy1[6];
                wfast(order, xh - fi1, sx0);
                wfast(order, yh - fj2, sy1);
                float vnx = vx[id] * inv_vnorm, vny = vy[id] * inv_vnorm;
                for (int i = 0; i < nw; i++)
                    for (int j = 0; j < nw; j++) {
                        jxp[(i1 + 1 + i - ngb) + (int64_t)(j1 + j - ngb) * (nx + 2)] +=
                            gpo_fx_from_float((sx0[i] * sy0[j]) * vnx, GPO_CUR_FRAC);
                        jyp[(i2 + i - ngb) + (int64_t)(j2 + 1 + j - ngb) * (nx + 1)] +=
                            gpo_fx_from_float((sx1[i] * sy1[j]) * vny, GPO_CUR_FRAC);
                    }
            }
        }
    }
    /* merge in thread order (integer sums: any order gives the same bits) */
    for (int t = 0; t < T; t++) {
        int64_t *r = priv + (size_t)t * per;
        for (int64_t k = 0; k < nrho; k++) rho_acc[k] += r[k];
        if (want_j) {
            for (int64_t k = 0; k < njx; k++) jx_acc[k] += r[nrho + k];
            for (int64_t k = 0; k < njy; k++) jy_acc[k] += r[nrho + njx + k];
        }
    }
    free(priv);
    return 0;
}

/* Sequential FP32 restatement of integrateChargeDensity,
 * src/Particle.mm:281-389 (one thread, particle order 0..n-1). */
int gpo_deposit_ref(const gpo_grid *g, int64_t n, const float *x, const float *y,
                    const float *vx, const float *vy, float scale,
                    float *rho, float *jx, float *jy)
{
    const int ngb = g->ngb, nx = g->ngx + 2 * g->ngb, order = g->order;
    for (int64_t id = 0; id < n; id++) {
        float px = x[id], py = y[id];
        if (!(px >= 0.0f)) continue;
        float xh = px - 0.5f, yh = py - 0.5f;
        int i1 = (int)floorf(xh), j1 = (int)floorf(py), i2 = (int)floorf(px), j2 = (int)floorf(yh);
        float hv[2][2] = {{xh - (float)i1, py - (float)j1}, {px - (float)i2, yh - (float)j2}};
        float sf[6][2][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                float w[6];
                gpo_weights_ref(order, hv[a][b], w);
                for (int k = 0; k <= order; k++) sf[k][a][b] = w[k];
            }
        for (int i = 0; i < order + 1; i++)
            for (int j = 0; j < order + 1; j++) {
                int ii = i2 + (i - ngb), jj = j1 + (j - ngb);
                rho[ii + jj * (nx + 1)] += sf[i][1][0] * sf[j][0][1] * scale;
                if (jx && jy) {
                    ii = i1 + (i - ngb + 1); jj = j1 + (j - ngb);
                    jx[ii + jj * (nx + 2)] += sf[i][0][0] * sf[j][0][1] * scale * vx[id];
                    ii = i2 + (i - ngb); jj = j2 + (j - ngb + 1);
                    jy[ii + jj * (nx + 1)] += sf[i][1][0] * sf[j][1][1] * scale * vy[id];
                }
            }
    }
    return 0;
}

int gpo_deposit_f64(const gpo_grid *g, int64_t n, const float *x, const float *y,
                    const float *vx, const float *vy, double scale,
                    double *rho, double *jx, double *jy)
{
    const int ngb = g->ngb, nx = g->ngx + 2 * g->ngb, order = g->order;
    for (int64_t id = 0; id < n; id++) {
        double px = x[id], py = y[id];
        if (!(px >= 0.0)) continue;
        double xh = px - 0.5, yh = py - 0.5;
        int i1 = (int)floor(xh), j1 = (int)floor(py), i2 = (int)floor(px), j2 = (int)floor(yh);
        double a00[6], a01[6], a10[6], a11[6];
        gpo_weights_f64(order, xh - i1, a00);
        gpo_weights_f64(order, py - j1, a01);
        gpo_weights_f64(order, px - i2, a10);
        gpo_weights_f64(order, yh - j2, a11);
        for (int i = 0; i <= order; i++)
            for (int j = 0; j <= order; j++) {
                rho[(i2 + i - ngb) + (j1 + j - ngb) * (nx + 1)] += a10[i] * a01[j] * scale;
                if (jx && jy) {
                    jx[(i1 + i - ngb + 1) + (j1 + j - ngb) * (nx + 2)] += a00[i] * a01[j] * scale * vx[id];
                    jy[(i2 + i - ngb) + (j2 + j - ngb + 1) * (nx + 1)] += a10[i] * a11[j] * scale * vy[id];
                }
            }
    }
    return 0;
}

/* Moments in FP64 (src/Moment.mm:62-129): node weights of int(x+1e-20). */
int gpo_moments_f64(const gpo_grid *g, int64_t n, const float *x, const float *y,
                    const float *vx, const float *vy, const float *vz, double scale, double *mom)
{
    const int ngb = g->ngb, nx = g->ngx + 2 * g->ngb, ny = g->ngy + 2 * g->ngb, order = g->order;
    const int64_t ng = (int64_t)(nx + 1) * (ny + 1);
    for (int64_t id = 0; id < n; id++) {
        float px = x[id], py = y[id];
        if (!(px >= 0.0f)) continue;
        int i1 = (int)(px + 1e-20f), j1 = (int)(py + 1e-20f);
        double wx[6], wy[6];
        gpo_weights_f64(order, (double)(px - (float)i1), wx);
        gpo_weights_f64(order, (double)(py - (float)j1), wy);
        double v[3] = {vx[id], vy[id], vz[id]};
        for (int i = 0; i <= order; i++)
            for (int j = 0; j <= order; j++) {
                int64_t k = (i1 + i - ngb) + (int64_t)(j1 + j - ngb) * (nx + 1);
                double s = wx[i] * wy[j];
                mom[0 * ng + k] += s * scale;
                mom[1 * ng + k] += s * v[0];
                mom[2 * ng + k] += s * v[1];
                mom[3 * ng + k] += s * v[2];
                mom[4 * ng + k] += s * v[0] * v[0];
                mom[5 * ng + k] += s * v[0] * v[1];
                mom[6 * ng + k] += s * v[0] * v[2];
                mom[7 * ng + k] += s * v[1] * v[1];
                mom[8 * ng + k] += s * v[1] * v[2];
                mom[9 * ng + k] += s * v[2] * v[2];
            }
    }
    return 0;
}

/* ======================================================================
 * Reduce: literal tail-swap compaction, src/Particle.mm:904-968.
 * counts[] are particle counts per wall (the reference multiplies by q/ec).
 * ==================================================================== */
int gpo_reduce_literal(int64_t *n, float *x, float *y, float *vx, float *vy, float *vz,
                       int32_t *piflag, int counts[4])
{
    int64_t kmax = *n, pulln = 0;
    counts[0] = counts[1] = counts[2] = counts[3] = 0;
#define CNT(f) do { if ((f) >= 1 && (f) <= 4) counts[(f) - 1]++; } while (0)
    for (int64_t k1 = 0; k1 < *n; k1++) {
        if (k1 >= kmax) {
            if (piflag[k1] > 0) { pulln++; CNT(piflag[k1]); }
            break;
        }
        if (piflag[k1] > 0) {
            CNT(piflag[k1]);
            int64_t kend = kmax;
            for (int64_t k2 = kend - 1; k2 >= k1; k2--) {
                kmax--; pulln++;
                if (piflag[k2] == 0) {
                    x[k1] = x[k2]; y[k1] = y[k2]; vx[k1] = vx[k2]; vy[k1] = vy[k2]; vz[k1] = vz[k2];
                    piflag[k1] = piflag[k2];
                    break;
                }
                if (k2 > k1) CNT(piflag[k2]);
            }
        }
    }
#undef CNT
    *n -= pulln;
    return 0;
}

double gpo_kinetic_energy(int64_t n, const float *x, const float *vx, const float *vy,
                          const float *vz, float m, float w)
{
    double e = 0.0;
    for (int64_t k = 0; k < n; k++) {
        if (!(x[k] >= 0.0f)) continue;
        float vv = vx[k] * vx[k] + vy[k] * vy[k] + vz[k] * vz[k];
        e += 0.5 * (double)m * (double)vv * (double)w;
    }
    return e;
}

/* ======================================================================
 * RNG and loading
 * ==================================================================== */

/* src/Collision.h:11-16 */
uint64_t gpo_splitmix64(uint64_t *seed)
{
    uint64_t z = (*seed += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
/* Half-Maxwellian flux injection through the x-min wall (no reference
 * function: the build's wall-flux source).  Particle k of the call uses the
 * PCG32 stream of id = stream + k seeded like the loader (state and inc are
 * the splitmix64(seed) outputs 2 id + 1 and 2 id + 2).  vx = u0 + vth
 * sqrt(-2 ln r1) is Rayleigh (the flux-weighted half-Maxwellian), vy, vz =
 * u + vth sqrt(-2 ln r2) (cos, sin)(2 pi r3); x = wall + vx dt/dx r4 (a
 * uniform fraction of the step flown inside), y uniform.  Grid units. */
void gpo_inject_wall(uint64_t seed, uint64_t stream, int64_t n, float vth, const float u[3], float dtdx, float lo,
                     float ngx, float ngy, float *x, float *y, float *vx, float *vy, float *vz)
{
    const float PIf = 3.14159265358979f;
    for (int64_t k = 0; k < n; k++) {
        const uint64_t id = stream + (uint64_t)k;
        uint64_t sd = seed + 2ULL * id * 0x9E3779B97F4A7C15ULL;
        uint64_t st = gpo_splitmix64(&sd);
        const uint64_t inc = gpo_splitmix64(&sd) | 1ULL;
        const float r1 = gpo_rand01(gpo_pcg32(&st, inc)), r2 = gpo_rand01(gpo_pcg32(&st, inc));
        const float r3 = gpo_rand01(gpo_pcg32(&st, inc)), r4 = gpo_rand01(gpo_pcg32(&st, inc));
        const float r5 = gpo_rand01(gpo_pcg32(&st, inc));
        const float v = u[0] + vth * sqrtf(-2.0f * logf(r1));
        const float am = vth * sqrtf(-2.0f * logf(r2));
        float xx = lo + v * dtdx * r4;
        const float hi = lo + ngx;
        if (xx >= hi) xx = nextafterf(hi, 0.0f);
        x[k] = xx;
        /* r5 = 1 - 2^-24 can round onto the upper face (B16) */
        float yy = lo + r5 * ngy;
        if (yy >= lo + ngy) yy = nextafterf(lo + ngy, 0.0f);
        y[k] = yy;
        vx[k] = v;
        vy[k] = u[1] + am * cosf(2.0f * PIf * r3);
        vz[k] = u[2] + am * sinf(2.0f * PIf * r3);
    }
}

/* src/Collision.mm:21-27 */
uint32_t gpo_pcg32(uint64_t *state, uint64_t inc)
{
    uint64_t old = *state;
    *state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
/* src/Collision.mm:28-30: float(u) * 2^-32, the conversion rounding to
 * nearest even.  Two guarded cases: u = 0 gives 2^-33 instead of 0 (the
 * loader and the sources take log of the draw), and the 128 largest u,
 * whose conversion rounds to 2^32, give the largest float below 1 instead
 * of 1 (a position drawn at the upper face would leave the domain). */
float gpo_rand01(uint32_t u)
{
    const float r = (float)u * 0x1p-32f;
    if (u == 0u) return 0x1p-33f;
    return r < 1.0f ? r : 0x1.fffffep-1f;
}

#define PIF ((float)PI_D)

static float load_pos_x(const gpo_gen *p, int64_t id, int64_t ntot, float rx, float *ry_io, uint64_t *st,
                        uint64_t inc, float *yout)
{
    float ry = *ry_io, x, y;
    switch (p->gen_type) {
    default:
    case 0: case 3:
        x = p->xmin + rx * (p->xmax - p->xmin);
        y = p->ymin + ry * (p->ymax - p->ymin);
        break;
    case 1:
        x = (p->xmin + p->xmax) / 2.0f + (p->xmax - p->xmin) / PIF * asinf(2.0f * rx - 1.0f);
        y = p->ymin + ry * (p->ymax - p->ymin);
        break;
    case 2:
        x = p->xmin + rx * (p->xmax - p->xmin);
        y = (p->ymin + p->ymax) / 2.0f + (p->ymax - p->ymin) / PIF * asinf(2.0f * ry - 1.0f);
        break;
    case 4: {
        int ax = (int)roundf(sqrtf((float)ntot * (p->xmax - p->xmin) / (p->ymax - p->ymin)));
        if (ax < 1) ax = 1;
        int ay = (int)(ntot / ax);
        if (ay < 1) ay = 1;
        float dxq = (p->xmax - p->xmin) / (float)ax, dyq = (p->ymax - p->ymin) / (float)ay;
        x = p->xmin + dxq * (0.5f + (float)(id % ax));
        y = p->ymin + dyq * (0.5f + (float)(id / ax));
        break; }
    case 5: {
        float rz = gpo_rand01(gpo_pcg32(st, inc));
        x = (p->xmin + p->xmax) / 2.0f + (p->xmax - p->xmin) * sqrtf(-logf(rx)) * cosf(2.0f * PIF * rz);
        y = p->ymin + ry * (p->ymax - p->ymin);
        break; }
    case 6: {
        float rz = gpo_rand01(gpo_pcg32(st, inc));
        x = p->xmin + rx * (p->xmax - p->xmin);
        y = (p->ymin + p->ymax) / 2.0f + (p->ymax - p->ymin) * sqrtf(-logf(rz)) * cosf(2.0f * PIF * ry);
        break; }
    case 7: {
        /* density 1 + delta cos(k (x - xmin)), k = 2 pi m / Lx: invert the CDF
         * F(s) = (s + delta/k sin(k s)) / Lx = rx by Newton in FP64 */
        double L = (double)p->xmax - (double)p->xmin;
        double k = 2.0 * PI_D * (double)p->pert_mode / L, d = p->pert_delta;
        double tgt = (double)rx * L, s = tgt;
        for (int it = 0; it < 8; it++) {
            double f = s + d / k * sin(k * s) - tgt, fp = 1.0 + d * cos(k * s);
            s -= f / fp;
        }
        x = (float)((double)p->xmin + s);
        y = p->ymin + ry * (p->ymax - p->ymin);
        break; }
    }
    *yout = y;
    return x;
}

/* one thread of the generation kernel src/Collision.mm:63-174 on stream
 * (st, inc): electron k, and ion k when ix is given and gen_type is not 3
 * (the hollow cathode, :136-139) */
static void gen_one(const gpo_gen *p, int64_t id, int64_t ntot, uint64_t *stp, uint64_t inc, int64_t k,
                    float *ex, float *ey, float *evx, float *evy, float *evz,
                    float *ix, float *iy, float *ivx, float *ivy, float *ivz)
{
    uint64_t st = *stp;
    {
        float rx = gpo_rand01(gpo_pcg32(&st, inc));
        float ry = gpo_rand01(gpo_pcg32(&st, inc));
        float y;
        float x = load_pos_x(p, id, ntot, rx, &ry, &st, inc, &y);
        float rv1 = gpo_rand01(gpo_pcg32(&st, inc)), rv2 = gpo_rand01(gpo_pcg32(&st, inc));
        float a = p->ele_vth * sqrtf(-logf(rv1));
        float vx = p->ele_u[0] + a * cosf(2.0f * PIF * rv2);
        float vy = p->ele_u[1] + a * sinf(2.0f * PIF * rv2);
        rv1 = gpo_rand01(gpo_pcg32(&st, inc)); rv2 = gpo_rand01(gpo_pcg32(&st, inc));
        float vz = p->ele_u[2] + p->ele_vth * sqrtf(-logf(rv1)) * cosf(2.0f * PIF * rv2);
        {
            const float hx = p->ngb + p->xdom, hy = p->ngb + p->ydom;
            const float xc = x, yc = y;
            x = x / p->dx + p->ngb;
            y = y / p->dy + p->ngb;
            if (p->xdom > 0 && xc <= p->xdom * p->dx && x >= hx) x = nextafterf(hx, 0.0f);
            if (p->ydom > 0 && yc <= p->ydom * p->dy && y >= hy) y = nextafterf(hy, 0.0f);
        }
        ex[k] = x; ey[k] = y; evx[k] = vx; evy[k] = vy; evz[k] = vz;
        if (ix && p->gen_type != 3) {
            rv1 = gpo_rand01(gpo_pcg32(&st, inc)); rv2 = gpo_rand01(gpo_pcg32(&st, inc));
            a = p->ion_vth * sqrtf(-logf(rv1));
            ivx[k] = p->ion_u[0] + a * cosf(2.0f * PIF * rv2);
            ivy[k] = p->ion_u[1] + a * sinf(2.0f * PIF * rv2);
            rv1 = gpo_rand01(gpo_pcg32(&st, inc)); rv2 = gpo_rand01(gpo_pcg32(&st, inc));
            ivz[k] = p->ion_u[2] + p->ion_vth * sqrtf(-logf(rv1)) * cosf(2.0f * PIF * rv2);
            ix[k] = x; iy[k] = y;
        }
    }
    *stp = st;
}

/* src/Collision.mm:280-294 (and :507-520 for later ids): the states are
 * drawn from one splitmix64 sequence in id order, state then inc|1, so id
 * takes outputs 2*id+1 and 2*id+2 of the master sequence. */
void gpo_stream_states(uint64_t master_seed, int64_t first_id, int64_t n, uint64_t *state, uint64_t *inc)
{
    for (int64_t k = 0; k < n; k++) {
        uint64_t sd = master_seed + (uint64_t)(2 * (first_id + k)) * 0x9E3779B97F4A7C15ULL;
        state[k] = gpo_splitmix64(&sd);
        inc[k] = gpo_splitmix64(&sd) | 1ULL;
    }
}

/* src/Collision.mm:63-174 with the seeding of src/Collision.mm:280-294:
 * particle id takes splitmix64 outputs 2*id+1 (state) and 2*id+2 (inc|1)
 * of the master sequence. */
int gpo_load_pair(const gpo_gen *p, uint64_t master_seed, int64_t first_id, int64_t n,
                  float *ex, float *ey, float *evx, float *evy, float *evz,
                  float *ix, float *iy, float *ivx, float *ivy, float *ivz)
{
    const int64_t ntot = first_id + n;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; k++) {
        uint64_t st, inc;
        gpo_stream_states(master_seed, first_id + k, 1, &st, &inc);
        gen_one(p, first_id + k, ntot, &st, inc, k, ex, ey, evx, evy, evz, ix, iy, ivx, ivy, ivz);
    }
    return 0;
}

/* -[Collision artificialIonization:] / -[Collision hollowCathode:] with
 * the reference's persistent RNGState array (src/Collision.mm:464-678): thread
 * id < addn reads state[id], draws, and writes it back (:136-139, :171-172). */
int gpo_generate(const gpo_gen *p, uint64_t *state, const uint64_t *inc, int64_t addn,
                 float *ex, float *ey, float *evx, float *evy, float *evz,
                 float *ix, float *iy, float *ivx, float *ivy, float *ivz)
{
    for (int64_t id = 0; id < addn; id++)
        gen_one(p, id, addn, &state[id], inc[id], id, ex, ey, evx, evy, evz, ix, iy, ivx, ivy, ivz);
    return 0;
}

/* ======================================================================
 * Field side
 * ==================================================================== */

void gpo_poisson_dims(const gpo_field *f, int *nkx, int *nky, int *ikmin, int *jkmin)
{
    int kx = f->ngx, ix = f->ngb, ky = f->ngy, jy = f->ngb;
    if (f->fbc[0] == GPO_FBC_DIRICHLET) { kx--; ix++; }
    if (f->fbc[1] == GPO_FBC_DIRICHLET) kx--;
    if (f->fbc[0] == GPO_FBC_PERIODIC) kx--;
    if (f->fbc[2] == GPO_FBC_DIRICHLET) { ky--; jy++; }
    if (f->fbc[3] == GPO_FBC_DIRICHLET) ky--;
    if (f->fbc[2] == GPO_FBC_PERIODIC) ky--;
    *nkx = kx; *nky = ky; *ikmin = ix; *jkmin = jy;
}

/* src/EMField.mm:749-850, in place on an FP64 copy of rho */
void gpo_fold_rho(const gpo_field *f, double *rho)
{
    const int ngb = f->ngb, ngx = f->ngx, ngy = f->ngy, nx = ngx + 2 * ngb, ny = ngy + 2 * ngb;
    const int P = nx + 1;
    if (f->fbc[0] != GPO_FBC_PERIODIC) {
        for (int j = 0; j <= ny; j++) {
            if (f->fbc[0] == GPO_FBC_DIRICHLET)
                for (int i = 0; i <= ngb; i++) rho[ngb + 1 + j * P] += rho[i + j * P];
            else if (f->fbc[0] == GPO_FBC_NEUMANN)
                for (int i = 0; i < ngb; i++) rho[ngb + j * P] += rho[i + j * P];
            if (f->fbc[1] == GPO_FBC_DIRICHLET)
                for (int i = 0; i <= ngb; i++) rho[ngb + ngx - 1 + j * P] += rho[ngb + ngx + i + j * P];
            else if (f->fbc[1] == GPO_FBC_NEUMANN)
                for (int i = 0; i < ngb; i++) rho[ngb + ngx + j * P] += rho[ngb + ngx + i + j * P];
        }
    }
    if (f->fbc[2] != GPO_FBC_PERIODIC) {
        for (int i = 0; i <= nx; i++) {
            if (f->fbc[2] == GPO_FBC_DIRICHLET)
                for (int j = 0; j <= ngb; j++) rho[i + (ngb + 1) * P] += rho[i + j * P];
            else if (f->fbc[2] == GPO_FBC_NEUMANN)
                for (int j = 0; j < ngb; j++) rho[i + ngb * P] += rho[i + j * P];
            if (f->fbc[3] == GPO_FBC_DIRICHLET)
                for (int j = 0; j <= ngb; j++) rho[i + (ngb + ngy - 1) * P] += rho[i + (ngb + ngy + j) * P];
            else if (f->fbc[3] == GPO_FBC_NEUMANN)
                for (int j = 0; j < ngb; j++) rho[i + (ngb + ngy) * P] += rho[i + (ngb + ngy + j) * P];
        }
    }
    if (f->fbc[0] == GPO_FBC_PERIODIC) {
        for (int j = 0; j <= ny; j++) {
            for (int i = 0; i < ngb; i++) rho[ngx + i + j * P] += rho[i + j * P];
            for (int i = 0; i <= ngb; i++) rho[ngb + i + j * P] += rho[ngb + ngx + i + j * P];
        }
    }
    if (f->fbc[2] == GPO_FBC_PERIODIC) {
        for (int i = 0; i <= nx; i++) {
            for (int j = 0; j < ngb; j++) rho[i + (ngy + j) * P] += rho[i + j * P];
            for (int j = 0; j <= ngb; j++) rho[i + (ngb + j) * P] += rho[i + (ngb + ngy + j) * P];
        }
    }
}

/* src/EMField.mm:236-405 (Dirichlet and periodic walls; the reference
 * rejects Neumann for the Poisson solver, :413-437). */
int64_t gpo_poisson_csr(const gpo_field *f, int32_t *ptr, int32_t *col, double *val, double *rhs_bc)
{
    int nkx, nky, ikmin, jkmin;
    gpo_poisson_dims(f, &nkx, &nky, &ikmin, &jkmin);
    const double dx = f->dx, dy = f->dy;
    const double Cd = -2 * (1.0 / dx / dx + 1.0 / dy / dy), Cx = 1.0 / dx / dx, Cy = 1.0 / dy / dy;
    int64_t nnz = 0;
    ptr[0] = 0;
    for (int j = 0; j <= nky; j++)
        for (int i = 0; i <= nkx; i++) {
            int64_t k = i + (int64_t)j * (nkx + 1);
            col[nnz] = (int32_t)k; val[nnz++] = Cd;
            rhs_bc[k] = 0.0;
            if (i == 0) {
                if (f->fbc[0] == GPO_FBC_DIRICHLET) rhs_bc[k] += -Cx * f->fval[0];
                else if (f->fbc[0] == GPO_FBC_PERIODIC) { col[nnz] = (int32_t)(k + nkx); val[nnz++] = Cx; }
            } else { col[nnz] = (int32_t)(k - 1); val[nnz++] = Cx; }
            if (i == nkx) {
                if (f->fbc[1] == GPO_FBC_DIRICHLET) rhs_bc[k] += -Cx * f->fval[1];
                else if (f->fbc[1] == GPO_FBC_PERIODIC) { col[nnz] = (int32_t)(k - nkx); val[nnz++] = Cx; }
            } else { col[nnz] = (int32_t)(k + 1); val[nnz++] = Cx; }
            if (j == 0) {
                if (f->fbc[2] == GPO_FBC_DIRICHLET) rhs_bc[k] += -Cy * f->fval[2];
                else if (f->fbc[2] == GPO_FBC_PERIODIC) { col[nnz] = (int32_t)(k + (int64_t)nky * (nkx + 1)); val[nnz++] = Cy; }
            } else { col[nnz] = (int32_t)(k - (nkx + 1)); val[nnz++] = Cy; }
            if (j == nky) {
                if (f->fbc[3] == GPO_FBC_DIRICHLET) rhs_bc[k] += -Cy * f->fval[3];
                else if (f->fbc[3] == GPO_FBC_PERIODIC) { col[nnz] = (int32_t)(k - (int64_t)nky * (nkx + 1)); val[nnz++] = Cy; }
            } else { col[nnz] = (int32_t)(k + (nkx + 1)); val[nnz++] = Cy; }
            ptr[k + 1] = (int32_t)nnz;
        }
    return nnz;
}

/* src/EMField.mm:853-861 */
void gpo_poisson_rhs(const gpo_field *f, const double *rho, const double *rhs_bc, double *rhs)
{
    int nkx, nky, ikmin, jkmin;
    gpo_poisson_dims(f, &nkx, &nky, &ikmin, &jkmin);
    const int P = f->ngx + 2 * f->ngb + 1;
    for (int j = 0; j <= nky; j++)
        for (int i = 0; i <= nkx; i++) {
            int64_t k = i + (int64_t)j * (nkx + 1);
            rhs[k] = rhs_bc[k] - 4 * PI_D * rho[(i + ikmin) + (int64_t)(j + jkmin) * P];
        }
}

static void spmv(int64_t n, const int32_t *ptr, const int32_t *col, const double *val,
                 const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; r++) {
        double s = 0.0;
        for (int32_t q = ptr[r]; q < ptr[r + 1]; q++) s += val[q] * x[col[q]];
        y[r] = s;
    }
}
static double dot(int64_t n, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+:s) schedule(static)
    for (int64_t k = 0; k < n; k++) s += a[k] * b[k];
    return s;
}
static void remove_mean(int64_t n, double *a)
{
    double s = 0.0;
    for (int64_t k = 0; k < n; k++) s += a[k];
    s /= (double)n;
    for (int64_t k = 0; k < n; k++) a[k] -= s;
}

double gpo_relres(int64_t n, const int32_t *ptr, const int32_t *col, const double *val,
                  const double *b, const double *x)
{
    double *r = (double *)malloc(sizeof(double) * n);
    spmv(n, ptr, col, val, x, r);
    double rr = 0.0, bb = 0.0;
    for (int64_t k = 0; k < n; k++) { double d = b[k] - r[k]; rr += d * d; bb += b[k] * b[k]; }
    free(r);
    return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}

/* Jacobi-PCG on -A x = -b (A is negative definite). */
int gpo_cg(int64_t n, const int32_t *ptr, const int32_t *col, const double *val,
           const double *rhs, double *x, double tol, int maxiter, int project, double *relres)
{
    double *b = (double *)malloc(sizeof(double) * n), *r = (double *)malloc(sizeof(double) * n);
    double *z = (double *)malloc(sizeof(double) * n), *p = (double *)malloc(sizeof(double) * n);
    double *q = (double *)malloc(sizeof(double) * n), *dinv = (double *)malloc(sizeof(double) * n);
    for (int64_t k = 0; k < n; k++) {
        b[k] = -rhs[k];
        double d = 0.0;
        for (int32_t e = ptr[k]; e < ptr[k + 1]; e++) if (col[e] == k) d += -val[e];
        dinv[k] = d != 0.0 ? 1.0 / d : 1.0;
    }
    if (project) { remove_mean(n, b); remove_mean(n, x); }
    spmv(n, ptr, col, val, x, q);
    for (int64_t k = 0; k < n; k++) r[k] = b[k] + q[k];   /* b - (-A)x */
    if (project) remove_mean(n, r);
    double bn = sqrt(dot(n, b, b));
    if (bn == 0.0) bn = 1.0;
    for (int64_t k = 0; k < n; k++) z[k] = dinv[k] * r[k];
    if (project) remove_mean(n, z);
    memcpy(p, z, sizeof(double) * n);
    double rz = dot(n, r, z), res = sqrt(dot(n, r, r)) / bn;
    int it = 0;
    while (it < maxiter && res > tol) {
        spmv(n, ptr, col, val, p, q);
        for (int64_t k = 0; k < n; k++) q[k] = -q[k];
        double alpha = rz / dot(n, p, q);
        for (int64_t k = 0; k < n; k++) { x[k] += alpha * p[k]; r[k] -= alpha * q[k]; }
        if (project) remove_mean(n, r);
        for (int64_t k = 0; k < n; k++) z[k] = dinv[k] * r[k];
        if (project) remove_mean(n, z);
        double rz2 = dot(n, r, z);
        double beta = rz2 / rz;
        rz = rz2;
        for (int64_t k = 0; k < n; k++) p[k] = z[k] + beta * p[k];
        res = sqrt(dot(n, r, r)) / bn;
        it++;
    }
    if (project) remove_mean(n, x);
    if (relres) *relres = res;
    free(b); free(r); free(z); free(p); free(q); free(dinv);
    return it;
}

/* src/EMField.mm:868-1061 for Dirichlet/periodic walls; the periodic-x
 * index uses (i-(ngb+1)+ngx)%ngx, the form the reference already uses in
 * y (:886) -- its x branch (:907) is off by one. */
void gpo_reconstruct_phi(const gpo_field *f, const double *sol, double *phi)
{
    int nkx, nky, ikmin, jkmin;
    gpo_poisson_dims(f, &nkx, &nky, &ikmin, &jkmin);
    const int ngb = f->ngb, ngx = f->ngx, ngy = f->ngy, nx = ngx + 2 * ngb, ny = ngy + 2 * ngb;
    const int K = nkx + 1;
    const double pXmin = f->fval[0], pXmax = f->fval[1], pYmin = f->fval[2], pYmax = f->fval[3];
    for (int j = 0; j <= ny + 2; j++) {
        int isB = 0, isT = 0, jn;
        if (f->fbc[2] == GPO_FBC_DIRICHLET) { isB = (j <= ngb + 1); jn = j - (ngb + 1) - 1; }
        else jn = (j - (ngb + 1) + ngy) % ngy;
        if (f->fbc[3] == GPO_FBC_DIRICHLET) isT = (j >= ngb + 1 + ngy);
        for (int i = 0; i <= nx + 2; i++) {
            int isL = 0, isR = 0, in;
            if (f->fbc[0] == GPO_FBC_DIRICHLET) { isL = (i <= ngb + 1); in = i - (ngb + 1) - 1; }
            else in = (i - (ngb + 1) + ngx) % ngx;
            if (f->fbc[1] == GPO_FBC_DIRICHLET) isR = (i >= ngb + 1 + ngx);
            double v = 0.0;
            const int di = i - (ngb + 1), dj = j - (ngb + 1);
            const int diR = i - (ngb + 1 + ngx), djT = j - (ngb + 1 + ngy);
            if (!isL && !isR && !isB && !isT) v = sol[in + (int64_t)jn * K];
            else if (isL && !isR && !isB && !isT) v = pXmin + di * (sol[0 + (int64_t)jn * K] - pXmin);
            else if (!isL && isR && !isB && !isT) v = pXmax + diR * (pXmax - sol[nkx + (int64_t)jn * K]);
            else if (!isL && !isR && isB && !isT) v = pYmin + dj * (sol[in] - pYmin);
            else if (!isL && !isR && !isB && isT) v = pYmax + djT * (pYmax - sol[in + (int64_t)nky * K]);
            else if (isL && isB) {
                double ddx = sol[1] - sol[0], ddy = sol[K] - sol[0];
                v = pXmin + di * ddx + dj * ddy;
            } else if (isR && isB) {
                double ddx = sol[nkx] - sol[nkx - 1], ddy = sol[nkx + K] - sol[nkx];
                v = pXmax + diR * ddx + dj * ddy;
            } else if (isL && isT) {
                double ddx = sol[1 + (int64_t)nky * K] - sol[(int64_t)nky * K];
                double ddy = sol[(int64_t)nky * K] - sol[(int64_t)(nky - 1) * K];
                v = pXmin + di * ddx + djT * ddy;
            } else if (isR && isT) {
                double ddx = sol[nkx + (int64_t)nky * K] - sol[nkx - 1 + (int64_t)nky * K];
                double ddy = sol[nkx + (int64_t)nky * K] - sol[nkx + (int64_t)(nky - 1) * K];
                v = pXmax + diR * ddx + djT * ddy;
            }
            phi[i + (int64_t)j * (nx + 3)] = v;
        }
    }
}

/* src/EMField.mm:1063-1077: the mean of the unknowns on column
 * (int)(pos/dx) over its nky+1 rows, summed in row order, then
 * phi -= mean*((i-ngb-2)*dx)/pos over the whole (nx+3)(ny+3) array, with dx
 * the float member and (i-ngb-2)*dx a float product. */
double gpo_cathode_ramp(const gpo_field *f, const double *sol, double pos, double *phi)
{
    int nkx, nky, ikmin, jkmin;
    gpo_poisson_dims(f, &nkx, &nky, &ikmin, &jkmin);
    const int nx = f->ngx + 2 * f->ngb, ny = f->ngy + 2 * f->ngb;
    double mean = 0.0;
    for (int j = 0; j <= nky; j++) mean += sol[(int)(pos / f->dx) + j * (nkx + 1)];
    mean /= nky + 1;
    for (int j = 0; j <= ny + 2; j++)
        for (int i = 0; i <= nx + 2; i++)
            phi[i + j * (nx + 3)] -= mean * ((i - f->ngb - 2) * f->dx) / pos;
    return mean;
}

/* src/EMField.mm:1080-1094 */
void gpo_gradient(const gpo_field *f, const double *phi, float *Ex, float *Ey)
{
    const int nx = f->ngx + 2 * f->ngb, ny = f->ngy + 2 * f->ngb;
    const double dx = f->dx, dy = f->dy;
    for (int i = 0; i <= nx + 1; i++)
        for (int j = 0; j <= ny; j++)
            Ex[i + j * (nx + 2)] = (float)(-(phi[i + 1 + (j + 1) * (nx + 3)] - phi[i + (j + 1) * (nx + 3)]) / dx);
    for (int i = 0; i <= nx; i++)
        for (int j = 0; j <= ny + 1; j++)
            Ey[i + j * (nx + 1)] = (float)(-(phi[i + 1 + (j + 1) * (nx + 3)] - phi[i + 1 + j * (nx + 3)]) / dy);
}

/* src/EMField.mm:1096-1116, plus the CGS form E^2/(8 pi) over the unique
 * staggered points (trapezoid weight 1/2 on wall nodes of a bounded axis). */
void gpo_field_energy(const gpo_field *f, const float *Ex, const float *Ey,
                      const double *jx, const double *jy, double dt,
                      double *we_ref, double *we_cgs, double *joule)
{
    const int ngb = f->ngb, ngx = f->ngx, ngy = f->ngy, nx = ngx + 2 * ngb, ny = ngy + 2 * ngb;
    double e = 0.0, jl = 0.0;
    for (int k = 0; k < (nx + 2) * (ny + 1); k++) { e += (double)Ex[k] * Ex[k]; if (jx) jl += (double)Ex[k] * jx[k]; }
    for (int k = 0; k < (nx + 1) * (ny + 2); k++) { e += (double)Ey[k] * Ey[k]; if (jy) jl += (double)Ey[k] * jy[k]; }
    if (we_ref) *we_ref = e * ((double)f->dx * f->dy / 2);
    if (joule) *joule = jl * (dt * f->dx * f->dy);
    if (we_cgs) {
        const int perx = (f->fbc[0] == GPO_FBC_PERIODIC), pery = (f->fbc[2] == GPO_FBC_PERIODIC);
        double s = 0.0;
        /* Ex index m is at x = m - 1/2: cell centres m = ngb+1 .. ngb+ngx;
         * y nodes ngb .. ngb+ngy (last one only if bounded) */
        for (int j = ngb; j <= ngb + ngy - (pery ? 1 : 0); j++) {
            double wy = (!pery && (j == ngb || j == ngb + ngy)) ? 0.5 : 1.0;
            for (int i = ngb + 1; i <= ngb + ngx; i++) s += wy * (double)Ex[i + j * (nx + 2)] * Ex[i + j * (nx + 2)];
        }
        for (int i = ngb; i <= ngb + ngx - (perx ? 1 : 0); i++) {
            double wx = (!perx && (i == ngb || i == ngb + ngx)) ? 0.5 : 1.0;
            for (int j = ngb + 1; j <= ngb + ngy; j++) s += wx * (double)Ey[i + j * (nx + 1)] * Ey[i + j * (nx + 1)];
        }
        *we_cgs = s * ((double)f->dx * f->dy) / (8.0 * PI_D);
    }
}

/* ---------------- Darwin ---------------- */
void gpo_darwin_dims(const gpo_field *f, int *nkxv, int *nkyv)
{
    *nkxv = f->ngx - (f->fbc[0] == GPO_FBC_PERIODIC ? 1 : 0);
    *nkyv = f->ngy - (f->fbc[2] == GPO_FBC_PERIODIC ? 1 : 0);
}

/* src/EMField.mm:483-623 */
int64_t gpo_darwin_csr(const gpo_field *f, int32_t *ptr, int32_t *col, double *val)
{
    int nkx, nky;
    gpo_darwin_dims(f, &nkx, &nky);
    const double dx = f->dx, dy = f->dy;
    const double Cd = -2 * (1.0 / dx / dx + 1.0 / dy / dy), Cx = 1.0 / dx / dx, Cy = 1.0 / dy / dy;
    const int perx = f->fbc[0] == GPO_FBC_PERIODIC, pery = f->fbc[2] == GPO_FBC_PERIODIC;
    int64_t nnz = 0;
    ptr[0] = 0;
    for (int j = 0; j <= nky; j++)
        for (int i = 0; i <= nkx; i++) {
            int64_t k = i + (int64_t)j * (nkx + 1);
            col[nnz] = (int32_t)k; val[nnz++] = Cd;
            if (i == 0) {
                if (perx) { col[nnz] = (int32_t)(k + nkx); val[nnz++] = Cx; col[nnz] = (int32_t)(k + 1); val[nnz++] = Cx; }
                else { col[nnz] = (int32_t)(k + 1); val[nnz++] = 2 * Cx; }
            } else if (i == nkx) {
                if (perx) { col[nnz] = (int32_t)(k - 1); val[nnz++] = Cx; col[nnz] = (int32_t)(k - nkx); val[nnz++] = Cx; }
                else { col[nnz] = (int32_t)(k - 1); val[nnz++] = 2 * Cx; }
            } else { col[nnz] = (int32_t)(k - 1); val[nnz++] = Cx; col[nnz] = (int32_t)(k + 1); val[nnz++] = Cx; }
            const int64_t R = nkx + 1;
            if (j == 0) {
                if (pery) { col[nnz] = (int32_t)(k + (int64_t)nky * R); val[nnz++] = Cy; col[nnz] = (int32_t)(k + R); val[nnz++] = Cy; }
                else { col[nnz] = (int32_t)(k + R); val[nnz++] = 2 * Cy; }
            } else if (j == nky) {
                if (pery) { col[nnz] = (int32_t)(k - R); val[nnz++] = Cy; col[nnz] = (int32_t)(k - (int64_t)nky * R); val[nnz++] = Cy; }
                else { col[nnz] = (int32_t)(k - R); val[nnz++] = 2 * Cy; }
            } else { col[nnz] = (int32_t)(k - R); val[nnz++] = Cy; col[nnz] = (int32_t)(k + R); val[nnz++] = Cy; }
            ptr[k + 1] = (int32_t)nnz;
        }
    return nnz;
}

/* src/EMField.mm:1147-1192 */
void gpo_darwin_rhs_chi(const gpo_field *f, const double *jx, const double *jy, double *rhs)
{
    int nkx, nky;
    gpo_darwin_dims(f, &nkx, &nky);
    const int ngb = f->ngb, nx = f->ngx + 2 * ngb;
    const double dx = f->dx, dy = f->dy;
    const int64_t N = (int64_t)(nkx + 1) * (nky + 1);
    double sum = 0.0;
    for (int j = 0; j <= nky; j++)
        for (int i = 0; i <= nkx; i++) {
            double r = 0.0;
            if (f->fbc[0] != GPO_FBC_PERIODIC) {
                if (i == 0) r += -2 * jx[(i + ngb) + (j + ngb) * (nx + 2)] / dx;
                else if (i == nkx) r += 2 * jx[(i + ngb) + (j + ngb) * (nx + 2)] / dx;
            }
            if (f->fbc[2] != GPO_FBC_PERIODIC) {
                if (j == 0) r += -2 * jy[(i + ngb) + (j + ngb) * (nx + 1)] / dy;
                else if (j == nky) r += 2 * jy[(i + ngb) + (j + ngb) * (nx + 1)] / dy;
            }
            double djxdx = (jx[(i + ngb + 1) + (j + ngb) * (nx + 2)] - jx[(i + ngb) + (j + ngb) * (nx + 2)]) / dx;
            double djydy = (jy[(i + ngb) + (j + ngb + 1) * (nx + 1)] - jy[(i + ngb) + (j + ngb) * (nx + 1)]) / dy;
            r += djxdx + djydy;
            sum += r;
            rhs[i + (int64_t)j * (nkx + 1)] = r;
        }
    sum /= (double)N;
    for (int64_t k = 0; k < N; k++) rhs[k] -= sum;
}

/* src/EMField.mm:1203-1275 (rhs parts) */
void gpo_darwin_rhs_a(const gpo_field *f, const double *jx, const double *jy, const double *chi,
                      double *rhs_x, double *rhs_y)
{
    int nkx, nky;
    gpo_darwin_dims(f, &nkx, &nky);
    const int ngb = f->ngb, nx = f->ngx + 2 * ngb;
    const double dx = f->dx, dy = f->dy;
    const int64_t N = (int64_t)(nkx + 1) * (nky + 1), R = nkx + 1;
    double sx = 0.0, sy = 0.0;
    for (int j = 0; j <= nky; j++)
        for (int i = 0; i <= nkx; i++) {
            int64_t k = i + j * R;
            double rx = 0.0, ry = 0.0;
            if (i < nkx) {
                double jl = (chi[k + 1] - chi[k]) / dx;
                rx = -4 * PI_D / C_LIGHT * (jx[(i + ngb + 1) + (j + ngb) * (nx + 2)] - jl);
            }
            if (j < nky) {
                double jl = (chi[k + R] - chi[k]) / dy;
                ry = -4 * PI_D / C_LIGHT * (jy[(i + ngb) + (j + ngb + 1) * (nx + 1)] - jl);
            }
            rhs_x[k] = rx; rhs_y[k] = ry; sx += rx; sy += ry;
        }
    sx /= (double)N; sy /= (double)N;
    for (int64_t k = 0; k < N; k++) { rhs_x[k] -= sx; rhs_y[k] -= sy; }
}
