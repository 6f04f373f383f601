// gp_api.hip -- the extern "C" boundary of libgample_amd (include/gample_amd.h).
//
// Every entry point validates its arguments on the host, converts C++
// failures into status codes and records the message for gp_last_error().
#include "gp_internal.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

static thread_local std::string g_err;

namespace {
struct DriftOut {
    unsigned long long *cnt[GP_MAX_DRIFT_SPECIES];
    int ns;
};
// each species' escaped, drifted and moved counters (counters[5..7]) to the
// pinned host array, then drifted and moved cleared for the next push: one
// launch for every species instead of a copy and a fill each
__global__ void k_drift_out(DriftOut d, unsigned long long *host)
{
    const int s = threadIdx.x;
    if (s < d.ns) {
        unsigned long long *c = d.cnt[s];
        host[3 * s] = c[5];
        host[3 * s + 1] = c[6];
        host[3 * s + 2] = c[7];
        c[6] = 0ull;
        c[7] = 0ull;
    }
    __threadfence_system();
}
}  // namespace

int gp_fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int gp_time_mark(gp_ctx *c, int which, int which2)
{
    hipEvent_t e;
    GP_HIP(hipEventCreate(&e));
    c->tev_all.push_back(e);
    GP_HIP(hipEventRecord(e, c->stream));
    c->tev[which].push_back(e);
    if (which2 >= 0 && which2 < GP_TIME_SLOTS) c->tev[which2].push_back(e);
    return GP_OK;
}

int gp_time_collect(gp_ctx *c)
{
    GP_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < GP_TIME_SLOTS; k++) {
        auto &v = c->tev[k];
        for (size_t i = 0; i + 1 < v.size(); i += 2) {
            float ms = 0.f;
            GP_HIP(hipEventElapsedTime(&ms, v[i], v[i + 1]));
            c->t_ms[k] += ms;
            c->t_n[k] += 1;
        }
        // an unpaired begin (a launch that failed after it) is dropped
        v.clear();
    }
    for (hipEvent_t e : c->tev_all) (void)hipEventDestroy(e);
    c->tev_all.clear();
    return GP_OK;
}

int gp_launch_ke(gp_ctx *c, int sid);
int gp_finalize_grids(gp_ctx *c);

#define GP_TRY(...)                                                      \
    try {                                                                \
        __VA_ARGS__                                                      \
    } catch (const std::bad_alloc &) {                                   \
        return gp_fail(GP_ERR_NOMEM, "host allocation failed");          \
    } catch (const std::exception &e) {                                  \
        return gp_fail(GP_ERR_STATE, e.what());                          \
    }

#define GP_CHECK_CTX(c) \
    do { if (!(c)) return gp_fail(GP_ERR_ARG, "null context"); } while (0)
#define GP_CHECK_SID(c, sid) \
    do { if ((sid) < 0 || (sid) >= (int)(c)->sp.size()) return gp_fail(GP_ERR_ARG, "bad species id"); } while (0)

extern "C" {

const char *gp_last_error(void) { return g_err.c_str(); }

int gp_ctx_create(const gp_config *cfg, gp_ctx **out)
{
    if (!cfg || !out) return gp_fail(GP_ERR_ARG, "null argument");
    *out = nullptr;
    const gp_config &k = *cfg;
    if (k.weight_order != 1 && k.weight_order != 3 && k.weight_order != 5)
        return gp_fail(GP_ERR_ARG, "WeightingOrder must be 1, 3 or 5");
    if (k.boris_order != 2 && k.boris_order != 4 && k.boris_order != 6)
        return gp_fail(GP_ERR_ARG, "BorisOrder must be 2, 4 or 6");
    if (k.ngx < 2 || k.ngy < 2 || !(k.dx > 0) || !(k.dy > 0)) return gp_fail(GP_ERR_ARG, "bad grid");
    for (int w = 0; w < 4; w++) {
        if (k.pbc[w] < -1 || k.pbc[w] > 2) return gp_fail(GP_ERR_ARG, "bad particle boundary code");
        if (k.fbc[w] < 0 || k.fbc[w] > 2) return gp_fail(GP_ERR_ARG, "bad field boundary code");
    }
    if (k.em_model == 2 && !k.cur_dens) return gp_fail(GP_ERR_ARG, "the Darwin model needs cur_dens = 1");
    GP_TRY({
        gp_ctx *c = new gp_ctx();
        c->cfg = k;
        c->order = k.weight_order;
        c->ngb = k.weight_order / 2;   // Init.mm:140
        c->ngx = k.ngx; c->ngy = k.ngy;
        c->nx = k.ngx + 2 * c->ngb; c->ny = k.ngy + 2 * c->ngb;
        c->dxf = (float)k.dx; c->dyf = (float)k.dy;
        c->ntx = (k.ngx + GP_TILE - 1) / GP_TILE;
        c->nty = (k.ngy + GP_TILE - 1) / GP_TILE;
        c->ntiles = c->ntx * c->nty;
        c->tile_bits = 1;
        while ((1 << c->tile_bits) < c->ntiles) c->tile_bits++;
        hipError_t e = hipSetDevice(k.device);
        if (e != hipSuccess) { delete c; return gp_fail(GP_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)); }
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { delete c; return gp_fail(GP_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
        // this device's copy of the static reduce scratch (gp_reduce_noalloc)
        int rs = gp_noalloc_setup(k.device);
        if (rs) { gp_ctx_destroy(c); return rs; }
        int rc = gp_fields_init(c);
        if (rc) { gp_ctx_destroy(c); return rc; }
        e = hipHostMalloc((void **)&c->h_drift, 3 * sizeof(unsigned long long) * GP_MAX_DRIFT_SPECIES);
        if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_counters, 8 * sizeof(unsigned long long) * GP_MAX_SPECIES);
        if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_ke, sizeof(double) * GP_MAX_SPECIES);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->drift_ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&c->d_drift, c->h_drift, 0);
        if (e != hipSuccess) { gp_ctx_destroy(c); return gp_fail(GP_ERR_HIP, std::string("drift counters: ") + hipGetErrorString(e)); }
        *out = c;
        return GP_OK;
    })
}

static int species_setup(gp_ctx *c, const gp_species_desc *d, gp_species &S);

static void species_free(gp_species &S)
{
    void *ps[] = {S.a.x, S.a.y, S.a.vx, S.a.vy, S.a.vz, S.b.x, S.b.y, S.b.vx, S.b.vy, S.b.vz, S.tile_off,
                  S.rho_acc, S.jx_acc, S.jy_acc, S.dead, S.dead_tmp, S.counters, S.ke_part, S.ke_out,
                  S.key, S.key2, S.val, S.val2};
    for (void *p : ps) if (p) (void)hipFree(p);
    S = gp_species();
}

int gp_ctx_destroy(gp_ctx *c)
{
    if (!c) return GP_OK;
    (void)hipDeviceSynchronize();
    for (auto &S : c->sp) species_free(S);
    for (hipEvent_t e : c->tev_all) (void)hipEventDestroy(e);
    gp_nn_destroy(c);
    gp_fields_free(c);
    if (c->stage) (void)hipFree(c->stage);
    if (c->stage_on) (void)hipFree(c->stage_on);
    if (c->h_drift) (void)hipHostFree(c->h_drift);
    if (c->h_counters) (void)hipHostFree(c->h_counters);
    if (c->h_ke) (void)hipHostFree(c->h_ke);
    if (c->drift_ev) (void)hipEventDestroy(c->drift_ev);
    if (c->boff) (void)hipFree(c->boff);
    if (c->rb_buf) (void)hipFree(c->rb_buf);
    if (c->rng_s) (void)hipFree(c->rng_s);
    if (c->rng_inc) (void)hipFree(c->rng_inc);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GP_OK;
}

int gp_sync(gp_ctx *c)
{
    GP_CHECK_CTX(c);
    GP_HIP(hipStreamSynchronize(c->stream));
    return GP_OK;
}

int gp_dims(const gp_ctx *c, int out[6])
{
    if (!c || !out) return gp_fail(GP_ERR_ARG, "null argument");
    out[0] = c->nx; out[1] = c->ny; out[2] = c->ps.nkx; out[3] = c->ps.nky; out[4] = c->ps.ikmin; out[5] = c->ps.jkmin;
    return GP_OK;
}

int gp_species_create(gp_ctx *c, const gp_species_desc *d, int *sid)
{
    GP_CHECK_CTX(c);
    if (!d || !sid) return gp_fail(GP_ERR_ARG, "null argument");
    if ((int)c->sp.size() >= GP_MAX_SPECIES) return gp_fail(GP_ERR_CAPACITY, "too many species");
    if (d->capacity < 1 || d->capacity > 0x7fffff00LL) return gp_fail(GP_ERR_ARG, "capacity must be in [1, 2^31)");
    if (!(d->m > 0)) return gp_fail(GP_ERR_ARG, "mass must be positive");
    GP_TRY({
        c->sp.emplace_back();
        // a failed set-up frees what it allocated and removes the species
        const int rc = species_setup(c, d, c->sp.back());
        if (rc) {
            (void)hipStreamSynchronize(c->stream);
            species_free(c->sp.back());
            c->sp.pop_back();
            return rc;
        }
        *sid = (int)c->sp.size() - 1;
        return GP_OK;
    })
}

static int species_setup(gp_ctx *c, const gp_species_desc *d, gp_species &S)
{
    {
        S.q = d->q; S.m = d->m; S.w = d->w;
        S.qf = (float)d->q; S.mf = (float)d->m; S.wf = (float)d->w;
        // Particle.mm:540: q*w in float, divided by dx*dy in double
        S.scale = (float)((double)(S.qf * S.wf) / (c->cfg.dx * c->cfg.dy));
        double vn = d->vnorm > 0 ? d->vnorm : (d->vth_hint > 0 ? 8.0 * d->vth_hint : 2.99792458e10);
        S.vnorm = std::ldexp(1.0, (int)std::ceil(std::log2(vn)));
        S.inv_vnorm = 1.0 / S.vnorm;
        S.cap = d->capacity;
        const size_t cap = (size_t)d->capacity;
        auto al = [&](void **p, size_t bytes) -> int {
            hipError_t e = hipMalloc(p, bytes);
            if (e != hipSuccess) return gp_fail(GP_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
            return GP_OK;
        };
        int rc = 0;
        float **fa[] = {&S.a.x, &S.a.y, &S.a.vx, &S.a.vy, &S.a.vz, &S.b.x, &S.b.y, &S.b.vx, &S.b.vy, &S.b.vz};
        for (float **p : fa) if ((rc = al((void **)p, cap * sizeof(float)))) break;
        const size_t nN = (size_t)(c->nx + 1) * (c->ny + 1), nEx = (size_t)(c->nx + 2) * (c->ny + 1),
                     nEy = (size_t)(c->nx + 1) * (c->ny + 2);
        if (!rc) rc = al((void **)&S.tile_off, sizeof(int32_t) * (c->ntiles + 1));
        if (!rc) rc = al((void **)&S.rho_acc, nN * 8);
        if (!rc && c->cfg.cur_dens) rc = al((void **)&S.jx_acc, nEx * 8);
        if (!rc && c->cfg.cur_dens) rc = al((void **)&S.jy_acc, nEy * 8);
        if (!rc) rc = al((void **)&S.dead, cap * 4);
        if (!rc) rc = al((void **)&S.dead_tmp, cap * 4);
        if (!rc) rc = al((void **)&S.counters, 8 * 8);
        S.ke_nblk = c->ntiles + gp_cdiv((int64_t)cap, 16 * GP_BLOCK) + 1024 + 1;
        if (!rc) rc = al((void **)&S.ke_part, (size_t)S.ke_nblk * 8);
        if (!rc) rc = al((void **)&S.ke_out, 16);   // [0] kinetic energy, [1] sum of vz^2 (see vz_gen)
        if (!rc) rc = al((void **)&S.key, (cap + 1) * 4);
        if (!rc) rc = al((void **)&S.key2, (cap + 2) * 4);
        if (!rc) rc = al((void **)&S.val, (cap + 1) * 4);
        if (!rc) rc = al((void **)&S.val2, (cap + 1) * 4);
        if (rc) return rc;
        GP_HIP(hipMemsetAsync(S.tile_off, 0, sizeof(int32_t) * (c->ntiles + 1), c->stream));
        GP_HIP(hipMemsetAsync(S.rho_acc, 0, nN * 8, c->stream));
        if (c->cfg.cur_dens) {
            GP_HIP(hipMemsetAsync(S.jx_acc, 0, nEx * 8, c->stream));
            GP_HIP(hipMemsetAsync(S.jy_acc, 0, nEy * 8, c->stream));
        }
        GP_HIP(hipMemsetAsync(S.counters, 0, 64, c->stream));
        GP_HIP(hipMemsetAsync(S.ke_out, 0, 16, c->stream));
        GP_HIP(hipStreamSynchronize(c->stream));
        return GP_OK;
    }
}

int gp_species_upload(gp_ctx *c, int sid, int64_t n, const float *x, const float *y, const float *vx,
                      const float *vy, const float *vz)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    gp_species &S = c->sp[sid];
    if (n < 0 || n > S.cap) return gp_fail(GP_ERR_CAPACITY, "pNum > pNumMax on upload");
    if (n > 0 && (!x || !y || !vx || !vy || !vz)) return gp_fail(GP_ERR_ARG, "null particle array");
    const size_t b = (size_t)n * 4;
    if (n > 0) {
        GP_HIP(hipMemcpyAsync(S.a.x, x, b, hipMemcpyHostToDevice, c->stream));
        GP_HIP(hipMemcpyAsync(S.a.y, y, b, hipMemcpyHostToDevice, c->stream));
        GP_HIP(hipMemcpyAsync(S.a.vx, vx, b, hipMemcpyHostToDevice, c->stream));
        GP_HIP(hipMemcpyAsync(S.a.vy, vy, b, hipMemcpyHostToDevice, c->stream));
        GP_HIP(hipMemcpyAsync(S.a.vz, vz, b, hipMemcpyHostToDevice, c->stream));
    }
    GP_HIP(hipStreamSynchronize(c->stream));
    S.n = n;
    S.vz_gen++;
    S.n_binned = 0;
    return GP_OK;
}

int gp_species_download(gp_ctx *c, int sid, int64_t cap, int64_t *n, float *x, float *y, float *vx, float *vy,
                        float *vz)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    gp_species &S = c->sp[sid];
    if (n) *n = S.n;
    if (cap < S.n) return gp_fail(GP_ERR_CAPACITY, "host buffer smaller than the particle count");
    const size_t b = (size_t)S.n * 4;
    if (S.n > 0) {
        if (x) GP_HIP(hipMemcpyAsync(x, S.a.x, b, hipMemcpyDeviceToHost, c->stream));
        if (y) GP_HIP(hipMemcpyAsync(y, S.a.y, b, hipMemcpyDeviceToHost, c->stream));
        if (vx) GP_HIP(hipMemcpyAsync(vx, S.a.vx, b, hipMemcpyDeviceToHost, c->stream));
        if (vy) GP_HIP(hipMemcpyAsync(vy, S.a.vy, b, hipMemcpyDeviceToHost, c->stream));
        if (vz) GP_HIP(hipMemcpyAsync(vz, S.a.vz, b, hipMemcpyDeviceToHost, c->stream));
    }
    GP_HIP(hipStreamSynchronize(c->stream));
    return GP_OK;
}

int gp_species_read(gp_ctx *c, int sid, int64_t first, int64_t count, int64_t *n, float *x, float *y, float *vx,
                    float *vy, float *vz)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    gp_species &S = c->sp[sid];
    if (first < 0 || count < 0) return gp_fail(GP_ERR_ARG, "negative range");
    const int64_t m = first >= S.n ? 0 : std::min(count, S.n - first);
    if (n) *n = m;
    const size_t b = (size_t)m * 4, o = (size_t)first;
    if (m > 0) {
        if (x) GP_HIP(hipMemcpyAsync(x, S.a.x + o, b, hipMemcpyDeviceToHost, c->stream));
        if (y) GP_HIP(hipMemcpyAsync(y, S.a.y + o, b, hipMemcpyDeviceToHost, c->stream));
        if (vx) GP_HIP(hipMemcpyAsync(vx, S.a.vx + o, b, hipMemcpyDeviceToHost, c->stream));
        if (vy) GP_HIP(hipMemcpyAsync(vy, S.a.vy + o, b, hipMemcpyDeviceToHost, c->stream));
        if (vz) GP_HIP(hipMemcpyAsync(vz, S.a.vz + o, b, hipMemcpyDeviceToHost, c->stream));
    }
    GP_HIP(hipStreamSynchronize(c->stream));
    return GP_OK;
}

int gp_species_count(gp_ctx *c, int sid, int64_t *n)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    if (!n) return gp_fail(GP_ERR_ARG, "null argument");
    *n = c->sp[sid].n;
    return GP_OK;
}

int gp_species_grid_get(gp_ctx *c, int sid, int which, int64_t *host, int64_t count)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    gp_species &S = c->sp[sid];
    const int64_t nN = (int64_t)(c->nx + 1) * (c->ny + 1), nEx = (int64_t)(c->nx + 2) * (c->ny + 1),
                  nEy = (int64_t)(c->nx + 1) * (c->ny + 2);
    const int64_t *src = nullptr;
    int64_t need = 0;
    if (which == GP_G_RHO_ACC) { src = S.rho_acc; need = nN; }
    else if (which == GP_G_JX_ACC) { src = S.jx_acc; need = nEx; }
    else if (which == GP_G_JY_ACC) { src = S.jy_acc; need = nEy; }
    else return gp_fail(GP_ERR_ARG, "bad accumulator id");
    if (!src) return gp_fail(GP_ERR_STATE, "current accumulators need cur_dens = 1");
    if (count != need || !host) return gp_fail(GP_ERR_ARG, "accumulator size mismatch");
    GP_HIP(hipMemcpyAsync(host, src, need * 8, hipMemcpyDeviceToHost, c->stream));
    GP_HIP(hipStreamSynchronize(c->stream));
    return GP_OK;
}

int gp_load_pair(gp_ctx *c, int se, int si, const gp_gen *g, uint64_t seed, int64_t first_id, int64_t n)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, se);
    if (si >= 0) GP_CHECK_SID(c, si);
    if (!g || n < 0 || first_id < 0) return gp_fail(GP_ERR_ARG, "bad generation arguments");
    if (g->gen_type < 0 || g->gen_type > 7) return gp_fail(GP_ERR_ARG, "unknown generation type");
    if (!(g->xmax > g->xmin) || !(g->ymax > g->ymin)) return gp_fail(GP_ERR_ARG, "empty generation region");
    int rc = gp_launch_load(c, se, si, g, seed, first_id, n);
    if (rc) return rc;
    c->sp[se].n_binned = std::min(c->sp[se].n_binned, c->sp[se].n);
    GP_HIP(hipStreamSynchronize(c->stream));
    return GP_OK;
}

int gp_generate(gp_ctx *c, int se, int si, const gp_gen *g, int64_t addn, int deposit)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, se);
    if (si >= 0) GP_CHECK_SID(c, si);
    if (!g || addn < 0) return gp_fail(GP_ERR_ARG, "bad generation arguments");
    if (g->gen_type < 0 || g->gen_type > 7) return gp_fail(GP_ERR_ARG, "unknown generation type");
    // a zero-width region is a line source (the hollow cathode's GeneratePosX)
    if (!(g->xmax >= g->xmin) || !(g->ymax >= g->ymin)) return gp_fail(GP_ERR_ARG, "inverted generation region");
    if (g->gen_type == 4 && !(g->xmax > g->xmin && g->ymax > g->ymin))
        return gp_fail(GP_ERR_ARG, "quiet start needs a region of positive area");
    GP_TRY({
        int rc = gp_launch_generate(c, se, si, g, addn, deposit);
        if (rc) return rc;
    })
    GP_HIP(hipStreamSynchronize(c->stream));
    return GP_OK;
}

int gp_set_cathode(gp_ctx *c, double x_cm)
{
    GP_CHECK_CTX(c);
    if (!(x_cm == x_cm)) return gp_fail(GP_ERR_ARG, "hollow-cathode position is NaN");
    if (x_cm > 0.0 && c->cfg.fbc[0] == GP_FBC_PERIODIC)
        return gp_fail(GP_ERR_ARG, "the hollow-cathode ramp needs non-periodic x walls");
    c->cathode_pos = x_cm > 0.0 ? x_cm : -1.0;
    return GP_OK;
}

int gp_inject_wall(gp_ctx *c, int sid, int64_t n, float vth, const float u[3], double dt, uint64_t seed,
                   uint64_t stream, int deposit)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    if (n < 0 || !u || !(vth >= 0) || !(dt > 0)) return gp_fail(GP_ERR_ARG, "bad injection arguments");
    return gp_launch_inject_wall(c, sid, n, vth, u, dt, seed, stream, deposit);
}

int gp_reset_grids(gp_ctx *c)
{
    GP_CHECK_CTX(c);
    const size_t nN = (size_t)(c->nx + 1) * (c->ny + 1), nEx = (size_t)(c->nx + 2) * (c->ny + 1),
                 nEy = (size_t)(c->nx + 1) * (c->ny + 2);
    for (auto &S : c->sp) {
        GP_HIP(hipMemsetAsync(S.rho_acc, 0, nN * 8, c->stream));
        if (S.jx_acc) GP_HIP(hipMemsetAsync(S.jx_acc, 0, nEx * 8, c->stream));
        if (S.jy_acc) GP_HIP(hipMemsetAsync(S.jy_acc, 0, nEy * 8, c->stream));
    }
    return GP_OK;
}

int gp_deposit(gp_ctx *c, int sid)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    return gp_launch_deposit(c, sid, 0, c->sp[sid].n);
}

int gp_solve_fields(gp_ctx *c, double dt, double tol, int maxiter, int solver, gp_solve_stats *st, gp_diag *diag)
{
    GP_CHECK_CTX(c);
    if (!(tol > 0) || maxiter < 1) return gp_fail(GP_ERR_ARG, "tol must be > 0 and maxiter >= 1");
    GP_TRY({
        if (c->timing) { int rc = gp_time_mark(c, GP_TIME_SOLVE); if (rc) return rc; }
        int rc = gp_solve_fields_impl(c, dt, tol, maxiter, solver, st, diag);
        if (c->timing && rc >= 0) { int r2 = gp_time_mark(c, GP_TIME_SOLVE); if (r2) return r2; }
        return rc;
    })
}

int gp_solve_rhs(gp_ctx *c, const double *b, double *x, double tol, int maxiter, int solver, gp_solve_stats *st)
{
    GP_CHECK_CTX(c);
    if (!b || !x) return gp_fail(GP_ERR_ARG, "null vector");
    if (!(tol > 0) || maxiter < 1) return gp_fail(GP_ERR_ARG, "tol must be > 0 and maxiter >= 1");
    GP_TRY({ return gp_solve_rhs_impl(c, b, x, tol, maxiter, solver, st); })
}

int gp_solve_darwin(gp_ctx *c, double dt, double tol, int maxiter, int solver, gp_solve_stats st[3])
{
    GP_CHECK_CTX(c);
    if (!st) return gp_fail(GP_ERR_ARG, "null stats");
    GP_TRY({
        const int rc = gp_solve_darwin_impl(c, dt, tol, maxiter, solver, st);
        if (rc >= 0)
            for (int k = 0; k < 3; k++) c->last_darwin[k] = st[k];
        return rc;
    })
}

int gp_darwin_stats(gp_ctx *c, gp_solve_stats out[3])
{
    GP_CHECK_CTX(c);
    if (!out) return gp_fail(GP_ERR_ARG, "null stats");
    for (int k = 0; k < 3; k++) out[k] = c->last_darwin[k];
    return GP_OK;
}

int gp_push(gp_ctx *c, int sid, double dt, int deposit)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    if (!(dt > 0)) return gp_fail(GP_ERR_ARG, "dt must be positive");
    return gp_launch_push(c, sid, dt, deposit);
}

int gp_reduce(gp_ctx *c, int sid, int64_t counts[4], double *kinetic)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    int64_t tmp[4];
    return gp_launch_reduce(c, sid, counts ? counts : tmp, kinetic);
}

int gp_reduce_noalloc(gp_ctx *c, int sid, int64_t counts[4], double *kinetic)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    gp_arena ar;
    const int rc = gp_noalloc_arena(c->cfg.device, &ar);
    if (rc) return rc;
    int64_t tmp[4];
    return gp_launch_reduce(c, sid, counts ? counts : tmp, kinetic, &ar);
}

int gp_sort(gp_ctx *c, int sid)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    int rc = gp_launch_sort(c, sid);
    if (rc) return rc;
    return GP_OK;
}

int gp_stream_probe(gp_ctx *c, int sid, int reps, double *ms_per_launch)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    if (reps < 1 || !ms_per_launch) return gp_fail(GP_ERR_ARG, "reps must be >= 1 and the output non-null");
    return gp_launch_stream_probe(c, sid, reps, ms_per_launch);
}

int gp_moments(gp_ctx *c, int sid, double *host10)
{
    GP_CHECK_CTX(c);
    GP_CHECK_SID(c, sid);
    if (!host10) return gp_fail(GP_ERR_ARG, "null output");
    return gp_launch_moments(c, sid, host10);
}

int gp_step(gp_ctx *c, double dt, double tol, int maxiter, int solver, int sort_interval, int cycle,
            gp_solve_stats *st, gp_diag *diag)
{
    GP_CHECK_CTX(c);
    if (!(dt > 0)) return gp_fail(GP_ERR_ARG, "dt must be positive");
    int status = gp_solve_fields(c, dt, tol, maxiter, solver, st, diag);
    if (status < 0) return status;
    if (c->cfg.em_model == 2) {
        gp_solve_stats ds[3];
        int rc = gp_solve_darwin(c, dt, tol, maxiter, solver, ds);
        if (rc < 0) return rc;
        if (rc == GP_NOT_CONVERGED) status = GP_NOT_CONVERGED;
    }
    bool has_delete = false;
    for (int w = 0; w < 4; w++) has_delete |= c->cfg.pbc[w] == GP_PBC_DELETE;
    const int nsp = (int)c->sp.size();
    if (c->drift_pending) {
        // drift counts of the previous step's pushes (copied after them; the
        // field solve has synchronised the stream since)
        GP_HIP(hipEventSynchronize(c->drift_ev));
        c->drift_pending = 0;
        for (int s = 0; s < nsp && s < GP_MAX_DRIFT_SPECIES; s++) {
            gp_species &S = c->sp[s];
            // a live particle that left the domain through a side with no
            // boundary condition (gp_reduce reports it too, when it runs)
            if (c->h_drift[3 * s])
                return gp_fail(GP_ERR_STATE, "particles outside the domain with no boundary condition: " +
                                                 std::to_string(c->h_drift[3 * s]));
            S.last_drift = (int64_t)c->h_drift[3 * s + 1];
            S.last_moved = (int64_t)c->h_drift[3 * s + 2] * (c->ntiles >= 256 ? GP_MOVED_SAMPLE : 1);
            if (c->sort_drift > 0.0 && (double)S.last_drift > c->sort_drift * (double)S.n) {
                int rc = gp_launch_sort(c, s);
                if (rc) return rc;
            }
        }
    }
    for (int s = 0; s < nsp; s++) {
        int rc = gp_launch_push(c, s, dt, 1);
        if (rc) return rc;
    }
    if (c->h_drift) {
        DriftOut d{};
        d.ns = std::min(nsp, GP_MAX_DRIFT_SPECIES);
        for (int s = 0; s < d.ns; s++) d.cnt[s] = c->sp[s].counters;
        k_drift_out<<<1, 64, 0, c->stream>>>(d, c->d_drift);
        GP_HIP(hipGetLastError());
        GP_HIP(hipEventRecord(c->drift_ev, c->stream));
        c->drift_pending = 1;
    }
    if (has_delete || diag) {
        // every species' counters in one synchronisation; the compactions
        // then run without any, and only the diagnostics wait for the stream
        for (int s = 0; s < nsp; s++)
            GP_HIP(hipMemcpyAsync(c->h_counters + 8 * s, c->sp[s].counters, 8 * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, c->stream));
        GP_HIP(hipStreamSynchronize(c->stream));
        int64_t cnt[GP_MAX_SPECIES][4] = {};
        for (int s = 0; s < nsp; s++) {
            int rc = gp_launch_reduce_from(c, s, c->h_counters + 8 * s, cnt[s], diag ? c->h_ke + s : nullptr);
            if (rc) return rc;
        }
        if (diag) {
            GP_HIP(hipStreamSynchronize(c->stream));
            for (int s = 0; s < nsp && s < 8; s++) {
                diag->kinetic[s] = c->h_ke[s];
                diag->count[s] = c->sp[s].n;
                for (int w = 0; w < 4; w++) diag->absorbed[s][w] = cnt[s][w];
            }
        }
    }
    if (sort_interval > 0 && cycle % sort_interval == 0)
        for (int s = 0; s < (int)c->sp.size(); s++) {
            int rc = gp_launch_sort(c, s);
            if (rc) return rc;
        }
    return status;
}

int gp_set_sort_drift(gp_ctx *c, double fraction)
{
    GP_CHECK_CTX(c);
    if (!(fraction >= 0.0) || fraction > 1.0) return gp_fail(GP_ERR_ARG, "sort drift fraction must lie in [0, 1]");
    c->sort_drift = fraction;
    return GP_OK;
}

int gp_sort_stats(gp_ctx *c, int sid, int64_t out[2])
{
    GP_CHECK_CTX(c);
    if (sid < 0 || sid >= (int)c->sp.size() || !out) return gp_fail(GP_ERR_ARG, "bad species or null output");
    out[0] = c->sp[sid].nsorts;
    out[1] = c->sp[sid].last_drift;
    return GP_OK;
}

int gp_field_get(gp_ctx *c, int which, void *host, int64_t bytes)
{
    GP_CHECK_CTX(c);
    if (!host) return gp_fail(GP_ERR_ARG, "null output");
    const int64_t nx = c->nx, ny = c->ny;
    const void *src = nullptr;
    int64_t need = 0;
    switch (which) {
    case GP_F_RHO: src = c->rho; need = (nx + 1) * (ny + 1) * 8; break;
    case GP_F_EX: src = c->Ex; need = (nx + 2) * (ny + 1) * 4; break;
    case GP_F_EY: src = c->Ey; need = (nx + 1) * (ny + 2) * 4; break;
    case GP_F_EZ: src = c->Ez; need = (nx + 1) * (ny + 1) * 4; break;
    case GP_F_BX: src = c->Bx; need = (nx + 1) * (ny + 2) * 4; break;
    case GP_F_BY: src = c->By; need = (nx + 2) * (ny + 1) * 4; break;
    case GP_F_BZ: src = c->Bz; need = (nx + 2) * (ny + 2) * 4; break;
    case GP_F_PHI: src = c->phi; need = (nx + 3) * (ny + 3) * 8; break;
    case GP_F_JX: src = c->jx; need = (nx + 2) * (ny + 1) * 8; break;
    case GP_F_JY: src = c->jy; need = (nx + 1) * (ny + 2) * 8; break;
    case GP_F_DEL_EX: src = c->dEx; need = (nx + 2) * (ny + 1) * 4; break;
    case GP_F_DEL_EY: src = c->dEy; need = (nx + 1) * (ny + 2) * 4; break;
    case GP_F_DEL_BZ: src = c->dBz; need = (nx + 2) * (ny + 2) * 4; break;
    case GP_F_SOL: src = c->ps.x64; need = (int64_t)c->ps.n * 8; break;
    case GP_F_CHI: src = c->dsol_chi; need = (int64_t)c->ds.n * 8; break;
    case GP_F_AX: src = c->dsol_ax; need = (int64_t)c->ds.n * 8; break;
    case GP_F_AY: src = c->dsol_ay; need = (int64_t)c->ds.n * 8; break;
    case GP_F_RHS: src = c->ps.b64; need = (int64_t)c->ps.n * 8; break;
    default: return gp_fail(GP_ERR_ARG, "unknown field id");
    }
    if (!src) return gp_fail(GP_ERR_STATE, "field not allocated (Darwin fields need em_model = 2)");
    if (bytes != need) return gp_fail(GP_ERR_ARG, "field size mismatch: expected " + std::to_string(need) + " bytes");
    GP_HIP(hipMemcpyAsync(host, src, need, hipMemcpyDeviceToHost, c->stream));
    GP_HIP(hipStreamSynchronize(c->stream));
    return GP_OK;
}

int gp_field_set(gp_ctx *c, int which, const void *host, int64_t bytes)
{
    GP_CHECK_CTX(c);
    if (!host) return gp_fail(GP_ERR_ARG, "null input");
    const int64_t nx = c->nx, ny = c->ny;
    void *dst = nullptr;
    int64_t need = 0;
    switch (which) {
    case GP_F_EX: dst = c->Ex; need = (nx + 2) * (ny + 1) * 4; c->e_periodic = 0; break;
    case GP_F_EY: dst = c->Ey; need = (nx + 1) * (ny + 2) * 4; c->e_periodic = 0; break;
    case GP_F_EZ: dst = c->Ez; need = (nx + 1) * (ny + 1) * 4; break;
    case GP_F_BX: dst = c->Bx; need = (nx + 1) * (ny + 2) * 4; break;
    case GP_F_BY: dst = c->By; need = (nx + 2) * (ny + 1) * 4; break;
    case GP_F_BZ: dst = c->Bz; need = (nx + 2) * (ny + 2) * 4; c->uniform_b = 0; break;
    case GP_F_SOL: dst = c->ps.x64; need = (int64_t)c->ps.n * 8; break;
    default: return gp_fail(GP_ERR_ARG, "field is not settable");
    }
    if (bytes != need) return gp_fail(GP_ERR_ARG, "field size mismatch: expected " + std::to_string(need) + " bytes");
    GP_HIP(hipMemcpyAsync(dst, host, need, hipMemcpyHostToDevice, c->stream));
    GP_HIP(hipStreamSynchronize(c->stream));
    return GP_OK;
}

int gp_field_energy(gp_ctx *c, double dt, double *we_ref, double *we_cgs, double *joule)
{
    GP_CHECK_CTX(c);
    return gp_energy_impl(c, dt, we_ref, we_cgs, joule);
}

int gp_timing(gp_ctx *c, int enable)
{
    GP_CHECK_CTX(c);
    int rc = gp_time_collect(c);
    if (rc) return rc;
    c->timing = enable ? 1 : 0;
    for (int k = 0; k < GP_TIME_SLOTS; k++) { c->t_ms[k] = 0; c->t_n[k] = 0; }
    return GP_OK;
}

int gp_kernel_time(gp_ctx *c, int which, double *total_ms, int64_t *launches)
{
    GP_CHECK_CTX(c);
    if (which < 0 || which >= GP_TIME_SLOTS || !total_ms || !launches) return gp_fail(GP_ERR_ARG, "bad timing query");
    int rc = gp_time_collect(c);
    if (rc) return rc;
    *total_ms = c->t_ms[which];
    *launches = c->t_n[which];
    return GP_OK;
}

int gp_nn_init(gp_ctx *c, int base_channels, uint64_t seed)
{
    GP_CHECK_CTX(c);
    GP_TRY({ return gp_nn_create(c, base_channels, seed); })
}
int gp_nn_train(gp_ctx *c, const double *rho, const double *phi, int nsamp, int epochs, int batch, float lr,
                float *loss_out)
{
    GP_CHECK_CTX(c);
    GP_TRY({ return gp_nn_train_impl(c, rho, phi, nsamp, epochs, batch, lr, loss_out); })
}
int gp_nn_infer(gp_ctx *c, const double *rhs, double *out)
{
    GP_CHECK_CTX(c);
    GP_TRY({ return gp_nn_infer_impl(c, rhs, out); })
}
int gp_nn_save(gp_ctx *c, const char *path)
{
    GP_CHECK_CTX(c);
    GP_TRY({ return gp_nn_save_impl(c, path); })
}
int gp_nn_load(gp_ctx *c, const char *path)
{
    GP_CHECK_CTX(c);
    GP_TRY({ return gp_nn_load_impl(c, path); })
}
int gp_nn_add_pair(gp_ctx *c)
{
    GP_CHECK_CTX(c);
    if (!c->nn) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    if (!c->have_solution) return gp_fail(GP_ERR_STATE, "no Poisson solution yet");
    if (gp_nn_train_system(c) != 0) return gp_fail(GP_ERR_STATE, "the training set is for the Darwin hierarchy");
    GP_TRY({ return gp_nn_add_pair_dev(c, c->ps.b64, c->ps.x64); })
}
int gp_nn_add_increment(gp_ctx *c, int *added)
{
    GP_CHECK_CTX(c);
    if (!c->nn) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    if (!c->have_solution) return gp_fail(GP_ERR_STATE, "no Poisson solution yet");
    if (gp_nn_train_system(c) != 0) return gp_fail(GP_ERR_STATE, "the training set is for the Darwin hierarchy");
    int tmp = 0;
    GP_TRY({ return gp_nn_add_increment_dev(c, 0, c->ps.b64, c->ps.x64, (int64_t)c->ps.n, added ? added : &tmp); })
}
int gp_nn_add_darwin_increments(gp_ctx *c, int *added)
{
    GP_CHECK_CTX(c);
    if (!c->nn) return gp_fail(GP_ERR_STATE, "gp_nn_init first");
    if (c->cfg.em_model != 2 || !c->dsol_chi) return gp_fail(GP_ERR_STATE, "no Darwin solve (em_model = 2)");
    if (gp_nn_train_system(c) != 1) return gp_fail(GP_ERR_STATE, "gp_nn_set_system(ctx, 1) first");
    GP_TRY({
        const int64_t m = (int64_t)(c->ds.nkx + 1) * (c->ds.nky + 1);
        const double *bs[3] = {c->dar_bchi, c->dar_bx, c->dar_by};
        const double *xs[3] = {c->dsol_chi, c->dsol_ax, c->dsol_ay};
        int tot = 0;
        for (int k = 0; k < 3; k++) {
            int a = 0;
            int rc = gp_nn_add_increment_dev(c, k + 1, bs[k], xs[k], m, &a);
            if (rc) return rc;
            tot += a;
        }
        if (added) *added = tot;
        return GP_OK;
    })
}
int gp_nn_set_system(gp_ctx *c, int sys)
{
    GP_CHECK_CTX(c);
    GP_TRY({ return gp_nn_set_system_impl(c, sys); })
}
int gp_nn_config(gp_ctx *c, int loss_kind, int smooth_sweeps, int fcg_cap)
{
    GP_CHECK_CTX(c);
    GP_TRY({ return gp_nn_config_impl(c, loss_kind, smooth_sweeps, fcg_cap); })
}
int gp_nn_params(gp_ctx *c, float *get, const float *set, int64_t *count)
{
    GP_CHECK_CTX(c);
    GP_TRY({ return gp_nn_params_impl(c, get, set, count); })
}
int gp_nn_dims(gp_ctx *c, int *H, int *W)
{
    GP_CHECK_CTX(c);
    if (!H || !W) return gp_fail(GP_ERR_ARG, "null output");
    const int mx = c->ps.lv[0].mx, my = c->ps.lv[0].my;
    *W = (mx + 7) / 8 * 8;
    *H = (my + 7) / 8 * 8;
    return GP_OK;
}
int gp_nn_forward(gp_ctx *c, const float *in, int B, float *out)
{
    GP_CHECK_CTX(c);
    if (!in || !out || B < 1) return gp_fail(GP_ERR_ARG, "bad forward arguments");
    GP_TRY({ return gp_nn_forward_impl(c, in, B, out); })
}
int gp_nn_grad(gp_ctx *c, const float *in, const float *tgt, int B, int kind, float *loss, float *grad)
{
    GP_CHECK_CTX(c);
    if (!in || !tgt || B < 1 || kind < 0 || kind > 1) return gp_fail(GP_ERR_ARG, "bad gradient arguments");
    GP_TRY({ return gp_nn_grad_impl(c, in, tgt, B, kind, loss, grad); })
}

}  // extern "C"
