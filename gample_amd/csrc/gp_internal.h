// gp_internal.h -- device-side layout and host-side context of libgample_amd.
//
// Particles: struct-of-arrays (x, y, vx, vy, vz float), positions in grid
// units shifted by the ghost width ngb as in the reference
// (Particle.mm:589-594).  Each species owns two SoA buffers (live + scratch)
// for the sort and the ping-pong of the gather step.  An absorbed particle
// is marked by x = -piflag (1..4) until gp_reduce removes it; positions in
// the domain are always >= 0, so no separate flag array is stored.
//
// Storage order is tile-binned: after gp_sort the particles of tile t
// (TILE x TILE cells) occupy [tile_off[t], tile_off[t+1]); a workgroup
// processes one tile with the tile's field stencil and charge accumulator
// staged in LDS.  Particles past tile_off[ntiles] (appended after the last
// sort) are processed by extra workgroups through global memory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/gample_amd.h"

#define GP_TILE 16           // tile edge, cells
#ifndef GP_DRIFT
#define GP_DRIFT 12          // cells a particle may leave its tile and still use LDS
#endif
#define GP_BLOCK 256         // threads per workgroup of the particle kernels
#define GP_MAX_DRIFT_SPECIES 64   // species followed by drift-triggered sorting
#define GP_MOVED_SAMPLE 8         // one tile in 8 counts cell changes (grids of 256+ tiles)
#define GP_MAX_SPECIES 8
// timing slots: all pushes, all sorts, field solves, all reduces, then per
// species s the push, reduce and sort at GP_TIME_SPECIES + 3 s + 0, 1, 2,
// then the push kernel alone
#define GP_TIME_PUSH 0
#define GP_TIME_SORT 1
#define GP_TIME_SOLVE 2
#define GP_TIME_REDUCE 3
#define GP_TIME_SPECIES 4
#define GP_TIME_PUSHK (GP_TIME_SPECIES + 3 * GP_MAX_SPECIES)   // the push kernel launch alone
#define GP_TIME_SLOTS (GP_TIME_PUSHK + 1)
#define GP_RHO_FRAC 46       // fractional bits of the rho accumulators
#define GP_CUR_FRAC 40       // fractional bits of current/moment accumulators

// status plumbing (gp_api.hip)
int gp_fail(int code, const std::string &msg);
#define GP_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return gp_fail(GP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

static inline size_t gp_align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bump allocator over a caller-provided device buffer: the scratch of the
// allocation-free entry points.  take() returns null when the request does
// not fit; a user restores `used` to release what it took.
struct gp_arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void *take(size_t bytes)
    {
        const size_t b = gp_align256(bytes);
        if (b > cap - used) return nullptr;
        void *p = base + used;
        used += b;
        return p;
    }
};

struct gp_soa {
    float *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
};

struct gp_species {
    double q = 0, m = 0, w = 0;
    float qf = 0, mf = 0, wf = 0;     // the reference keeps these as float members
    float scale = 0;                  // q*w/(dx*dy) as Particle.mm:540 computes it
    double vnorm = 1, inv_vnorm = 1;  // power of two
    int64_t cap = 0, n = 0;
    gp_soa a, b;                      // live, scratch
    int64_t n_binned = 0;             // particles covered by tile_off (<= n)
    int32_t *tile_off = nullptr;      // ntiles+1
    // accumulators (int64 fixed point), reference layouts
    int64_t *rho_acc = nullptr, *jx_acc = nullptr, *jy_acc = nullptr;
    // absorption bookkeeping
    uint32_t *dead = nullptr;         // indices of absorbed particles (cap)
    uint32_t *dead_tmp = nullptr;
    unsigned long long *counters = nullptr;  // [0] n_dead, [1..4] per wall, [5] escaped, [6] drifted
    int64_t nsorts = 0;               // sorts so far
    int64_t last_drift = -1;          // drifted particles of the last push read back (-1: none)
    int64_t last_moved = -1;          // particles that changed cell in that push (-1: none)
    // vz never changes in a push (Ez = Bx = By = 0), so the particles-periodic
    // hot push does not read it: its kinetic energy adds the sum of vz^2 over
    // the pushed range (ke_out[1]), recomputed when vz_gen (bumped by every
    // change of the particle set or of vz) or the range differs
    int64_t vz_gen = 0, vz2_gen = -1, vz2_first = -1, vz2_end = -1;
    // the last push's kinetic-energy partials (ke_part) not yet summed into
    // ke_out: their count (0: none), whether ke_out[1] adds in, the factor;
    // summed when the energy is read, dropped by the next push
    int ke_pending_n = 0, ke_pending_novz = 0;
    double ke_pending_coef = 0.0;
    double *ke_part = nullptr;        // per-workgroup kinetic-energy partials
    int ke_nblk = 0;
    double *ke_out = nullptr;         // device scalar
    // sort scratch
    uint32_t *key = nullptr, *key2 = nullptr, *val = nullptr, *val2 = nullptr;
};

// multigrid axis types
#define GP_AX_PER 0
#define GP_AX_DIR 1
#define GP_AX_NEU 2

struct gp_mg_level {
    int mx, my;          // unknowns per axis
    int perx, pery;      // axis types GP_AX_* of x and y
    double hx2i, hy2i;   // 1/hx^2, 1/hy^2 at this level
    float *u = nullptr;  // correction (output of the level)
    float *f = nullptr;  // rhs
    float *r = nullptr;  // pre-smoothed correction (down leg -> up leg)
};

struct gp_solver {
    // unknown grid of the fine problem
    int nkx = 0, nky = 0, ikmin = 0, jkmin = 0, n = 0;
    int singular = 0;
    std::vector<gp_mg_level> lv;
    // FP64 fine vectors
    double *b64 = nullptr, *x64 = nullptr, *r64 = nullptr, *rhs_bc = nullptr;
    // FP32 PCG vectors
    float *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *q2 = nullptr, *e = nullptr;
    int bottom = 0;            // first level solved inside one workgroup
    double *red = nullptr;     // reduction partials
    double *scal = nullptr;    // device scalars
    float *dinv = nullptr;     // dense inverse of the coarsest level (column-major), or null
    // fast diagonalisation of the coarsest level when it is too large for a
    // dense inverse: per-axis eigenvectors V (row-major m x m) and the
    // per-mode inverse eigenvalues 1 / (cx lx_i + cy ly_j) (0 for a null mode)
    float *fd_vx = nullptr, *fd_vy = nullptr, *fd_inv = nullptr;
    // MFMA fast diagonalisation of a coarsest level of at most 64 x 64
    // unknowns (or null): Vx, Vy and the inverse eigenvalues, each zero-padded
    // to 64 x 64 row-major, in one allocation
    float *fd64 = nullptr;
    double *hscal = nullptr;   // pinned host mirror
    int red_blocks = 0;
    int last_it = 4;           // MG-PCG iterations of the previous solve (its first batch)
    int rep_every = 4;         // MG-PCG iterations between residual replacements
    int mg_ref_it = 0;         // iterations of the last plain MG-PCG solve
};

struct gp_nn;   // gp_nn.hip

struct gp_ctx {
    gp_config cfg;
    int ngx, ngy, ngb, order, nx, ny;
    float dxf, dyf;                 // EMField keeps dx, dy as float (EMField.mm:66-67)
    int ntx, nty, ntiles, tile_bits;
    hipStream_t stream = nullptr;
    std::vector<gp_species> sp;
    // fields, reference layouts (EMField.mm:133-141)
    float *Ex = nullptr, *Ey = nullptr, *Ez = nullptr, *Bx = nullptr, *By = nullptr, *Bz = nullptr;
    float *dEx = nullptr, *dEy = nullptr, *dBz = nullptr;
    double *rho = nullptr, *jx = nullptr, *jy = nullptr, *phi = nullptr;
    int uniform_b = 1;
    float bz0 = 0;
    // Ex, Ey exactly periodic on the periodic axes (uniform at set-up, or the
    // gradient of the periodic potential); cleared when a caller sets them
    int e_periodic = 1;
    gp_solver ps;                   // Poisson
    gp_solver ds;                   // Darwin (Neumann/periodic, singular)
    double *dsol_chi = nullptr, *dsol_ax = nullptr, *dsol_ay = nullptr, *last_ax = nullptr, *last_ay = nullptr;
    double *dar_bx = nullptr, *dar_by = nullptr;   // Ax, Ay right-hand sides
    double *dar_bchi = nullptr;                    // chi right-hand side of the last Darwin solve
    gp_solve_stats last_darwin[3] = {};   // chi, Ax, Ay of the last Darwin solve
    double *red = nullptr;          // generic reduction scratch
    double *hred = nullptr;         // pinned
    int have_solution = 0;
    // timing: event pairs recorded on the context stream around each launch
    // of a slot (GP_TIME_* below), read back on demand; an event may close a
    // pair in two slots (a species' push and all pushes), so the events are
    // owned by tev_all
    int timing = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> tev[GP_TIME_SLOTS];
    std::vector<hipEvent_t> tev_all;
    double t_ms[GP_TIME_SLOTS] = {};
    int64_t t_n[GP_TIME_SLOTS] = {};
    gp_nn *nn = nullptr;
    // staged tile windows of the deposit (gp_particles.hip)
    void *stage = nullptr;
    uint8_t *stage_on = nullptr;
    size_t stage_bytes = 0;
    uint32_t *boff = nullptr;       // sort: first position of each (tile, class) bucket
    uint32_t *rb_buf = nullptr;     // neighbour re-bin: per-tile and per-segment counts
    int64_t rb_words = 0;
    // drift-triggered sorting: a species is sorted before its push when the
    // previous push gathered more than sort_drift * n of its particles
    // outside their tile windows (0: off)
    double sort_drift = 0.0;
    unsigned long long *h_drift = nullptr;   // pinned, three per species: escaped, drifted, moved
    unsigned long long *d_drift = nullptr;   // h_drift as the device addresses it
    unsigned long long *h_counters = nullptr;   // pinned, the eight counters of each species (reduce)
    double *h_ke = nullptr;                     // pinned, each species' kinetic energy (reduce)
    hipEvent_t drift_ev = nullptr;
    int drift_pending = 0;
    // generation streams (Collision.mm:189-336, 464-678): thread id of every
    // generation call draws from one PCG32 stream, seeded by splitmix64
    // outputs 2 id + 1 and 2 id + 2 of the master seed and continued from
    // call to call.  States are materialised on the device only for the ids
    // that gp_generate has used (rng_n); a seeded load records its id range
    // and draw count, and the states of ids it covered but that are not yet
    // materialised are recomputed from these records when first needed.
    uint64_t *rng_s = nullptr, *rng_inc = nullptr;
    int64_t rng_cap = 0, rng_n = 0;
    uint64_t rng_seed = 0;
    int rng_seeded = 0;
    struct rng_load { int64_t first, n; int draws; };
    std::vector<rng_load> rng_loads;
    // hollow-cathode potential ramp (EMField.mm:1063-1077): x position in cm
    // (<= 0: off) and the last solve's mean potential on that column (statV)
    double cathode_pos = -1.0;
    double cathode_mean = 0.0;
};

// ---- particle side (gp_particles.hip) ----
int gp_launch_push(gp_ctx *c, int sid, double dt, int deposit);
int gp_launch_deposit(gp_ctx *c, int sid, int64_t first, int64_t count);
int gp_launch_stream_probe(gp_ctx *c, int sid, int reps, double *ms);
int gp_launch_reduce(gp_ctx *c, int sid, int64_t counts[4], double *ke, gp_arena *ar = nullptr);
int gp_launch_reduce_from(gp_ctx *c, int sid, const unsigned long long *h, int64_t counts[4], double *ke,
                          gp_arena *ar = nullptr);
// the device's static reduce scratch (set up once per device by gp_ctx_create)
int gp_noalloc_setup(int device);
int gp_noalloc_arena(int device, gp_arena *ar);
int gp_launch_sort(gp_ctx *c, int sid);
int gp_launch_load(gp_ctx *c, int se, int si, const gp_gen *g, uint64_t seed, int64_t first_id, int64_t n);
int gp_launch_generate(gp_ctx *c, int se, int si, const gp_gen *g, int64_t addn, int deposit);
int gp_launch_inject_wall(gp_ctx *c, int sid, int64_t n, float vth, const float u[3], double dt,
                          uint64_t seed, uint64_t stream, int deposit);
int gp_launch_moments(gp_ctx *c, int sid, double *host10);
int gp_radix_sort_u32(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp,
                      int64_t n, int bits, int init_vals, uint32_t **keys_out, uint32_t **vals_out,
                      gp_arena *ar = nullptr);

// ---- field side (gp_fields.hip) ----
int gp_fields_init(gp_ctx *c);
void gp_fields_free(gp_ctx *c);
int gp_solve_fields_impl(gp_ctx *c, double dt, double tol, int maxiter, int solver,
                         gp_solve_stats *st, gp_diag *diag);
int gp_solve_darwin_impl(gp_ctx *c, double dt, double tol, int maxiter, int solver, gp_solve_stats st[3]);
int gp_energy_impl(gp_ctx *c, double dt, double *we_ref, double *we_cgs, double *joule);

// ---- elliptic solver (gp_mg.hip) ----
int gp_mg_build(gp_solver &s, int tx, int ty, int mx, int my, double cx, double cy);
void gp_mg_free(gp_solver &s);
// learned hooks of an elliptic solve (null members: plain MG-PCG)
struct gp_hooks {
    int (*warm)(void *, float *e, const float *r) = nullptr;      // first correction
    int (*precond)(void *, float *z, const float *r) = nullptr;   // flexible-CG preconditioner
    void *ctx = nullptr;
    int cap = 0;   // flexible-CG iterations before the fallback (0: last plain MG-PCG count)
};
int gp_solve_spd(hipStream_t st, gp_solver &s, double tol, int maxiter, int singular, gp_solve_stats *stats,
                 const gp_hooks *hk);
int gp_mg_scale(hipStream_t st, gp_solver &s, float *e, const float *r);
int gp_mg_smooth(hipStream_t st, gp_solver &s, float *z, const float *r, int nsweeps);
// context of the learned hooks: the network and the hierarchy it serves
struct gp_nn_bind {
    gp_ctx *c;
    gp_solver *s;
};
int gp_mg_sub_mean(hipStream_t st, gp_solver &s, double *a);
int gp_mg_apply(hipStream_t st, gp_solver &s, const float *in, float *out);

// ---- learned preconditioner (gp_nn.hip) ----
int gp_nn_create(gp_ctx *c, int base_ch, uint64_t seed);
void gp_nn_destroy(gp_ctx *c);
int gp_nn_train_impl(gp_ctx *c, const double *rho, const double *phi, int nsamp, int epochs, int batch,
                     float lr, float *loss_out);
int gp_nn_apply(gp_ctx *c, gp_solver *s, const float *in_dev, float *out_dev);
int gp_nn_infer_impl(gp_ctx *c, const double *rhs, double *out);
int gp_nn_add_pair_dev(gp_ctx *c, const double *b, const double *x);
// stream k of time-loop increments: 0 the Poisson solve, 1..3 the Darwin chi,
// Ax, Ay solves; pairs go to the training set of the current training system
int gp_nn_add_increment_dev(gp_ctx *c, int k, const double *b, const double *x, int64_t m, int *added);
int gp_nn_set_system_impl(gp_ctx *c, int sys);
int gp_nn_train_system(const gp_ctx *c);   // 0 Poisson, 1 Darwin (0 without a network)
int gp_nn_grad_impl(gp_ctx *c, const float *in, const float *tgt, int B, int kind, float *loss, float *grad);
int gp_nn_forward_impl(gp_ctx *c, const float *in, int B, float *out);
int gp_nn_params_impl(gp_ctx *c, float *get, const float *set, int64_t *count);
int gp_solve_rhs_impl(gp_ctx *c, const double *b, double *x, double tol, int maxiter, int solver, gp_solve_stats *st);
int gp_nn_precorrect_cb(void *ctx, float *e, const float *r);
int gp_nn_precond_cb(void *ctx, float *z, const float *r);
int gp_nn_config_impl(gp_ctx *c, int loss_kind, int smooth, int fcg_cap);
int gp_nn_fcg_cap(const gp_ctx *c);
int gp_nn_save_impl(gp_ctx *c, const char *path);
int gp_nn_load_impl(gp_ctx *c, const char *path);

int gp_time_mark(gp_ctx *c, int which, int which2 = -1);   // one event of a begin/end pair in one or two slots
int gp_time_collect(gp_ctx *c);             // synchronise and fold pairs into t_ms

static inline int gp_cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
