/*
 * gp_oracle.h -- CPU oracle for the gample PIC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is called from
 * tests/, from __graft_entry__.smoke() and from bench.py's cpu_baseline leg,
 * never from the product path (gample_amd/), which must fail loudly when its
 * HIP library is missing rather than fall back to anything here.
 *
 * It restates, in plain C with OpenMP, what the reference computes on its
 * hot path (citations are into the reference tree):
 *   - B-spline weights           src/Particle.mm:97-131
 *   - field gather               src/Particle.mm:83-94,133-154
 *   - Boris push (2/4/6)         src/Particle.mm:156-209
 *   - position update + BCs      src/Particle.mm:211-269
 *   - charge/current deposit     src/Particle.mm:281-389 (scale :540)
 *   - absorption compaction, KE  src/Particle.mm:893-1002
 *   - ghost-charge fold          src/EMField.mm:749-850
 *   - Poisson CSR + rhs          src/EMField.mm:236-405, 853-861
 *   - phi reconstruction, E      src/EMField.mm:868-1094
 *   - field energy / Joule       src/EMField.mm:1096-1116
 *   - Darwin rhs + CSR           src/EMField.mm:483-623, 1138-1329
 *   - moments (n, nv, nvv)       src/Moment.mm:33-135
 *   - splitmix64 / PCG32 loader  src/Collision.h:11-16, src/Collision.mm:21-174,280-336
 *
 * Two push variants exist:
 *   gpo_push_ref   -- the reference's literal FP32 expression order (powf
 *                     weights, i-outer/j-inner gather of (sx*sy)*F, no FMA).
 *   gpo_push_fast  -- the op order the HIP kernel uses (Horner weights,
 *                     factorised gather, explicit fmaf).  GPU results must be
 *                     bit-identical to this one; it in turn must agree with
 *                     gpo_push_ref to 1e-5 relative.
 * The whole file is compiled with -ffp-contract=off so that every mul/add
 * not written as fmaf() rounds separately, exactly as in the kernels.
 */
#ifndef GP_ORACLE_H
#define GP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* particle boundary codes (src/Particle.mm:461-509) */
#define GPO_PBC_NONE     (-1)
#define GPO_PBC_PERIODIC   0
#define GPO_PBC_DELETE     1
#define GPO_PBC_REFLECT    2
/* field boundary codes (src/EMField.mm:200-233) */
#define GPO_FBC_PERIODIC   0
#define GPO_FBC_DIRICHLET  1
#define GPO_FBC_NEUMANN    2

typedef struct {
    int ngx, ngy;        /* interior cells */
    int ngb;             /* ghost width = order/2 (src/Init.mm:140) */
    int order;           /* 1, 3 or 5 */
    int boris_order;     /* 2, 4 or 6 */
    int pbc[4];          /* particle BC: Xmin, Xmax, Ymin, Ymax */
    float dx, dy;        /* cell sizes as EMField stores them (FP32) */
} gpo_grid;

/* ---------------- weights ---------------- */
void   gpo_weights_fast(int order, float h, float *w);          /* kernel op order */
void   gpo_weights_ref(int order, float h, float *w);           /* literal powf form */
void   gpo_weights_f64(int order, double h, double *w);         /* FP64 literal form */

/* ---------------- push ----------------
 * x,y,vx,vy,vz: SoA, length n.  Fields in the reference layouts:
 *   Ex (nx+2)*(ny+1), Ey (nx+1)*(ny+2), Bz (nx+2)*(ny+2), nx = ngx+2ngb.
 * uniform_bz != 0: Bz is uniform and static, the gather of Bz is replaced by
 * the constant bz0 (fast variant only; the kernel does the same).
 * Absorbed particles (Delete BC) are marked by x = -piflag (1..4).
 * Returns 0. */
int gpo_push_fast(const gpo_grid *g, const float *Ex, const float *Ey, const float *Bz,
                  int uniform_bz, float bz0,
                  float constE, float constB, float constX, float constY,
                  int64_t n, float *x, float *y, float *vx, float *vy, float *vz,
                  int nthreads);
int gpo_push_ref(const gpo_grid *g, const float *Ex, const float *Ey, const float *Bz,
                 float constE, float constB, float constX, float constY,
                 int64_t n, float *x, float *y, float *vx, float *vy, float *vz,
                 int32_t *piflag, int nthreads);

/* ---------------- deposit ----------------
 * Build scheme: per-axis weights quantised to u32 with 23 fractional bits,
 * product accumulated as int64 with 46 fractional bits (rho).  Current:
 * float product (sx*sy)*(v*inv_vnorm) converted exactly (round to nearest
 * even) to int64 with 40 fractional bits.  Particles with x < 0 are skipped.
 * Accumulators are ADDED to (not cleared). */
#define GPO_RHO_FRAC 46
#define GPO_CUR_FRAC 40
int gpo_deposit_fx(const gpo_grid *g, int64_t n, const float *x, const float *y,
                   const float *vx, const float *vy, float inv_vnorm,
                   int64_t *rho_acc, int64_t *jx_acc, int64_t *jy_acc, int nthreads);
/* reference scheme: sequential FP32 sums of sf*sf*scale (and *vx, *vy) */
int gpo_deposit_ref(const gpo_grid *g, int64_t n, const float *x, const float *y,
                    const float *vx, const float *vy, float scale,
                    float *rho, float *jx, float *jy);
/* FP64 exact-ish accumulation of the reference weights (error statements) */
int gpo_deposit_f64(const gpo_grid *g, int64_t n, const float *x, const float *y,
                    const float *vx, const float *vy, double scale,
                    double *rho, double *jx, double *jy);
int64_t gpo_fx_from_float(float c, int frac_bits);    /* llrint(c * 2^frac) */

/* moments n, nvx, nvy, nvz, nvxvx, nvxvy, nvxvz, nvyvy, nvyvz, nvzvz
 * (src/Moment.mm:62-129) in FP64, node-centred weights of floor(x+1e-20). */
int gpo_moments_f64(const gpo_grid *g, int64_t n, const float *x, const float *y,
                    const float *vx, const float *vy, const float *vz, double scale,
                    double *mom /* 10 * (nx+1)*(ny+1) */);

/* ---------------- reduce (compaction) ----------------
 * Literal restatement of the reference's tail-swap loop on SoA + flags. */
int gpo_reduce_literal(int64_t *n, float *x, float *y, float *vx, float *vy, float *vz,
                       int32_t *piflag, int counts[4]);
/* Kinetic energy sum 0.5*m*|v|^2*w over particles with x >= 0 (FP64). */
double gpo_kinetic_energy(int64_t n, const float *x, const float *vx, const float *vy,
                          const float *vz, float m, float w);

/* ---------------- RNG / loading ---------------- */
uint64_t gpo_splitmix64(uint64_t *seed);
uint32_t gpo_pcg32(uint64_t *state, uint64_t inc);
float    gpo_rand01(uint32_t u);
/* Seeded electron/ion pair loading (genType 0 uniform, 1 Xsinusoidal,
 * 2 Ysinusoidal, 4 quiet start, 5 Xgaussian, 6 Ygaussian, 7 small-amplitude
 * cosine perturbation 1+delta*cos(2*pi*m*x/Lx)).  Writes n particles of each
 * species starting at index 0 (src/Collision.mm:63-174, 280-336).  ion arrays
 * may be NULL (electrons only). */
typedef struct {
    int gen_type;
    float xmin, xmax, ymin, ymax;  /* physical cm */
    float ele_u[3], ele_vth;
    float ion_u[3], ion_vth;
    float dx, dy, ngb;
    float pert_delta; int pert_mode;  /* gen_type 7 */
    float xdom, ydom;  /* domain extent in cells (0: no clamp); a position
                          inside the domain in cm that rounds onto the upper
                          face in grid units is moved just below it */
} gpo_gen;
void gpo_inject_wall(uint64_t seed, uint64_t stream, int64_t n, float vth, const float u[3], float dtdx, float lo,
                     float ngx, float ngy, float *x, float *y, float *vx, float *vy, float *vz);
int gpo_load_pair(const gpo_gen *p, uint64_t master_seed, int64_t first_id, int64_t n,
                  float *ex, float *ey, float *evx, float *evy, float *evz,
                  float *ix, float *iy, float *ivx, float *ivy, float *ivz);
void gpo_stream_states(uint64_t master_seed, int64_t first_id, int64_t n, uint64_t *state, uint64_t *inc);
int gpo_generate(const gpo_gen *p, uint64_t *state, const uint64_t *inc, int64_t addn,
                 float *ex, float *ey, float *evx, float *evy, float *evz,
                 float *ix, float *iy, float *ivx, float *ivy, float *ivz);

/* ---------------- field side ---------------- */
typedef struct {
    int ngx, ngy, ngb;
    float dx, dy;
    int fbc[4];            /* Xmin, Xmax, Ymin, Ymax */
    double fval[4];        /* Dirichlet phi (statV) or Neumann slope */
} gpo_field;
/* unknown-grid geometry (src/EMField.mm:247-271) */
void gpo_poisson_dims(const gpo_field *f, int *nkx, int *nky, int *ikmin, int *jkmin);
/* ghost-charge fold in place on the (nx+1)*(ny+1) node array (FP64 copy) */
void gpo_fold_rho(const gpo_field *f, double *rho);
/* CSR of the reference (row-major k = i + j*(nkx+1)) and rhs_BC.
 * Returns nnz; arrays must hold 5 entries per row. */
int64_t gpo_poisson_csr(const gpo_field *f, int32_t *ptr, int32_t *col, double *val, double *rhs_bc);
/* rhs = rhs_bc - 4 pi rho[node] */
void gpo_poisson_rhs(const gpo_field *f, const double *rho, const double *rhs_bc, double *rhs);
/* FP64 Jacobi-PCG on a CSR system (negated to SPD internally);
 * project != 0 removes the mean of rhs and iterates (singular periodic case). */
int gpo_cg(int64_t n, const int32_t *ptr, const int32_t *col, const double *val,
           const double *rhs, double *x, double tol, int maxiter, int project, double *relres);
/* relative residual ||b - A x|| / ||b|| in FP64 */
double gpo_relres(int64_t n, const int32_t *ptr, const int32_t *col, const double *val,
                  const double *b, const double *x);
/* phi (nx+3)*(ny+3) from the solution (corrected periodic-x mapping) */
void gpo_reconstruct_phi(const gpo_field *f, const double *sol, double *phi);
/* Ex (nx+2)*(ny+1), Ey (nx+1)*(ny+2) = -grad phi, stored FP32 */
double gpo_cathode_ramp(const gpo_field *f, const double *sol, double pos, double *phi);
void gpo_gradient(const gpo_field *f, const double *phi, float *Ex, float *Ey);
/* energies: reference form sum(E^2)*dx*dy/2 over whole arrays, CGS form
 * sum(E^2)*dx*dy/(8 pi) over unique interior points, Joule sum(E.J)*dt*dx*dy */
void gpo_field_energy(const gpo_field *f, const float *Ex, const float *Ey,
                      const double *jx, const double *jy, double dt,
                      double *we_ref, double *we_cgs, double *joule);

/* Darwin (src/EMField.mm:483-623, 1138-1329): node grid (nkx_vec+1)*(nky_vec+1),
 * Neumann where not periodic. */
void gpo_darwin_dims(const gpo_field *f, int *nkxv, int *nkyv);
int64_t gpo_darwin_csr(const gpo_field *f, int32_t *ptr, int32_t *col, double *val);
/* rhs of the chi solve from jx, jy (double copies of the FP32 grids), zero-mean */
void gpo_darwin_rhs_chi(const gpo_field *f, const double *jx, const double *jy, double *rhs);
/* rhs of Ax / Ay solves from chi, zero-mean */
void gpo_darwin_rhs_a(const gpo_field *f, const double *jx, const double *jy, const double *chi,
                      double *rhs_x, double *rhs_y);

#ifdef __cplusplus
}
#endif
#endif
