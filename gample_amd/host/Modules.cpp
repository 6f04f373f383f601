// Modules.cpp -- EMField / Particle / Moment / Collision over libgample_amd.
#include "Modules.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

#include "Constant.h"

namespace gample {

void check(int rc, const char *what)
{
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + gp_last_error());
}

// ---------------------------------------------------------------- EMField
static int fbc_code(const std::string &t)
{
    if (t == "periodic") return GP_FBC_PERIODIC;
    if (t == "Dirichlet") return GP_FBC_DIRICHLET;
    if (t == "Neumann") return GP_FBC_NEUMANN;
    throw std::runtime_error("invalid field boundary type '" + t + "'");
}
static int pbc_code(const std::string &t)
{
    if (t == "periodic") return GP_PBC_PERIODIC;
    if (t == "Delete") return GP_PBC_DELETE;
    if (t == "Reflect") return GP_PBC_REFLECT;
    throw std::runtime_error("[FatalError] invalid particle boundary");   // Particle.mm:505-508
}
static int wall_index(const std::string &p)
{
    if (p == "Xmin") return 0;
    if (p == "Xmax") return 1;
    if (p == "Ymin") return 2;
    if (p == "Ymax") return 3;
    return -1;
}

// 1D B profile file (EMField.mm:654-687): int n, float d, n+1 floats (T)
static std::vector<float> load1d(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("Unable to open file " + path + " for reading");
    int n = 0;
    float d = 0;
    std::vector<float> v;
    if (std::fread(&n, 4, 1, f) == 1 && std::fread(&d, 4, 1, f) == 1 && n >= 0) {
        v.resize((size_t)n + 1);
        if (std::fread(v.data(), 4, v.size(), f) != v.size()) v.clear();
    }
    std::fclose(f);
    if (v.empty()) throw std::runtime_error("Unable to read 1D field file " + path);
    return v;
}

EMField::EMField(const Init &init, int device)
{
    const ParamForField &F = init.paramForField;
    gp_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.ngx = F.ngx; cfg.ngy = F.ngy; cfg.dx = F.dx; cfg.dy = F.dy;
    cfg.weight_order = F.weightOrder; cfg.boris_order = F.BorisOrder;
    for (int w = 0; w < 4; w++) { cfg.pbc[w] = GP_PBC_NONE; cfg.fbc[w] = GP_FBC_PERIODIC; }
    for (auto &b : init.particleBoundaries) {
        int w = wall_index(b.position);
        if (w >= 0) cfg.pbc[w] = pbc_code(b.type);
    }
    for (auto &b : init.fieldBoundaries) {
        int w = wall_index(b.position);
        if (w < 0) continue;   // the hollow-cathode entry: below
        cfg.fbc[w] = fbc_code(b.type);
        if (cfg.fbc[w] == GP_FBC_DIRICHLET) cfg.fbc_value[w] = b.val * VtosV;   // EMField.mm:206
    }
    if (F.InitTypeE == "Uniform")
        for (int k = 0; k < 3 && k < (int)F.ampE.size(); k++) cfg.e0[k] = F.ampE[k];
    if (F.InitTypeB == "Uniform")
        for (int k = 0; k < 3 && k < (int)F.ampB.size(); k++) cfg.b0[k] = F.ampB[k];
    cfg.em_model = init.flagForEquation.EMField == 2 ? 2 : 1;
    cfg.cur_dens = (F.checkCurDens == 1 || init.flagForEquation.EMField > 1) ? 1 : 0;   // Particle.mm:523-525
    cfg.device = device;
    check(gp_ctx_create(&cfg, &ctx), "EMField init");
    int d[6];
    check(gp_dims(ctx, d), "gp_dims");
    ngx = F.ngx; ngy = F.ngy; ngb = F.ngb; nx = d[0]; ny = d[1];
    dx = F.dx; dy = F.dy;
    tol = init.paramForComputing.tolerance;
    maxiter = init.paramForComputing.maxiter;
    solver = init.paramForComputing.solver;
    check(gp_set_sort_drift(ctx, init.paramForComputing.sortDrift), "SortDriftFraction");
    for (auto &b : init.fieldBoundaries)   // EMField.mm:231-232
        if (b.position == "hollow-cathode") check(gp_set_cathode(ctx, b.val), "hollow-cathode");
    if (solver != 0) check(gp_nn_load(ctx, init.paramForComputing.networkFile.c_str()), "NetworkFile");
    if (F.InitTypeB == "From1dXFile") {   // EMField.mm:175-197, T -> G
        auto bx = load1d(F.FilePathBx), by = load1d(F.FilePathBy), bz = load1d(F.FilePathBz);
        if ((int)bz.size() < nx + 2 || (int)bx.size() < nx + 1 || (int)by.size() < nx + 2)
            throw std::runtime_error("1D B file shorter than the grid");
        std::vector<float> Bx((size_t)(nx + 1) * (ny + 2)), By((size_t)(nx + 2) * (ny + 1)), Bz((size_t)(nx + 2) * (ny + 2));
        for (int i = 0; i <= nx; i++) for (int j = 0; j <= ny + 1; j++) Bx[i + j * (nx + 1)] = (float)(bx[i] * TtoG);
        for (int i = 0; i <= nx + 1; i++) for (int j = 0; j <= ny; j++) By[i + j * (nx + 2)] = (float)(by[i] * TtoG);
        for (int i = 0; i <= nx + 1; i++) for (int j = 0; j <= ny + 1; j++) Bz[i + j * (nx + 2)] = (float)(bz[i] * TtoG);
        check(gp_field_set(ctx, GP_F_BX, Bx.data(), (int64_t)Bx.size() * 4), "set Bx");
        check(gp_field_set(ctx, GP_F_BY, By.data(), (int64_t)By.size() * 4), "set By");
        check(gp_field_set(ctx, GP_F_BZ, Bz.data(), (int64_t)Bz.size() * 4), "set Bz");
    }
}
EMField::~EMField() { gp_ctx_destroy(ctx); }
void EMField::resetChargeDensity() { check(gp_reset_grids(ctx), "resetChargeDensity"); }
gp_solve_stats EMField::solvePoisson(double dt, gp_diag *diag)
{
    gp_solve_stats st;
    std::memset(&st, 0, sizeof(st));
    check(gp_solve_fields(ctx, dt, tol, maxiter, solver, &st, diag), "solvePoisson");
    return st;
}
void EMField::solveVectorPotential(double dt, gp_solve_stats st[3])
{
    check(gp_solve_darwin(ctx, dt, tol, maxiter, solver, st), "solveVectorPotential");
}
void EMField::probe(double &phi, double &ex, double &ey)
{
    const int pi = ngx / 4 + ngb, pj = ngy / 2 + ngb;
    std::vector<double> p((size_t)(nx + 3) * (ny + 3));
    std::vector<float> Ex((size_t)(nx + 2) * (ny + 1)), Ey((size_t)(nx + 1) * (ny + 2));
    check(gp_field_get(ctx, GP_F_PHI, p.data(), (int64_t)p.size() * 8), "phi");
    check(gp_field_get(ctx, GP_F_EX, Ex.data(), (int64_t)Ex.size() * 4), "Ex");
    check(gp_field_get(ctx, GP_F_EY, Ey.data(), (int64_t)Ey.size() * 4), "Ey");
    phi = p[(pi + 2) + (size_t)(pj + 2) * (nx + 3)];
    ex = Ex[(pi + 1) + (size_t)pj * (nx + 2)];
    ey = Ey[pi + (size_t)(pj + 1) * (nx + 1)];
}

// ---------------------------------------------------------------- Particle
Particle::Particle(EMField &f, const Init &init, int s) : fld(f)
{
    const ParamForParticle &P = init.paramForParticle.at(s);
    pName = P.pName;
    m = (float)P.m; q = (float)P.q; w = (float)P.w;
    const long long cap = init.paramForComputing.pNumMax;
    if (P.pNum > cap)
        throw std::runtime_error("pNum > pNumMax: pName = " + pName + ", pNum = " + std::to_string(P.pNum) +
                                 ", pNumMax = " + std::to_string(cap));   // Particle.mm:544-547
    gp_species_desc d;
    std::memset(&d, 0, sizeof(d));
    d.q = P.q; d.m = P.m; d.w = P.w; d.capacity = cap;
    d.vth_hint = std::sqrt(kb * std::max(P.genT, 1.0) / P.m);
    check(gp_species_create(fld.ctx, &d, &sid), "Particle init");
}
void Particle::integrateChargeDensity() { check(gp_deposit(fld.ctx, sid), "integrateChargeDensity"); }
void Particle::update(double dt, bool fused) { check(gp_push(fld.ctx, sid, dt, fused ? 1 : 0), "update"); }
long long Particle::count() const
{
    int64_t n = 0;
    check(gp_species_count(fld.ctx, sid, &n), "count");
    return n;
}
std::map<std::string, std::string> Particle::reduce()
{
    int64_t cnt[4] = {0, 0, 0, 0};
    double ke = 0;
    check(gp_reduce(fld.ctx, sid, cnt, &ke), "reduce");
    const int qe = (int)(q / (float)ec);      // charge units, Particle.mm:914-961
    pinum_Xmin = (int)cnt[0] * qe; pinum_Xmax = (int)cnt[1] * qe;
    pinum_Ymin = (int)cnt[2] * qe; pinum_Ymax = (int)cnt[3] * qe;
    return flowout(cnt[0] + cnt[1] + cnt[2] + cnt[3], ke);
}

std::map<std::string, std::string> Particle::flowout(int64_t pulled, double ke)
{
    const double diff = ke - lastKE;
    lastKE = ke;
    return {{"particleNumber", std::to_string(count())},
            {"pulledPtclNum", std::to_string(pulled)},
            {"Xmin", std::to_string(pinum_Xmin)}, {"Xmax", std::to_string(pinum_Xmax)},
            {"Ymin", std::to_string(pinum_Ymin)}, {"Ymax", std::to_string(pinum_Ymax)},
            {"totalKineticEnergy", fmtSci((float)(ke * ergtoev), 6)},
            {"totalKEincrement", fmtSci((float)(diff * ergtoev), 6)}};
}

// ---------------------------------------------------------------- Moment
std::vector<double> Moment::integrateMoments(const Particle &p)
{
    const size_t ng = (size_t)(fld.nx + 1) * (fld.ny + 1);
    std::vector<double> a(10 * ng);
    check(gp_moments(fld.ctx, p.sid, a.data()), "integrateMoments");
    const double scale = (double)p.w / ((double)(float)fld.dx * (double)(float)fld.dy);
    for (size_t i = 0; i < ng; i++) {
        const double n = a[i];
        double *u = &a[ng + i];
        if (n > 1e-20) {
            const double cnt = n / scale;
            double ux = a[1 * ng + i] / cnt, uy = a[2 * ng + i] / cnt, uz = a[3 * ng + i] / cnt;
            a[1 * ng + i] = ux; a[2 * ng + i] = uy; a[3 * ng + i] = uz;
            a[4 * ng + i] = p.m * (a[4 * ng + i] / cnt - ux * ux) * n;
            a[5 * ng + i] = p.m * (a[5 * ng + i] / cnt - ux * uy) * n;
            a[6 * ng + i] = p.m * (a[6 * ng + i] / cnt - ux * uz) * n;
            a[7 * ng + i] = p.m * (a[7 * ng + i] / cnt - uy * uy) * n;
            a[8 * ng + i] = p.m * (a[8 * ng + i] / cnt - uy * uz) * n;
            a[9 * ng + i] = p.m * (a[9 * ng + i] / cnt - uz * uz) * n;
        } else {
            for (int k = 1; k < 10; k++) a[k * ng + i] = 0.0;
        }
        (void)u;
    }
    return a;
}

// ---------------------------------------------------------------- Collision
static int gen_code(const std::string &t)
{
    if (t == "uniform-Gaussian") return 0;
    if (t == "Xsinusoidal-Gaussian") return 1;
    if (t == "Ysinusoidal-Gaussian") return 2;
    if (t == "quietStart") return 4;
    if (t == "Xgaussian-Gaussian") return 5;
    if (t == "Ygaussian-Gaussian") return 6;
    throw std::runtime_error("unknown GenerateType '" + t + "'");
}

static uint64_t splitmix64(uint64_t *x)   // Collision.h:11-16
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

Collision::Collision(EMField &f, const Init &init, std::vector<Particle> &ptcl_) : fld(f), ptcl(ptcl_)
{
    std::vector<Particle> &ptcl = ptcl_;
    const auto &P = init.paramForParticle;
    const auto &F = init.paramForField;
    if (P.empty()) return;
    gp_gen g;
    std::memset(&g, 0, sizeof(g));
    g.gen_type = gen_code(P[0].genType);
    if (F.pertDelta != 0.0) { g.gen_type = 7; g.pert_delta = (float)F.pertDelta; g.pert_mode = F.pertMode; }
    if (P[0].genX.size() < 2 || P[0].genX[0] < 0) { g.xmin = 0.0f; g.xmax = (float)(F.dx * F.ngx); }
    else { g.xmin = (float)P[0].genX[0]; g.xmax = (float)P[0].genX[1]; }
    if (P[0].genY.size() < 2 || P[0].genY[0] < 0) { g.ymin = 0.0f; g.ymax = (float)(F.dy * F.ngy); }
    else { g.ymin = (float)P[0].genY[0]; g.ymax = (float)P[0].genY[1]; }
    int se = -1, si = -1;
    for (size_t s = 0; s < P.size(); s++) {
        float u[3] = {0, 0, 0};
        for (int k = 0; k < 3 && k < (int)P[s].genU.size(); k++) u[k] = (float)P[s].genU[k];
        const float vth = (float)std::sqrt(2 * kb * P[s].genT / P[s].m);   // Collision.mm:264,269
        if (P[s].pName == "electron") {
            se = (int)s;
            std::memcpy(g.ele_u, u, sizeof u);
            g.ele_vth = vth;
        } else {
            si = (int)s;
            std::memcpy(g.ion_u, u, sizeof u);
            g.ion_vth = vth;
        }
    }
    if (se < 0) { se = si; si = -1; }
    // Collision.mm:280-336: master seed 0, ids 0..addn-1, ions co-located
    check(gp_load_pair(fld.ctx, ptcl[se].sid, si >= 0 ? ptcl[si].sid : -1, &g, 0ULL, 0, P[0].pNum), "initial load");
    for (auto &p : ptcl) check(gp_sort(fld.ctx, p.sid), "sort");

    // sources (Collision.mm:366-458)
    auto species = [&](const std::string &name) {
        for (size_t s = 0; s < ptcl.size(); s++)
            if (ptcl[s].pName == name) return (int)s;
        return -1;
    };
    auto region = [&](const SourceForParticle &src, gp_gen &q) {
        if (src.genX.size() < 2 || src.genX[0] < 0) { q.xmin = 0.0f; q.xmax = (float)(F.dx * F.ngx); }
        else { q.xmin = (float)src.genX[0]; q.xmax = (float)src.genX[1]; }
        if (src.genY.size() < 2 || src.genY[0] < 0) { q.ymin = 0.0f; q.ymax = (float)(F.dy * F.ngy); }
        else { q.ymin = (float)src.genY[0]; q.ymax = (float)src.genY[1]; }
    };
    for (const auto &src : init.particleSources) {
        int t;
        if (src.genType == "uniformIonization") t = 0;
        else if (src.genType == "XsinusoidalIonization") t = 1;
        else if (src.genType == "YsinusoidalIonization") t = 2;
        else if (src.genType == "hollow-cathode") t = 3;
        else throw std::runtime_error("unknown source GenerateType '" + src.genType + "'");
        if (t != 3) {
            existAI = true;
            std::memset(&ai, 0, sizeof ai);
            ai.gen_type = t;
            region(src, ai);
            for (int k = 0; k < 3 && k < (int)src.genU_ele.size(); k++) ai.ele_u[k] = (float)src.genU_ele[k];
            for (int k = 0; k < 3 && k < (int)src.genU.size(); k++) ai.ion_u[k] = (float)src.genU[k];
            ai.ele_vth = (float)std::sqrt(2 * kb * src.genT_ele / me);
            aiEle = species("electron");
            aiIon = species(src.pName);
            if (aiEle < 0 || aiIon < 0) throw std::runtime_error("ionization source needs species electron and " + src.pName);
            ai.ion_vth = (float)std::sqrt(2 * kb * src.genT / init.paramForParticle[aiIon].m);
            aiSrc = (float)src.src;
        } else {
            existHC = true;
            std::memset(&hc, 0, sizeof hc);
            hc.gen_type = 3;
            region(src, hc);
            for (int k = 0; k < 3 && k < (int)src.genU.size(); k++) hc.ele_u[k] = (float)src.genU[k];
            hc.ele_vth = (float)std::sqrt(2 * kb * src.genT / me);
            hcEle = species("electron");
            if (hcEle < 0) throw std::runtime_error("hollow-cathode source needs species electron");
            hcWall = (int)src.src;   // 0 Xmin, 1 Xmax, 2 Ymin, 3 Ymax
        }
    }
    uint64_t sd = 0x5eedULL;
    rndState = splitmix64(&sd);
    rndInc = splitmix64(&sd) | 1ULL;
}

double Collision::uniform01()
{
    const uint64_t old = rndState;
    rndState = old * 6364136223846793005ULL + rndInc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u), rot = (uint32_t)(old >> 59u);
    const uint32_t r = (xs >> rot) | (xs << ((-rot) & 31u));
    return (double)r * (1.0 / 4294967296.0);
}

int Collision::artificialIonization(double dt, bool deposit)
{
    if (!existAI) return 0;
    // float src over the float weight, then times dt (Collision.mm:482)
    const double addn_d = (double)(aiSrc / ptcl[aiEle].w) * dt;
    const int base = (int)addn_d;
    const int addn = uniform01() < addn_d - base ? base + 1 : base;
    check(gp_generate(fld.ctx, ptcl[aiEle].sid, ptcl[aiIon].sid, &ai, addn, deposit ? 1 : 0), "artificialIonization");
    return addn;
}

int Collision::hollowCathode(bool deposit, int &neglected)
{
    neglected = 0;
    if (!existHC) return 0;
    int addn = 0;
    for (auto &p : ptcl) {
        const int pin[4] = {p.pinum_Xmin, p.pinum_Xmax, p.pinum_Ymin, p.pinum_Ymax};
        addn -= pin[hcWall];
    }
    if (addn <= 0) {
        neglected = -addn;
        return 0;
    }
    check(gp_generate(fld.ctx, ptcl[hcEle].sid, -1, &hc, addn, deposit ? 1 : 0), "hollowCathode");
    return addn;
}

}  // namespace gample
