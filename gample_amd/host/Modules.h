// Modules.h -- the reference's physics modules as C++ classes over the C-ABI.
//
//   EMField   <- src/EMField.h:11-45   (grids, Poisson / Darwin solves)
//   Particle  <- src/Particle.h:57-79  (species store, update, reduce, deposit)
//   Moment    <- src/Moment.h:21-54    (n, u, P moments)
//   Collision <- src/Collision.h:42-58 (seeded particle generation)
// Method names and roles follow the reference; all state lives on the device
// behind gp_ctx.  Failures throw std::runtime_error with gp_last_error().
#pragma once
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/gample_amd.h"
#include "Init.h"
#include "XmlLogger.h"

namespace gample {

void check(int rc, const char *what);

class EMField {
public:
    EMField(const Init &init, int device);
    ~EMField();
    EMField(const EMField &) = delete;
    EMField &operator=(const EMField &) = delete;
    void resetChargeDensity();
    gp_solve_stats solvePoisson(double dt, gp_diag *diag);        // EMField.mm:745-1136
    void solveVectorPotential(double dt, gp_solve_stats st[3]);   // EMField.mm:1138-1361
    void probe(double &phi, double &ex, double &ey);              // EMField.mm:1130-1135
    gp_ctx *ctx = nullptr;
    int ngx, ngy, ngb, nx, ny;
    double dx, dy, tol;
    int maxiter, solver;
};

class Particle {
public:
    Particle(EMField &fld, const Init &init, int s);
    void integrateChargeDensity();                  // Particle.mm:1004-1130
    void update(double dt, bool fused_deposit);     // Particle.mm:615-891
    // Particle.mm:893-1002; returns the flowout_<species> log section
    std::map<std::string, std::string> reduce();
    // the flowout_<species> section (Particle.mm:989-1001) from pulled
    // particles and the kinetic energy (erg), pinum_* already set
    std::map<std::string, std::string> flowout(int64_t pulled, double ke);
    long long count() const;
    std::string pName;
    float m, q, w;
    int sid = -1;
    int pinum_Xmin = 0, pinum_Xmax = 0, pinum_Ymin = 0, pinum_Ymax = 0;
    double lastKE = 0.0;
    EMField &fld;
};

class Moment {
public:
    explicit Moment(EMField &fld) : fld(fld) {}
    // n, ux, uy, uz, Pxx, Pxy, Pxz, Pyy, Pyz, Pzz after the host
    // normalisation of Moment.mm:275-316
    std::vector<double> integrateMoments(const Particle &p);
    EMField &fld;
};

class Collision {
public:
    // initial seeded loading, Collision.mm:189-336, and the set-up of the
    // SourceForParticle entries (:366-458): the last artificial-ionization
    // entry and the last hollow-cathode entry are the ones run
    Collision(EMField &fld, const Init &init, std::vector<Particle> &ptcl);
    // Collision.mm:464-566: addn = src / w_e * dt pairs, rounded up with the
    // probability of its fraction; returns addn.  deposit: deposit the new
    // particles at once (the fused step's accumulators)
    int artificialIonization(double dt, bool deposit);
    // Collision.mm:568-678: addn = -(sum over species of pinum at the
    // source's wall) electrons on the source line; a non-positive addn is
    // neglected and reported in `neglected`; returns the injected count
    int hollowCathode(bool deposit, int &neglected);
    bool existAI = false, existHC = false;
    EMField &fld;

private:
    std::vector<Particle> &ptcl;
    gp_gen ai{}, hc{};
    int aiEle = -1, aiIon = -1, hcEle = -1, hcWall = 0;
    float aiSrc = 0.0f;
    // the reference draws the rounding from an unseeded std engine; here a
    // PCG32 stream of a fixed seed makes runs repeatable
    uint64_t rndState = 0, rndInc = 1;
    double uniform01();
};

}  // namespace gample
