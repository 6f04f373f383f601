// Init.h -- condition-file parameters (reference: src/Init.h:9-99).
//
// The same sections, keys and defaults as -[Init parseInputFile:]
// (src/Init.mm:27-394): defaults are merged under the user's keys per
// section, unknown keys are ignored, a missing section is an error.
#pragma once
#include <string>
#include <vector>

#include "json.h"

namespace gample {

struct FlagForEquation {
    int Particle = 0, EMField = 0, MCCollision = 0;
};

struct ParamForTimeIntegration {
    int Start = 0, End = 0;
    double TimeStep = 0;
    int LogOutput = 0, ParticleOutput = 0, FieldOutput = 0, ProgressOutput = 0;
    std::string ProjectName, RestartName;
};

struct ParamForComputing {
    long long pNumMax = 0;
    int threadGroupSize = 256, integrationChunkSize = 256;
    double aggrThreshold = 0.1;
    int chebDegree = 5, amgCycleType = 1, maxiter = 200;
    double tolerance = 1e-7;
    // keys added by this build (ignored by the reference): solver, sort
    int solver = 0;            // "PoissonSolver": "MG-CG" (0), "NN-FPCG" (1), "NN-WARM-MG" (2), "NN-FCG" (3)
    std::string networkFile;   // "NetworkFile": trained network for the network solvers
    int sortInterval = 20;     // "SortInterval": fixed tile-sort interval (0: none)
    double sortDrift = 0.0;    // "SortDriftFraction": sort a species once this fraction drifted out of its tile windows (0: off)
};

struct ParamForParticle {
    std::string pName;
    long long pNum = 0;
    double q = 0, m = -1, w = 0;
    std::string genType;
    std::vector<double> genX, genY, genU;
    double genT = 0;           // K
};

struct BoundaryConditionForParticle {
    std::string position, type;
};

struct SourceForParticle {
    std::string pName, genType;
    double src = 0;
    std::vector<double> genX, genY, genU, genU_ele;
    double genT = 0, genT_ele = 0;
};

struct ParamForField {
    int ngx = 0, ngy = 0, ngb = 0;
    double dx = 0, dy = 0;
    std::string InitTypeE, InitTypeB;
    std::vector<double> ampE, ampB;
    std::string FilePathEx, FilePathEy, FilePathEz, FilePathBx, FilePathBy, FilePathBz;
    int weightOrder = 5, BorisOrder = 2, checkCurDens = 0, diff_ngb = 0;
    // keys added by this build: a small cosine density perturbation
    double pertDelta = 0;      // "PerturbationAmplitude"
    int pertMode = 0;          // "PerturbationMode"
};

struct BoundaryConditionForField {
    std::string position, type;
    double val = 0;
};

class Init {
public:
    // throws std::runtime_error with the reference's failure reason
    void parseInputFile(const std::string &path);
    void parse(const Json &root);

    FlagForEquation flagForEquation;
    ParamForTimeIntegration paramForTimeIntegration;
    ParamForComputing paramForComputing;
    std::vector<ParamForParticle> paramForParticle;
    std::vector<BoundaryConditionForParticle> particleBoundaries;
    std::vector<SourceForParticle> particleSources;
    ParamForField paramForField;
    std::vector<BoundaryConditionForField> fieldBoundaries;
};

}  // namespace gample
