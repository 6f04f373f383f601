// Init.cpp -- condition-file parsing with the reference's keys, defaults and
// derived quantities (src/Init.mm:27-394).
#include "Init.h"

#include <cmath>
#include <stdexcept>

#include "Constant.h"

namespace gample {

namespace {

Json merged(const Json *user, Json defaults)
{
    for (auto &kv : user->obj) defaults.set(kv.first, kv.second);
    return defaults;
}
// NSNumber integerValue truncates toward zero; strings parse like atoi
long long ival(const Json &o, const char *k)
{
    const Json *v = o.find(k);
    if (!v) return 0;
    if (v->type == Json::Number) return (long long)std::trunc(v->num);
    if (v->type == Json::Bool) return v->b ? 1 : 0;
    if (v->type == Json::String) return std::atoll(v->str.c_str());
    return 0;
}
double dval(const Json &o, const char *k)
{
    const Json *v = o.find(k);
    if (!v) return 0;
    if (v->type == Json::Number) return v->num;
    if (v->type == Json::Bool) return v->b ? 1 : 0;
    if (v->type == Json::String) return std::atof(v->str.c_str());
    return 0;
}
std::string sval(const Json &o, const char *k)
{
    const Json *v = o.find(k);
    return (v && v->type == Json::String) ? v->str : std::string();
}
std::vector<double> vval(const Json &o, const char *k)
{
    std::vector<double> out;
    const Json *v = o.find(k);
    if (v && v->type == Json::Array)
        for (auto &e : v->arr) out.push_back(e.type == Json::Number ? e.num : 0.0);
    return out;
}
Json obj(std::initializer_list<std::pair<const char *, Json>> kv)
{
    Json j = Json::object();
    for (auto &p : kv) j.set(p.first, p.second);
    return j;
}
Json N(double v) { return Json::number(v); }
Json S(const char *s) { return Json::string(s); }
Json A(std::initializer_list<double> v) { return Json::array(v); }

// default dictionaries, src/Init.mm:304-394
Json defFlag() { return obj({{"Particle", N(0)}, {"EMField", N(0)}, {"MCCollision", N(0)}}); }
Json defTime()
{
    return obj({{"Start", N(0)}, {"End", N(0)}, {"TimeStep", N(0)}, {"LogOutput", N(0)}, {"ParticleOutput", N(0)},
                {"FieldOutput", N(0)}, {"ProgressOutput", N(0)}, {"ProjectName", S("undefined")},
                {"RestartName", S("undefined")}});
}
Json defComp()
{
    return obj({{"MaximumParticleNumber", N(0.08)}, {"ThreadGroupSize", N(256)}, {"IntegrationChunkSize", N(256)},
                {"AggregationThreshold", N(0.1)}, {"ChebyshevDegree", N(5)}, {"AMGCycleType", N(1)},
                {"MaxiterForPoisson", N(200)}, {"ToleranceForPoisson", N(1e-7)},
                {"PoissonSolver", S("MG-CG")}, {"SortInterval", N(20)}, {"SortDriftFraction", N(0)},
                {"NetworkFile", S("")}});
}
Json defParticle()
{
    return obj({{"ParticleName", S("undefined")}, {"pNumSetMethod", S("undefined")}, {"pNumValue", N(0)},
                {"Charge", N(0)}, {"Mass", N(-1)}, {"WeightSetMethod", S("undefined")}, {"WeightValue", N(0)},
                {"GenerateType", S("undefined")}, {"InitialPosX", A({-1, -1})}, {"InitialPosY", A({-1, -1})},
                {"InitialVel", A({0, 0, 0})}, {"InitialTemp", N(0)}});
}
Json defBC() { return obj({{"RegionName", S("undefined")}, {"BCType", S("undefined")}, {"Value", N(0)}}); }
Json defSource()
{
    return obj({{"ParticleName", S("undefined")}, {"GenerateType", S("undefined")}, {"SrcSetMethod", S("undefined")},
                {"SrcVal", N(0)}, {"GeneratePosX", A({-1, -1})}, {"GeneratePosY", A({-1, -1})},
                {"GenerateVel", A({0, 0, 0})}, {"GenerateTemp", N(0)}, {"GenVelForElectron", A({0, 0, 0})},
                {"GenTempForElectron", N(0)}});
}
Json defField()
{
    return obj({{"NumberOfGridX", N(0)}, {"NumberOfGridY", N(0)}, {"GridSizeOfX", N(0)}, {"GridSizeOfY", N(0)},
                {"InitializeTypeOfE", S("undefined")}, {"AmplitudeOfE", A({0, 0, 0})},
                {"FilePathOfEx", S("undefined")}, {"FilePathOfEy", S("undefined")}, {"FilePathOfEz", S("undefined")},
                {"InitializeTypeOfB", S("undefined")}, {"AmplitudeOfB", A({0, 0, 0})},
                {"FilePathOfBx", S("undefined")}, {"FilePathOfBy", S("undefined")}, {"FilePathOfBz", S("undefined")},
                {"WeightingOrder", N(5)}, {"BorisOrder", N(2)}, {"FlagForCurDens", N(0)}, {"PtclPosOffset", N(0)},
                {"PerturbationAmplitude", N(0)}, {"PerturbationMode", N(0)}});
}

const Json &section(const Json &root, const char *name, Json::Type t)
{
    const Json *v = root.find(name);
    if (!v || v->type != t)
        throw std::runtime_error(std::string(name) + (t == Json::Object ? " is not a dictionary" : " is not an array"));
    return *v;
}

}  // namespace

void Init::parseInputFile(const std::string &path) { parse(json_load(path)); }

void Init::parse(const Json &root)
{
    if (!root.is_object()) throw std::runtime_error("top level is not a dictionary");
    // FlagForEquation (Init.mm:50-63)
    Json f = merged(&section(root, "FlagForEquation", Json::Object), defFlag());
    flagForEquation.Particle = (int)ival(f, "Particle");
    flagForEquation.EMField = (int)ival(f, "EMField");
    flagForEquation.MCCollision = (int)ival(f, "MCCollision");
    // ParamForTimeIntegration (Init.mm:65-85)
    Json t = merged(&section(root, "ParamForTimeIntegration", Json::Object), defTime());
    auto &T = paramForTimeIntegration;
    T.Start = (int)ival(t, "Start"); T.End = (int)ival(t, "End"); T.TimeStep = dval(t, "TimeStep");
    T.LogOutput = (int)ival(t, "LogOutput"); T.ParticleOutput = (int)ival(t, "ParticleOutput");
    T.FieldOutput = (int)ival(t, "FieldOutput"); T.ProgressOutput = (int)ival(t, "ProgressOutput");
    T.ProjectName = sval(t, "ProjectName"); T.RestartName = sval(t, "RestartName");
    // ParamForComputing (Init.mm:87-104)
    Json c = merged(&section(root, "ParamForComputing", Json::Object), defComp());
    auto &C = paramForComputing;
    C.pNumMax = ival(c, "MaximumParticleNumber");
    C.threadGroupSize = (int)ival(c, "ThreadGroupSize");
    C.integrationChunkSize = (int)ival(c, "IntegrationChunkSize");
    C.aggrThreshold = dval(c, "AggregationThreshold");
    C.chebDegree = (int)ival(c, "ChebyshevDegree");
    C.amgCycleType = (int)ival(c, "AMGCycleType");
    C.maxiter = (int)ival(c, "MaxiterForPoisson");
    C.tolerance = dval(c, "ToleranceForPoisson");
    const std::string solver = sval(c, "PoissonSolver");
    if (solver == "MG-CG") C.solver = 0;
    else if (solver == "NN-FPCG") C.solver = 1;
    else if (solver == "NN-WARM-MG") C.solver = 2;
    else if (solver == "NN-FCG") C.solver = 3;
    else throw std::runtime_error("PoissonSolver must be MG-CG, NN-FPCG, NN-WARM-MG or NN-FCG: " + solver);
    C.networkFile = sval(c, "NetworkFile");
    if (C.solver != 0 && C.networkFile.empty())
        throw std::runtime_error("PoissonSolver " + solver + " needs NetworkFile (a trained network, gp_nn_save)");
    C.sortInterval = (int)ival(c, "SortInterval");
    C.sortDrift = dval(c, "SortDriftFraction");
    if (!(C.sortDrift >= 0.0) || C.sortDrift > 1.0) throw std::runtime_error("SortDriftFraction must lie in [0, 1]");
    // ParamForField (Init.mm:106-149)
    Json d = merged(&section(root, "ParamForField", Json::Object), defField());
    auto &F = paramForField;
    F.ngx = (int)ival(d, "NumberOfGridX"); F.ngy = (int)ival(d, "NumberOfGridY");
    F.dx = dval(d, "GridSizeOfX"); F.dy = dval(d, "GridSizeOfY");
    F.InitTypeE = sval(d, "InitializeTypeOfE"); F.ampE = vval(d, "AmplitudeOfE");
    F.FilePathEx = sval(d, "FilePathOfEx"); F.FilePathEy = sval(d, "FilePathOfEy"); F.FilePathEz = sval(d, "FilePathOfEz");
    F.InitTypeB = sval(d, "InitializeTypeOfB"); F.ampB = vval(d, "AmplitudeOfB");
    F.FilePathBx = sval(d, "FilePathOfBx"); F.FilePathBy = sval(d, "FilePathOfBy"); F.FilePathBz = sval(d, "FilePathOfBz");
    F.weightOrder = (int)ival(d, "WeightingOrder");
    F.ngb = F.weightOrder / 2;                       // Init.mm:140
    F.diff_ngb = (int)ival(d, "PtclPosOffset");
    F.BorisOrder = (int)ival(d, "BorisOrder");
    F.checkCurDens = (int)ival(d, "FlagForCurDens");
    F.pertDelta = dval(d, "PerturbationAmplitude");
    F.pertMode = (int)ival(d, "PerturbationMode");
    if (F.BorisOrder != 2 && F.BorisOrder != 4 && F.BorisOrder != 6)
        throw std::runtime_error("BorisOrder must be 2, 4 or 6");   // the reference leaves fN1/fN2 unset
    // BoundaryConditionForField (Init.mm:151-169)
    fieldBoundaries.clear();
    for (auto &e : section(root, "BoundaryConditionForField", Json::Array).arr) {
        if (!e.is_object()) continue;
        Json m = merged(&e, defBC());
        BoundaryConditionForField b;
        b.position = sval(m, "RegionName"); b.type = sval(m, "BCType"); b.val = dval(m, "Value");
        fieldBoundaries.push_back(b);
    }
    // ParamForParticle (Init.mm:171-223)
    paramForParticle.clear();
    for (auto &e : section(root, "ParamForParticle", Json::Array).arr) {
        if (!e.is_object()) continue;
        Json m = merged(&e, defParticle());
        ParamForParticle p;
        p.pName = sval(m, "ParticleName");
        const std::string pm = sval(m, "pNumSetMethod");
        if (pm == "InitialParticleNumber") p.pNum = ival(m, "pNumValue");
        else if (pm == "PtclNumPerCell") p.pNum = ival(m, "pNumValue") * (long long)F.ngx * F.ngy;
        else throw std::runtime_error("pNumSet failed.");
        p.q = dval(m, "Charge");
        p.m = dval(m, "Mass");
        const std::string wm = sval(m, "WeightSetMethod");
        if (wm == "WeightValue") p.w = dval(m, "WeightValue");
        else if (wm == "InitialNumDens") {
            const double ppc = (double)p.pNum / (double)((long long)F.ngx * F.ngy);
            p.w = dval(m, "WeightValue") * F.dx * F.dy / ppc;
        } else throw std::runtime_error("WeightSet failed.");
        p.genType = sval(m, "GenerateType");
        p.genX = vval(m, "InitialPosX"); p.genY = vval(m, "InitialPosY"); p.genU = vval(m, "InitialVel");
        p.genT = dval(m, "InitialTemp") * eVtoK;
        paramForParticle.push_back(p);
    }
    // SourceForParticle (Init.mm:226-276)
    particleSources.clear();
    for (auto &e : section(root, "SourceForParticle", Json::Array).arr) {
        if (!e.is_object()) continue;
        Json m = merged(&e, defSource());
        SourceForParticle s;
        s.pName = sval(m, "ParticleName");
        s.genType = sval(m, "GenerateType");
        const std::string sm = sval(m, "SrcSetMethod");
        if (sm == "Source[1/cm/s]") s.src = dval(m, "SrcValue");
        else if (sm == "CurrentDensity[A/m2]") s.src = dval(m, "SrcValue") * JtosJ * F.ngy * F.dy / ec;
        else if (sm == "BoundaryCurrentAtXmin") s.src = 0;
        else if (sm == "BoundaryCurrentAtXmax") s.src = 1;
        else if (sm == "BoundaryCurrentAtYmin") s.src = 2;
        else if (sm == "BoundaryCurrentAtYmax") s.src = 3;
        else throw std::runtime_error("SrcSet failed.");
        s.genX = vval(m, "GeneratePosX"); s.genY = vval(m, "GeneratePosY"); s.genU = vval(m, "GenerateVel");
        s.genT = dval(m, "GenerateTemp") * eVtoK;
        s.genU_ele = vval(m, "GenVelForElectron");
        s.genT_ele = dval(m, "GenTempForElectron") * eVtoK;
        particleSources.push_back(s);
    }
    // BoundaryConditionForParticle (Init.mm:279-298): RegionName/BCType
    // are read from the user dictionary itself
    particleBoundaries.clear();
    for (auto &e : section(root, "BoundaryConditionForParticle", Json::Array).arr) {
        if (!e.is_object()) continue;
        BoundaryConditionForParticle b;
        b.position = sval(e, "RegionName"); b.type = sval(e, "BCType");
        particleBoundaries.push_back(b);
    }
}

}  // namespace gample
