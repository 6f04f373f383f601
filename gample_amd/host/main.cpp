// main.cpp -- `gample -i condition.json [-o tag] [--unfused] [--dump-config]`; binary outputs and
// restart files in bin/ as the reference writes them (FileIO.cpp)
//
// The driver loop of the reference (src/main.mm:28-263): parse the
// condition file, build the modules, load particles, initial deposit and
// Poisson solve, then for each cycle: deposit -> Poisson [-> Darwin] ->
// push -> reduce, with the Kahan-compensated time (main.mm:150-163).  By
// default the deposit of the next cycle is fused into the push kernel
// (gp_step); --unfused runs the pieces in the reference's order.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <map>
#include <string>
#include <vector>

#include "Constant.h"
#include "Init.h"
#include "FileIO.h"
#include "Modules.h"

using namespace gample;

static std::map<std::string, std::string> parseArgs(int argc, char **argv)   // main.mm:15-26
{
    std::map<std::string, std::string> o;
    for (int i = 1; i < argc; i++) {
        std::string k = argv[i];
        if (k == "--unfused" || k == "--dump-config") { o[k] = "1"; continue; }
        if (k[0] == '-' && i + 1 < argc) { o[k] = argv[i + 1]; i++; }
    }
    return o;
}

static void dump(const Init &in)
{
    const auto &F = in.paramForField;
    std::printf("{\"Particle\":%d,\"EMField\":%d,\"Start\":%d,\"End\":%d,\"TimeStep\":%.17g,"
                "\"pNumMax\":%lld,\"maxiter\":%d,\"tolerance\":%.17g,\"ngx\":%d,\"ngy\":%d,\"ngb\":%d,"
                "\"dx\":%.17g,\"dy\":%.17g,\"weightOrder\":%d,\"BorisOrder\":%d,\"checkCurDens\":%d,"
                "\"SortInterval\":%d,\"SortDriftFraction\":%.17g,\"species\":[",
                in.flagForEquation.Particle, in.flagForEquation.EMField, in.paramForTimeIntegration.Start,
                in.paramForTimeIntegration.End, in.paramForTimeIntegration.TimeStep, in.paramForComputing.pNumMax,
                in.paramForComputing.maxiter, in.paramForComputing.tolerance, F.ngx, F.ngy, F.ngb, F.dx, F.dy,
                F.weightOrder, F.BorisOrder, F.checkCurDens, in.paramForComputing.sortInterval,
                in.paramForComputing.sortDrift);
    for (size_t s = 0; s < in.paramForParticle.size(); s++) {
        const auto &p = in.paramForParticle[s];
        std::printf("%s{\"name\":\"%s\",\"pNum\":%lld,\"q\":%.17g,\"m\":%.17g,\"w\":%.17g,\"genType\":\"%s\","
                    "\"genT\":%.17g}", s ? "," : "", p.pName.c_str(), p.pNum, p.q, p.m, p.w, p.genType.c_str(), p.genT);
    }
    std::printf("],\"pbc\":[");
    for (size_t i = 0; i < in.particleBoundaries.size(); i++)
        std::printf("%s[\"%s\",\"%s\"]", i ? "," : "", in.particleBoundaries[i].position.c_str(),
                    in.particleBoundaries[i].type.c_str());
    std::printf("],\"fbc\":[");
    for (size_t i = 0; i < in.fieldBoundaries.size(); i++)
        std::printf("%s[\"%s\",\"%s\",%.17g]", i ? "," : "", in.fieldBoundaries[i].position.c_str(),
                    in.fieldBoundaries[i].type.c_str(), in.fieldBoundaries[i].val);
    std::printf("],\"sources\":%zu}\n", in.particleSources.size());
}

int main(int argc, char **argv)
{
    auto args = parseArgs(argc, argv);
    if (!args.count("-i")) {
        std::fprintf(stderr, "Usage: %s -i inputfile.json [-o tag] [--unfused]\n", argv[0]);
        return 1;
    }
    Init init;
    try {
        init.parseInputFile(args["-i"]);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[Fatal] parseInputFile Failed: %s\n", e.what());
        return -1;
    }
    if (args.count("--dump-config")) { dump(init); return 0; }
    std::string tag;
    if (args.count("-o")) tag = args["-o"];
    else {
        char b[32];
        std::time_t t = std::time(nullptr);
        std::strftime(b, sizeof b, "%Y%m%d%H%M%S", std::localtime(&t));
        tag = b;
    }
    const auto &T = init.paramForTimeIntegration;
    const bool fused = !args.count("--unfused");
    try {
        Logger logger(T.ProjectName + "_" + tag, true);
        EMField fld(init, 0);
        std::vector<Particle> ptcl;
        ptcl.reserve(init.flagForEquation.Particle);
        for (int s = 0; s < init.flagForEquation.Particle; s++) ptcl.emplace_back(fld, init, s);
        Collision col(fld, init, ptcl);
        bool has_cathode = false;
        for (auto &b : init.fieldBoundaries) has_cathode |= b.position == "hollow-cathode";
        Moment mom(fld);
        const bool out_field = T.FieldOutput != 0, out_ptcl = T.ParticleOutput != 0;
        // the reference tests FieldOutput before taking the cycle modulo
        // ProgressOutput (main.mm:229); a zero ProgressOutput writes nothing
        const bool out_prog = T.FieldOutput != 0 && T.ProgressOutput != 0;
        ensureOutputDir();
        // restart (main.mm:125-133): particle states of cycle Start
        if (T.Start != 0) {
            loadProgress(T.Start, ptcl, init);
            for (auto &p : ptcl) check(gp_sort(fld.ctx, p.sid), "sort");
            std::printf("restarted from cycle %d\n", T.Start);
        }
        // initial field (main.mm:135-147) and its output
        fld.resetChargeDensity();
        for (auto &p : ptcl) p.integrateChargeDensity();
        gp_solve_stats st0 = fld.solvePoisson(0.0, nullptr);
        std::printf("initial solve: %d iterations, relres %.3e\n", st0.iterations, st0.relres);
        outputField(T.Start, fld, init);
        for (auto &p : ptcl) outputMoments(T.Start, fld, mom.integrateMoments(p), p.pName, init);
        if (fused) {   // accumulators for cycle 1 (the solve consumed them)
            fld.resetChargeDensity();
            for (auto &p : ptcl) p.integrateChargeDensity();
        }
        const double dt = T.TimeStep;
        double time = T.Start * dt, comp = 0.0;
        const int sort_every = init.paramForComputing.sortInterval;
        auto t_start = std::chrono::steady_clock::now();
        for (int cycle = T.Start + 1; cycle <= T.End; cycle++) {
            const double y = dt - comp, t = time + y;
            comp = (t - time) - y;
            time = t;
            const bool logc = T.LogOutput > 0 && cycle % T.LogOutput == 0;
            auto c0 = std::chrono::steady_clock::now();
            gp_solve_stats st;
            gp_diag dg;
            std::memset(&dg, 0, sizeof dg);
            std::vector<std::map<std::string, std::string>> flow;
            if (fused) {
                // the hollow cathode needs every cycle's wall counts
                const bool want = logc || col.existHC;
                check(gp_step(fld.ctx, dt, fld.tol, fld.maxiter, fld.solver, sort_every, cycle, &st, want ? &dg : nullptr),
                      "gp_step");
                if (want)
                    for (size_t s = 0; s < ptcl.size() && s < 8; s++) {   // Particle.mm:914-961
                        Particle &p = ptcl[s];
                        const int qe = (int)(p.q / (float)ec);
                        p.pinum_Xmin = (int)dg.absorbed[s][0] * qe; p.pinum_Xmax = (int)dg.absorbed[s][1] * qe;
                        p.pinum_Ymin = (int)dg.absorbed[s][2] * qe; p.pinum_Ymax = (int)dg.absorbed[s][3] * qe;
                    }
            } else {
                fld.resetChargeDensity();
                for (auto &p : ptcl) p.integrateChargeDensity();
                st = fld.solvePoisson(dt, logc ? &dg : nullptr);
                if (init.flagForEquation.EMField == 2) {
                    gp_solve_stats ds[3];
                    fld.solveVectorPotential(dt, ds);
                }
                for (auto &p : ptcl) {
                    p.update(dt, false);
                    flow.push_back(p.reduce());
                }
                if (sort_every > 0 && cycle % sort_every == 0)
                    for (auto &p : ptcl) check(gp_sort(fld.ctx, p.sid), "sort");
            }
            // particle sources after the reduce (main.mm:219-232); the fused
            // step deposits the new particles into the next solve's charge
            const int ai_addn = col.artificialIonization(dt, fused);
            int hc_neglected = 0;
            const int hc_addn = col.hollowCathode(fused, hc_neglected);
            // outputs of main.mm:190-233.  The fields after the push are those
            // of the solve (the push changes no field), so the fused step
            // writes them here too.
            if (out_field && cycle % T.FieldOutput == 0) outputField(cycle, fld, init);
            for (auto &p : ptcl) {
                if (out_ptcl && cycle % T.ParticleOutput == 0) outputPhaseSpace(cycle, p, init);
                if (out_field && cycle % T.FieldOutput == 0)
                    outputMoments(cycle, fld, mom.integrateMoments(p), p.pName, init);
            }
            if (out_prog && cycle % T.ProgressOutput == 0) saveProgress(cycle, ptcl, init);
            auto c1 = std::chrono::steady_clock::now();
            if (logc) {
                double phi, ex, ey;
                fld.probe(phi, ex, ey);
                logger.probe(cycle, time, phi, ex, ey);
                char b[64];
                std::map<std::string, std::string> row;
                row["cycle"] = std::to_string(cycle);
                std::snprintf(b, sizeof b, "%.9e", time); row["time"] = b;
                row["iteration"] = std::to_string(st.iterations);
                std::snprintf(b, sizeof b, "%.3e", st.relres); row["error"] = b;
                if (init.flagForEquation.EMField == 2) {
                    // the solveVectorPotential section's keys (EMField.mm:1349-1360)
                    gp_solve_stats ds[3];
                    check(gp_darwin_stats(fld.ctx, ds), "gp_darwin_stats");
                    static const char *key[3] = {"decompose", "solveAx", "solveAy"};
                    for (int k = 0; k < 3; k++) {
                        row[std::string(key[k]) + "_iteration"] = std::to_string(ds[k].iterations);
                        std::snprintf(b, sizeof b, "%.6e", ds[k].relres);
                        row[std::string(key[k]) + "_error"] = b;
                    }
                }
                if (col.existAI) row["artificialIonization_addn_injected"] = std::to_string(ai_addn);
                if (col.existHC) {
                    row["hollowCathode_addn_injected"] = std::to_string(hc_addn);
                    row["hollowCathode_addn_neglected"] = std::to_string(hc_neglected);
                }
                if (has_cathode) {
                    std::snprintf(b, sizeof b, "%.6e", dg.mean_cathode * sVtoV); row["meanCathode"] = b;
                }
                std::snprintf(b, sizeof b, "%.6e", dg.we_ref * ergtoev); row["totalElectricEnergy"] = b;
                std::snprintf(b, sizeof b, "%.9e", dg.we_cgs); row["fieldEnergyCGS"] = b;
                std::snprintf(b, sizeof b, "%.9e", dg.kinetic[0] + dg.kinetic[1]); row["kineticEnergy"] = b;
                row["elapsed_us"] = std::to_string(
                    std::chrono::duration_cast<std::chrono::microseconds>(c1 - c0).count());
                for (size_t s = 0; s < ptcl.size(); s++) {
                    const Particle &p = ptcl[s];
                    row["pNum_" + p.pName] = std::to_string(p.count());
                    // the flowout_<species> wall counts in charge units (Particle.mm:990-1001)
                    row[p.pName + "_Xmin"] = std::to_string(p.pinum_Xmin);
                    row[p.pName + "_Xmax"] = std::to_string(p.pinum_Xmax);
                    row[p.pName + "_Ymin"] = std::to_string(p.pinum_Ymin);
                    row[p.pName + "_Ymax"] = std::to_string(p.pinum_Ymax);
                }
                logger.row(row);
                std::printf("Frame %d completed\n", cycle);
            }
        }
        auto t_end = std::chrono::steady_clock::now();
        std::printf("calculation end. total elapsed time: %.3f sec.\n",
                    std::chrono::duration<double>(t_end - t_start).count());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[Fatal] %s\n", e.what());
        return 1;
    }
    return 0;
}
