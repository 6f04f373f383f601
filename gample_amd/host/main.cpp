// main.cpp -- `gample -i condition.json [-o tag] [--unfused] [--dump-config]`; binary outputs and
// restart files in bin/ as the reference writes them (FileIO.cpp)
//
// The driver loop of the reference (src/main.mm:28-263): parse the
// condition file, build the modules, load particles, initial deposit and
// Poisson solve, then for each cycle: deposit -> Poisson [-> Darwin] ->
// push -> reduce, with the Kahan-compensated time (main.mm:150-163).  By
// default the deposit of the next cycle is fused into the push kernel
// (gp_step); --unfused runs the pieces in the reference's order.  Every
// LogOutput cycles the run log <ProjectName>_<tag>_log.xml receives the
// reference's sections (solvePoisson, solveVectorPotential, outputField,
// flowout_<species>, outputMoment_<species>, artificialIonization,
// hollowCathode, elapsedTime, memoryUsage) and _timeseries.csv a probe row.
// Under the fused step elapsedTime holds gp_step, the sources and the
// step's phases as the device timed them with stream events: solvePoisson,
// update_<s> (the push with the next cycle's deposit fused in, so
// integCDens_<s> is 0), reduce_<s> and sort_<s>; the reference-order pieces
// time integCDens_, solvePoisson, update_, reduce_ on the host.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
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

static std::string dump(const Init &in)
{
    const auto &F = in.paramForField;
    std::string o;
    char b[1024];
    std::snprintf(b, sizeof b,
                  "{\"Particle\":%d,\"EMField\":%d,\"Start\":%d,\"End\":%d,\"TimeStep\":%.17g,"
                  "\"pNumMax\":%lld,\"maxiter\":%d,\"tolerance\":%.17g,\"ngx\":%d,\"ngy\":%d,\"ngb\":%d,"
                  "\"dx\":%.17g,\"dy\":%.17g,\"weightOrder\":%d,\"BorisOrder\":%d,\"checkCurDens\":%d,"
                  "\"SortInterval\":%d,\"SortDriftFraction\":%.17g,\"species\":[",
                  in.flagForEquation.Particle, in.flagForEquation.EMField, in.paramForTimeIntegration.Start,
                  in.paramForTimeIntegration.End, in.paramForTimeIntegration.TimeStep, in.paramForComputing.pNumMax,
                  in.paramForComputing.maxiter, in.paramForComputing.tolerance, F.ngx, F.ngy, F.ngb, F.dx, F.dy,
                  F.weightOrder, F.BorisOrder, F.checkCurDens, in.paramForComputing.sortInterval,
                  in.paramForComputing.sortDrift);
    o += b;
    for (size_t s = 0; s < in.paramForParticle.size(); s++) {
        const auto &p = in.paramForParticle[s];
        std::snprintf(b, sizeof b,
                      "%s{\"name\":\"%s\",\"pNum\":%lld,\"q\":%.17g,\"m\":%.17g,\"w\":%.17g,\"genType\":\"%s\","
                      "\"genT\":%.17g}",
                      s ? "," : "", p.pName.c_str(), p.pNum, p.q, p.m, p.w, p.genType.c_str(), p.genT);
        o += b;
    }
    o += "],\"pbc\":[";
    for (size_t i = 0; i < in.particleBoundaries.size(); i++) {
        std::snprintf(b, sizeof b, "%s[\"%s\",\"%s\"]", i ? "," : "", in.particleBoundaries[i].position.c_str(),
                      in.particleBoundaries[i].type.c_str());
        o += b;
    }
    o += "],\"fbc\":[";
    for (size_t i = 0; i < in.fieldBoundaries.size(); i++) {
        std::snprintf(b, sizeof b, "%s[\"%s\",\"%s\",%.17g]", i ? "," : "", in.fieldBoundaries[i].position.c_str(),
                      in.fieldBoundaries[i].type.c_str(), in.fieldBoundaries[i].val);
        o += b;
    }
    std::snprintf(b, sizeof b, "],\"sources\":%zu}", in.particleSources.size());
    return o + b;
}

// resident set of this process in kB (/proc/self/status VmRSS)
static long rss_kb()
{
    FILE *f = std::fopen("/proc/self/status", "r");
    if (!f) return 0;
    char line[256];
    long kb = 0;
    while (std::fgets(line, sizeof line, f))
        if (std::strncmp(line, "VmRSS:", 6) == 0) { kb = std::atol(line + 6); break; }
    std::fclose(f);
    return kb;
}

// the solvePoisson section (EMField.mm:1118-1127)
static std::map<std::string, std::string> poisson_section(const gp_solve_stats &st, const gp_diag &dg, double &lastEE)
{
    const double diff = dg.we_ref - lastEE;
    lastEE = dg.we_ref;
    return {{"iteration", std::to_string(st.iterations)},
            {"error", fmtSci((float)st.relres, 6)},
            {"meanCathode", fmtSci((float)(dg.mean_cathode * sVtoV), 6)},
            {"totalElectricEnergy", fmtSci((float)(dg.we_ref * ergtoev), 6)},
            {"totalEEincrement", fmtSci((float)(diff * ergtoev), 6)},
            {"jouleHeating", fmtSci((float)(dg.joule * ergtoev), 6)}};
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
    if (args.count("--dump-config")) { std::printf("%s\n", dump(init).c_str()); return 0; }
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
        // main.mm:80-88: <ProjectName>_<tag>_log.xml and _timeseries.csv
        XmlLogger logger(T.ProjectName + "_" + tag, "cycle,time,phi1,Ex1,Ey1");
        logger.logComment("parsed condition: " + dump(init) + "\n");
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
        const int LO = T.LogOutput;
        // device phase times of a logged fused step (HIP events on the
        // context stream, recorded for that step only)
        auto device_us = [&](int which) {
            double ms = 0.0;
            int64_t n = 0;
            check(gp_kernel_time(fld.ctx, which, &ms, &n), "gp_kernel_time");
            return std::to_string((long long)std::llround(ms * 1000.0));
        };
        double lastEE = 0.0;
        for (int cycle = T.Start + 1; cycle <= T.End; cycle++) {
            const bool logc = LO > 0 && cycle % LO == 0;
            logger.switchLog(logc);
            const double y = dt - comp, t = time + y;
            comp = (t - time) - y;
            time = t;
            logger.logCycleStart(cycle, time);
            // the increments of the logged energies are against the previous
            // cycle, so the diagnostics are also taken the cycle before a
            // logged one; the hollow cathode needs every cycle's wall counts
            const bool want = logc || (LO > 0 && (cycle + 1) % LO == 0) || col.existHC;
            const bool fieldc = out_field && cycle % T.FieldOutput == 0;
            std::map<std::string, std::string> elapsed;
            auto timed = [&](const std::string &name, auto &&fn) {   // ResourceUtils.h:40-48
                auto a0 = std::chrono::steady_clock::now();
                fn();
                auto a1 = std::chrono::steady_clock::now();
                elapsed[name] = std::to_string(std::chrono::duration_cast<std::chrono::microseconds>(a1 - a0).count());
            };
            gp_solve_stats st;
            gp_diag dg;
            std::memset(&st, 0, sizeof st);
            std::memset(&dg, 0, sizeof dg);
            // the solve's log sections on logged cycles (the energy of the cycle
            // before one is kept for the increment), then the field output
            auto field_sections = [&]() {
                if (logc) {
                    logger.logSection("solvePoisson", poisson_section(st, dg, lastEE));
                    double phi, ex, ey;
                    fld.probe(phi, ex, ey);
                    logger.csvOutput(fmtSci((float)phi, 6) + "," + fmtSci((float)ex, 6) + "," + fmtSci((float)ey, 6));
                    if (init.flagForEquation.EMField == 2) {
                        // the solveVectorPotential section (EMField.mm:1349-1360)
                        gp_solve_stats ds[3];
                        check(gp_darwin_stats(fld.ctx, ds), "gp_darwin_stats");
                        static const char *key[3] = {"decompose", "solveAx", "solveAy"};
                        std::map<std::string, std::string> d;
                        for (int k = 0; k < 3; k++) {
                            d[std::string(key[k]) + "_iteration"] = std::to_string(ds[k].iterations);
                            d[std::string(key[k]) + "_error"] = fmtSci((float)ds[k].relres, 6);
                        }
                        logger.logSection("solveVectorPotential", d);
                    }
                } else if (want) {
                    lastEE = dg.we_ref;
                }
                if (fieldc) logger.logSection("outputField", outputField(cycle, fld, init));
            };
            auto species_outputs = [&](Particle &p, const std::map<std::string, std::string> &flow) {
                logger.logSection("flowout_" + p.pName, flow);
                if (out_ptcl && cycle % T.ParticleOutput == 0) outputPhaseSpace(cycle, p, init);
                if (fieldc)
                    logger.logSection("outputMoment_" + p.pName,
                                      outputMoments(cycle, fld, mom.integrateMoments(p), p.pName, init));
            };
            if (fused) {
                if (logc) check(gp_timing(fld.ctx, 1), "gp_timing");
                timed("gp_step", [&] {
                    check(gp_step(fld.ctx, dt, fld.tol, fld.maxiter, fld.solver, sort_every, cycle, &st,
                                  want ? &dg : nullptr), "gp_step");
                });
                if (logc) {
                    elapsed["solvePoisson"] = device_us(2);
                    for (size_t s = 0; s < ptcl.size(); s++) {
                        const int b = 4 + 3 * (int)s;
                        elapsed["integCDens_" + ptcl[s].pName] = "0";
                        elapsed["update_" + ptcl[s].pName] = device_us(b);
                        elapsed["reduce_" + ptcl[s].pName] = device_us(b + 1);
                        elapsed["sort_" + ptcl[s].pName] = device_us(b + 2);
                    }
                    check(gp_timing(fld.ctx, 0), "gp_timing");
                }
                // the fields after the push are those of the solve (the push
                // changes no field), so the field outputs follow the step
                field_sections();
                for (size_t s = 0; s < ptcl.size(); s++) {
                    Particle &p = ptcl[s];
                    if (!want || s >= 8) {
                        if (fieldc || (out_ptcl && cycle % T.ParticleOutput == 0)) species_outputs(p, {});
                        continue;
                    }
                    const int qe = (int)(p.q / (float)ec);   // Particle.mm:914-961
                    const int64_t *a = dg.absorbed[s];
                    p.pinum_Xmin = (int)a[0] * qe; p.pinum_Xmax = (int)a[1] * qe;
                    p.pinum_Ymin = (int)a[2] * qe; p.pinum_Ymax = (int)a[3] * qe;
                    species_outputs(p, p.flowout(a[0] + a[1] + a[2] + a[3], dg.kinetic[s]));
                }
            } else {
                fld.resetChargeDensity();
                for (auto &p : ptcl) timed("integCDens_" + p.pName, [&] { p.integrateChargeDensity(); });
                timed("solvePoisson", [&] { st = fld.solvePoisson(dt, want ? &dg : nullptr); });
                if (init.flagForEquation.EMField == 2) {
                    gp_solve_stats ds[3];
                    timed("solveVectorPotential", [&] { fld.solveVectorPotential(dt, ds); });
                }
                field_sections();
                for (auto &p : ptcl) {
                    std::map<std::string, std::string> flow;
                    timed("update_" + p.pName, [&] { p.update(dt, false); });
                    timed("reduce_" + p.pName, [&] { flow = p.reduce(); });
                    species_outputs(p, flow);
                }
                if (sort_every > 0 && cycle % sort_every == 0)
                    for (auto &p : ptcl) check(gp_sort(fld.ctx, p.sid), "sort");
            }
            // particle sources after the reduce and the species outputs
            // (main.mm:219-232); the fused step deposits the new particles
            // into the next solve's charge
            int ai_addn = 0, hc_addn = 0, hc_neglected = 0;
            timed("artificialIonization", [&] { ai_addn = col.artificialIonization(dt, fused); });
            if (col.existAI) logger.logSection("artificialIonization", {{"addn_injected", std::to_string(ai_addn)}});
            timed("hollowCathode", [&] { hc_addn = col.hollowCathode(fused, hc_neglected); });
            if (col.existHC)
                logger.logSection("hollowCathode", {{"addn_neglected", std::to_string(hc_neglected)},
                                                    {"addn_injected", std::to_string(hc_addn)}});
            if (out_prog && cycle % T.ProgressOutput == 0) saveProgress(cycle, ptcl, init);
            // resources (main.mm:242-248): times in us, memory in kB (the
            // resident set for both; the device memory is not counted)
            logger.logSection("elapsedTime", elapsed);
            const std::string rss = std::to_string(rss_kb());
            logger.logSection("memoryUsage", {{"physicalFootprint", rss}, {"residentSetSize", rss}});
            if (logc) std::printf("Frame %d completed\n", cycle);
            logger.logCycleEnd();
        }
        auto t_end = std::chrono::steady_clock::now();
        const double secs = std::chrono::duration<double>(t_end - t_start).count();
        std::printf("calculation end. total elapsed time: %.3f sec.\n", secs);
        logger.logComment("culculation end.\ntotal elapsed time: " + std::to_string((long)secs) + " sec.\n");
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[Fatal] %s\n", e.what());
        return 1;
    }
    return 0;
}
