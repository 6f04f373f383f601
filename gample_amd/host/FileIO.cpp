// FileIO.cpp -- the reference's binary outputs and restart files (FileIO.mm).
#include "FileIO.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <stdexcept>

#include "Constant.h"
#include "XmlLogger.h"

namespace gample {

namespace {

std::string fmt_cycle(int cycle)
{
    char b[16];
    std::snprintf(b, sizeof b, "%08d", cycle);
    return b;
}

struct File {
    FILE *f = nullptr;
    File(const std::string &path, const char *mode) : f(std::fopen(path.c_str(), mode)) {}
    ~File() { if (f) std::fclose(f); }
    void put(const void *p, size_t bytes, const std::string &path)
    {
        if (bytes && std::fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("write failed: " + path);
    }
};

// header of a field or moment file (FileIO.mm:117-121)
void header(File &o, const std::string &path, const EMField &fld)
{
    const int h[3] = {fld.ngx, fld.ngy, fld.ngb};
    const float d[2] = {(float)fld.dx, (float)fld.dy};
    o.put(h, sizeof h, path);
    o.put(d, sizeof d, path);
}

// one record (FileIO.mm:228-250 writeField): 32-byte name, type id, the
// array times the unit factor, both in FP32 as the reference multiplies them
template <typename T>
void record(File &o, const std::string &path, const std::string &name, int type, const T *a, size_t n, float scale)
{
    char nb[32] = {};
    std::strncpy(nb, name.c_str(), sizeof nb - 1);
    o.put(nb, sizeof nb, path);
    o.put(&type, sizeof type, path);
    std::vector<float> v(n);
    for (size_t i = 0; i < n; i++) v[i] = (float)a[i] * scale;
    o.put(v.data(), v.size() * sizeof(float), path);
}

// min and max of an array as floats (FileIO.mm:68-88 keeps float arrays)
template <typename T>
std::pair<float, float> minmax(const T *a, size_t n)
{
    float lo = 1e20f, hi = -1e20f;
    for (size_t i = 0; i < n; i++) {
        const float v = (float)a[i];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    return {lo, hi};
}

void put_minmax(std::map<std::string, std::string> &d, const std::string &key, std::pair<float, float> mm, double scale)
{
    d[key + "_min"] = fmtSci((float)(mm.first * scale), 6);
    d[key + "_max"] = fmtSci((float)(mm.second * scale), 6);
}

template <typename T>
std::vector<T> get(EMField &fld, int which, size_t n)
{
    std::vector<T> v(n);
    check(gp_field_get(fld.ctx, which, v.data(), (int64_t)(n * sizeof(T))), "field output");
    return v;
}

}  // namespace

void ensureOutputDir()
{
    if (mkdir("bin", 0755) != 0 && errno != EEXIST) throw std::runtime_error("cannot create the output directory bin");
}

std::map<std::string, std::string> outputField(int cycle, EMField &fld, const Init &init)
{
    const std::string path = "bin/" + init.paramForTimeIntegration.ProjectName + "_EMField_" + fmt_cycle(cycle) + ".bin";
    File o(path, "wb");
    if (!o.f) throw std::runtime_error("Unable to open file " + path + " for writing");
    const size_t nx = fld.nx, ny = fld.ny;
    const size_t nN = (nx + 1) * (ny + 1), nEx = (nx + 2) * (ny + 1), nEy = (nx + 1) * (ny + 2);
    const size_t nB = (nx + 2) * (ny + 2), nP = (nx + 3) * (ny + 3);
    header(o, path, fld);
    auto rho = get<double>(fld, GP_F_RHO, nN);
    auto phi = get<double>(fld, GP_F_PHI, nP);
    auto ex = get<float>(fld, GP_F_EX, nEx), ey = get<float>(fld, GP_F_EY, nEy), bz = get<float>(fld, GP_F_BZ, nB);
    record(o, path, "rho", 0, rho.data(), nN, 1.0f);
    record(o, path, "phi", 4, phi.data(), nP, (float)sVtoV);
    record(o, path, "Ex", 1, ex.data(), nEx, (float)GtoV);
    record(o, path, "Ey", 2, ey.data(), nEy, (float)GtoV);
    record(o, path, "Bz", 3, bz.data(), nB, (float)GtoT);
    if (init.flagForEquation.EMField == 2) {
        auto jx = get<double>(fld, GP_F_JX, nEx), jy = get<double>(fld, GP_F_JY, nEy);
        auto dex = get<float>(fld, GP_F_DEL_EX, nEx), dey = get<float>(fld, GP_F_DEL_EY, nEy);
        auto dbz = get<float>(fld, GP_F_DEL_BZ, nB);
        record(o, path, "jx", 1, jx.data(), nEx, (float)sJtoJ);
        record(o, path, "jy", 2, jy.data(), nEy, (float)sJtoJ);
        record(o, path, "delEx", 1, dex.data(), nEx, (float)GtoV);
        record(o, path, "delEy", 2, dey.data(), nEy, (float)GtoV);
        record(o, path, "delBz", 3, dbz.data(), nB, (float)GtoT);
    }
    // the outputField log section (FileIO.mm:89-101)
    std::map<std::string, std::string> d;
    put_minmax(d, "rho", minmax(rho.data(), nN), 1.0);
    put_minmax(d, "phi", minmax(phi.data(), nP), sVtoV);
    put_minmax(d, "Ex", minmax(ex.data(), nEx), GtoV);
    put_minmax(d, "Ey", minmax(ey.data(), nEy), GtoV);
    return d;
}

std::map<std::string, std::string> outputMoments(int cycle, EMField &fld, const std::vector<double> &mom,
                                                 const std::string &pName, const Init &init)
{
    const std::string path = "bin/" + init.paramForTimeIntegration.ProjectName + "_Moments_" + pName + "_" +
                             fmt_cycle(cycle) + ".bin";
    File o(path, "wb");
    if (!o.f) throw std::runtime_error("Unable to open file " + path + " for writing");
    const size_t ng = (size_t)(fld.nx + 1) * (fld.ny + 1);
    if (mom.size() != 10 * ng) throw std::runtime_error("moment array size");
    header(o, path, fld);
    static const char *suffix[10] = {"_n", "_ux", "_uy", "_uz", "_Pxx", "_Pxy", "_Pxz", "_Pyy", "_Pyz", "_Pzz"};
    for (int k = 0; k < 10; k++)   // 1/cm3, cm/s, and 0.1 Pa -> Pa
        record(o, path, pName + suffix[k], 0, mom.data() + k * ng, ng, k < 4 ? 1.0f : 0.1f);
    // the outputMoment_<species> log section (FileIO.mm:181-197); uz is
    // computed here, where the reference's loop repeats uy (FileIO.mm:178)
    std::map<std::string, std::string> d;
    for (int k = 0; k < 10; k++) put_minmax(d, suffix[k] + 1, minmax(mom.data() + k * ng, ng), k < 4 ? 1.0 : 0.1);
    return d;
}

void outputPhaseSpace(int cycle, Particle &p, const Init &init)
{
    constexpr int NS = 20;   // sample particles 0..19 of the store
    float x[NS], y[NS], vx[NS], vy[NS], vz[NS];
    int64_t n = 0;
    check(gp_species_read(p.fld.ctx, p.sid, 0, NS, &n, x, y, vx, vy, vz), "outputPhaseSpace");
    const auto &F = init.paramForField;
    for (int idx = 0; idx < (int)n; idx++) {
        const std::string path = "bin/" + init.paramForTimeIntegration.ProjectName + "_PhaseSpace_" + p.pName + "_" +
                                 std::to_string(idx) + ".bin";
        File o(path, "ab");
        if (!o.f) { std::fprintf(stderr, "Failed to open file: %s\n", path.c_str()); continue; }
        // back to physical units (FileIO.mm:23-25)
        const float px = (float)((double)(x[idx] - (float)F.ngb) * F.dx);
        const float py = (float)((double)(y[idx] - (float)F.ngb) * F.dy);
        o.put(&cycle, sizeof cycle, path);
        const float rec[5] = {px, py, vx[idx], vy[idx], vz[idx]};
        o.put(rec, sizeof rec, path);
    }
}

void saveProgress(int cycle, std::vector<Particle> &ptcl, const Init &init)
{
    for (auto &p : ptcl) {
        const std::string path = "bin/" + init.paramForTimeIntegration.ProjectName + "_ProgressData_" + p.pName + "_" +
                                 fmt_cycle(cycle) + ".bin";
        int64_t n = 0;
        check(gp_species_count(p.fld.ctx, p.sid, &n), "saveProgress");
        std::vector<float> x(n), y(n), vx(n), vy(n), vz(n);
        int64_t got = 0;
        check(gp_species_download(p.fld.ctx, p.sid, n, &got, x.data(), y.data(), vx.data(), vy.data(), vz.data()),
              "saveProgress");
        File o(path, "wb");
        if (!o.f) throw std::runtime_error("[Error] file open failed: " + path);
        struct Rec { float x, y, vx, vy, vz; int32_t flag; };   // ParticleState, Particle.h:14-21
        static_assert(sizeof(Rec) == 24, "24-byte particle record");
        std::vector<Rec> r((size_t)got);
        for (int64_t k = 0; k < got; k++) r[k] = Rec{x[k], y[k], vx[k], vy[k], vz[k], 0};
        o.put(r.data(), r.size() * sizeof(Rec), path);
    }
}

void loadProgress(int cycle, std::vector<Particle> &ptcl, const Init &init)
{
    const int shift = init.paramForField.diff_ngb;   // PtclPosOffset = ngb_new - ngb_old
    for (auto &p : ptcl) {
        const std::string path = "bin/" + init.paramForTimeIntegration.RestartName + "_ProgressData_" + p.pName + "_" +
                                 fmt_cycle(cycle) + ".bin";
        File in(path, "rb");
        if (!in.f) throw std::runtime_error("[Error] file open failed: " + path);
        std::fseek(in.f, 0, SEEK_END);
        const long bytes = std::ftell(in.f);
        std::rewind(in.f);
        if (bytes < 0 || bytes % 24) throw std::runtime_error("[Error] progress file size is not a multiple of 24: " + path);
        const size_t n = (size_t)bytes / 24;
        std::vector<unsigned char> raw((size_t)bytes);
        if (bytes && std::fread(raw.data(), 1, raw.size(), in.f) != raw.size())
            throw std::runtime_error("[Error] read failed: " + path);
        std::vector<float> a[5];
        for (auto &v : a) v.resize(n);
        for (size_t k = 0; k < n; k++)
            for (int c = 0; c < 5; c++) std::memcpy(&a[c][k], raw.data() + 24 * k + 4 * c, 4);
        if (shift != 0)
            for (size_t k = 0; k < n; k++) { a[0][k] += (float)shift; a[1][k] += (float)shift; }
        check(gp_species_upload(p.fld.ctx, p.sid, (int64_t)n, a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                                a[4].data()),
              "loadProgress");
    }
}

}  // namespace gample
