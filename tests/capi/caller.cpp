// A C++ caller of the C ABI, with no Python: reads a model, the trajectory
// parameters and a hook plan written by the test, runs
// lmx::MpdRdmeSolver::generateTrajectory (include/lmx_mpd_solver.hpp, the
// reference's output and hook cadence over include/lmx_mpd.h) and writes
// every record the trajectory produced.  Test infrastructure (built by
// tests/capi/build.py); it is the path a Lattice Microbes C++ shim takes
// (MpdRdmeSolver::initialize / buildModel / buildDiffusionModel /
// generateTrajectory / hookCheckSimulation, MpdRdmeSolver.cu:107-668,
// 958-985).
//
// Input (little endian): uint32 nS, R, nST, X, Y, Z, P, kCols, replicate,
// seedTop, boundary, overflowHandler (0 classic, 1 relaxed), hookEnabled;
// double tau, lambda, maxTime, writeInterval, latticeWriteInterval,
// hookInterval; uint32 initialCounts[nS], types[R]; double K[R*kCols]; int32
// S[nS*R]; uint32 D[nS*R]; double DF[nST*nST*nS]; uint32 RL[R*nST]; uint8
// lattice[X*Y*Z*P], sites[X*Y*Z]; the onBeginTrajectory edit (int32 x, y, z,
// species, siteType; x < 0: none); uint32 nplan and per hook step: uint32
// step; int32 x, y, z, species, siteType (x < 0: no edit; siteType < 0:
// unchanged), code, abort; double rate of reaction 0 (NaN: unchanged).
// An edit empties the site and puts the species in slot 0.
//
// Output: records of uint32 kind, uint32 step, double time, uint64 bytes and
// the payload: 1 species counts (step = entries; times then counts), 2
// lattice frame (row-major), 3 site frame (row-major), 4 MaxParticleCounts,
// 5 MaxSiteCounts, 6 the lattice onEndTrajectory sees (native words), 7 a
// hook call.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "lmx_mpd_solver.hpp"

namespace {

template <typename T>
bool read_vec(std::FILE *f, std::vector<T> &v, size_t n)
{
    v.resize(n);
    return n == 0 || std::fread(v.data(), sizeof(T), n, f) == n;
}

struct Edit {
    int32_t x = -1, y = 0, z = 0, species = 0, siteType = -1;
    void apply(lmx::HostLattice &lat) const
    {
        if (x < 0) return;
        lat.removeParticles((uint32_t)x, (uint32_t)y, (uint32_t)z);
        lat.setParticle((uint32_t)x, (uint32_t)y, (uint32_t)z, 0, (uint32_t)species);
        if (siteType >= 0) lat.setSiteType((uint32_t)x, (uint32_t)y, (uint32_t)z, (uint8_t)siteType);
    }
};

struct PlanEntry {
    Edit edit;
    int32_t code = 0, abort = 0;
    double rate = NAN;
};

class Records : public lmx::TrajectoryOutput {
public:
    explicit Records(std::FILE *f, double tau) : f_(f), tau_(tau) {}
    void put(uint32_t kind, uint32_t step, double time, const void *data, uint64_t bytes)
    {
        std::fwrite(&kind, 4, 1, f_);
        std::fwrite(&step, 4, 1, f_);
        std::fwrite(&time, 8, 1, f_);
        std::fwrite(&bytes, 8, 1, f_);
        if (bytes) std::fwrite(data, 1, bytes, f_);
    }
    uint32_t step_of(double time) const { return (uint32_t)std::llround(time / tau_); }
    void speciesCounts(uint32_t, const std::vector<double> &times, const std::vector<uint32_t> &counts,
                       uint32_t) override
    {
        std::vector<uint8_t> buf(times.size() * 8 + counts.size() * 4);
        std::memcpy(buf.data(), times.data(), times.size() * 8);
        std::memcpy(buf.data() + times.size() * 8, counts.data(), counts.size() * 4);
        put(1, (uint32_t)times.size(), 0.0, buf.data(), buf.size());
    }
    void lattice(uint32_t, double time, const lmx::HostLattice &lat) override
    {
        const std::vector<uint8_t> rm = lat.rowMajorParticles();
        put(2, step_of(time), time, rm.data(), rm.size());
    }
    void sites(uint32_t, double time, const lmx::HostLattice &lat) override
    {
        const std::vector<uint8_t> rm = lat.rowMajorSites();
        put(3, step_of(time), time, rm.data(), rm.size());
    }
    void maxCounts(uint32_t, const std::vector<uint32_t> &particles, const std::vector<uint32_t> &sites) override
    {
        put(4, 0, 0.0, particles.data(), particles.size() * 4);
        put(5, 0, 0.0, sites.data(), sites.size() * 4);
    }

private:
    std::FILE *f_;
    double tau_;
};

// A solver whose hooks follow the plan.
class PlannedSolver : public lmx::MpdRdmeSolver {
public:
    PlannedSolver(Edit begin, std::map<uint32_t, PlanEntry> plan, Records *rec)
        : lmx::MpdRdmeSolver(0), begin_(begin), plan_(std::move(plan)), rec_(rec)
    {
    }

protected:
    int onBeginTrajectory(lmx::HostLattice &lat) override
    {
        begin_.apply(lat);
        return 0;
    }
    int hookSimulation(double time, lmx::HostLattice &lat) override
    {
        const uint32_t step = rec_->step_of(time);
        rec_->put(7, step, time, &time, 0);
        auto it = plan_.find(step);
        if (it == plan_.end()) return 0;
        const PlanEntry &e = it->second;
        e.edit.apply(lat);
        if (!std::isnan(e.rate)) setReactionRate(0, (float)e.rate);
        if (e.abort) lmx::request_abort();
        return e.code;
    }
    int onEndTrajectory(lmx::HostLattice &lat) override
    {
        rec_->put(6, 0, 0.0, lat.words.data(), lat.words.size() * 4);
        return 0;
    }

private:
    Edit begin_;
    std::map<uint32_t, PlanEntry> plan_;
    Records *rec_;
};

std::string num(double v)
{
    char b[64];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s model.bin out.bin\n", argv[0]);
        return 1;
    }
    std::FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 1;
    uint32_t h[13];
    double fp[6];
    if (std::fread(h, sizeof(uint32_t), 13, f) != 13 || std::fread(fp, sizeof(double), 6, f) != 6) return 1;
    const uint32_t nS = h[0], R = h[1], nST = h[2], X = h[3], Y = h[4], Z = h[5], P = h[6], kCols = h[7];
    const uint32_t replicate = h[8], seedTop = h[9], boundary = h[10], overflow = h[11], hookEnabled = h[12];
    const size_t N = (size_t)X * Y * Z;
    std::vector<uint32_t> init, types, D, RL;
    std::vector<double> K, DF;
    std::vector<int32_t> S;
    std::vector<uint8_t> lattice, sites;
    bool ok = read_vec(f, init, nS) && read_vec(f, types, R) && read_vec(f, K, (size_t)R * kCols) &&
              read_vec(f, S, (size_t)nS * R) && read_vec(f, D, (size_t)nS * R) &&
              read_vec(f, DF, (size_t)nST * nST * nS) && read_vec(f, RL, (size_t)R * nST) &&
              read_vec(f, lattice, N * P) && read_vec(f, sites, N);
    Edit begin;
    ok = ok && std::fread(&begin, sizeof(int32_t), 5, f) == 5;
    uint32_t nplan = 0;
    ok = ok && std::fread(&nplan, sizeof nplan, 1, f) == 1;
    std::map<uint32_t, PlanEntry> plan;
    for (uint32_t k = 0; ok && k < nplan; k++) {
        uint32_t step = 0;
        int32_t v[7];
        double rate = 0.0;
        ok = std::fread(&step, 4, 1, f) == 1 && std::fread(v, 4, 7, f) == 7 && std::fread(&rate, 8, 1, f) == 1;
        PlanEntry e;
        e.edit = Edit{v[0], v[1], v[2], v[3], v[4]};
        e.code = v[5];
        e.abort = v[6];
        e.rate = rate;
        plan[step] = e;
    }
    std::fclose(f);
    if (!ok) return 1;

    std::FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 1;
    Records rec(o, fp[0]);
    try {
        PlannedSolver s(begin, plan, &rec);
        std::map<std::string, std::string> prm = {
            {"timestep", num(fp[0])}, {"seed", std::to_string(seedTop)}, {"maxTime", num(fp[2])},
            {"writeInterval", num(fp[3])}, {"latticeWriteInterval", num(fp[4])},
            {"rdme.mpd.overflowhandler", overflow ? "relaxed" : "classic"}};
        if (hookEnabled) prm["hookInterval"] = num(fp[5]);
        s.initialize(replicate, prm, (int)boundary);
        s.buildModel(nS, R, init.data(), types.data(), K.data(), S.data(), D.data(), kCols);
        s.buildDiffusionModel(nST, DF.data(), RL.empty() ? nullptr : RL.data(), X, Y, Z, P, 1, fp[1],
                              lattice.data(), sites.data(), true);
        s.generateTrajectory(&rec);
    } catch (const lmx::Exception &e) {
        std::fprintf(stderr, "lmx error %d: %s\n", e.code, e.what());
        std::fclose(o);
        return 2;
    }
    std::fclose(o);
    std::printf("capi caller: trajectory to %s on %ux%ux%u\n", fp[2] > 0 ? num(fp[2]).c_str() : "0", X, Y, Z);
    return 0;
}
