/*
 * lmx_mpd_solver.hpp: lm::rdme::MpdRdmeSolver's host loop in C++, over the C
 * ABI of lmx_mpd.h (header-only; link liblmx_mpd.so).
 *
 * lmx::MpdRdmeSolver keeps the reference class's protocol
 * (src/rdme/MpdRdmeSolver.h:66-156): initialize with the string parameter
 * map, buildModel, buildDiffusionModel, setReactionRate, generateTrajectory,
 * and the overridable hooks hookSimulation / onBeginTrajectory /
 * onEndTrajectory, which see a host copy of the lattice (HostLattice, in
 * place of CudaByteLattice / CudaIntLattice).  generateTrajectory follows
 * MpdRdmeSolver::generateTrajectory (MpdRdmeSolver.cu:485-668):
 *
 *   - "writeInterval", "latticeWriteInterval" and "hookInterval" are step
 *     counts (read as uint32_t from atof, :491-520); an interval of 0 fires
 *     after every step; hooks run only if "hookInterval" is given;
 *   - steps run while time < maxTime with time = step * tau (:558-611), so the
 *     last step is the first whose time reaches maxTime;
 *   - at time 0: onBeginTrajectory (its changes are copied to the device),
 *     counts, lattice frame, max counts, then the hook check (:539-557);
 *   - at a later event: the hook check, then the lattice frame, then the
 *     counts (:612-659); the lattice comes to the host only for a hook or a
 *     frame, counts alone are taken on the device (the same numbers);
 *   - hook return codes (hookCheckSimulation, :958-985): 0 unchanged, 1 copy
 *     the host lattice back to the device, 2 copy back and record the sites,
 *     3 record the sites without copying back and without stopping; any
 *     other value is an error;
 *   - a stop request (request_abort(), the SIGINT/SIGTERM handler of
 *     install_abort_handler(); globalAbort, src/core/SignalHandler.cpp:109,
 *     read before every step at :561-565) ends the loop; the end of the
 *     trajectory still runs (:662-667): the lattice is copied back,
 *     onEndTrajectory sees it, the buffered counts and the max counts are
 *     written;
 *   - diagnostics: "%d uses of the particle overflow list in the last 1000
 *     timesteps" when more than 10 of a window of 1000 steps overflowed
 *     (classic handler, :1180-1186), "%d unresolved overflows" (relaxed
 *     handler, :1094-1101), and the progress line every
 *     "perfPrintInterval" seconds (default 60, :526-603).
 *
 * The steps between two events run as one lmx_run_steps batch (at most 1000
 * steps, so a stop request and the overflow window are seen within 1000
 * steps) with no host synchronisation per step.  Records go to a
 * TrajectoryOutput, the counterpart of the reference's output queue
 * (DataOutputQueue::pushDataSet): species counts in buffers of 1000 entries
 * (TUNE_SPECIES_COUNTS_BUFFER_SIZE), lattice and site frames, and
 * MaxParticleCounts / MaxSiteCounts at the end.
 */
#ifndef LMX_MPD_SOLVER_HPP
#define LMX_MPD_SOLVER_HPP

#include <signal.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "lmx_mpd.h"

namespace lmx {

/* lm::Exception / InvalidArgException / CUDAException analogue: an lmx status. */
class Exception : public std::runtime_error {
public:
    Exception(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
    int code;
};

inline void check(int rc)
{
    if (rc != LMX_OK) throw Exception(rc, lmx_last_error());
}

/* globalAbort. */
inline std::atomic<bool> &abort_flag()
{
    static std::atomic<bool> flag{false};
    return flag;
}
inline void request_abort() { abort_flag().store(true); }
inline void clear_abort() { abort_flag().store(false); }
extern "C" inline void lmx_abort_signal_handler(int) { abort_flag().store(true); }
/* SIGINT and SIGTERM set the stop request (SignalHandler.cpp:109). */
inline void install_abort_handler()
{
    signal(SIGINT, lmx_abort_signal_handler);
    signal(SIGTERM, lmx_abort_signal_handler);
}

/* Host copy of a solver's lattice: particles in the native layout (W planes
 * of N words, N = X*Y*Z, site i = x + X*y + X*Y*z; a byte lattice keeps four
 * slots per word, slot 4w+b in byte b of word w, the uint32 lattice one
 * particle per word) and one site-type byte per site. */
class HostLattice {
public:
    uint32_t X = 0, Y = 0, Z = 0, P = 0, W = 0, bytesPerParticle = 1;
    std::vector<uint32_t> words;
    std::vector<uint8_t> siteTypes;

    size_t numberSites() const { return (size_t)X * Y * Z; }
    size_t index(uint32_t x, uint32_t y, uint32_t z) const { return x + (size_t)X * y + (size_t)X * Y * z; }
    uint32_t getParticle(uint32_t x, uint32_t y, uint32_t z, uint32_t slot) const
    {
        const size_t i = index(x, y, z), N = numberSites();
        if (bytesPerParticle == 4) return words[(size_t)slot * N + i];
        return (words[(size_t)(slot >> 2) * N + i] >> (8 * (slot & 3u))) & 0xFFu;
    }
    void setParticle(uint32_t x, uint32_t y, uint32_t z, uint32_t slot, uint32_t particle)
    {
        const size_t i = index(x, y, z), N = numberSites();
        if (bytesPerParticle == 4) {
            words[(size_t)slot * N + i] = particle;
            return;
        }
        uint32_t &w = words[(size_t)(slot >> 2) * N + i];
        const uint32_t sh = 8 * (slot & 3u);
        w = (w & ~(0xFFu << sh)) | ((particle & 0xFFu) << sh);
    }
    void removeParticles(uint32_t x, uint32_t y, uint32_t z)
    {
        for (uint32_t j = 0; j < P; j++) setParticle(x, y, z, j, 0u);
    }
    uint8_t getSiteType(uint32_t x, uint32_t y, uint32_t z) const { return siteTypes[index(x, y, z)]; }
    void setSiteType(uint32_t x, uint32_t y, uint32_t z, uint8_t t) { siteTypes[index(x, y, z)] = t; }

    /* ByteLattice::getParticleCounts (ByteLattice.cpp:495-520: within a word,
     * counting stops at the first empty slot) or IntLattice's (every nonzero
     * word); counts[s-1] for species s = 1..nS. */
    std::vector<uint32_t> speciesCounts(uint32_t nS) const
    {
        std::vector<uint32_t> c(nS, 0u);
        for (uint32_t w : words) {
            if (bytesPerParticle == 4) {
                if (w > 0 && w <= nS) c[w - 1]++;
                continue;
            }
            for (int b = 0; b < 4; b++) {
                const uint32_t p = (w >> (8 * b)) & 0xFFu;
                if (p == 0) break;
                if (p <= nS) c[p - 1]++;
            }
        }
        return c;
    }
    /* Particles of each type 0..n-1 over every slot (type 0 not counted). */
    std::vector<uint32_t> typeCounts(size_t n) const
    {
        std::vector<uint32_t> c(n, 0u);
        for (uint32_t w : words) {
            if (bytesPerParticle == 4) {
                if (w > 0 && w < n) c[w]++;
                continue;
            }
            for (int b = 0; b < 4; b++) {
                const uint32_t p = (w >> (8 * b)) & 0xFFu;
                if (p > 0 && p < n) c[p]++;
            }
        }
        return c;
    }
    /* Row-major (x, y, z, p) particles, bytesPerParticle bytes each (the
     * file layout, ByteLattice::copyNativeToRowMajorByte). */
    std::vector<uint8_t> rowMajorParticles() const
    {
        std::vector<uint8_t> out(numberSites() * P * bytesPerParticle);
        size_t o = 0;
        for (uint32_t x = 0; x < X; x++)
            for (uint32_t y = 0; y < Y; y++)
                for (uint32_t z = 0; z < Z; z++)
                    for (uint32_t j = 0; j < P; j++) {
                        const uint32_t v = getParticle(x, y, z, j);
                        for (uint32_t b = 0; b < bytesPerParticle; b++) out[o++] = (uint8_t)(v >> (8 * b));
                    }
        return out;
    }
    /* Row-major (x, y, z) site types. */
    std::vector<uint8_t> rowMajorSites() const
    {
        std::vector<uint8_t> out(numberSites());
        size_t o = 0;
        for (uint32_t x = 0; x < X; x++)
            for (uint32_t y = 0; y < Y; y++)
                for (uint32_t z = 0; z < Z; z++) out[o++] = getSiteType(x, y, z);
        return out;
    }
};

/* What the reference pushes to its output queue during a trajectory
 * (MpdRdmeSolver.cu:824-956).  counts holds entries x nS values. */
class TrajectoryOutput {
public:
    virtual ~TrajectoryOutput() = default;
    virtual void speciesCounts(uint32_t replicate, const std::vector<double> &times,
                               const std::vector<uint32_t> &counts, uint32_t nS)
    {
        (void)replicate; (void)times; (void)counts; (void)nS;
    }
    virtual void lattice(uint32_t replicate, double time, const HostLattice &lattice)
    {
        (void)replicate; (void)time; (void)lattice;
    }
    virtual void sites(uint32_t replicate, double time, const HostLattice &lattice)
    {
        (void)replicate; (void)time; (void)lattice;
    }
    virtual void maxCounts(uint32_t replicate, const std::vector<uint32_t> &particles,
                           const std::vector<uint32_t> &sites)
    {
        (void)replicate; (void)particles; (void)sites;
    }
};

class MpdRdmeSolver {
public:
    explicit MpdRdmeSolver(int device = 0) { check(lmx_create(device, &h_)); }
    virtual ~MpdRdmeSolver()
    {
        if (h_) lmx_destroy(h_);
    }
    MpdRdmeSolver(const MpdRdmeSolver &) = delete;
    MpdRdmeSolver &operator=(const MpdRdmeSolver &) = delete;

    lmx_solver *handle() const { return h_; }
    uint32_t replicate() const { return replicate_; }
    double timestep() const { return tau_; }

    /* MpdRdmeSolver::initialize (MpdRdmeSolver.cu:107-162). */
    void initialize(unsigned replicate, const std::map<std::string, std::string> &parameters,
                    int boundary = LMX_BOUNDARY_PERIODIC)
    {
        params_ = parameters;
        replicate_ = replicate;
        tau_ = std::atof(param("timestep").c_str());
        const uint32_t seedTop = (uint32_t)std::strtoull(param("seed").c_str(), nullptr, 10);
        const std::string &oh = param("rdme.mpd.overflowhandler");
        int mode = LMX_OVERFLOW_CLASSIC;
        if (oh == "relaxed") mode = LMX_OVERFLOW_RELAXED;
        else if (!oh.empty() && oh != "classic") warning("Unknown overflow handler requested: '" + oh + "'");
        relaxed_ = mode == LMX_OVERFLOW_RELAXED;
        check(lmx_initialize(h_, replicate, seedTop, tau_, mode, boundary));
    }

    /* CMESolver::buildModel + MpdRdmeSolver::buildModel. */
    void buildModel(uint32_t nS, uint32_t R, const uint32_t *initialCounts, const uint32_t *types,
                    const double *K, const int32_t *S, const uint32_t *D, uint32_t kCols)
    {
        check(lmx_build_model(h_, nS, R, initialCounts, types, K, S, D, kCols));
        nS_ = nS;
    }

    /* RDMESolver::buildDiffusionModel + MpdRdmeSolver::buildDiffusionModel. */
    void buildDiffusionModel(uint32_t nST, const double *DF, const uint32_t *RL, uint32_t X, uint32_t Y,
                             uint32_t Z, uint32_t P, uint32_t bytesPerParticle, double lambda,
                             const uint8_t *lattice, const uint8_t *sites, bool rowMajor = true)
    {
        check(lmx_build_diffusion_model(h_, nST, DF, RL, X, Y, Z, P, bytesPerParticle, lambda, lattice, sites,
                                        rowMajor ? 1 : 0));
        nST_ = nST;
        host_.X = X; host_.Y = Y; host_.Z = Z; host_.P = P;
        host_.bytesPerParticle = bytesPerParticle;
        host_.W = bytesPerParticle == 4 ? P : P / 4;
    }

    /* MpdRdmeSolver::setReactionRate (MpdRdmeSolver.cu:670-695). */
    void setReactionRate(uint32_t rxid, float rate) { check(lmx_set_reaction_rate(h_, rxid, rate)); }

    /* MpdRdmeSolver::generateTrajectory (MpdRdmeSolver.cu:485-668). */
    void generateTrajectory(TrajectoryOutput *out)
    {
        TrajectoryOutput none;
        out_ = out ? out : &none;
        const uint32_t wi = (uint32_t)std::atof(param("writeInterval").c_str());
        const uint32_t lwi = (uint32_t)std::atof(param("latticeWriteInterval").c_str());
        const double maxTime = std::atof(param("maxTime").c_str());
        const bool hookEnabled = !param("hookInterval").empty();
        const uint32_t hi = hookEnabled ? (uint32_t)std::atof(param("hookInterval").c_str()) : 0u;
        double perfInterval = 60.0;
        if (!param("perfPrintInterval").empty()) perfInterval = std::atof(param("perfPrintInterval").c_str());
        counts_t_.clear();
        counts_.clear();
        maxParticles_.assign((size_t)nS_ + 1, 0u);
        maxSites_.assign(nST_, 0u);

        copyFromGPU();
        onBeginTrajectory(host_);
        copyToGPU();
        double time = 0.0;
        uint64_t t = 0;
        recordCounts(time, true);
        writeLattice(time);
        initMaxCounts();
        hookCheck(time);

        const uint64_t last = lastStep(tau_, maxTime);
        uint64_t nextC = wi, nextL = lwi, nextH = hi;
        uint64_t windowEnd = 1000;  // overflow-use window (classic handler)
        uint64_t ovfStepsAtWindow = overflowSteps();
        auto perfStart = std::chrono::steady_clock::now();
        uint64_t perfSteps = 0;
        while (time < maxTime) {
            if (abort_flag().load()) {
                info("Global abort: terminating solver");
                break;
            }
            uint64_t due = std::min(nextC, nextL);
            if (hookEnabled) due = std::min(due, nextH);
            uint64_t stop = std::min(std::max(due, t + 1), std::max(last, t + 1));
            stop = std::min({stop, windowEnd, t + 1000});
            check(lmx_run_steps(h_, (uint32_t)(stop - t)));
            check(lmx_synchronize(h_));
            perfSteps += stop - t;
            t = stop;
            time = (double)t * tau_;
            if (relaxed_) {
                uint32_t unresolved = 0;
                check(lmx_get_unresolved_overflows(h_, &unresolved));
                if (unresolved > 0) warning(std::to_string(unresolved) + " unresolved overflows");
            }
            if (t >= windowEnd) {
                const uint64_t now = overflowSteps();
                if (!relaxed_ && now - ovfStepsAtWindow > 10)
                    warning(std::to_string(now - ovfStepsAtWindow) +
                            " uses of the particle overflow list in the last 1000 timesteps, performance may be "
                            "degraded.");
                ovfStepsAtWindow = now;
                windowEnd += 1000;
            }
            const double elapsed =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - perfStart).count();
            if (elapsed >= perfInterval && perfSteps > 0) {
                char buf[256];
                const double stepTime = elapsed / (double)perfSteps;
                std::snprintf(buf, sizeof buf,
                              "Average walltime per timestep: %.2f ms. Progress: %.4fs/%.4fs (%.3g%% done / %.2g "
                              "seconds walltime remaining)",
                              1000.0 * stepTime, time, maxTime, 100.0 * time / maxTime,
                              stepTime * (maxTime - time) / tau_);
                info(buf);
                perfStart = std::chrono::steady_clock::now();
                perfSteps = 0;
            }
            const bool hookDue = hookEnabled && t >= nextH;
            const bool frameDue = t >= nextL;
            bool host = false;
            if (hookDue || frameDue) {
                copyFromGPU();
                host = true;
            }
            if (hookDue) {
                hookCheck(time);
                nextH += hi;
            }
            if (frameDue) {
                writeLattice(time);
                nextL += lwi;
            }
            if (t >= nextC) {
                recordCounts(time, host);
                nextC += wi;
            }
        }
        copyFromGPU();
        onEndTrajectory(host_);
        flushCounts();
        out_->maxCounts(replicate_, maxParticles_, maxSites_);
        out_ = nullptr;
    }

    /* The number of steps the reference's loop runs: the first step whose
     * time = step * tau reaches maxTime. */
    static uint64_t lastStep(double tau, double maxTime)
    {
        if (!(maxTime > 0.0)) return 0;
        uint64_t k = std::max<uint64_t>((uint64_t)std::ceil(maxTime / tau), 1);
        while (k > 1 && (double)(k - 1) * tau >= maxTime) k--;
        while ((double)k * tau < maxTime) k++;
        return k;
    }

protected:
    /* Overridable hooks (MpdRdmeSolver.cu:789-821). */
    virtual int hookSimulation(double time, HostLattice &lattice)
    {
        (void)time; (void)lattice;
        return 0;
    }
    virtual int onBeginTrajectory(HostLattice &lattice)
    {
        (void)lattice;
        return 0;
    }
    virtual int onEndTrajectory(HostLattice &lattice)
    {
        (void)lattice;
        return 0;
    }
    virtual void warning(const std::string &msg) { std::fprintf(stderr, "WARNING: %s\n", msg.c_str()); }
    virtual void info(const std::string &msg) { std::fprintf(stderr, "INFO: %s\n", msg.c_str()); }

    std::string param(const std::string &key) const
    {
        auto it = params_.find(key);
        return it == params_.end() ? std::string() : it->second;
    }

    uint32_t numberSpecies() const { return nS_; }

private:
    void copyFromGPU()
    {
        host_.words.resize(host_.numberSites() * host_.W);
        check(lmx_get_lattice_native(h_, host_.words.data(), host_.words.size()));
        std::vector<uint8_t> rm(host_.numberSites());
        check(lmx_get_sites(h_, rm.data(), rm.size()));
        host_.siteTypes.resize(host_.numberSites());
        size_t o = 0;
        for (uint32_t x = 0; x < host_.X; x++)
            for (uint32_t y = 0; y < host_.Y; y++)
                for (uint32_t z = 0; z < host_.Z; z++) host_.siteTypes[host_.index(x, y, z)] = rm[o++];
    }
    void copyToGPU()
    {
        check(lmx_set_lattice_native(h_, host_.words.data(), host_.words.size()));
        const std::vector<uint8_t> rm = host_.rowMajorSites();
        check(lmx_set_sites(h_, rm.data(), rm.size()));
    }
    uint64_t overflowSteps()
    {
        uint64_t entries = 0, steps = 0;
        check(lmx_get_overflow_stats(h_, &entries, &steps));
        return steps;
    }
    /* hookCheckSimulation (MpdRdmeSolver.cu:958-985). */
    void hookCheck(double time)
    {
        switch (hookSimulation(time, host_)) {
        case 0: break;
        case 1: copyToGPU(); break;
        case 2:
            copyToGPU();
            writeSites(time);
            break;
        case 3:
            info("hook return value is 3, force to stop.");
            writeSites(time);
            break;
        default: throw Exception(LMX_ERR_ARG, "Unknown hook return value");
        }
    }
    /* recordSpeciesCounts (:933-942): from the host copy when it is current,
     * else on the device. */
    void recordCounts(double time, bool fromHost)
    {
        std::vector<uint32_t> c;
        if (fromHost) {
            c = host_.speciesCounts(nS_);
        } else {
            c.assign(nS_, 0u);
            if (nS_) check(lmx_get_species_counts(h_, c.data(), nS_));
        }
        counts_t_.push_back(time);
        counts_.insert(counts_.end(), c.begin(), c.end());
        if (counts_t_.size() >= 1000) flushCounts();  // TUNE_SPECIES_COUNTS_BUFFER_SIZE
    }
    void flushCounts()
    {
        if (counts_t_.empty()) return;
        out_->speciesCounts(replicate_, counts_t_, counts_, nS_);
        counts_t_.clear();
        counts_.clear();
    }
    void raise(std::vector<uint32_t> &mx, const std::vector<uint32_t> &c)
    {
        for (size_t i = 0; i < mx.size() && i < c.size(); i++) mx[i] = std::max(mx[i], c[i]);
    }
    std::vector<uint32_t> siteCounts() const
    {
        std::vector<uint32_t> c(nST_, 0u);
        for (uint8_t s : host_.siteTypes)
            if (s > 0 && s < nST_) c[s]++;
        return c;
    }
    /* writeLatticeData (:851-880): the frame, and the max particle counts. */
    void writeLattice(double time)
    {
        raise(maxParticles_, host_.typeCounts(maxParticles_.size()));
        out_->lattice(replicate_, time, host_);
    }
    /* writeLatticeSites (:883-910): the site frame, and the max site counts. */
    void writeSites(double time)
    {
        raise(maxSites_, siteCounts());
        out_->sites(replicate_, time, host_);
    }
    /* initMaxCounts (:824-849). */
    void initMaxCounts()
    {
        raise(maxSites_, siteCounts());
        raise(maxParticles_, host_.typeCounts(maxParticles_.size()));
    }

    lmx_solver *h_ = nullptr;
    std::map<std::string, std::string> params_;
    uint32_t replicate_ = 0, nS_ = 0, nST_ = 0;
    double tau_ = 0.0;
    bool relaxed_ = false;
    HostLattice host_;
    TrajectoryOutput *out_ = nullptr;
    std::vector<double> counts_t_;
    std::vector<uint32_t> counts_;
    std::vector<uint32_t> maxParticles_, maxSites_;
};

}  // namespace lmx

#endif
