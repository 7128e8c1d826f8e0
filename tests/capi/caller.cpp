// A C++ caller of the C ABI, with no Python: reads a model file written by
// the test, runs the MPD-RDME steps through include/lmx_mpd.h and writes the
// native lattice and species counts.  Test infrastructure (built by
// tests/capi/build.py); it is the path a Lattice Microbes C++ shim would
// take (MpdRdmeSolver::initialize / buildModel / buildDiffusionModel /
// runTimestep, MpdRdmeSolver.cu:107-483, 996-1193).
//
// Input (little endian): uint32 nS, R, nST, X, Y, Z, P, kCols, replicate,
// seedTop, steps, boundary; double tau, lambda; uint32 initialCounts[nS],
// types[R]; double K[R*kCols]; int32 S[nS*R]; uint32 D[nS*R]; double
// DF[nST*nST*nS]; uint32 RL[R*nST]; uint8 lattice[X*Y*Z*P], sites[X*Y*Z].
// Output: uint32 native[(P/4)*X*Y*Z], uint32 counts[nS].
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "lmx_mpd.h"

namespace {

template <typename T>
bool read_vec(std::FILE *f, std::vector<T> &v, size_t n)
{
    v.resize(n);
    return n == 0 || std::fread(v.data(), sizeof(T), n, f) == n;
}

int check(int rc, const char *what)
{
    if (rc != LMX_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, lmx_last_error());
        std::exit(2);
    }
    return rc;
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
    uint32_t h[12];
    double fp[2];
    if (std::fread(h, sizeof(uint32_t), 12, f) != 12 || std::fread(fp, sizeof(double), 2, f) != 2) return 1;
    const uint32_t nS = h[0], R = h[1], nST = h[2], X = h[3], Y = h[4], Z = h[5], P = h[6], kCols = h[7];
    const uint32_t replicate = h[8], seedTop = h[9], steps = h[10], boundary = h[11];
    const size_t N = (size_t)X * Y * Z;
    std::vector<uint32_t> init, types, D, RL;
    std::vector<double> K, DF;
    std::vector<int32_t> S;
    std::vector<uint8_t> lattice, sites;
    bool ok = read_vec(f, init, nS) && read_vec(f, types, R) && read_vec(f, K, (size_t)R * kCols) &&
              read_vec(f, S, (size_t)nS * R) && read_vec(f, D, (size_t)nS * R) &&
              read_vec(f, DF, (size_t)nST * nST * nS) && read_vec(f, RL, (size_t)R * nST) &&
              read_vec(f, lattice, N * P) && read_vec(f, sites, N);
    std::fclose(f);
    if (!ok) return 1;

    lmx_solver *s = nullptr;
    check(lmx_create(0, &s), "lmx_create");
    check(lmx_initialize(s, replicate, seedTop, fp[0], 0, (int)boundary), "lmx_initialize");
    check(lmx_build_model(s, nS, R, init.data(), types.data(), K.data(), S.data(), D.data(), kCols),
          "lmx_build_model");
    check(lmx_build_diffusion_model(s, nST, DF.data(), RL.empty() ? nullptr : RL.data(), X, Y, Z, P, 1, fp[1],
                                    lattice.data(), sites.data(), 1),
          "lmx_build_diffusion_model");
    check(lmx_run_steps(s, steps), "lmx_run_steps");
    check(lmx_synchronize(s), "lmx_synchronize");
    std::vector<uint32_t> native(N * (P / 4)), counts(nS);
    check(lmx_get_lattice_native(s, native.data(), native.size()), "lmx_get_lattice_native");
    check(lmx_get_species_counts(s, counts.data(), nS), "lmx_get_species_counts");
    check(lmx_destroy(s), "lmx_destroy");

    std::FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 1;
    std::fwrite(native.data(), sizeof(uint32_t), native.size(), o);
    std::fwrite(counts.data(), sizeof(uint32_t), counts.size(), o);
    std::fclose(o);
    std::printf("capi caller: %u steps on %ux%ux%u\n", steps, X, Y, Z);
    return 0;
}
