// C-ABI implementation of the MPD-RDME solver (declared in include/lmx_mpd.h).
//
// Host-side responsibilities of lm::rdme::MpdRdmeSolver restated for this
// build: model ingestion and table construction (MpdRdmeSolver.cu:232-483,
// :697-774), lattice layout (RDMESolver.cpp:144-191, ByteLattice.cpp:371-435),
// the step sequence (runTimestep, :996-1193) and lattice readback.  The step
// itself runs entirely on the device; the host never touches the lattice
// between steps (overflow resolution is a device kernel).
#include "lmx_mpd.h"

#include <time.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "lmx_batch.h"
#include "lmx_device.h"
#include "lmx_kernels.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg)
{
    g_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(LMX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Entry points that allocate host memory run their body through guarded():
// a C++ exception must not cross the C ABI.
template <typename F>
int guarded(F &&body)
{
    try {
        return body();
    } catch (const std::bad_alloc &) {
        g_error = "host memory allocation failed";
        return LMX_ERR_NOMEM;
    } catch (const std::exception &e) {
        g_error = std::string("internal error: ") + e.what();
        return LMX_ERR_STATE;
    }
}

#define LMX_HIP(call)                                  \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

template <typename T>
int dev_upload(T **dptr, const std::vector<T> &host)
{
    if (*dptr) {
        hipFree(*dptr);
        *dptr = nullptr;
    }
    size_t bytes = host.size() * sizeof(T);
    if (bytes == 0) bytes = sizeof(T);
    LMX_HIP(hipMalloc((void **)dptr, bytes));
    if (!host.empty()) LMX_HIP(hipMemcpy(*dptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
    return LMX_OK;
}

template <typename T>
void dev_free(T *&p)
{
    if (p) hipFree(p);
    p = nullptr;
}

}  // namespace

struct lmx_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    bool initialized = false, model_built = false, diffusion_built = false;
    uint32_t replicate = 0, seed32 = 0;
    double tau = 0.0;
    int overflow_mode = 0, boundary = 0, path = LMX_PATH_SWEEPS;

    // Reaction model (CMESolver members).
    uint32_t nS = 0, R = 0;
    std::vector<uint32_t> initialCounts, types, si, s2i, order, D1, D2;
    std::vector<double> k;
    std::vector<int32_t> S_sm;
    std::vector<int8_t> S_rm;

    // Diffusion model and lattice.
    uint32_t nST = 0, X = 0, Y = 0, Z = 0, P = 0, W = 0;
    uint32_t bpp = 1;  // bytes per particle: 1 (byte lattice) or 4 (uint32 lattice, one particle per word)
    uint32_t counts_cap = 256;  // entries of d_counts
    // Every site of the current lattice holds its particles in its first
    // slots.  The kernels keep this; a lattice written from outside (native
    // words, device pointers) may break it until the next step.
    bool packed = true;
    // The chunk flags of both buffers are valid (kept by every kernel since
    // they were last zeroed); cleared by host writes and by steps whose
    // kernels do not keep them.
    bool flags_fresh = false;
    size_t flag_words = 0;  // per buffer, after the planes
    size_t N = 0;
    double lambda = 0.0;
    std::vector<double> DF;
    std::vector<uint32_t> RL;
    std::vector<float> T, rate, qp0, qp1, qp2;

    // Device state.
    uint32_t *d_lat[2] = {nullptr, nullptr};
    int cur = 0;
    uint8_t *d_sites = nullptr;
    float *d_T = nullptr, *d_qp0 = nullptr, *d_qp1 = nullptr, *d_qp2 = nullptr, *d_rate = nullptr;
    float *d_prob1 = nullptr, *d_prob0 = nullptr;
    uint32_t *d_order = nullptr, *d_D1 = nullptr, *d_D2 = nullptr;
    int8_t *d_S = nullptr;
    uint8_t *d_RL = nullptr;
    uint32_t *d_rx_off = nullptr, *d_rx_list = nullptr, *d_sto_off = nullptr, *d_sto_sp = nullptr;
    int32_t *d_sto_val = nullptr;
    lmx::Control *d_ctl = nullptr;
    uint32_t *d_counts = nullptr;
    bool tables_dirty = false;
    uint32_t timestep = 0;
    bool any_qp0 = false;
    // launch options
    uint32_t march_seg = 8;
    bool force_generic = false;
    int yz_kernel = 0;  // 0: LDS tiles, 1: register march, 2: LDS tiles with compacted work lists
    lmx_batch *batch = nullptr;  // ensemble this solver is a member of, if any
    uint32_t fused_zseg = 0;  // 0: automatic
    bool fuse_reaction = true;  // the tiled z sweep lists its occupied output sites for the reaction phase
    uint4 *d_rl = nullptr;      // the reaction list: one record of list_units(W) 16-byte units per entry
    lmx::ReactionList rlist{};
    uint32_t num_cu = 256;
};

namespace {

int select_device(lmx_solver *s)
{
    LMX_HIP(hipSetDevice(s->device));
    return LMX_OK;
}

// MpdRdmeSolver::computePropensities (MpdRdmeSolver.cu:697-774).  The float
// tables are updated through double arithmetic exactly as the compound
// assignments there evaluate.
void compute_propensities(lmx_solver *s)
{
    const uint32_t nsites32 = s->X * s->Y * s->Z;  // unsigned product, as the reference
    s->rate.assign(s->R, 0.0f);
    for (uint32_t i = 0; i < s->R; i++) {
        const double k = s->k[i];
        switch (s->types[i]) {
        case 0: s->rate[i] = (float)(k * s->tau / (double)nsites32); break;
        case 1: s->rate[i] = (float)(k * s->tau); break;
        default: s->rate[i] = (float)(k * s->tau * (double)s->X * (double)s->Y * (double)s->Z); break;
        }
    }
    const uint32_t nS = s->nS, nST = s->nST;
    s->qp0.assign(nST, 0.0f);
    s->qp1.assign((size_t)nST * nS, 0.0f);
    s->qp2.assign((size_t)nST * nS * nS, 0.0f);
    const float scale = (float)nsites32;
    for (uint32_t st = 0; st < nST; st++) {
        const size_t o1 = (size_t)st * nS, o2 = (size_t)st * nS * nS;
        for (uint32_t i = 0; i < s->R; i++) {
            if (!s->RL[(size_t)i * nST + st]) continue;
            const double k = s->k[i];
            switch (s->types[i]) {
            case 0: s->qp0[st] = (float)((double)s->qp0[st] + k * s->tau / (double)scale); break;
            case 1: s->qp1[o1 + s->si[i]] = (float)((double)s->qp1[o1 + s->si[i]] + k * s->tau); break;
            case 2: {
                const float v = (float)(k * s->tau * (double)scale);
                s->qp2[o2 + (size_t)s->si[i] * nS + s->s2i[i]] = v;
                s->qp2[o2 + (size_t)s->s2i[i] * nS + s->si[i]] = v;
            } break;
            case 3: {
                float &e = s->qp2[o2 + (size_t)s->si[i] * nS + s->si[i]];
                e = (float)((double)e + k * s->tau * (double)scale * 2);
            } break;
            }
        }
    }
    s->any_qp0 = false;
    for (float v : s->qp0) s->any_qp0 |= (v != 0.0f);
    s->tables_dirty = true;
}

lmx::Params make_params(const lmx_solver *s);

int upload_tables(lmx_solver *s)
{
    int rc;
    if ((rc = dev_upload(&s->d_rate, s->rate))) return rc;
    if ((rc = dev_upload(&s->d_qp0, s->qp0))) return rc;
    if ((rc = dev_upload(&s->d_qp1, s->qp1))) return rc;
    if ((rc = dev_upload(&s->d_qp2, s->qp2))) return rc;
    // Single-particle reaction probabilities, built on the device with the
    // kernels' own exp.
    std::vector<float> zeros1((size_t)s->nST * s->nS * 2, 0.0f), zeros0(s->nST, 0.0f);
    if ((rc = dev_upload(&s->d_prob1, zeros1))) return rc;
    if ((rc = dev_upload(&s->d_prob0, zeros0))) return rc;
    const lmx::Params p = make_params(s);
    hipError_t e = lmx::launch_prob_tables(p, s->d_prob1, s->d_prob0, s->stream);
    if (e != hipSuccess) return hip_fail(e, "probability table launch");
    LMX_HIP(hipStreamSynchronize(s->stream));
    s->tables_dirty = false;
    return LMX_OK;
}

lmx::Params make_params(const lmx_solver *s)
{
    lmx::Params p;
    p.X = s->X; p.Y = s->Y; p.Z = s->Z;
    p.XY = s->X * s->Y;
    p.N = (uint32_t)s->N;
    p.P = s->P; p.W = s->W;
    p.word_particles = s->bpp == 4 ? 1u : 0u;
    p.packed = s->packed ? 1u : 0u;
    p.plane_flags = 0u;
    p.relaxed = s->overflow_mode == LMX_OVERFLOW_RELAXED ? 1u : 0u;
    p.pcb = s->P <= 4 ? 2u : (s->P <= 8 ? 3u : 4u);  // MPD_PARTICLE_COUNT_BITS
    p.nS = s->nS; p.nST = s->nST; p.R = s->R;
    p.boundary = (uint32_t)s->boundary;
    p.seed32 = s->seed32;
    p.any_qp0 = s->any_qp0 ? 1u : 0u;
    p.T = s->d_T;
    p.qp0 = s->d_qp0; p.qp1 = s->d_qp1; p.qp2 = s->d_qp2;
    p.rate = s->d_rate;
    p.prob1 = s->d_prob1;
    p.prob0 = s->d_prob0;
    p.order = s->d_order; p.D1 = s->d_D1; p.D2 = s->d_D2;
    p.S = s->d_S;
    p.RL = s->d_RL;
    p.rx_off = s->d_rx_off;
    p.rx_list = s->d_rx_list;
    p.sto_off = s->d_sto_off;
    p.sto_sp = s->d_sto_sp;
    p.sto_val = s->d_sto_val;
    p.sites = s->d_sites;
    p.ctl = s->d_ctl;
    return p;
}

int require_ready(lmx_solver *s)
{
    if (!s) return fail(LMX_ERR_ARG, "null solver");
    if (!s->diffusion_built) return fail(LMX_ERR_STATE, "the diffusion model and lattice have not been set");
    return select_device(s);
}

// Row-major (x,y,z,p) -> native planes, inserting each nonzero particle at
// the first empty slot of its site (RDMESolver::setLatticeData + addParticle;
// IntLattice::setFromRowMajorData for the uint32 lattice).  rm holds N*P
// elements of bpp bytes.
int row_major_to_native(const lmx_solver *s, const void *rm, std::vector<uint32_t> &native)
{
    const size_t N = s->N;
    const bool words = s->bpp == 4;
    const uint8_t *rb = static_cast<const uint8_t *>(rm);
    const uint32_t *rw = static_cast<const uint32_t *>(rm);
    native.assign(N * s->W, 0u);
    std::vector<uint8_t> fill(N, 0);
    size_t idx = 0;
    for (uint32_t x = 0; x < s->X; x++)
        for (uint32_t y = 0; y < s->Y; y++)
            for (uint32_t z = 0; z < s->Z; z++) {
                const size_t i = (size_t)x + (size_t)s->X * y + (size_t)s->X * s->Y * z;
                for (uint32_t l = 0; l < s->P; l++, idx++) {
                    const uint32_t v = words ? rw[idx] : rb[idx];
                    if (v == 0) continue;
                    if (v > s->nS) return fail(LMX_ERR_ARG, "latticeData: an invalid species was found: " + std::to_string(v));
                    const uint32_t j = fill[i]++;
                    if (words)
                        native[(size_t)j * N + i] = v;
                    else
                        native[(size_t)(j >> 2) * N + i] |= v << (8 * (j & 3));
                }
            }
    return LMX_OK;
}

void native_to_row_major(const lmx_solver *s, const uint32_t *native, void *rm)
{
    const size_t N = s->N;
    for (uint32_t x = 0; x < s->X; x++)
        for (uint32_t y = 0; y < s->Y; y++)
            for (uint32_t z = 0; z < s->Z; z++) {
                const size_t i = (size_t)x + (size_t)s->X * y + (size_t)s->X * s->Y * z;
                const size_t o = (((size_t)x * s->Y + y) * s->Z + z) * s->P;
                if (s->bpp == 4) {
                    uint32_t *d = static_cast<uint32_t *>(rm) + o;
                    for (uint32_t j = 0; j < s->P; j++) d[j] = native[(size_t)j * N + i];
                } else {
                    uint8_t *d = static_cast<uint8_t *>(rm) + o;
                    for (uint32_t j = 0; j < s->P; j++) d[j] = (uint8_t)(native[(size_t)(j >> 2) * N + i] >> (8 * (j & 3)));
                }
            }
}

int sites_row_major_to_native(const lmx_solver *s, const uint8_t *rm, std::vector<uint8_t> &native)
{
    native.assign(s->N, 0);
    size_t idx = 0;
    for (uint32_t x = 0; x < s->X; x++)
        for (uint32_t y = 0; y < s->Y; y++)
            for (uint32_t z = 0; z < s->Z; z++, idx++) {
                const uint8_t v = rm[idx];
                if (v == 0) continue;
                if (v >= s->nST) return fail(LMX_ERR_ARG, "latticeSitesData: an invalid site type was found: " + std::to_string(v));
                native[(size_t)x + (size_t)s->X * y + (size_t)s->X * s->Y * z] = v;
            }
    return LMX_OK;
}

int check_device_error(lmx_solver *s)
{
    uint32_t hdr[4];
    LMX_HIP(hipMemcpy(hdr, s->d_ctl, sizeof(hdr), hipMemcpyDeviceToHost));
    switch (hdr[2]) {
    case 0: return LMX_OK;
    case lmx::kErrOverflowCap:
        return fail(LMX_ERR_OVERFLOW_CAP, "Too many particle overflows for the available buffer at timestep " + std::to_string(hdr[3]));
    case lmx::kErrUnplaceable:
        return fail(LMX_ERR_OVERFLOW_UNPLACEABLE, "Unable to find an available site to handle a particle overflow at timestep " + std::to_string(hdr[3]));
    case lmx::kErrOutOfRange:
        return fail(LMX_ERR_OVERFLOW_OUT_OF_RANGE, "Overflow search left the lattice at timestep " + std::to_string(hdr[3]));
    case lmx::kErrReactionUnderflow:
        return fail(LMX_ERR_REACTION_UNDERFLOW, "Reaction selected without its reactants at timestep " + std::to_string(hdr[3]));
    case lmx::kErrInvalidLattice:
        return fail(LMX_ERR_INVALID_LATTICE, "latticeData: an invalid species or site type was written on the device before timestep " + std::to_string(hdr[3]));
    }
    return fail(LMX_ERR_HIP, "unknown device error " + std::to_string(hdr[2]));
}

// Planes per block of the fused step: the option's value, else enough
// z segments that the grid covers every compute unit once (one block per CU
// fits the kernel's LDS).
uint32_t fused_zseg(lmx_solver *s)
{
    if (s->fused_zseg) return s->fused_zseg;
    const uint32_t ty = lmx::fused_tile_rows(s->X, s->W);
    const uint32_t nty = (s->Y + ty - 1) / ty;
    uint32_t nzs = s->num_cu > nty ? s->num_cu / nty : 1u;
    if (nzs > s->Z) nzs = s->Z;
    return (s->Z + nzs - 1) / nzs;
}

// One timestep (runTimestep, MpdRdmeSolver.cu:996-1193).  ev, if given,
// receives events around each launch (ev[0] before the first launch, then one
// after every launch); kinds receives the kernel class of each launch.
// rl_in: the reaction list the z sweep fills (default: the solver's own).
int enqueue_step(lmx_solver *s, const lmx::Params &p0, hipEvent_t *ev, int *kinds, int *nlaunch,
                 const lmx::ReactionList *rl_in = nullptr)
{
    const lmx::ReactionList &rlist = rl_in ? *rl_in : s->rlist;
    // The step writes packed sites whatever it reads, so only its first
    // kernel may see an unpacked lattice; every launch of this step takes the
    // current flag, and the next step may rely on packing.
    lmx::Params p = p0;
    p.packed = s->packed ? 1u : 0u;
    hipError_t e;
    int n = 0;
    auto mark = [&](int kind) -> int {
        if (ev) {
            hipError_t r = hipEventRecord(ev[n + 1], s->stream);
            if (r != hipSuccess) return hip_fail(r, "hipEventRecord");
            kinds[n] = kind;
        }
        n++;
        return LMX_OK;
    };
    const bool fast = !s->force_generic && lmx::fast_path_ok(p);
    // Chunk flags: kept by the default sweeps-path kernels of the byte
    // lattice (x4, LDS tiles, the reaction phases, the resolver); any other
    // step leaves them stale, so they are zeroed before the next step that
    // uses them.
    const bool flags_on = fast && s->path == LMX_PATH_SWEEPS && s->yz_kernel == 0 && s->flag_words > 0;
    if (flags_on) {
        if (!s->flags_fresh) {
            for (int b = 0; b < 2; b++) {
                e = lmx::launch_fill(s->d_lat[b] + s->N * s->W, s->flag_words, 0u, s->stream);
                if (e != hipSuccess) return hip_fail(e, "flag fill launch");
            }
            s->flags_fresh = true;
        }
        p.plane_flags = 1u;
    } else {
        p.plane_flags = 0u;
        s->flags_fresh = false;
    }
    if (ev) {
        e = hipEventRecord(ev[0], s->stream);
        if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    }
    if (fast && s->path == LMX_PATH_FUSED && lmx::fused_ok(p)) {
        e = lmx::launch_fused(p, s->d_lat[s->cur], s->d_lat[1 - s->cur], fused_zseg(s), s->stream);
        if (e != hipSuccess) return hip_fail(e, "fused step launch");
        s->cur = 1 - s->cur;
        if (int rc = mark(LMX_K_FUSED)) return rc;
        e = lmx::launch_resolve(p, s->d_lat[s->cur], s->stream);
        if (e != hipSuccess) return hip_fail(e, "resolve launch");
        if (int rc = mark(LMX_K_RESOLVE)) return rc;
        if (nlaunch) *nlaunch = n;
        s->timestep++;
        s->packed = true;
        return LMX_OK;
    }
    // The tiled z sweeps list their occupied output sites, and the reaction
    // phase evaluates only those (not with zero-order reactions, which make
    // every site a candidate).
    const bool z_reacts = fast && (s->yz_kernel == 0 || s->yz_kernel == 2) && s->fuse_reaction && s->R > 0 &&
                          !s->any_qp0 && rlist.rec;
    for (int axis = 0; axis < 3; axis++) {
        const bool react = axis == 2 && z_reacts;
        if (!fast)
            e = lmx::launch_sweep(p, s->d_lat[s->cur], s->d_lat[1 - s->cur], axis, s->stream);
        else if (axis > 0 && s->yz_kernel == 3 && !react && lmx::col_ok(p, axis, s->march_seg))
            e = lmx::launch_sweep_col(p, s->d_lat[s->cur], s->d_lat[1 - s->cur], axis, s->march_seg, s->stream);
        else if (axis > 0 && (s->yz_kernel == 0 || s->yz_kernel == 3))
            e = lmx::launch_sweep_tile(p, s->d_lat[s->cur], s->d_lat[1 - s->cur], axis, react, rlist, s->stream);
        else if (axis > 0 && s->yz_kernel == 2)
            e = lmx::launch_sweep_sparse(p, s->d_lat[s->cur], s->d_lat[1 - s->cur], axis, react, rlist,
                                         s->stream);
        else
            e = lmx::launch_sweep_fast(p, s->d_lat[s->cur], s->d_lat[1 - s->cur], axis, s->march_seg, s->stream);
        if (e != hipSuccess) return hip_fail(e, "sweep launch");
        s->cur = 1 - s->cur;
        if (int rc = mark(LMX_K_SWEEP_X + axis)) return rc;
    }
    if (z_reacts) {
        e = lmx::launch_reaction_list(p, s->d_lat[s->cur], rlist, (int)s->num_cu, s->stream);
        if (e != hipSuccess) return hip_fail(e, "reaction launch");
        if (int rc = mark(LMX_K_REACTION)) return rc;
    } else if (s->R > 0) {
        e = fast ? lmx::launch_reaction_fast(p, s->d_lat[s->cur], s->stream)
                 : lmx::launch_reaction(p, s->d_lat[s->cur], s->stream);
        if (e != hipSuccess) return hip_fail(e, "reaction launch");
        if (int rc = mark(LMX_K_REACTION)) return rc;
    }
    e = lmx::launch_resolve(p, s->d_lat[s->cur], s->stream);
    if (e != hipSuccess) return hip_fail(e, "resolve launch");
    if (int rc = mark(LMX_K_RESOLVE)) return rc;
    if (nlaunch) *nlaunch = n;
    s->timestep++;
    s->packed = true;
    return LMX_OK;
}

void free_device(lmx_solver *s)
{
    dev_free(s->d_lat[0]);
    dev_free(s->d_lat[1]);
    dev_free(s->d_sites);
    dev_free(s->d_rl);
    s->rlist = lmx::ReactionList{};
    dev_free(s->d_T);
    dev_free(s->d_qp0);
    dev_free(s->d_qp1);
    dev_free(s->d_qp2);
    dev_free(s->d_rate);
    dev_free(s->d_prob1);
    dev_free(s->d_prob0);
    dev_free(s->d_order);
    dev_free(s->d_D1);
    dev_free(s->d_D2);
    dev_free(s->d_S);
    dev_free(s->d_RL);
    dev_free(s->d_rx_off);
    dev_free(s->d_rx_list);
    dev_free(s->d_sto_off);
    dev_free(s->d_sto_sp);
    dev_free(s->d_sto_val);
}

}  // namespace

extern "C" {

const char *lmx_last_error(void) { return g_error.c_str(); }

int lmx_abi_version(void) { return LMX_ABI_VERSION; }

int lmx_device_count(int *count)
{
    if (!count) return fail(LMX_ERR_ARG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return LMX_OK;
}

int lmx_create(int device, lmx_solver **out)
{
    return guarded([&]() -> int {
        if (!out) return fail(LMX_ERR_ARG, "null output handle");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(LMX_ERR_HIP, "no HIP device is available");
        if (device < 0 || device >= n) return fail(LMX_ERR_ARG, "device index out of range");
        lmx_solver *s = new lmx_solver();
        s->device = device;
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipMalloc((void **)&s->d_ctl, sizeof(lmx::Control));
        if (e == hipSuccess) e = hipMemset(s->d_ctl, 0, sizeof(lmx::Control));
        if (e == hipSuccess) e = hipMalloc((void **)&s->d_counts, 256 * sizeof(uint32_t));
        int ncu = 0;
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
        if (ncu > 0) s->num_cu = (uint32_t)ncu;
        if (e != hipSuccess) {
            int rc = hip_fail(e, "lmx_create");
            lmx_destroy(s);
            return rc;
        }
        *out = s;
        return LMX_OK;
    });
}

int lmx_destroy(lmx_solver *s)
{
    if (!s) return LMX_OK;
    if (s->batch) lmx_batch_remove(s->batch, s);
    hipSetDevice(s->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    free_device(s);
    dev_free(s->d_ctl);
    dev_free(s->d_counts);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
    return LMX_OK;
}

int lmx_initialize(lmx_solver *s, uint32_t replicate, uint32_t seedTop, double timestep, int overflow_mode,
                   int boundary)
{
    if (!s) return fail(LMX_ERR_ARG, "null solver");
    if (!(timestep > 0.0)) return fail(LMX_ERR_ARG, "timestep: A positive timestep must be specified for the solver.");
    if (overflow_mode != LMX_OVERFLOW_CLASSIC && overflow_mode != LMX_OVERFLOW_RELAXED)
        return fail(LMX_ERR_ARG, "rdme.mpd.overflowhandler: unknown handler");
    if (boundary != LMX_BOUNDARY_PERIODIC && boundary != LMX_BOUNDARY_REFLECTING)
        return fail(LMX_ERR_ARG, "unknown boundary mode");
    if (seedTop == 0) {
        struct timespec ts;
        clock_gettime(CLOCK_REALTIME, &ts);
        seedTop = (uint32_t)ts.tv_nsec;
    }
    s->replicate = replicate;
    s->seed32 = (seedTop << 16) | (replicate & 0x0000FFFFu);
    s->tau = timestep;
    s->overflow_mode = overflow_mode;
    s->boundary = boundary;
    s->initialized = true;
    return LMX_OK;
}

int lmx_get_seed(lmx_solver *s, uint32_t *seed32)
{
    if (!s || !seed32) return fail(LMX_ERR_ARG, "null argument");
    *seed32 = s->seed32;
    return LMX_OK;
}

int lmx_build_model(lmx_solver *s, uint32_t nS, uint32_t R, const uint32_t *initialCounts, const uint32_t *types,
                    const double *K, const int32_t *S, const uint32_t *D, uint32_t kCols)
{
    return guarded([&]() -> int {
        if (!s) return fail(LMX_ERR_ARG, "null solver");
        if (!s->initialized) return fail(LMX_ERR_STATE, "initialize must be called before buildModel");
        if (R > 0 && kCols == 0) return fail(LMX_ERR_ARG, "K: must have at least 1 column");
        if (R > 16384) return fail(LMX_ERR_ARG, "The number of reaction table entries exceeds the maximum supported by the solver (16384).");
        if ((R > 0 || nS > 0) && (!types || !K || !S || !D) && R > 0) return fail(LMX_ERR_ARG, "null model array");
        if (s->batch) return fail(LMX_ERR_STATE, "the solver is an ensemble member; remove it before rebuilding its model");
        // The fields below are overwritten before every check has passed, so
        // until this call succeeds the solver has no model and no lattice:
        // a failed rebuild leaves nothing that a step or a readback could use
        // with the old device tables.
        s->model_built = false;
        s->diffusion_built = false;
        s->nS = nS;
        s->R = R;
        s->initialCounts.assign(initialCounts ? initialCounts : nullptr, initialCounts ? initialCounts + nS : nullptr);
        s->types.assign(types, types + R);
        s->k.assign(R, 0.0);
        s->si.assign(R, 0);
        s->s2i.assign(R, 0);
        s->order.assign(R, 0);
        s->D1.assign(R, 0);
        s->D2.assign(R, 0);
        s->S_sm.assign(S, S + (size_t)nS * R);
        s->S_rm.assign((size_t)nS * R, 0);
        for (uint32_t i = 0; i < R; i++) {
            uint32_t deps[2] = {0, 0}, nd = 0;
            for (uint32_t j = 0; j < nS; j++)
                if (D[(size_t)j * R + i] == 1) {
                    if (nd < 2) deps[nd] = j;
                    nd++;
                }
            s->k[i] = K[(size_t)i * kCols];
            switch (types[i]) {
            case 0:
                if (nd != 0) return fail(LMX_ERR_ARG, "D: zeroth order reaction cannot have any dependencies");
                s->order[i] = lmx::kZeroOrder;
                break;
            case 1:
                if (nd != 1) return fail(LMX_ERR_ARG, "D: first order reaction had invalid number of dependencies");
                s->order[i] = lmx::kFirstOrder;
                s->si[i] = deps[0];
                s->D1[i] = deps[0] + 1;
                break;
            case 2:
                if (nd != 2) return fail(LMX_ERR_ARG, "D: second order reaction had invalid number of dependencies");
                s->order[i] = lmx::kSecondOrder;
                s->si[i] = deps[0];
                s->s2i[i] = deps[1];
                s->D1[i] = deps[0] + 1;
                s->D2[i] = deps[1] + 1;
                break;
            case 3:
                if (nd != 1) return fail(LMX_ERR_ARG, "D: second order self reaction had invalid number of dependencies");
                s->order[i] = lmx::kSecondOrderSelf;
                s->si[i] = deps[0];
                s->D1[i] = deps[0] + 1;
                break;
            default:
                return fail(LMX_ERR_ARG, "reactionTypeA: the reaction type was not supported by the solver");
            }
        }
        // S transposed to reaction-major int8 (MpdRdmeSolver.cu:317-325).
        for (uint32_t rx = 0; rx < R; rx++)
            for (uint32_t p = 0; p < nS; p++) s->S_rm[(size_t)rx * nS + p] = (int8_t)S[(size_t)R * p + rx];
        s->model_built = true;
        s->diffusion_built = false;
        return LMX_OK;
    });
}

int lmx_build_diffusion_model(lmx_solver *s, uint32_t nST, const double *DF, const uint32_t *RL, uint32_t X,
                              uint32_t Y, uint32_t Z, uint32_t P, uint32_t bytesPerParticle, double lambda,
                              const uint8_t *lattice, const uint8_t *sites, int rowMajor)
{
    return guarded([&]() -> int {
        if (!s) return fail(LMX_ERR_ARG, "null solver");
        if (!s->model_built) return fail(LMX_ERR_STATE, "buildModel must be called before buildDiffusionModel");
        if (!rowMajor) return fail(LMX_ERR_ARG, "rowMajorData: columns major lattice data is not currently supported");
        if (bytesPerParticle != 1 && bytesPerParticle != 4)
            return fail(LMX_ERR_ARG, "bytes_per_particle must be 1 (byte lattice) or 4 (uint32 lattice)");
        if (bytesPerParticle == 1 && s->nS > 255) return fail(LMX_ERR_ARG, "a byte lattice holds at most 255 species");
        if ((unsigned long long)nST * s->nS * s->nS >= (1ull << 32))
            return fail(LMX_ERR_ARG, "numberSiteTypes * numberSpecies^2 must be below 2^32 (second-order propensity table)");
        if (P != 4 && P != 8 && P != 16) return fail(LMX_ERR_ARG, "particlesPerSite must be 4, 8 or 16");
        if (nST == 0 || nST > 256) return fail(LMX_ERR_ARG, "numberSiteTypes must be in 1..256");
        if (X == 0 || Y == 0 || Z == 0) return fail(LMX_ERR_ARG, "incorrect lattice dimensions");
        const unsigned long long N = (unsigned long long)X * Y * Z;
        if (N >= (1ull << 32)) return fail(LMX_ERR_ARG, "lattice has 2^32 or more sites");
        if (!DF || !lattice || !sites || (s->R > 0 && !RL)) return fail(LMX_ERR_ARG, "null model array");
        if (s->batch) return fail(LMX_ERR_STATE, "the solver is an ensemble member; remove it before rebuilding its lattice");
        int rc = select_device(s);
        if (rc) return rc;
        LMX_HIP(hipStreamSynchronize(s->stream));
        // Shape, tables and buffers are replaced below before every check and
        // allocation has passed; the solver has no lattice until this call
        // succeeds, so a failed rebuild cannot launch a step over buffers of
        // another size.
        s->diffusion_built = false;

        s->nST = nST;
        s->X = X; s->Y = Y; s->Z = Z; s->P = P;
        s->bpp = bytesPerParticle;
        s->W = bytesPerParticle == 4 ? P : P / 4;
        s->N = (size_t)N;
        s->lambda = lambda;
        const size_t nT = (size_t)s->nS * nST * nST;
        s->DF.assign(DF, DF + nT);
        s->RL.assign(RL ? RL : nullptr, RL ? RL + (size_t)s->R * nST : nullptr);
        // Transition probabilities (MpdRdmeSolver.cu:408-421).
        s->T.assign(nT, 0.0f);
        for (size_t i = 0; i < nT; i++) {
            const float q = (float)(DF[i] * s->tau / pow(lambda, 2));
            if (q > 0.50f) return fail(LMX_ERR_ARG, "D: The specified diffusion coefficient is too high for the diffusion model.");
            s->T[i] = q;
        }
        std::vector<uint32_t> native;
        if ((rc = row_major_to_native(s, lattice, native))) return rc;
        std::vector<uint8_t> nsites;
        if ((rc = sites_row_major_to_native(s, sites, nsites))) return rc;
        compute_propensities(s);

        free_device(s);
        const size_t lat_bytes = s->N * s->W * sizeof(uint32_t);
        // the byte lattice's chunk flags follow the planes of each buffer
        s->flag_words = (s->bpp == 1 && s->W > 1) ? (size_t)(s->W - 1) * lmx::chunk_count((uint32_t)s->N) : 0;
        s->flags_fresh = false;
        const size_t buf_bytes = lat_bytes + s->flag_words * sizeof(uint32_t);
        hipError_t e = hipMalloc((void **)&s->d_lat[0], buf_bytes);
        if (e == hipSuccess) e = hipMalloc((void **)&s->d_lat[1], buf_bytes);
        if (e == hipSuccess) e = hipMalloc((void **)&s->d_sites, s->N);
        if (e == hipSuccess && s->nS > s->counts_cap) {
            hipFree(s->d_counts);
            s->d_counts = nullptr;
            e = hipMalloc((void **)&s->d_counts, (size_t)s->nS * sizeof(uint32_t));
            s->counts_cap = e == hipSuccess ? s->nS : 0u;
        }
        // The reaction list: kListShards shards of cap records (index, W words,
        // type).  Any capacity is correct (an overflowing shard makes the
        // reaction phase a full pass); N/8 entries in total hold every occupied
        // site up to 12.5% occupancy.
        const size_t rcap = std::max<size_t>(64, (s->N / 8 + lmx::kListShards - 1) / lmx::kListShards);
        const size_t rtotal = rcap * lmx::kListShards;
        // (the byte lattice's vector kernels only)
        if (e == hipSuccess && s->bpp == 1) e = hipMalloc((void **)&s->d_rl, rtotal * lmx::list_units(s->W) * sizeof(uint4));
        if (e == hipSuccess && s->d_rl) {
            s->rlist.rec = s->d_rl;
            s->rlist.cap = (uint32_t)rcap;
        }
        if (e != hipSuccess) {
            free_device(s);
            return fail(LMX_ERR_NOMEM, std::string("Failed to allocate the device lattice: ") + hipGetErrorString(e));
        }
        s->cur = 0;
        LMX_HIP(hipMemcpy(s->d_lat[0], native.data(), lat_bytes, hipMemcpyHostToDevice));
        LMX_HIP(hipMemset(s->d_lat[1], 0, lat_bytes));
        LMX_HIP(hipMemcpy(s->d_sites, nsites.data(), s->N, hipMemcpyHostToDevice));
        if ((rc = dev_upload(&s->d_T, s->T))) return rc;
        if ((rc = dev_upload(&s->d_order, s->order))) return rc;
        if ((rc = dev_upload(&s->d_D1, s->D1))) return rc;
        if ((rc = dev_upload(&s->d_D2, s->D2))) return rc;
        if ((rc = dev_upload(&s->d_S, s->S_rm))) return rc;
        std::vector<uint8_t> rl8(s->RL.size());
        for (size_t i = 0; i < rl8.size(); i++) rl8[i] = (uint8_t)s->RL[i];  // tmpRL (MpdRdmeSolver.cu:424-428)
        if ((rc = dev_upload(&s->d_RL, rl8))) return rc;
        // Reactions listed by first reactant (zero-order ones under nS), and the
        // nonzero stoichiometry of each reaction, both in ascending order.
        {
            std::vector<uint32_t> rx_off(s->nS + 2, 0), rx_list(s->R), sto_off(s->R + 1, 0), sto_sp;
            std::vector<int32_t> sto_val;
            auto bucket = [&](uint32_t r) { return s->order[r] == lmx::kZeroOrder ? s->nS : s->D1[r] - 1; };
            for (uint32_t r = 0; r < s->R; r++) rx_off[bucket(r) + 1]++;
            for (uint32_t b = 0; b <= s->nS; b++) rx_off[b + 1] += rx_off[b];
            std::vector<uint32_t> fill(rx_off.begin(), rx_off.end() - 1);
            for (uint32_t r = 0; r < s->R; r++) rx_list[fill[bucket(r)]++] = r;
            for (uint32_t r = 0; r < s->R; r++) {
                for (uint32_t sp = 0; sp < s->nS; sp++) {
                    const int8_t v = s->S_rm[(size_t)r * s->nS + sp];
                    if (v != 0) {
                        sto_sp.push_back(sp);
                        sto_val.push_back(v);
                    }
                }
                sto_off[r + 1] = (uint32_t)sto_sp.size();
            }
            if ((rc = dev_upload(&s->d_rx_off, rx_off))) return rc;
            if ((rc = dev_upload(&s->d_rx_list, rx_list))) return rc;
            if ((rc = dev_upload(&s->d_sto_off, sto_off))) return rc;
            if ((rc = dev_upload(&s->d_sto_sp, sto_sp))) return rc;
            if ((rc = dev_upload(&s->d_sto_val, sto_val))) return rc;
        }
        if ((rc = upload_tables(s))) return rc;
        LMX_HIP(hipMemset(s->d_ctl, 0, sizeof(lmx::Control)));
        s->timestep = 0;
        s->packed = true;  // row_major_to_native fills each site's first slots
        s->diffusion_built = true;
        return LMX_OK;
    });
}

int lmx_set_reaction_rate(lmx_solver *s, uint32_t rxid, float rate)
{
    return guarded([&]() -> int {
        if (!s) return fail(LMX_ERR_ARG, "null solver");
        if (!s->model_built || rxid >= s->R) return fail(LMX_ERR_ARG, "reaction index is out of bounds");
        s->k[rxid] = rate;
        if (s->diffusion_built) compute_propensities(s);
        return LMX_OK;
    });
}

int lmx_set_path(lmx_solver *s, int path)
{
    if (!s) return fail(LMX_ERR_ARG, "null solver");
    if (path != LMX_PATH_SWEEPS && path != LMX_PATH_FUSED) return fail(LMX_ERR_ARG, "unsupported step path");
    s->path = path;
    return LMX_OK;
}

int lmx_set_launch_option(lmx_solver *s, int option, int value)
{
    if (!s) return fail(LMX_ERR_ARG, "null solver");
    switch (option) {
    case LMX_OPT_MARCH_SEGMENT:
        if (value < 1 || value > 65536) return fail(LMX_ERR_ARG, "march segment must be in 1..65536");
        s->march_seg = (uint32_t)value;
        return LMX_OK;
    case LMX_OPT_FORCE_GENERIC:
        s->force_generic = value != 0;
        return LMX_OK;
    case LMX_OPT_YZ_KERNEL:
        if (value < 0 || value > 3)
            return fail(LMX_ERR_ARG,
                        "y/z kernel must be 0 (LDS tiles), 1 (march), 2 (compacted tiles) or 3 (register columns)");
        s->yz_kernel = value;
        return LMX_OK;
    case LMX_OPT_FUSE_REACTION:
        s->fuse_reaction = value != 0;
        return LMX_OK;
    case LMX_OPT_FUSED_ZSEG:
        if (value < 0 || value > 65536) return fail(LMX_ERR_ARG, "fused z segment must be in 0..65536");
        s->fused_zseg = (uint32_t)value;
        return LMX_OK;
    }
    return fail(LMX_ERR_ARG, "unknown launch option");
}

int lmx_run_steps(lmx_solver *s, uint32_t nsteps)
{
    int rc = require_ready(s);
    if (rc) return rc;
    if (s->tables_dirty) {
        LMX_HIP(hipStreamSynchronize(s->stream));
        if ((rc = upload_tables(s))) return rc;
    }
    const lmx::Params p = make_params(s);
    for (uint32_t i = 0; i < nsteps; i++)
        if ((rc = enqueue_step(s, p, nullptr, nullptr, nullptr))) return rc;
    return LMX_OK;
}

int lmx_run_steps_static(lmx_solver *s, uint32_t nsteps)
{
    int rc = require_ready(s);
    if (rc) return rc;
    if (s->N > lmx::kStaticListMaxSites)
        return fail(LMX_ERR_ARG, "lmx_run_steps_static: the lattice has " + std::to_string(s->N) +
                                     " sites; its static reaction list serves at most " +
                                     std::to_string(lmx::kStaticListMaxSites));
    if (s->tables_dirty)
        return fail(LMX_ERR_STATE, "lmx_run_steps_static: reaction rates changed since the tables were uploaded; "
                                   "lmx_run_steps uploads them (it allocates)");
    lmx::ReactionList rl{};
    if (s->bpp == 1) {
        const hipError_t e = lmx::static_reaction_list(s->device, s->W, &rl);
        if (e != hipSuccess) return hip_fail(e, "static reaction list");
    }
    const lmx::Params p = make_params(s);
    for (uint32_t i = 0; i < nsteps; i++)
        if ((rc = enqueue_step(s, p, nullptr, nullptr, nullptr, &rl))) return rc;
    return LMX_OK;
}

int lmx_synchronize(lmx_solver *s)
{
    if (!s) return fail(LMX_ERR_ARG, "null solver");
    int rc = select_device(s);
    if (rc) return rc;
    LMX_HIP(hipStreamSynchronize(s->stream));
    if (s->d_ctl) return check_device_error(s);
    return LMX_OK;
}

int lmx_get_timestep(lmx_solver *s, uint32_t *t)
{
    if (!s || !t) return fail(LMX_ERR_ARG, "null argument");
    *t = s->timestep;
    return LMX_OK;
}

int lmx_set_timestep(lmx_solver *s, uint32_t t)
{
    int rc = require_ready(s);
    if (rc) return rc;
    LMX_HIP(hipStreamSynchronize(s->stream));
    LMX_HIP(hipMemcpy(&s->d_ctl->timestep, &t, sizeof(t), hipMemcpyHostToDevice));
    s->timestep = t;
    return LMX_OK;
}

int lmx_get_lattice_native(lmx_solver *s, uint32_t *native, size_t words)
{
    int rc = require_ready(s);
    if (rc) return rc;
    if (!native || words != s->N * s->W) return fail(LMX_ERR_ARG, "the buffer size was not equal to the lattice data size");
    LMX_HIP(hipStreamSynchronize(s->stream));
    LMX_HIP(hipMemcpy(native, s->d_lat[s->cur], words * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return LMX_OK;
}

int lmx_get_lattice(lmx_solver *s, uint8_t *rm, size_t size)
{
    return guarded([&]() -> int {
        int rc = require_ready(s);
        if (rc) return rc;
        if (!rm || size != s->N * s->P * s->bpp) return fail(LMX_ERR_ARG, "bufferSize: the buffer size was not equal to the lattice data size");
        std::vector<uint32_t> native(s->N * s->W);
        if ((rc = lmx_get_lattice_native(s, native.data(), native.size()))) return rc;
        native_to_row_major(s, native.data(), rm);
        return LMX_OK;
    });
}

int lmx_get_sites(lmx_solver *s, uint8_t *rm, size_t size)
{
    return guarded([&]() -> int {
        int rc = require_ready(s);
        if (rc) return rc;
        if (!rm || size != s->N) return fail(LMX_ERR_ARG, "bufferSize: the buffer size was not equal to the lattice site size");
        std::vector<uint8_t> native(s->N);
        LMX_HIP(hipStreamSynchronize(s->stream));
        LMX_HIP(hipMemcpy(native.data(), s->d_sites, s->N, hipMemcpyDeviceToHost));
        size_t idx = 0;
        for (uint32_t x = 0; x < s->X; x++)
            for (uint32_t y = 0; y < s->Y; y++)
                for (uint32_t z = 0; z < s->Z; z++, idx++)
                    rm[idx] = native[(size_t)x + (size_t)s->X * y + (size_t)s->X * s->Y * z];
        return LMX_OK;
    });
}

int lmx_set_lattice_native(lmx_solver *s, const uint32_t *native, size_t words)
{
    int rc = require_ready(s);
    if (rc) return rc;
    if (!native || words != s->N * s->W) return fail(LMX_ERR_ARG, "the buffer size was not equal to the lattice data size");
    for (size_t i = 0; i < words; i++) {
        if (s->bpp == 4) {
            if (native[i] > s->nS) return fail(LMX_ERR_ARG, "latticeData: an invalid species was found");
            continue;
        }
        for (int b = 0; b < 4; b++)
            if (((native[i] >> (8 * b)) & 0xFFu) > s->nS)
                return fail(LMX_ERR_ARG, "latticeData: an invalid species was found");
    }
    // Whether every site holds its particles in its first slots.
    bool packed = true;
    for (size_t i = 0; i < s->N && packed; i++) {
        bool gap = false;
        for (uint32_t j = 0; j < s->P; j++) {
            const uint32_t v = s->bpp == 4 ? native[(size_t)j * s->N + i]
                                           : (native[(size_t)(j >> 2) * s->N + i] >> (8 * (j & 3))) & 0xFFu;
            if (v == 0) {
                gap = true;
            } else if (gap) {
                packed = false;
                break;
            }
        }
    }
    LMX_HIP(hipStreamSynchronize(s->stream));
    LMX_HIP(hipMemcpy(s->d_lat[s->cur], native, words * sizeof(uint32_t), hipMemcpyHostToDevice));
    s->packed = packed;
    s->flags_fresh = false;
    return LMX_OK;
}

int lmx_device_lattice(lmx_solver *s, void **particles, void **sites)
{
    int rc = require_ready(s);
    if (rc) return rc;
    if (!particles || !sites) return fail(LMX_ERR_ARG, "null argument");
    LMX_HIP(hipStreamSynchronize(s->stream));
    if ((rc = check_device_error(s))) return rc;
    *particles = s->d_lat[s->cur];
    *sites = s->d_sites;
    s->packed = false;  // the caller may write any slot
    s->flags_fresh = false;
    return LMX_OK;
}

int lmx_validate_lattice(lmx_solver *s)
{
    int rc = require_ready(s);
    if (rc) return rc;
    const lmx::Params p = make_params(s);
    hipError_t e = lmx::launch_validate(p, s->d_lat[s->cur], s->stream);
    if (e != hipSuccess) return hip_fail(e, "validate launch");
    LMX_HIP(hipStreamSynchronize(s->stream));
    return check_device_error(s);
}

int lmx_set_lattice(lmx_solver *s, const uint8_t *rm, size_t size)
{
    return guarded([&]() -> int {
        int rc = require_ready(s);
        if (rc) return rc;
        if (!rm || size != s->N * s->P * s->bpp) return fail(LMX_ERR_ARG, "bufferSize: the buffer size was not equal to the lattice data size");
        std::vector<uint32_t> native;
        if ((rc = row_major_to_native(s, rm, native))) return rc;
        LMX_HIP(hipStreamSynchronize(s->stream));
        LMX_HIP(hipMemcpy(s->d_lat[s->cur], native.data(), native.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        s->packed = true;  // each site's particles inserted at its first empty slot
        s->flags_fresh = false;
        return LMX_OK;
    });
}

int lmx_set_sites(lmx_solver *s, const uint8_t *rm, size_t size)
{
    return guarded([&]() -> int {
        int rc = require_ready(s);
        if (rc) return rc;
        if (!rm || size != s->N) return fail(LMX_ERR_ARG, "bufferSize: the buffer size was not equal to the lattice site size");
        std::vector<uint8_t> native;
        if ((rc = sites_row_major_to_native(s, rm, native))) return rc;
        LMX_HIP(hipStreamSynchronize(s->stream));
        LMX_HIP(hipMemcpy(s->d_sites, native.data(), s->N, hipMemcpyHostToDevice));
        return LMX_OK;
    });
}

int lmx_get_species_counts(lmx_solver *s, uint32_t *counts, uint32_t n)
{
    return guarded([&]() -> int {
        int rc = require_ready(s);
        if (rc) return rc;
        if (!counts || n < s->nS) return fail(LMX_ERR_ARG, "counts buffer smaller than the number of species");
        LMX_HIP(lmx::launch_fill(s->d_counts, s->counts_cap, 0u, s->stream));
        hipError_t e = lmx::launch_counts(s->d_lat[s->cur], s->N * s->W, s->nS, s->bpp == 4, s->d_counts, s->stream);
        if (e != hipSuccess) return hip_fail(e, "count launch");
        LMX_HIP(hipStreamSynchronize(s->stream));
        std::vector<uint32_t> h(s->counts_cap);
        LMX_HIP(hipMemcpy(h.data(), s->d_counts, s->counts_cap * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; i++) counts[i] = i < s->nS ? h[i] : 0u;
        return LMX_OK;
    });
}

int lmx_get_overflow_stats(lmx_solver *s, uint64_t *entries, uint64_t *steps)
{
    if (!s) return fail(LMX_ERR_ARG, "null solver");
    int rc = select_device(s);
    if (rc) return rc;
    LMX_HIP(hipStreamSynchronize(s->stream));
    lmx::Control c;
    LMX_HIP(hipMemcpy(&c, s->d_ctl, offsetof(lmx::Control, ovf_keys), hipMemcpyDeviceToHost));
    if (entries) *entries = c.ovf_total;
    if (steps) *steps = c.ovf_steps;
    return LMX_OK;
}

int lmx_get_unresolved_overflows(lmx_solver *s, uint32_t *count)
{
    if (!s || !count) return fail(LMX_ERR_ARG, "null argument");
    int rc = select_device(s);
    if (rc) return rc;
    LMX_HIP(hipStreamSynchronize(s->stream));
    lmx::Control c;
    LMX_HIP(hipMemcpy(&c, s->d_ctl, offsetof(lmx::Control, ovf_keys), hipMemcpyDeviceToHost));
    *count = c.carry_count;
    return LMX_OK;
}

int lmx_autotune(lmx_solver *s, uint32_t nsteps, double *best_ms, int *chosen)
{
    return guarded([&]() -> int {
        int rc = require_ready(s);
        if (rc) return rc;
        if (s->batch) return fail(LMX_ERR_STATE, "the solver is an ensemble member; tune it before adding it");
        if (nsteps == 0 || nsteps > 10000) return fail(LMX_ERR_ARG, "nsteps must be in 1..10000");
        LMX_HIP(hipStreamSynchronize(s->stream));
        if (s->tables_dirty && (rc = upload_tables(s))) return rc;

        // Everything a step reads or advances: both buffers (planes and chunk
        // flags), the control block (step counter, overflow statistics and
        // carried entries) and the host-side bookkeeping.
        const size_t buf_bytes = (s->N * s->W + s->flag_words) * sizeof(uint32_t);
        struct Scratch {
            void *p = nullptr;
            ~Scratch() { if (p) (void)hipFree(p); }
        } snap;
        LMX_HIP(hipMalloc(&snap.p, 2 * buf_bytes + sizeof(lmx::Control)));
        char *base = (char *)snap.p;
        const uint32_t t0 = s->timestep;
        const int cur0 = s->cur;
        const bool packed0 = s->packed, fresh0 = s->flags_fresh;
        for (int i = 0; i < 2; i++)
            LMX_HIP(hipMemcpyAsync(base + i * buf_bytes, s->d_lat[i], buf_bytes, hipMemcpyDeviceToDevice, s->stream));
        LMX_HIP(hipMemcpyAsync(base + 2 * buf_bytes, s->d_ctl, sizeof(lmx::Control), hipMemcpyDeviceToDevice,
                               s->stream));
        auto restore = [&]() -> int {
            for (int i = 0; i < 2; i++)
                LMX_HIP(hipMemcpyAsync(s->d_lat[i], base + i * buf_bytes, buf_bytes, hipMemcpyDeviceToDevice,
                                       s->stream));
            LMX_HIP(hipMemcpyAsync(s->d_ctl, base + 2 * buf_bytes, sizeof(lmx::Control), hipMemcpyDeviceToDevice,
                                   s->stream));
            s->timestep = t0;
            s->cur = cur0;
            s->packed = packed0;
            s->flags_fresh = fresh0;
            return LMX_OK;
        };

        struct Config {
            int path, yz;
            uint32_t seg;
            bool fuse_reaction;
        };
        // The y/z kernels and their window shapes; options a shape cannot use
        // fall back to the sweeps path and are timed as such.
        const Config cands[] = {
            {LMX_PATH_SWEEPS, 0, 8, true},  {LMX_PATH_SWEEPS, 0, 8, false}, {LMX_PATH_SWEEPS, 1, 4, true},
            {LMX_PATH_SWEEPS, 1, 8, true},  {LMX_PATH_SWEEPS, 2, 8, true},  {LMX_PATH_SWEEPS, 3, 4, true},
            {LMX_PATH_SWEEPS, 3, 8, true},  {LMX_PATH_FUSED, 0, 8, true},
        };
        const Config before{s->path, s->yz_kernel, s->march_seg, s->fuse_reaction};
        // With LMX_OPT_FORCE_GENERIC the caller chose the one-site-per-thread
        // kernels, which have no y/z shapes to choose among: only the current
        // configuration is timed, and every option stays as it is.
        const bool generic0 = s->force_generic;
        hipEvent_t e0, e1;
        LMX_HIP(hipEventCreate(&e0));
        LMX_HIP(hipEventCreate(&e1));
        struct Events {
            hipEvent_t a, b;
            ~Events() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
        } events{e0, e1};
        // The current configuration first; two interleaved passes, the
        // minimum of each candidate; another configuration replaces the
        // current one only if it is at least 2% faster, so that timing noise
        // between near-equal shapes does not flip the choice.
        std::vector<Config> list{before};
        for (const Config &c : cands)
            if (!generic0 && (c.path != before.path || c.yz != before.yz || c.seg != before.seg ||
                c.fuse_reaction != before.fuse_reaction))
                list.push_back(c);
        std::vector<double> times(list.size(), -1.0);
        for (int pass = 0; pass < 2 && !rc; pass++) {
            for (size_t k = 0; k < list.size() && !rc; k++) {
                const Config &c = list[k];
                s->path = c.path;
                s->yz_kernel = c.yz;
                s->march_seg = c.seg;
                s->fuse_reaction = c.fuse_reaction;
                if ((rc = restore())) break;
                lmx::Params p = make_params(s);
                if ((rc = enqueue_step(s, p, nullptr, nullptr, nullptr))) break;  // untimed first step
                p = make_params(s);
                LMX_HIP(hipEventRecord(e0, s->stream));
                for (uint32_t i = 0; i < nsteps && !rc; i++) rc = enqueue_step(s, p, nullptr, nullptr, nullptr);
                if (rc) break;
                LMX_HIP(hipEventRecord(e1, s->stream));
                LMX_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                LMX_HIP(hipEventElapsedTime(&ms, e0, e1));
                const double per = (double)ms / nsteps;
                if (times[k] < 0.0 || per < times[k]) times[k] = per;
            }
        }
        size_t win = 0;
        for (size_t k = 1; k < times.size(); k++)
            if (times[k] >= 0.0 && times[k] < times[win]) win = k;
        if (times[win] > 0.98 * times[0]) win = 0;
        const double best = times[win];
        const Config pick = list[win];
        const Config keep = rc ? before : pick;
        s->path = keep.path;
        s->yz_kernel = keep.yz;
        s->march_seg = keep.seg;
        s->fuse_reaction = keep.fuse_reaction;
        s->force_generic = generic0;
        const int rc2 = restore();
        LMX_HIP(hipStreamSynchronize(s->stream));
        if (rc) return rc;
        if (rc2) return rc2;
        if (best_ms) *best_ms = best;
        if (chosen) {
            chosen[0] = keep.path;
            chosen[1] = keep.yz;
            chosen[2] = (int)keep.seg;
            chosen[3] = keep.fuse_reaction ? 1 : 0;
        }
        return LMX_OK;
    });
}

int lmx_time_steps(lmx_solver *s, uint32_t nsteps, double *total_ms, double *kernel_ms, uint64_t *kernel_launches,
                   uint32_t nkernels)
{
    return guarded([&]() -> int {
        int rc = require_ready(s);
        if (rc) return rc;
        if (s->tables_dirty) {
            LMX_HIP(hipStreamSynchronize(s->stream));
            if ((rc = upload_tables(s))) return rc;
        }
        const lmx::Params p = make_params(s);
        if (kernel_ms)
            for (uint32_t k = 0; k < nkernels; k++) kernel_ms[k] = 0.0;
        if (kernel_launches)
            for (uint32_t k = 0; k < nkernels; k++) kernel_launches[k] = 0;
        hipEvent_t t0, t1;
        LMX_HIP(hipEventCreate(&t0));
        LMX_HIP(hipEventCreate(&t1));
        const int kMaxLaunches = 8;
        const uint32_t chunk = 256;
        std::vector<hipEvent_t> ev;
        std::vector<int> kinds;
        if (kernel_ms) {
            ev.resize((size_t)chunk * (kMaxLaunches + 1));
            // Timing-only events: without the system-scope fence (a cache
            // write-back and invalidate) each record adds less between the
            // kernels it separates.
            for (auto &e : ev) LMX_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            kinds.resize((size_t)chunk * kMaxLaunches);
        }
        LMX_HIP(hipStreamSynchronize(s->stream));
        LMX_HIP(hipEventRecord(t0, s->stream));
        for (uint32_t base = 0; base < nsteps; base += chunk) {
            const uint32_t m = (nsteps - base) < chunk ? (nsteps - base) : chunk;
            std::vector<int> nl(m, 0);
            for (uint32_t i = 0; i < m; i++) {
                if (kernel_ms)
                    rc = enqueue_step(s, p, &ev[(size_t)i * (kMaxLaunches + 1)], &kinds[(size_t)i * kMaxLaunches], &nl[i]);
                else
                    rc = enqueue_step(s, p, nullptr, nullptr, nullptr);
                if (rc) return rc;
            }
            if (kernel_ms) {
                LMX_HIP(hipEventSynchronize(ev[(size_t)(m - 1) * (kMaxLaunches + 1) + nl[m - 1]]));
                for (uint32_t i = 0; i < m; i++) {
                    hipEvent_t *e = &ev[(size_t)i * (kMaxLaunches + 1)];
                    for (int j = 0; j < nl[i]; j++) {
                        float ms = 0.0f;
                        LMX_HIP(hipEventElapsedTime(&ms, e[j], e[j + 1]));
                        const int kind = kinds[(size_t)i * kMaxLaunches + j];
                        if ((uint32_t)kind < nkernels) {
                            kernel_ms[kind] += ms;
                            if (kernel_launches) kernel_launches[kind]++;
                        }
                    }
                }
            }
        }
        LMX_HIP(hipEventRecord(t1, s->stream));
        LMX_HIP(hipEventSynchronize(t1));
        float ms = 0.0f;
        LMX_HIP(hipEventElapsedTime(&ms, t0, t1));
        if (total_ms) *total_ms = ms;
        for (auto &e : ev) hipEventDestroy(e);
        hipEventDestroy(t0);
        hipEventDestroy(t1);
        return check_device_error(s);
    });
}

int lmx_get_table_f32(lmx_solver *s, int which, float *out, size_t n, size_t *written)
{
    return guarded([&]() -> int {
        if (!s) return fail(LMX_ERR_ARG, "null solver");
        const std::vector<float> *v = nullptr;
        switch (which) {
        case 0: v = &s->T; break;
        case 1: v = &s->qp0; break;
        case 2: v = &s->qp1; break;
        case 3: v = &s->qp2; break;
        case 4: v = &s->rate; break;
        default: return fail(LMX_ERR_ARG, "unknown table");
        }
        if (written) *written = v->size();
        if (out) {
            if (n < v->size()) return fail(LMX_ERR_ARG, "table buffer too small");
            memcpy(out, v->data(), v->size() * sizeof(float));
        }
        return LMX_OK;
    });
}

int lmx_copy_bandwidth(int device, size_t bytes, int reps, double *gbps)
{
    if (!gbps || bytes < (1u << 20) || reps < 1) return fail(LMX_ERR_ARG, "bytes must be at least 1 MiB and reps positive");
    LMX_HIP(hipSetDevice(device));
    int num_cu = 256;
    hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, device);
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipMemset(a, 1, bytes);
    if (e == hipSuccess) e = hipStreamCreate(&st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = lmx::launch_copy(a, b, bytes, num_cu, st);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = lmx::launch_copy(r & 1 ? b : a, r & 1 ? a : b, bytes, num_cu, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    if (a) hipFree(a);
    if (b) hipFree(b);
    if (e != hipSuccess) return hip_fail(e, "copy bandwidth");
    *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return LMX_OK;
}

int lmx_device_exp(int device, const double *x, double *y, size_t n)
{
    return guarded([&]() -> int {
        if (!x || !y) return fail(LMX_ERR_ARG, "null argument");
        LMX_HIP(hipSetDevice(device));
        double *dx = nullptr, *dy = nullptr;
        LMX_HIP(hipMalloc((void **)&dx, (n ? n : 1) * sizeof(double)));
        LMX_HIP(hipMalloc((void **)&dy, (n ? n : 1) * sizeof(double)));
        LMX_HIP(hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
        hipError_t e = lmx::launch_exp(dx, dy, n, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(y, dy, n * sizeof(double), hipMemcpyDeviceToHost);
        hipFree(dx);
        hipFree(dy);
        if (e != hipSuccess) return hip_fail(e, "lmx_device_exp");
        return LMX_OK;
    });
}

int lmx_device_reaction_probability(int device, const float *a, float *prob, size_t n)
{
    return guarded([&]() -> int {
        if (!a || !prob) return fail(LMX_ERR_ARG, "null argument");
        LMX_HIP(hipSetDevice(device));
        float *da = nullptr, *dp = nullptr;
        LMX_HIP(hipMalloc((void **)&da, (n ? n : 1) * sizeof(float)));
        LMX_HIP(hipMalloc((void **)&dp, (n ? n : 1) * sizeof(float)));
        LMX_HIP(hipMemcpy(da, a, n * sizeof(float), hipMemcpyHostToDevice));
        hipError_t e = lmx::launch_prob(da, dp, n, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(prob, dp, n * sizeof(float), hipMemcpyDeviceToHost);
        hipFree(da);
        hipFree(dp);
        if (e != hipSuccess) return hip_fail(e, "lmx_device_reaction_probability");
        return LMX_OK;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Ensembles
// ---------------------------------------------------------------------------
struct lmx_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<lmx_solver *> members;
    uint32_t W = 0;
    bool use_graph = true;
    bool use_list = true;  // reaction phases over the members' lists (LMX_BATCH_REACTION_LIST)
    uint32_t tile_threads = 0;  // threads of the tiled sweeps' blocks (LMX_BATCH_TILE_THREADS; 0: the default)
    // device descriptors and block prefixes (x, y, z, reaction), sized for cap members
    uint32_t cap = 0;
    lmx::Member *d_members = nullptr, *h_members = nullptr;
    uint32_t *d_start = nullptr, *h_start = nullptr;
    uint32_t total[4] = {0, 0, 0, 0};
    // member of every block of the four launches, concatenated (map_cap entries)
    size_t map_cap = 0;
    uint32_t *d_map = nullptr, *h_map = nullptr;
    // captured step, valid for the member list and totals it was made with
    hipGraphExec_t exec = nullptr;
    std::vector<lmx_solver *> graph_members;
    uint32_t graph_total[4] = {0, 0, 0, 0};
    // recorded after the last upload from h_members/h_start: the pinned
    // buffers are rewritten only once that upload has left them
    hipEvent_t staged = nullptr;
    bool staged_pending = false;
};

namespace {

void batch_drop_graph(lmx_batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->graph_members.clear();
}

// A member's own work is done, its tables are current and its chunk flags
// zeroed.  Uploading changed tables allocates; without may_alloc that is an
// error instead.
int member_ready(lmx_solver *s, bool may_alloc)
{
    if (s->tables_dirty) {
        if (!may_alloc)
            return fail(LMX_ERR_STATE, "a member's reaction rates changed since its tables were uploaded; an allocating "
                                       "call (lmx_batch_run_steps or lmx_run_steps) uploads them");
        int rc = upload_tables(s);
        if (rc) return rc;
    }
    LMX_HIP(hipStreamSynchronize(s->stream));
    if (s->flag_words > 0 && !s->flags_fresh) {
        for (int bb = 0; bb < 2; bb++)
            LMX_HIP(hipMemset(s->d_lat[bb] + s->N * s->W, 0, s->flag_words * sizeof(uint32_t)));
        s->flags_fresh = true;
    }
    return LMX_OK;
}

uint32_t batch_tile_threads(const lmx_batch *b)
{
    return b->tile_threads ? b->tile_threads : lmx::kNarrowTileThreads;
}

// Descriptor of one member from its host-side state (buffer parity, tables)
// and the grid sizes of its share of the four batched launches.  With
// use_list, a member without zero-order reactions runs its reaction phase
// over the list its z sweep fills, as a single solver does by default.
void fill_member(const lmx_solver *s, bool use_list, uint32_t threads, lmx::Member &m, uint32_t blocks[4])
{
    m.p = make_params(s);
    if (s->flag_words > 0) m.p.plane_flags = 1u;  // the batched kernels are the default ones
    m.lat[0] = s->d_lat[s->cur];
    m.lat[1] = s->d_lat[1 - s->cur];
    for (uint32_t k = 0; k < 3; k++) m.h[k] = lmx::timestep_hash(m.p.seed32, (uint32_t)s->timestep, k);
    m.list_by_block = lmx::batch_list_by_block((uint32_t)s->N) ? 1u : 0u;
    m.geo[0] = lmx::tile_geom(m.p, 1, threads);
    m.geo[1] = lmx::tile_geom(m.p, 2, threads);
    blocks[0] = lmx::batch_blocks(lmx::x4_blocks(m.p));
    blocks[1] = lmx::batch_blocks(lmx::tile_blocks(m.geo[0]));
    blocks[2] = lmx::batch_blocks(lmx::tile_blocks(m.geo[1]));
    m.rl = lmx::ReactionList{};
    m.nshards = 0;
    if (s->R == 0) {
        m.react = lmx::kReactNone;
        blocks[3] = 0;
    } else if (use_list && !s->any_qp0 && s->rlist.rec) {
        m.react = lmx::kReactList;
        m.rl = s->rlist;
        m.nshards = lmx::batch_list_shards(m.geo[1], threads);
        blocks[3] = lmx::batch_list_blocks(m.nshards, (uint32_t)s->N);
    } else {
        m.react = lmx::kReactFull;
        blocks[3] = lmx::reaction_blocks(m.p);
    }
}

// After nsteps batched steps were enqueued: the members' host-side state
// advances, and later work on a member's own stream starts after the batch's.
int batch_finish(lmx_batch *b, uint32_t nsteps)
{
    LMX_HIP(hipEventRecord(b->done, b->stream));
    for (lmx_solver *s : b->members) {
        if (nsteps & 1u) s->cur = 1 - s->cur;
        s->timestep += nsteps;
        s->packed = true;
        LMX_HIP(hipStreamWaitEvent(s->stream, b->done, 0));
    }
    return LMX_OK;
}

// Descriptors of every member from its host-side state (buffer parity,
// tables); the device copy is refreshed before each run.
int batch_prepare(lmx_batch *b)
{
    if (b->staged_pending) {
        LMX_HIP(hipEventSynchronize(b->staged));
        b->staged_pending = false;
    }
    const uint32_t n = (uint32_t)b->members.size();
    if (n > b->cap) {
        uint32_t cap = b->cap ? b->cap : 16;
        while (cap < n) cap *= 2;
        dev_free(b->d_members);
        dev_free(b->d_start);
        if (b->h_members) hipHostFree(b->h_members);
        if (b->h_start) hipHostFree(b->h_start);
        b->h_members = nullptr;
        b->h_start = nullptr;
        b->cap = 0;
        LMX_HIP(hipMalloc((void **)&b->d_members, sizeof(lmx::Member) * cap));
        LMX_HIP(hipMalloc((void **)&b->d_start, sizeof(uint32_t) * 4 * (cap + 1)));
        LMX_HIP(hipHostMalloc((void **)&b->h_members, sizeof(lmx::Member) * cap, hipHostMallocDefault));
        LMX_HIP(hipHostMalloc((void **)&b->h_start, sizeof(uint32_t) * 4 * (cap + 1), hipHostMallocDefault));
        b->cap = cap;
        batch_drop_graph(b);
    }
    uint32_t acc[4] = {0, 0, 0, 0};
    for (uint32_t r = 0; r < n; r++) {
        lmx_solver *s = b->members[r];
        int rc = member_ready(s, true);
        if (rc) return rc;
        uint32_t blocks[4];
        fill_member(s, b->use_list, batch_tile_threads(b), b->h_members[r], blocks);
        for (int k = 0; k < 4; k++) {
            b->h_start[k * (b->cap + 1) + r] = acc[k];
            acc[k] += blocks[k];
        }
    }
    size_t blocks = 0;
    for (int k = 0; k < 4; k++) {
        b->h_start[k * (b->cap + 1) + n] = acc[k];
        b->total[k] = acc[k];
        blocks += acc[k];
    }
    if (blocks > b->map_cap) {
        size_t cap = b->map_cap ? b->map_cap : 4096;
        while (cap < blocks) cap *= 2;
        dev_free(b->d_map);
        if (b->h_map) hipHostFree(b->h_map);
        b->h_map = nullptr;
        b->map_cap = 0;
        LMX_HIP(hipMalloc((void **)&b->d_map, sizeof(uint32_t) * cap));
        LMX_HIP(hipHostMalloc((void **)&b->h_map, sizeof(uint32_t) * cap, hipHostMallocDefault));
        b->map_cap = cap;
        batch_drop_graph(b);
    }
    uint32_t *map = b->h_map;
    for (int k = 0; k < 4; k++) {
        const uint32_t *st = b->h_start + k * (b->cap + 1);
        for (uint32_t r = 0; r < n; r++) std::fill(map + st[r], map + st[r + 1], r);
        map += acc[k];
    }
    LMX_HIP(hipMemcpyAsync(b->d_members, b->h_members, sizeof(lmx::Member) * n, hipMemcpyHostToDevice, b->stream));
    LMX_HIP(hipMemcpyAsync(b->d_start, b->h_start, sizeof(uint32_t) * 4 * (b->cap + 1), hipMemcpyHostToDevice,
                           b->stream));
    if (blocks) LMX_HIP(hipMemcpyAsync(b->d_map, b->h_map, sizeof(uint32_t) * blocks, hipMemcpyHostToDevice, b->stream));
    LMX_HIP(hipEventRecord(b->staged, b->stream));
    b->staged_pending = true;
    return LMX_OK;
}

int batch_enqueue_step(lmx_batch *b)
{
    const uint32_t n = (uint32_t)b->members.size();
    const uint32_t stride = b->cap + 1;
    const uint32_t *start[4], *map[4];
    size_t off = 0;
    for (int k = 0; k < 4; k++) {
        start[k] = b->d_start + k * stride;
        map[k] = b->d_map + off;
        off += b->total[k];
    }
    hipError_t e = lmx::launch_batch_sweeps(b->d_members, start, map, n, b->total, b->W, batch_tile_threads(b),
                                            b->stream);
    if (e != hipSuccess) return hip_fail(e, "batched sweep launch");
    e = lmx::launch_batch_resolve(b->d_members, n, b->stream);
    if (e != hipSuccess) return hip_fail(e, "batched resolve launch");
    return LMX_OK;
}

int batch_capture(lmx_batch *b)
{
    if (b->exec && b->graph_members == b->members && std::memcmp(b->graph_total, b->total, sizeof(b->total)) == 0)
        return LMX_OK;
    batch_drop_graph(b);
    hipGraph_t g = nullptr;
    LMX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue_step(b);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->exec = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    b->graph_members = b->members;
    std::memcpy(b->graph_total, b->total, sizeof(b->total));
    return LMX_OK;
}

}  // namespace

int lmx_batch_create(int device, lmx_batch **out)
{
    return guarded([&]() -> int {
        if (!out) return fail(LMX_ERR_ARG, "null output handle");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(LMX_ERR_HIP, "no HIP device is available");
        if (device < 0 || device >= n) return fail(LMX_ERR_ARG, "device index out of range");
        lmx_batch *b = new lmx_batch();
        b->device = device;
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming);
        if (e != hipSuccess) {
            int rc = hip_fail(e, "lmx_batch_create");
            lmx_batch_destroy(b);
            return rc;
        }
        *out = b;
        return LMX_OK;
    });
}

int lmx_batch_destroy(lmx_batch *b)
{
    if (!b) return LMX_OK;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    for (lmx_solver *s : b->members) s->batch = nullptr;
    batch_drop_graph(b);
    dev_free(b->d_members);
    dev_free(b->d_start);
    dev_free(b->d_map);
    if (b->h_members) hipHostFree(b->h_members);
    if (b->h_start) hipHostFree(b->h_start);
    if (b->h_map) hipHostFree(b->h_map);
    if (b->done) hipEventDestroy(b->done);
    if (b->staged) hipEventDestroy(b->staged);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return LMX_OK;
}

int lmx_batch_add(lmx_batch *b, lmx_solver *s)
{
    return guarded([&]() -> int {
        if (!b || !s) return fail(LMX_ERR_ARG, "null batch or solver");
        if (s->batch) return fail(LMX_ERR_ARG, "the solver is already a member of a batch");
        if (s->device != b->device) return fail(LMX_ERR_ARG, "the solver lives on another device");
        int rc = require_ready(s);
        if (rc) return rc;
        const lmx::Params p = make_params(s);
        if (s->bpp != 1) return fail(LMX_ERR_ARG, "batched solvers need the byte lattice");
        if (!lmx::fast_path_ok(p)) return fail(LMX_ERR_ARG, "batched solvers need X % 4 == 0");
        if (!b->members.empty() && s->W != b->W) return fail(LMX_ERR_ARG, "batched solvers need the same particles per site");
        if (b->tile_threads > lmx::tile_threads_for(s->W))
            return fail(LMX_ERR_ARG, "the batch's tile threads exceed what solvers with this many slots per site allow");
        b->W = s->W;
        b->members.push_back(s);
        s->batch = b;
        return LMX_OK;
    });
}

int lmx_batch_remove(lmx_batch *b, lmx_solver *s)
{
    return guarded([&]() -> int {
        if (!b || !s) return fail(LMX_ERR_ARG, "null batch or solver");
        auto it = std::find(b->members.begin(), b->members.end(), s);
        if (it == b->members.end()) return fail(LMX_ERR_ARG, "the solver is not a member of this batch");
        hipSetDevice(b->device);
        hipStreamSynchronize(b->stream);
        b->members.erase(it);
        s->batch = nullptr;
        return LMX_OK;
    });
}

int lmx_batch_size(const lmx_batch *b, uint32_t *n)
{
    if (!b || !n) return fail(LMX_ERR_ARG, "null argument");
    *n = (uint32_t)b->members.size();
    return LMX_OK;
}

int lmx_batch_set_option(lmx_batch *b, int option, int value)
{
    if (!b) return fail(LMX_ERR_ARG, "null batch");
    if (option == LMX_BATCH_GRAPH) {
        b->use_graph = value != 0;
        if (!b->use_graph) batch_drop_graph(b);
        return LMX_OK;
    }
    if (option == LMX_BATCH_REACTION_LIST) {
        if (b->use_list != (value != 0)) batch_drop_graph(b);
        b->use_list = value != 0;
        return LMX_OK;
    }
    if (option == LMX_BATCH_TILE_THREADS) {
        const uint32_t t = (uint32_t)value;
        if (value != 0 && (value < 0 || t % 64u || t < 256u || (b->W && t > lmx::tile_threads_for(b->W)) || t > 1024u))
            return fail(LMX_ERR_ARG, "tile threads: 0 (default, 256) or a multiple of 64 from 256 up to 1024 (512 for P = 16)");
        if (b->tile_threads != t) batch_drop_graph(b);
        b->tile_threads = t;
        return LMX_OK;
    }
    return fail(LMX_ERR_ARG, "unknown batch option");
}

int lmx_batch_run_steps(lmx_batch *b, uint32_t nsteps)
{
    return guarded([&]() -> int {
        if (!b) return fail(LMX_ERR_ARG, "null batch");
        if (nsteps == 0 || b->members.empty()) return LMX_OK;
        LMX_HIP(hipSetDevice(b->device));
        int rc = batch_prepare(b);
        if (rc) return rc;
        if (b->use_graph) {
            if ((rc = batch_capture(b))) return rc;
            for (uint32_t i = 0; i < nsteps; i++) LMX_HIP(hipGraphLaunch(b->exec, b->stream));
        } else {
            for (uint32_t i = 0; i < nsteps; i++)
                if ((rc = batch_enqueue_step(b))) return rc;
        }
        return batch_finish(b, nsteps);
    });
}


int lmx_batch_synchronize(lmx_batch *b)
{
    if (!b) return fail(LMX_ERR_ARG, "null batch");
    LMX_HIP(hipSetDevice(b->device));
    LMX_HIP(hipStreamSynchronize(b->stream));
    for (lmx_solver *s : b->members) {
        int rc = check_device_error(s);
        if (rc) return rc;
    }
    return LMX_OK;
}

int lmx_batch_autotune(lmx_batch *b, uint32_t nsteps, double *best_ms, int *chosen)
{
    return guarded([&]() -> int {
        if (!b) return fail(LMX_ERR_ARG, "null batch");
        if (nsteps == 0 || nsteps > 10000) return fail(LMX_ERR_ARG, "nsteps must be in 1..10000");
        if (b->members.empty()) return fail(LMX_ERR_STATE, "the batch has no members");
        LMX_HIP(hipSetDevice(b->device));
        int rc = batch_prepare(b);  // members' own work done, tables current
        if (rc) return rc;
        LMX_HIP(hipStreamSynchronize(b->stream));

        // Every member's state a step reads or advances: both buffers
        // (planes and chunk flags), the control block, the host bookkeeping.
        struct Saved {
            size_t off, buf_bytes;
            uint32_t timestep;
            int cur;
            bool packed, fresh;
        };
        std::vector<Saved> saved;
        size_t total = 0;
        for (lmx_solver *s : b->members) {
            const size_t bb = (s->N * s->W + s->flag_words) * sizeof(uint32_t);
            saved.push_back({total, bb, s->timestep, s->cur, s->packed, s->flags_fresh});
            total += 2 * bb + sizeof(lmx::Control);
            total = (total + 255) & ~(size_t)255;
        }
        struct Scratch {
            void *p = nullptr;
            ~Scratch() { if (p) (void)hipFree(p); }
        } snap;
        LMX_HIP(hipMalloc(&snap.p, total));
        char *base = (char *)snap.p;
        auto copy_state = [&](bool save) -> int {
            for (size_t k = 0; k < b->members.size(); k++) {
                lmx_solver *s = b->members[k];
                const Saved &v = saved[k];
                char *o = base + v.off;
                for (int i = 0; i < 2; i++) {
                    void *dev = s->d_lat[i], *cp = o + i * v.buf_bytes;
                    LMX_HIP(hipMemcpyAsync(save ? cp : dev, save ? dev : cp, v.buf_bytes, hipMemcpyDeviceToDevice,
                                           b->stream));
                }
                void *cc = o + 2 * v.buf_bytes;
                LMX_HIP(hipMemcpyAsync(save ? cc : (void *)s->d_ctl, save ? (void *)s->d_ctl : cc,
                                       sizeof(lmx::Control), hipMemcpyDeviceToDevice, b->stream));
                if (!save) {
                    s->timestep = v.timestep;
                    s->cur = v.cur;
                    s->packed = v.packed;
                    s->flags_fresh = v.fresh;
                }
            }
            return LMX_OK;
        };
        if ((rc = copy_state(true))) return rc;

        struct Config {
            bool list;
            uint32_t threads;
        };
        const uint32_t full = lmx::tile_threads_for(b->W), narrow = lmx::kNarrowTileThreads;
        const Config before{b->use_list, batch_tile_threads(b)};
        std::vector<Config> list{before};
        for (bool l : {true, false})
            for (uint32_t t : {narrow, full / 2, full}) {
                bool seen = false;
                for (const Config &c : list) seen = seen || (c.list == l && c.threads == t);
                if (!seen) list.push_back({l, t});
            }
        hipEvent_t e0, e1;
        LMX_HIP(hipEventCreate(&e0));
        LMX_HIP(hipEventCreate(&e1));
        struct Events {
            hipEvent_t a, b;
            ~Events() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
        } events{e0, e1};
        std::vector<double> times(list.size(), -1.0);
        for (int pass = 0; pass < 2 && !rc; pass++) {
            for (size_t k = 0; k < list.size() && !rc; k++) {
                if (b->use_list != list[k].list || batch_tile_threads(b) != list[k].threads) batch_drop_graph(b);
                b->use_list = list[k].list;
                b->tile_threads = list[k].threads == narrow ? 0u : list[k].threads;
                if ((rc = copy_state(false))) break;
                if ((rc = lmx_batch_run_steps(b, 1))) break;  // untimed first step (captures the graph)
                LMX_HIP(hipEventRecord(e0, b->stream));
                if ((rc = lmx_batch_run_steps(b, nsteps))) break;
                LMX_HIP(hipEventRecord(e1, b->stream));
                LMX_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                LMX_HIP(hipEventElapsedTime(&ms, e0, e1));
                const double per = (double)ms / nsteps;
                if (times[k] < 0.0 || per < times[k]) times[k] = per;
            }
        }
        size_t win = 0;
        for (size_t k = 1; k < times.size(); k++)
            if (times[k] >= 0.0 && times[k] < times[win]) win = k;
        if (times[win] > 0.98 * times[0]) win = 0;
        const Config keep = rc ? before : list[win];
        if (b->use_list != keep.list || batch_tile_threads(b) != keep.threads) batch_drop_graph(b);
        b->use_list = keep.list;
        b->tile_threads = keep.threads == narrow ? 0u : keep.threads;
        const int rc2 = copy_state(false);
        LMX_HIP(hipStreamSynchronize(b->stream));
        for (lmx_solver *s : b->members) LMX_HIP(hipStreamWaitEvent(s->stream, b->done, 0));
        if (rc) return rc;
        if (rc2) return rc2;
        if (best_ms) *best_ms = times[win];
        if (chosen) {
            chosen[0] = keep.list ? 1 : 0;
            chosen[1] = (int)keep.threads;
        }
        return lmx_batch_synchronize(b);
    });
}

int lmx_batch_time_steps(lmx_batch *b, uint32_t nsteps, double *total_ms)
{
    return guarded([&]() -> int {
        if (!b || !total_ms) return fail(LMX_ERR_ARG, "null argument");
        LMX_HIP(hipSetDevice(b->device));
        hipEvent_t t0, t1;
        LMX_HIP(hipEventCreate(&t0));
        LMX_HIP(hipEventCreate(&t1));
        int rc = batch_prepare(b);
        if (!rc && b->use_graph) rc = batch_capture(b);
        if (!rc) {
            LMX_HIP(hipStreamSynchronize(b->stream));
            LMX_HIP(hipEventRecord(t0, b->stream));
            rc = lmx_batch_run_steps(b, nsteps);
            LMX_HIP(hipEventRecord(t1, b->stream));
            LMX_HIP(hipEventSynchronize(t1));
            float ms = 0.0f;
            LMX_HIP(hipEventElapsedTime(&ms, t0, t1));
            *total_ms = ms;
        }
        hipEventDestroy(t0);
        hipEventDestroy(t1);
        if (rc) return rc;
        return lmx_batch_synchronize(b);
    });
}
