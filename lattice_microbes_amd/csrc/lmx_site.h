// Per-site MPD-RDME update rules shared by every kernel variant, so that all
// launch shapes and fusions evaluate exactly the same operations.
// Reference paths are relative to the Lattice Microbes 2.5.0 tree.
#pragma once

#include "lmx_device.h"

namespace lmx {

// ---------------------------------------------------------------------------
// One site of a diffusion sweep.  The new content of site i is its own
// stayers, then the minus neighbour's PLUS movers, then the plus neighbour's
// MINUS movers, each in slot order (performPropagation,
// byte_diffusion_1d_dev.cu:199-270).  A particle moves minus iff u < qm and
// plus iff 0.5 <= u < qp + 0.5 (makeDiffusionChoices, :86-121); q <= 0.5 makes
// the two exclusive, so a site only evaluates the decisions that can move a
// particle into it.  wo/wm/wp are the W particle words of the site and its
// neighbours; has_m/has_p are false at a reflecting edge.  Slots beyond P are
// recorded as overflow entries of this site and phase.  L is the slot layout
// (ByteSites<W> or WordSites<P>); the RNG key of slot j is (site << pcb) | j
// in both, as word_diffusion_1d_dev.cu:106 keys word w of the uint32 lattice.
// ---------------------------------------------------------------------------
template <class L>
__device__ inline void pull_site_l(const Params &p, uint32_t i, uint32_t im, uint32_t ip, const uint32_t *wo,
                                   const uint32_t *wm, const uint32_t *wp, uint32_t st, uint32_t stm, uint32_t stp,
                                   bool has_m, bool has_p, unsigned long long h, uint32_t phase, uint32_t *out)
{
    constexpr int W = L::W, Q = L::kPerWord;
    typename L::Buf b;
    b.clear();
#pragma unroll
    for (int w = 0; w < W; w++) {
        if (wo[w] == 0) continue;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const uint32_t part = L::part(wo[w], q);
            if (part == 0) continue;
            const float u = uniform01(i, p.pcb, Q * w + q, h);
            const float qm = lookup_T_if(p, p.T, has_m, part, st, stm);
            const float qp = lookup_T_if(p, p.T, has_p, part, st, stp);
            const bool leaves = (u < qm) || (u >= 0.5f && u < (qp + 0.5f));
            if (!leaves && !b.push(part)) record_overflow(p.ctl, phase, i, b.n - 1 - L::P, part);
        }
    }
    if (has_m) {
#pragma unroll
        for (int w = 0; w < W; w++) {
            if (wm[w] == 0) continue;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const uint32_t part = L::part(wm[w], q);
                if (part == 0) continue;
                const float u = uniform01(im, p.pcb, Q * w + q, h);
                const float qp = lookup_T(p, part, stm, st);
                if (u >= 0.5f && u < (qp + 0.5f) && !b.push(part))
                    record_overflow(p.ctl, phase, i, b.n - 1 - L::P, part);
            }
        }
    }
    if (has_p) {
#pragma unroll
        for (int w = 0; w < W; w++) {
            if (wp[w] == 0) continue;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const uint32_t part = L::part(wp[w], q);
                if (part == 0) continue;
                const float u = uniform01(ip, p.pcb, Q * w + q, h);
                const float qm = lookup_T(p, part, stp, st);
                if (u < qm && !b.push(part)) record_overflow(p.ctl, phase, i, b.n - 1 - L::P, part);
            }
        }
    }
#pragma unroll
    for (int w = 0; w < W; w++) out[w] = b.w[w];
}

template <int W>
__device__ inline void pull_site(const Params &p, uint32_t i, uint32_t im, uint32_t ip, const uint32_t *wo,
                                 const uint32_t *wm, const uint32_t *wp, uint32_t st, uint32_t stm, uint32_t stp,
                                 bool has_m, bool has_p, unsigned long long h, uint32_t phase, uint32_t *out)
{
    pull_site_l<ByteSites<W>>(p, i, im, ip, wo, wm, wp, st, stm, stp, has_m, has_p, h, phase, out);
}

// calculateReactionPropensity, global-R branch (byte_reaction_dev.cu:77-100).
template <class L>
__device__ inline float reaction_propensity(const Params &p, uint32_t st, const uint32_t *w, uint32_t r)
{
    if (!p.RL[r * p.nST + st]) return 0.0f;
    const uint32_t d1 = p.D1[r], d2 = p.D2[r];
    float n1 = 0.0f, n2 = 0.0f;
#pragma unroll
    for (int j = 0; j < L::P; j++) {
        const uint32_t s = L::slot(w, j);
        if (s > 0) {
            n1 += (s == d1) ? 1.0f : 0.0f;
            n2 += (s == d2) ? 1.0f : 0.0f;
        }
    }
    const float k = p.rate[r];
    switch (p.order[r]) {
    case kZeroOrder: return k;
    case kFirstOrder: return k * n1;
    case kSecondOrder: return k * n1 * n2;
    case kSecondOrderSelf: return k * n1 * (n1 - 1.0f);
    }
    return 0.0f;
}

// ---------------------------------------------------------------------------
// Reaction phase at one site, in place on w[W] (precomp_reaction_kernel,
// MpdRdmeSolver.cu:1592-1675, and byte_reaction_dev.cu:130-257).  Returns
// true if the site's particles changed.  Needs h3 = timestep hash of substep 3.
// ---------------------------------------------------------------------------
// Total propensity in the reference's summation order (:1617-1639).
template <class L>
__device__ inline float total_propensity(const Params &p, uint32_t st, const uint32_t *w)
{
    const uint32_t nS = p.nS;
    float a = p.qp0[st];
    uint32_t any = 0;
#pragma unroll
    for (int k = 0; k < L::W; k++) any |= w[k];
    if (any != 0) {
#pragma unroll
        for (int m = 0; m < L::P; m++) {
            const uint32_t p1 = L::slot(w, m);
            if (p1 > 0) {
                a += p.qp1[st * nS + (p1 - 1)];
#pragma unroll
                for (int n = m + 1; n < L::P; n++) {
                    const uint32_t p2 = L::slot(w, n);
                    if (p2 > 0) a += p.qp2[(st * nS + (p1 - 1)) * nS + (p2 - 1)];
                }
            }
        }
    }
    return a;
}

template <class L>
__device__ inline bool react_apply(const Params &p, uint32_t i, uint32_t st, uint32_t *w, float a, float prob,
                                   unsigned long long h3);

template <class L>
__device__ inline bool react_site_l(const Params &p, uint32_t i, uint32_t st, uint32_t *w, unsigned long long h3)
{
    const float a = total_propensity<L>(p, st, w);
    if (a == 0.0f) return false;
    return react_apply<L>(p, i, st, w, a, reaction_probability(a), h3);
}

template <int W>
__device__ inline bool react_site(const Params &p, uint32_t i, uint32_t st, uint32_t *w, unsigned long long h3)
{
    return react_site_l<ByteSites<W>>(p, i, st, w, h3);
}

// Same result as react_site.  A site holding at most one particle takes its
// total propensity as the same float sum and its probability from the table
// precomputed on the device with the same exp (prob1[st][species],
// prob0[st]), so the double-precision exp runs only for multi-particle sites.
template <class L>
__device__ inline bool react_site_table_l(const Params &p, uint32_t i, uint32_t st, uint32_t *w,
                                          unsigned long long h3, const float2 *tab)
{
    uint32_t n = 0, first = 0;
#pragma unroll
    for (int j = L::P - 1; j >= 0; j--) {
        const uint32_t s = L::slot(w, j);
        if (s) {
            n++;
            first = s;
        }
    }
    if (n > 1) return react_site_l<L>(p, i, st, w, h3);
    float a, prob;
    if (n == 1) {
        // tab (prob1, or its copy in LDS) holds (qp0[st] + qp1[st][s], its
        // probability) as one float2.
        const float2 ap = tab[st * p.nS + (first - 1)];
        a = ap.x;
        prob = ap.y;
    } else {
        a = p.qp0[st];
        prob = p.prob0[st];
    }
    if (a == 0.0f) return false;
    return react_apply<L>(p, i, st, w, a, prob, h3);
}

template <int W>
__device__ inline bool react_site_table(const Params &p, uint32_t i, uint32_t st, uint32_t *w,
                                        unsigned long long h3, const float2 *tab)
{
    return react_site_table_l<ByteSites<W>>(p, i, st, w, h3, tab);
}

// checkForReaction, determineReactionIndex and evaluateReaction
// (byte_reaction_dev.cu:139-257) for a site with total propensity a and
// reaction probability prob.
template <class L>
__device__ inline bool react_apply(const Params &p, uint32_t i, uint32_t st, uint32_t *w, float a, float prob,
                                   unsigned long long h3)
{
    const uint32_t nS = p.nS;
    // checkForReaction (byte_reaction_dev.cu:139-144).
    if (!(prob > 0.0f && uniform01(i, 1, 0, h3) <= prob)) return false;
    // determineReactionIndex (:146-177): reactions in ascending index, each
    // with a positive propensity moving rp.  A reaction whose first reactant
    // is absent has propensity 0 (or -0) and changes nothing, so only the
    // reactions listed under a species present at the site, plus the
    // zero-order ones, are visited: a merge of those ascending lists.
    constexpr int P = L::P, W = L::W;
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    float rp = uniform01(i, 1, 1, h3) * a;
    uint32_t idx = 0;
    {
        uint32_t cur[P + 1], end[P + 1], head[P + 1];
#pragma unroll
        for (int j = 0; j < P; j++) {
            const uint32_t s = L::slot(w, j);
            bool first = s != 0;
#pragma unroll
            for (int q = 0; q < j; q++) first = first && (L::slot(w, q) != s);
            const uint32_t sl = first ? s - 1 : 0u;
            cur[j] = first ? p.rx_off[sl] : 0u;
            end[j] = first ? p.rx_off[sl + 1] : 0u;
        }
        cur[P] = p.rx_off[nS];
        end[P] = p.rx_off[nS + 1];
#pragma unroll
        for (int j = 0; j <= P; j++) head[j] = cur[j] < end[j] ? p.rx_list[cur[j]] : kNone;
        for (;;) {
            uint32_t best = kNone;
            int bj = -1;
#pragma unroll
            for (int j = 0; j <= P; j++)
                if (head[j] < best) {
                    best = head[j];
                    bj = j;
                }
            if (best == kNone) break;
#pragma unroll
            for (int j = 0; j <= P; j++)
                if (j == bj) {
                    cur[j]++;
                    head[j] = cur[j] < end[j] ? p.rx_list[cur[j]] : kNone;
                }
            const float pr = reaction_propensity<L>(p, st, w, best);
            if (pr > 0.0f) {
                if (rp > 0.0f) idx = best;
                rp -= pr;
            }
        }
    }
    // evaluateReaction (:179-257) without the device-heap copy of the S row:
    // the m-th particle of species s is consumed iff fewer than -S[s] earlier
    // particles of s were seen.  S[idx] comes from its nonzero entries.
    const uint32_t k0 = p.sto_off[idx], k1 = p.sto_off[idx + 1];
    typename L::Buf out;
    out.clear();
#pragma unroll
    for (int m = 0; m < P; m++) {
        const uint32_t s = L::slot(w, m);
        if (s == 0) continue;
        int sv = 0;
        for (uint32_t k = k0; k < k1; k++)
            if (p.sto_sp[k] == s - 1) sv = p.sto_val[k];
        if (sv >= 0) {
            out.push(s);
        } else {
            uint32_t earlier = 0;
#pragma unroll
            for (int q = 0; q < m; q++) earlier += (L::slot(w, q) == s) ? 1u : 0u;
            if ((int)earlier >= -sv) out.push(s);
        }
    }
    // Products in species order; a reactant that was not present leaves a
    // negative residue, which the reference turns into an unbounded loop.
    bool underflow = false;
    for (uint32_t k = k0; k < k1; k++) {
        const uint32_t s = p.sto_sp[k];
        const int sv = p.sto_val[k];
        if (sv < 0) {
            uint32_t present = 0;
#pragma unroll
            for (int q = 0; q < P; q++) present += (L::slot(w, q) == s + 1) ? 1u : 0u;
            if ((int)present < -sv) underflow = true;
            continue;
        }
        for (int j = 0; j < sv; j++) {
            if (!out.push(s + 1)) record_overflow(p.ctl, 3, i, out.n - 1 - P, s + 1);
        }
    }
    if (underflow) {
        if (atomicCAS(&p.ctl->error, 0u, kErrReactionUnderflow) == 0u) p.ctl->error_step = p.ctl->timestep;
    }
#pragma unroll
    for (int k = 0; k < W; k++) w[k] = out.w[k];
    return true;
}

}  // namespace lmx
