// Driver around the reference's header-only event queue
// (src/reaction/ReactionQueue.h, included from the reference tree): runs a
// sequence of updates and reports the queue top and every index's heap
// position after each one.  Built into oracle/_ref/ as a checker for the
// next-subvolume counterpart's heap; never part of the product.
#include <cstdlib>
#include <sys/types.h>

#include "reaction/ReactionQueue.h"

namespace {
class Probe : public lm::reaction::ReactionQueue {
public:
    explicit Probe(uint n) : ReactionQueue(n) {}
    uint position(uint i) const { return reactionPositions[i]; }
};
}  // namespace

extern "C" int ref_heap_sequence(unsigned n, unsigned nupd, const unsigned *idx, const double *times,
                                 unsigned *top_out, unsigned *pos_out)
{
    Probe q(n);
    for (unsigned u = 0; u < nupd; u++) {
        if (idx[u] >= n) return 1;
        q.updateReactionEvent(idx[u], times[u], 1.0);
        top_out[u] = q.getNextReaction();
        if (pos_out)
            for (unsigned i = 0; i < n; i++) pos_out[(size_t)u * n + i] = q.position(i);
    }
    return 0;
}
