// Host-side schedule of the batched operator launch: which workgroup runs
// which z-span of which design's column.  Plain C++ (no HIP calls), so the
// schedule can be checked without a GPU.
#pragma once
#include <algorithm>
#include <cstdint>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

namespace qms {

// One design as the planner sees it: workgroup columns of one plane (ntx *
// nty), planes, and the z-band of its natural pieces (the design's z-chunk).
struct PlanDesign {
    unsigned ncol;
    int nz;
    int band;
};

struct PlanSeg {
    unsigned d, col;
    int kb, ke;
};

struct BatchPlan {
    unsigned slots = 8;                 // workgroups, a multiple of 8
    std::vector<PlanSeg> segs;          // slot s runs segs[off[s] .. off[s+1])
    std::vector<unsigned> off;          // slots + 1 entries
};

// Cost model of a segment, in plane-equivalents: its planes plus the start of
// the z-march (two extra planes staged, the first coefficients loaded).
constexpr int kSegStart = 2;
// Smallest span a split produces.
constexpr int kMinSpan = 2;

// Natural pieces: for each design, z-band by z-band, every column of the
// band (x fastest, then y).  Consecutive pieces are neighbouring columns at
// the same depth.
//
// blocks == 0: one workgroup per natural piece (slots = pieces rounded up to
// a multiple of 8; XCD x runs the x-th contiguous eighth of the slots).
//
// blocks > 0: slots = blocks rounded up to a multiple of 8.  The pieces are
// cut into 8 contiguous runs of equal cost, one per XCD.  Within an XCD the
// pieces are dealt in order, each to the slot with the least work so far
// (ties to the lowest slot), and a piece that would take that slot past the
// XCD's mean is split in z.  So the slots of an XCD start on consecutive
// pieces (neighbouring columns, which share the XCD's L2 for their halos),
// and each slot ends with about the mean cost.
inline BatchPlan plan_batch(const std::vector<PlanDesign> &ds, unsigned blocks) {
    std::vector<PlanSeg> nat;
    for (size_t d = 0; d < ds.size(); ++d) {
        const int band = ds[d].band >= 1 ? ds[d].band : 1;
        for (int kb = 0; kb < ds[d].nz; kb += band)
            for (unsigned c = 0; c < ds[d].ncol; ++c)
                nat.push_back(PlanSeg{(unsigned)d, c, kb, std::min(kb + band, ds[d].nz)});
    }
    BatchPlan p;
    if (blocks == 0) {
        const size_t n = nat.size();
        p.slots = (unsigned)std::max<size_t>(8, (n + 7) / 8 * 8);
        p.segs = nat;
        p.off.resize(p.slots + 1);
        for (unsigned s = 0; s <= p.slots; ++s) p.off[s] = (unsigned)std::min<size_t>(s, n);
        return p;
    }
    p.slots = (blocks + 7u) / 8u * 8u;
    const unsigned per = p.slots / 8u;
    auto cost = [](const PlanSeg &s) { return (int64_t)(s.ke - s.kb) + kSegStart; };
    int64_t total = 0;
    for (const PlanSeg &s : nat) total += cost(s);
    // XCD runs: piece i goes to the XCD whose eighth of the total cost holds
    // the piece's midpoint.
    std::vector<size_t> run(9, nat.size());
    run[0] = 0;
    {
        int64_t cum = 0;
        unsigned x = 0;
        for (size_t i = 0; i < nat.size(); ++i) {
            const int64_t mid2 = 2 * cum + cost(nat[i]);          // twice the midpoint
            while (x < 7 && mid2 * 8 >= 2 * total * (int64_t)(x + 1)) run[++x] = i;
            cum += cost(nat[i]);
        }
        while (x < 7) run[++x] = nat.size();
    }
    std::vector<std::vector<PlanSeg>> lists(p.slots);
    for (unsigned x = 0; x < 8; ++x) {
        int64_t cx = 0;
        for (size_t i = run[x]; i < run[x + 1]; ++i) cx += cost(nat[i]);
        const int64_t target = (cx + per - 1) / per;              // mean cost per slot, rounded up
        typedef std::pair<int64_t, unsigned> LoadSlot;
        std::priority_queue<LoadSlot, std::vector<LoadSlot>, std::greater<LoadSlot>> heap;
        for (unsigned s = 0; s < per; ++s) heap.push(LoadSlot(0, s));
        for (size_t i = run[x]; i < run[x + 1]; ++i) {
            PlanSeg piece = nat[i];
            while (piece.kb < piece.ke) {
                const LoadSlot ls = heap.top();
                heap.pop();
                const int planes = piece.ke - piece.kb;
                const int64_t room = target - ls.first;
                int take = planes;
                if (planes + kSegStart > room) {
                    const int64_t q = std::max<int64_t>(kMinSpan, room - kSegStart);
                    if (planes - q >= kMinSpan) take = (int)q;
                }
                lists[x * per + ls.second].push_back(PlanSeg{piece.d, piece.col, piece.kb, piece.kb + take});
                heap.push(LoadSlot(ls.first + take + kSegStart, ls.second));
                piece.kb += take;
            }
        }
    }
    p.off.resize(p.slots + 1);
    p.off[0] = 0;
    for (unsigned s = 0; s < p.slots; ++s) {
        p.segs.insert(p.segs.end(), lists[s].begin(), lists[s].end());
        p.off[s + 1] = (unsigned)p.segs.size();
    }
    return p;
}

}  // namespace qms
