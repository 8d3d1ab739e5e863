// Device-resident plan of a compiled DAG context: record layouts shared by the
// host compiler (mcdp_host.cpp) and the gfx950 kernels (mcdp_kernels.hip).
//
// The reference walks events one sample at a time in Kahn order over a CSR
// keyed by event id (_core.cpp:266-285, 332-350).  Here the same predecessor
// lists are grouped into *steps*: every event of a step depends only on events
// of earlier steps, and each step is split over the waves of a workgroup.  A
// workgroup owns 64 samples (one per lane), so all control flow and all plan
// records are wave-uniform and come through scalar loads; only the per-sample
// values (realized times, durations, random draws) live in vector registers.
#pragma once
#include <stdint.h>

namespace mcdp {

constexpr int kLanes = 64;  // wave64: one sample per lane

// One event evaluated by one wave in one step.  32 bytes, 8-byte aligned so a
// wave fetches it with a single scalar load.
struct EventRec {
    double earliest;   // EventTimestamp.earliest (initial realized value)
    double upper;      // earliest + max_delay, rounded once on the host
    int32_t event;     // event id: output row of realized / cause
    int32_t slot;      // LDS slot receiving this event's realized, or -1
    int32_t pred_begin;  // first PredRec
    int32_t pred_count;  // number of PredRecs (input order of the winning entry)
};

// Where a predecessor's realized time is read from.
enum : int32_t {
    kSrcGlobal = -1,  // from the realized output row (written in an earlier step)
    kSrcStale = -2,   // the reference reads this predecessor before it is evaluated:
                      // use its earliest time (held in src_value)
};

// Flags packed in PredRec::info above the 24-bit distribution index.
enum : uint32_t {
    kInfoDistMask = 0x00FFFFFFu,  // distribution index; kNoDist = no distribution
    kNoDist = 0x00FFFFFFu,
    kInfoOwner = 1u << 30,        // this record writes durations[edge]
    kInfoEdgeValid = 1u << 29,    // edge in [0, link_count); otherwise duration 0.0
};

// One predecessor entry (src, edge) of an event.  32 bytes.
struct PredRec {
    int32_t src_event;  // predecessor event id (the cause_event value)
    int32_t src_slot;   // LDS slot (>= 0), kSrcGlobal or kSrcStale
    int32_t edge;       // activity index (link) -- durations row and RNG counter
    uint32_t info;      // distribution index | flags
    double dmin;        // Activity.minimal_duration of the edge (0.0 for gaps)
    double src_value;   // earliest time of src (used when src_slot == kSrcStale)
};

// An activity that no predecessor entry references still has a duration row
// (the reference samples every link, _core.cpp:323-329).
struct OrphanRec {
    int32_t edge;
    uint32_t info;  // distribution index (kNoDist if none)
    double dmin;
};

// Sampling methods.  Each registered distribution compiles to one of these.
enum DistKind : int32_t {
    kConstant = 0,         // extra = d * p0
    kExpInf = 1,           // X = -p0 * log(u), u in (0,1]                       (mean p0)
    kExpTrunc = 2,         // X = -p0 * log(p1 + p3 u), u in (0,1], clamp at p2  (p1 = e^{-m/p0}, p3 = 1 - p1)
    kGammaErlang2 = 3,     // X = -p0 * log(u1 * u2)                              (shape 2, scale p0)
    kGammaHalf = 4,        // X = p0 * (-log u1) * cos^2(2 pi u2)                 (shape 1/2, scale p0)
    kGammaMT = 5,          // Marsaglia-Tsang, shape p1 >= 1 (or p1 + 1 with boost), scale p0
    kGammaPow = 6,         // truncated gamma on [0, m]: X = m' U^{1/k}, accept w.p. e^{-X}
    kEmpiricalAbs = 7,     // extra = v[alias draw]
    kEmpiricalRel = 8,     // extra = f[alias draw] * d
    kEmpiricalAbs1 = 9,    // single bin: extra = p0 (no draw)
    kEmpiricalRel1 = 10,   // single bin: extra = p0 * d (no draw)
};

// A compiled distribution.  64 bytes.  Parameter meaning depends on kind:
//   kGammaMT:  p0 = scale, p1 = d = k' - 1/3, p2 = c = 1/sqrt(9 d), p3 = max_scale,
//              p4 = 1/k (boost exponent, used when boost != 0)
//   kGammaPow: p0 = scale, p1 = m' = max_scale/scale, p2 = 1/k, p3 = max_scale
//   kGammaErlang2 / kGammaHalf: p0 = scale, p3 = max_scale
struct DistRec {
    int32_t kind;
    int32_t n;          // empirical: number of bins
    int32_t table_off;  // empirical: first AliasRec
    int32_t boost;      // kGammaMT: 1 when shape < 1 (sample shape+1, multiply by U^{1/k})
    double p[6];
};

// One bin of a Vose alias table with both candidate values resolved, so a
// draw needs one gather: take v_self when frac < prob, else v_alias.
struct AliasRec {
    double prob;
    double v_self;
    double v_alias;
    double pad;
};

// Philox counter tags (second counter word) so that separate draws of one
// activity never share a counter.
constexpr uint32_t kTagMain = 0u;
constexpr uint32_t kTagBoost = 1u;

// Maximum attempts of a rejection loop before the draw is clamped to the
// truncation bound.  The host only selects a rejection method whose
// acceptance probability is >= kMinAcceptance, so reaching this bound has
// probability < (1 - 1e-4)^(2^20) ~ e^-105.
constexpr uint32_t kMaxAttempts = 1u << 20;
constexpr double kMinAcceptance = 1e-4;

}  // namespace mcdp
