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
//
// Each (step, wave) owns one contiguous stream of 32-byte records: an event
// header followed by that event's predecessor records, for each of the wave's
// events of the step.  A wave walks its stream with one scalar load per
// record, the next record in flight while the current one is sampled.
#pragma once
#include <stdint.h>

namespace mcdp {

constexpr int kLanes = 64;  // wave64: one sample per lane

// Where a predecessor's realized time is read from.
enum : int32_t {
    kSrcGlobal = -1,  // from the realized output row (written in an earlier step)
    kSrcStale = -2,   // the reference reads this predecessor before it is evaluated:
                      // use its earliest time (read from the earliest array)
};

// Sampling methods.  Each registered distribution compiles to one of these.
// Parameters: p0 travels in the record; p1..p4 in the distribution table.
enum DistKind : uint32_t {
    // no draw
    kNone = 0,            // no distribution: duration = minimal_duration
    kConstant = 1,        // extra = d * p0
    kEmpiricalAbs1 = 2,   // single bin: extra = p0
    kEmpiricalRel1 = 3,   // single bin: extra = p0 * d
    // one alias-table draw
    kEmpiricalAbs = 4,    // extra = v[alias draw]; p0 holds (n, table_off) as two uint32
    kEmpiricalRel = 5,    // extra = f[alias draw] * d
    // exponential inversion
    kExpInf = 6,          // X = -p0 log u, u in (0,1]                           (mean p0)
    kExpTrunc = 7,        // X = -p0 log(p1 + p3 u), u in (0,1], clamp at p2  (p1 = e^{-m/p0}, p3 = 1 - p1)
    // gamma family
    kGammaErlang2 = 8,    // X = -p0 log(u1 u2)                                  (shape 2, scale p0)
    kGammaHalf = 9,       // X = p0 (-log u1) cos^2(2 pi u2)                     (shape 1/2, scale p0)
    kGammaMT = 10,        // Marsaglia-Tsang: scale p0, d = p1, c = p2, 1/k = p4 (boost flag)
    kGammaPow = 11,       // truncated on [0, m]: X = scale m' U^{1/k}, accept w.p. e^{-m' U^{1/k}}
};

// Record::info.
enum : uint32_t {
    kInfoKindMask = 0xFFu,       // DistKind (predecessor records; kNone in padding records)
    kInfoBoost = 1u << 8,        // gamma shape < 1: Gamma(k + 1) U^{1/k}
    kInfoTruncated = 1u << 9,    // gamma family: rejection on max_scale (table p3)
    kInfoEdgeValid = 1u << 10,   // edge in [0, link_count); otherwise duration 0.0 (kind kNone, a = 0)
    kInfoOwner = 1u << 11,       // this record writes durations[edge]
    kInfoLast = 1u << 12,        // the event's last record: its realized time is final
    kInfoDistShift = 13,         // bits 13..30: distribution table index
    kInfoDistMask = 0x3FFFFu,
    kInfoHeader = 1u << 31,      // an event header, not a predecessor
};

// A 32-byte plan record, read with one s_load_dwordx8.  Two shapes share it:
//   event header (info & kInfoHeader):
//     a = earliest, b = upper (earliest + max_delay, rounded once on the host),
//     i0 = event id, i1 = LDS slot receiving its realized time (or -1),
//     i2 = number of predecessor records that follow (kInfoLast if none);
//   predecessor (src, edge) of the event above, in the entry's input order:
//     a = minimal_duration of the edge, b = p0 of its distribution,
//     i0 = src event id (the cause_event value), i1 = src LDS slot or
//     kSrcGlobal / kSrcStale, i2 = edge (activity index: durations row and
//     RNG counter); the event's last predecessor carries kInfoLast.
struct Record {
    double a;
    double b;
    int32_t i0;
    int32_t i1;
    int32_t i2;
    uint32_t info;
};
static_assert(sizeof(Record) == 32, "Record is one s_load_dwordx8");

// A compiled distribution: kind, flags and parameters.  Records carry p[0];
// kinds that need more read this table (a scalar load, off the common path).
struct DistParams {
    uint32_t kind;   // DistKind
    uint32_t flags;  // kInfoBoost | kInfoTruncated
    double p[5];
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
