// Host side of libmcdp: the delay-generator registry, the context compiler
// that turns a DAG into a levelled, wave-scheduled device plan, and the C ABI
// declared in include/mcdp.h.
//
// Reference semantics followed here (WonJayne/mc_dagprop,
// src/mc_dagprop/monte_carlo/_core.cpp):
//   * validation order and messages: :196-201, :262-264;
//   * link table: link_count = max(idx) + 1, gaps are {0.0, type -1} (:213-229);
//   * the last precedence entry of a target supplies its predecessor list and
//     indegree, every entry adds adjacency, Kahn order with a FIFO (:232-264);
//   * per-event predecessor order is the input order (:266-285);
//   * links without a distribution keep minimal_duration (:303-306).
#include <hip/hip_runtime_api.h>

#include <fcntl.h>
#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mcdp.h"
#include "mcdp_internal.hpp"
#include "mcdp_kernels.hpp"
#include "mcdp_plan.hpp"

using namespace mcdp;

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace mcdp {
int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}
int hip_fail(hipError_t e, const char *what) {
    return fail(MCDP_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace mcdp

namespace {

// API-level distribution kinds (the reference's DistVar alternatives, :143).
enum ApiKind { kApiConstant, kApiExponential, kApiGamma, kApiEmpAbs, kApiEmpRel };

struct DistSpec {
    ApiKind kind;
    double a = 0.0, b = 0.0, c = 0.0;
    std::vector<double> values, weights;
};

// ---- regularized lower incomplete gamma P(a, x) ---------------------------
// Series  P = e^{-x} x^a / Gamma(a+1) * sum_n x^n / ((a+1)...(a+n))  for x < a+1,
// continued fraction for Q = 1 - P otherwise (Abramowitz & Stegun 6.5.29,
// 6.5.31; modified Lentz evaluation).  Used only to choose a sampling method.
double reg_lower_gamma(double a, double x) {
    if (!(x > 0.0)) return 0.0;
    if (std::isinf(x)) return 1.0;
    const double lg = std::lgamma(a);
    if (x < a + 1.0) {
        double term = 1.0 / a, sum = term, ap = a;
        for (int n = 0; n < 100000; ++n) {
            ap += 1.0;
            term *= x / ap;
            sum += term;
            if (std::fabs(term) < std::fabs(sum) * 1e-17) break;
        }
        return std::min(1.0, sum * std::exp(-x + a * std::log(x) - lg));
    }
    const double tiny = 1e-300;
    double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i < 100000; ++i) {
        const double an = -i * (i - a);
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (std::fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (std::fabs(del - 1.0) < 1e-16) break;
    }
    const double q = std::exp(-x + a * std::log(x) - lg) * h;
    return std::max(0.0, 1.0 - q);
}

// ---- Vose alias table (M. D. Vose, IEEE TSE 17(9), 1991) -------------------
// P(bin j) = (prob[j] + sum_{i: alias[i] = j} (1 - prob[i])) / n = w_j / sum(w).
// Zero-weight bins get prob 0 and are never an alias target.
void build_alias(const std::vector<double> &w, std::vector<double> &prob, std::vector<int> &alias) {
    const size_t n = w.size();
    long double total = 0.0L;
    for (double x : w) total += x;
    std::vector<long double> p(n);
    std::vector<size_t> small, large;
    size_t heaviest = 0;
    for (size_t i = 0; i < n; ++i) {
        p[i] = (long double)w[i] * (long double)n / total;
        if (w[i] > w[heaviest]) heaviest = i;
    }
    // Zero-weight bins are pushed last so they pair first (stack order).
    for (size_t i = 0; i < n; ++i)
        if (p[i] < 1.0L && w[i] != 0.0) small.push_back(i);
    for (size_t i = 0; i < n; ++i)
        if (w[i] == 0.0) small.push_back(i);
    for (size_t i = 0; i < n; ++i)
        if (p[i] >= 1.0L) large.push_back(i);
    prob.assign(n, 1.0);
    alias.resize(n);
    for (size_t i = 0; i < n; ++i) alias[i] = (int)i;
    while (!small.empty() && !large.empty()) {
        const size_t l = small.back();
        small.pop_back();
        const size_t g = large.back();
        large.pop_back();
        prob[l] = (double)p[l];
        alias[l] = (int)g;
        p[g] = (p[g] + p[l]) - 1.0L;
        if (p[g] < 1.0L)
            small.push_back(g);
        else
            large.push_back(g);
    }
    for (size_t g : large) prob[g] = 1.0;
    for (size_t l : small) {  // rounding leftovers: p[l] is 1 up to rounding
        if (w[l] == 0.0) {
            prob[l] = 0.0;
            alias[l] = (int)heaviest;
        } else {
            prob[l] = 1.0;
        }
    }
}

bool all_finite_nonneg(const std::vector<double> &w, double &sum) {
    sum = 0.0;
    for (double x : w) {
        if (!std::isfinite(x) || x < 0.0) return false;
        sum += x;
    }
    return std::isfinite(sum) && sum > 0.0;
}

int compile_exponential(double lam, double m, DistRec &D) {
    if (std::isnan(lam) || lam < 0.0) return fail(MCDP_E_INVALID, "exponential lambda_ must be non-negative");
    if (std::isnan(m)) m = std::numeric_limits<double>::infinity();  // x > NaN never rejects (:86)
    if (m < 0.0) return fail(MCDP_E_INVALID, "exponential max_scale must be non-negative");
    if (lam == 0.0 || m == 0.0) {  // X == 0 always
        D.kind = kConstant;
        D.p[0] = 0.0;
    } else if (std::isinf(m)) {
        D.kind = kExpInf;
        D.p[0] = lam;
    } else {
        D.kind = kExpTrunc;
        D.p[0] = lam;
        D.p[1] = std::exp(-m / lam);     // q
        D.p[2] = m;
        D.p[3] = -std::expm1(-m / lam);  // 1 - q
    }
    return MCDP_OK;
}

int compile_gamma(double k, double th, double m, DistRec &D) {
    if (std::isnan(k) || !(k > 0.0) || std::isinf(k)) return fail(MCDP_E_INVALID, "gamma shape must be positive and finite");
    if (std::isnan(th) || !(th > 0.0) || std::isinf(th)) return fail(MCDP_E_INVALID, "gamma scale must be positive and finite");
    if (std::isnan(m)) m = std::numeric_limits<double>::infinity();
    if (m < 0.0) return fail(MCDP_E_INVALID, "gamma max_scale must be non-negative");
    if (m == 0.0) {
        D.kind = kConstant;
        D.p[0] = 0.0;
        return MCDP_OK;
    }
    if (k == 1.0) return compile_exponential(th, m, D);  // Gamma(1, th) = Exp(mean th)
    const double mp = m / th;
    const double p_rej = std::isinf(m) ? 1.0 : reg_lower_gamma(k, mp);
    double p_pow = 0.0;
    if (!std::isinf(m)) {
        const double lp = std::lgamma(k + 1.0) + std::log(std::max(p_rej, 1e-300)) - k * std::log(mp);
        p_pow = std::min(1.0, std::exp(lp));
    }
    if (p_rej >= 0.25 || p_rej >= p_pow) {
        if (p_rej < kMinAcceptance) return fail(MCDP_E_INVALID, "gamma max_scale truncation too severe to sample");
        D.p[0] = th;
        D.p[3] = m;
        D.flags = std::isinf(m) ? 0u : kInfoTruncated;
        if (k == 2.0) {
            D.kind = kGammaErlang2;
        } else if (k == 0.5) {
            D.kind = kGammaHalf;
        } else {
            const double kp = k < 1.0 ? k + 1.0 : k;
            D.kind = kGammaMT;
            D.flags |= k < 1.0 ? kInfoBoost : 0u;
            D.p[1] = kp - 1.0 / 3.0;
            D.p[2] = 1.0 / std::sqrt(9.0 * D.p[1]);
            D.p[4] = 1.0 / k;
        }
    } else {
        if (p_pow < kMinAcceptance) return fail(MCDP_E_INVALID, "gamma max_scale truncation too severe to sample");
        D.kind = kGammaPow;
        D.flags = kInfoTruncated;
        D.p[0] = th;
        D.p[1] = mp;
        D.p[2] = 1.0 / k;
        D.p[3] = m;
    }
    return MCDP_OK;
}

int compile_empirical(const DistSpec &s, DistRec &D, std::vector<AliasRec> &table) {
    const size_t n = s.values.size();
    const bool abs = s.kind == kApiEmpAbs;
    if (n == 0) return fail(MCDP_E_INVALID, "empirical distribution needs at least one value");
    if (n == 1) {  // libstdc++ discrete_distribution with < 2 weights always yields 0
        D.kind = abs ? kEmpiricalAbs1 : kEmpiricalRel1;
        D.p[0] = s.values[0];
        return MCDP_OK;
    }
    double sum;
    if (!all_finite_nonneg(s.weights, sum))
        return fail(MCDP_E_INVALID, "empirical weights must be finite, non-negative and not all zero");
    if (n > (size_t)std::numeric_limits<int32_t>::max() / 2) return fail(MCDP_E_INVALID, "empirical table too large");
    std::vector<double> prob;
    std::vector<int> alias;
    build_alias(s.weights, prob, alias);
    if (table.size() + n > (size_t)std::numeric_limits<int32_t>::max()) return fail(MCDP_E_INVALID, "empirical tables too large");
    D.kind = abs ? kEmpiricalAbs : kEmpiricalRel;
    // p0 carries (n, first alias record) as two uint32 (low, high).
    const uint64_t packed = (uint64_t)(uint32_t)n | ((uint64_t)(uint32_t)table.size() << 32);
    std::memcpy(&D.p[0], &packed, sizeof(double));
    for (size_t i = 0; i < n; ++i) table.push_back(AliasRec{prob[i], s.values[i], s.values[(size_t)alias[i]], 0.0});
    return MCDP_OK;
}

int compile_dist(const DistSpec &s, DistRec &D, std::vector<AliasRec> &table) {
    D = DistRec{};
    D.kind = kNone;
    switch (s.kind) {
        case kApiConstant:
            D.kind = kConstant;
            D.p[0] = s.a;
            return MCDP_OK;
        case kApiExponential: return compile_exponential(s.a, s.b, D);
        case kApiGamma: return compile_gamma(s.a, s.b, s.c, D);
        default: return compile_empirical(s, D, table);
    }
}

// Relative cost of one draw (VALU instructions, measured per method), for
// balancing waves within a step.
int draw_cost(const DistRec *D) {
    if (!D) return 2;
    switch (D->kind) {
        case kConstant:
        case kEmpiricalAbs1:
        case kEmpiricalRel1: return 5;
        case kEmpiricalAbs:
        case kEmpiricalRel: return 45;
        case kExpInf: return 75;
        case kExpTrunc: return 80;
        case kGammaErlang2: return (D->flags & kInfoTruncated) ? 160 : 110;
        case kGammaHalf: return (D->flags & kInfoTruncated) ? 180 : 130;
        default: return 250;
    }
}
constexpr int kPredCost = 20;   // propagation work per predecessor record
constexpr int kEventCost = 15;  // per event: header, final cap, stores
// A wide step's launch gap and predecessor loads, in the same units
// (calibrated on the network workload: a 200-step tail run alone costs about
// 12 ns per unit either way).
constexpr int kWideStepCost = 300;
// Wide steps for a partial round of at most this many groups of 64 samples
// (each step launches events x groups single-wave workgroups; up to 128
// groups they were measured to end inside the tiles' time on the network and
// the small workloads).
constexpr int64_t kWideMaxGroups = 128;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

struct mcdp_generator {
    std::map<int32_t, DistSpec> dists;  // last add_* per type wins
};


namespace {

constexpr size_t kLdsBudgetBytes = 64 * 1024;  // LDS per workgroup for predecessor slots (one sample per lane)
// Sample indices and leading dimensions are 32-bit in the kernel; tiles of up
// to 128 samples must not overflow them.
constexpr int64_t kMaxSamplesPerLaunch = ((int64_t)1 << 31) - 256;

// A predecessor record for edge l: its minimal duration, sampling kind and p0
// inline; the distribution index for the kinds that read more parameters.
Record edge_record(const mcdp_propagator &P, int32_t l) {
    Record r{};
    r.i2 = l;
    if (l >= 0 && l < P.L) {
        r.a = P.dmin[(size_t)l];
        r.info = kInfoEdgeValid;
        const uint32_t di = P.ldist[(size_t)l];
        if (di == kNoDistIndex) {
            r.info |= kNone;
        } else {
            const DistRec &D = P.dist[di];
            r.info |= D.kind | D.flags | (di << kInfoDistShift);
            r.b = D.p[0];
        }
    } else {
        r.info = kNone;  // undefined in the reference: contributes 0.0
    }
    return r;
}

// Steps, wave assignment, LDS slots and record streams for W waves per
// workgroup and at most K events per wave and step.
int build_schedule(mcdp_propagator &P, int32_t W, int32_t K) {
    const int32_t E = P.E, L = P.L;
    P.waves = W;
    P.k = K;
    P.step_off.assign(1, 0);
    P.ev_off.assign(1, 0);
    P.slot_list.clear();
    P.ev_rec.clear();
    P.recs.clear();
    P.orph.clear();
    P.max_per_wave = 0;
    P.n_global = 0;
    P.tile_cost = 0;
    P.wide_cost = 0;

    auto event_cost = [&](int32_t e) {
        int c = kEventCost;
        for (int64_t k = P.pred_off[(size_t)e]; k < P.pred_off[(size_t)e + 1]; ++k) {
            const int32_t l = P.preds[(size_t)k].edge;
            const DistRec *D = (l >= 0 && l < L && P.ldist[(size_t)l] != kNoDistIndex) ? &P.dist[P.ldist[(size_t)l]] : nullptr;
            c += kPredCost + draw_cost(D);
        }
        return c;
    };

    // (a) steps: each level in as few steps as W x K events allow, split evenly;
    // within a step, events go to the least-loaded wave (longest first).
    std::vector<int32_t> ev_step((size_t)E, -1);
    std::vector<int32_t> flat_events;  // stream order
    int32_t step = 0;
    for (const auto &lev : P.levels) {
        if (lev.empty()) continue;
        const size_t cap = (size_t)W * (size_t)K;
        const size_t nsteps = (lev.size() + cap - 1) / cap;
        for (size_t c = 0; c < nsteps; ++c) {
            const size_t c0 = lev.size() * c / nsteps, c1 = lev.size() * (c + 1) / nsteps;
            std::vector<std::pair<int, int32_t>> items;
            for (size_t i = c0; i < c1; ++i) items.emplace_back(event_cost(lev[i]), lev[i]);
            std::stable_sort(items.begin(), items.end(), [](auto &a, auto &b) { return a.first > b.first; });
            std::vector<std::vector<int32_t>> per((size_t)W);
            std::vector<int64_t> load((size_t)W, 0);
            for (auto &it : items) {
                int best = -1;
                for (int w = 0; w < W; ++w)
                    if ((int)per[(size_t)w].size() < K && (best < 0 || load[(size_t)w] < load[(size_t)best])) best = w;
                per[(size_t)best].push_back(it.second);
                load[(size_t)best] += it.first;
            }
            for (int w = 0; w < W; ++w) {
                for (int32_t e : per[(size_t)w]) {
                    flat_events.push_back(e);
                    ev_step[(size_t)e] = step;
                }
                P.max_per_wave = std::max<int32_t>(P.max_per_wave, (int32_t)per[(size_t)w].size());
                P.ev_off.push_back((int32_t)flat_events.size());
            }
            ++step;
        }
    }
    P.n_steps = step;

    // (b) LDS slots: an event read by a later step occupies a slot from the
    // commit of its step to the end of the last step that reads it.
    std::vector<int32_t> last_use((size_t)E, -1);
    for (int32_t e = 0; e < E; ++e)
        for (int64_t k = P.pred_off[(size_t)e]; k < P.pred_off[(size_t)e + 1]; ++k) {
            const PredEntry &pe = P.preds[(size_t)k];
            if (!pe.stale) last_use[(size_t)pe.src] = std::max(last_use[(size_t)pe.src], ev_step[(size_t)e]);
        }
    const int32_t slot_cap = (int32_t)(kLdsBudgetBytes / (kLanes * sizeof(double)));
    std::vector<int32_t> slot_of((size_t)E, -1);
    {
        std::vector<std::vector<int32_t>> release((size_t)std::max(P.n_steps, 1));
        std::vector<int32_t> free_slots;
        int32_t used = 0;
        for (int32_t t = 0; t < P.n_steps; ++t) {
            for (int32_t e : release[(size_t)t]) free_slots.push_back(slot_of[(size_t)e]);
            const int32_t b = P.ev_off[(size_t)t * W], en = P.ev_off[(size_t)(t + 1) * W];
            for (int32_t i = b; i < en; ++i) {
                const int32_t e = flat_events[(size_t)i];
                if (last_use[(size_t)e] < 0) continue;
                int32_t s = -1;
                if (!free_slots.empty()) {
                    s = free_slots.back();
                    free_slots.pop_back();
                } else if (used < slot_cap) {
                    s = used++;
                }
                slot_of[(size_t)e] = s;
                if (s >= 0) release[(size_t)last_use[(size_t)e]].push_back(e);
            }
        }
        P.slots = used;
    }

    // (c) record streams: for each (step, wave), each event's header then its
    // predecessors in input order.  The first predecessor record of an edge
    // writes its duration row.
    std::vector<uint8_t> referenced((size_t)L, 0);
    P.recs.reserve(flat_events.size() + P.preds.size() + 1);
    P.step_off.assign(1, 0);
    for (size_t g = 0; g + 1 < P.ev_off.size(); ++g) {
        for (int32_t i = P.ev_off[g]; i < P.ev_off[g + 1]; ++i) {
            const int32_t e = flat_events[(size_t)i];
            P.slot_list.push_back(slot_of[(size_t)e]);
            Record h{};
            h.a = P.earliest[(size_t)e];
            h.b = P.earliest[(size_t)e] + P.max_delay;
            h.i0 = e;
            h.i1 = slot_of[(size_t)e];
            h.i2 = (int32_t)(P.pred_off[(size_t)e + 1] - P.pred_off[(size_t)e]);
            h.info = kInfoHeader | (h.i2 == 0 ? kInfoLast : 0u);
            P.ev_rec.push_back((int32_t)P.recs.size());
            P.recs.push_back(h);
            for (int64_t k = P.pred_off[(size_t)e]; k < P.pred_off[(size_t)e + 1]; ++k) {
                const PredEntry &pe = P.preds[(size_t)k];
                Record r = edge_record(P, pe.edge);
                r.i0 = pe.src;
                if (pe.stale) {
                    r.i1 = kSrcStale;
                } else {
                    r.i1 = slot_of[(size_t)pe.src] >= 0 ? slot_of[(size_t)pe.src] : kSrcGlobal;
                    if (r.i1 == kSrcGlobal) ++P.n_global;
                }
                if (pe.edge >= 0 && pe.edge < L && !referenced[(size_t)pe.edge]) {
                    referenced[(size_t)pe.edge] = 1;
                    r.info |= kInfoOwner;
                }
                if (k + 1 == P.pred_off[(size_t)e + 1]) r.info |= kInfoLast;
                P.recs.push_back(r);
            }
        }
        P.step_off.push_back((int32_t)P.recs.size());
    }
    // padding: the walk kernel prefetches one record ahead, a wide step loads
    // kWideBatch predecessor records at a time
    for (int i = 0; i < kWideBatch; ++i) P.recs.push_back(Record{});
    if (P.recs.size() > (size_t)std::numeric_limits<int32_t>::max()) return fail(MCDP_E_INVALID, "plan too large");
    for (int32_t t = 0; t < P.n_steps; ++t) {
        int64_t busiest = 0, costliest = 0;
        for (int32_t w = 0; w < W; ++w) {
            int64_t c = 0;
            for (int32_t i = P.ev_off[(size_t)t * W + w]; i < P.ev_off[(size_t)t * W + w + 1]; ++i) {
                const int64_t ce = event_cost(flat_events[(size_t)i]);
                c += ce;
                costliest = std::max(costliest, ce);
            }
            busiest = std::max(busiest, c);
        }
        P.tile_cost += busiest;
        P.wide_cost += kWideStepCost + costliest;
    }
    for (int32_t l = 0; l < L; ++l)
        if (!referenced[(size_t)l]) P.orph.push_back(edge_record(P, l));
    return MCDP_OK;
}

// Default (waves, events per wave) for a DAG: enough waves to spread the widest
// level (up to 8), and the smallest compiled bound that holds it in one step.
void default_shape(const mcdp_propagator &P, int32_t &W, int32_t &K) {
    size_t max_width = 0;
    for (const auto &v : P.levels) max_width = std::max(max_width, v.size());
    W = 1;
    while (W < 8 && (size_t)W * 4 < max_width) W *= 2;
    K = kMaxEventsPerWave;
    for (int k : kEventsPerWaveChoices)
        if ((size_t)k * (size_t)W >= max_width) {
            K = k;
            break;
        }
}

int compile_propagator(mcdp_propagator &P, int32_t n_events, const double *earliest, int32_t n_acts,
                       const int32_t *act_idx, const double *act_dur, const int32_t *act_type, int32_t n_entries,
                       const int32_t *entry_target, const int64_t *entry_off, const int32_t *pred_event,
                       const int32_t *pred_act, double max_delay, const mcdp_generator &gen) {
    // (1) validation, in the reference's order.
    if (gen.dists.count(-1)) return fail(MCDP_E_RESERVED_TYPE, "Activity type -1 is reserved for no delay");
    if (max_delay < 0.0) return fail(MCDP_E_MAX_DELAY, "max_delay must be non-negative");
    if (n_events < 0 || n_acts < 0 || n_entries < 0) return fail(MCDP_E_INVALID, "negative count");
    const int32_t E = n_events;
    P.max_delay = max_delay;

    // (2) distributions.
    std::map<int32_t, int32_t> type_to_dist;
    for (const auto &kv : gen.dists) {
        DistRec D;
        int rc = compile_dist(kv.second, D, P.alias);
        if (rc) return fail(rc, g_last_error + " (activity_type " + std::to_string(kv.first) + ")");
        type_to_dist[kv.first] = (int32_t)P.dist.size();
        P.dist.push_back(D);
    }
    if (P.dist.size() > (size_t)kInfoDistMask) return fail(MCDP_E_INVALID, "too many distributions");

    // (3) link table.
    int64_t max_idx = -1;
    for (int32_t i = 0; i < n_acts; ++i) {
        if (act_idx[i] < 0) return fail(MCDP_E_INVALID, "activity index must be non-negative");
        max_idx = std::max<int64_t>(max_idx, act_idx[i]);
    }
    const int32_t L = (int32_t)(max_idx + 1);
    P.dmin.assign((size_t)L, 0.0);
    std::vector<int32_t> ltype((size_t)L, -1);
    std::vector<uint8_t> seen((size_t)L, 0);
    for (int32_t i = 0; i < n_acts; ++i) {
        const int32_t l = act_idx[i];
        if (seen[(size_t)l] &&
            (std::memcmp(&P.dmin[(size_t)l], &act_dur[i], sizeof(double)) != 0 || ltype[(size_t)l] != act_type[i]))
            return fail(MCDP_E_INVALID, "two activities share index " + std::to_string(l) + " with different attributes");
        seen[(size_t)l] = 1;
        P.dmin[(size_t)l] = act_dur[i];
        ltype[(size_t)l] = act_type[i];
    }
    P.ldist.assign((size_t)L, kNoDistIndex);
    for (int32_t l = 0; l < L; ++l) {
        auto it = type_to_dist.find(ltype[(size_t)l]);
        if (it != type_to_dist.end()) P.ldist[(size_t)l] = (uint32_t)it->second;
    }

    // (4) topology.
    if (n_entries > 0 && entry_off[0] != 0) return fail(MCDP_E_INVALID, "entry_offsets[0] must be 0");
    for (int32_t i = 0; i < n_entries; ++i) {
        if (entry_off[i + 1] < entry_off[i]) return fail(MCDP_E_INVALID, "entry_offsets must be non-decreasing");
        if (entry_target[i] < 0 || entry_target[i] >= E) return fail(MCDP_E_INVALID, "precedence target out of range");
        for (int64_t k = entry_off[i]; k < entry_off[i + 1]; ++k)
            if (pred_event[k] < 0 || pred_event[k] >= E) return fail(MCDP_E_INVALID, "predecessor event out of range");
    }
    std::vector<int32_t> win((size_t)E, -1), indeg((size_t)E, 0);
    std::vector<int64_t> adj_off((size_t)E + 1, 0);
    for (int32_t i = 0; i < n_entries; ++i) {
        win[(size_t)entry_target[i]] = i;
        indeg[(size_t)entry_target[i]] = (int32_t)(entry_off[i + 1] - entry_off[i]);
        for (int64_t k = entry_off[i]; k < entry_off[i + 1]; ++k) adj_off[(size_t)pred_event[k] + 1]++;
    }
    for (int32_t i = 0; i < E; ++i) adj_off[(size_t)i + 1] += adj_off[(size_t)i];
    std::vector<int32_t> adj((size_t)adj_off[(size_t)E]);
    {
        std::vector<int64_t> fill(adj_off.begin(), adj_off.end() - 1);
        for (int32_t i = 0; i < n_entries; ++i)
            for (int64_t k = entry_off[i]; k < entry_off[i + 1]; ++k)
                adj[(size_t)fill[(size_t)pred_event[k]]++] = entry_target[i];
    }
    std::vector<int32_t> order;
    order.reserve((size_t)E);
    {
        std::vector<int32_t> q;
        q.reserve((size_t)E);
        for (int32_t i = 0; i < E; ++i)
            if (indeg[(size_t)i] == 0) q.push_back(i);
        size_t head = 0;
        while (head < q.size()) {
            const int32_t n = q[head++];
            order.push_back(n);
            for (int64_t k = adj_off[(size_t)n]; k < adj_off[(size_t)n + 1]; ++k)
                if (--indeg[(size_t)adj[(size_t)k]] == 0) q.push_back(adj[(size_t)k]);
        }
    }
    if ((int32_t)order.size() != E) return fail(MCDP_E_CYCLE, "Invalid DAG: cycle detected in precedence list");

    // (5) predecessor lists and levels.  A predecessor evaluated after its
    // consumer in the reference's order is read before it is written there,
    // i.e. at its earliest time ("stale"); every other predecessor is a true
    // dependency.
    std::vector<int32_t> pos((size_t)E), level((size_t)E, 0);
    for (int32_t i = 0; i < E; ++i) pos[(size_t)order[(size_t)i]] = i;
    P.pred_off.assign((size_t)E + 1, 0);
    for (int32_t e = 0; e < E; ++e) {
        const int32_t w = win[(size_t)e];
        P.pred_off[(size_t)e + 1] = P.pred_off[(size_t)e] + (w >= 0 ? entry_off[w + 1] - entry_off[w] : 0);
    }
    P.preds.resize((size_t)P.pred_off[(size_t)E]);
    for (int32_t e = 0; e < E; ++e) {
        const int32_t w = win[(size_t)e];
        if (w < 0) continue;
        int64_t o = P.pred_off[(size_t)e];
        for (int64_t k = entry_off[w]; k < entry_off[w + 1]; ++k, ++o) {
            const int32_t s = pred_event[k];
            const bool stale = pos[(size_t)s] >= pos[(size_t)e];
            P.preds[(size_t)o] = PredEntry{s, pred_act[k], stale};
            if (stale) ++P.n_stale;
        }
    }
    int32_t n_levels = E > 0 ? 1 : 0;
    for (int32_t e : order) {
        int32_t lv = 0;
        for (int64_t k = P.pred_off[(size_t)e]; k < P.pred_off[(size_t)e + 1]; ++k)
            if (!P.preds[(size_t)k].stale) lv = std::max(lv, level[(size_t)P.preds[(size_t)k].src] + 1);
        level[(size_t)e] = lv;
        n_levels = std::max(n_levels, lv + 1);
    }
    P.levels.assign((size_t)n_levels, {});
    for (int32_t e : order) P.levels[(size_t)level[(size_t)e]].push_back(e);

    P.E = E;
    P.L = L;
    P.earliest.assign(earliest, earliest + E);

    // (6) default schedule.
    int32_t W, K;
    default_shape(P, W, K);
    return build_schedule(P, W, K);
}

// Uploads the plan of `P` to the current device (once per device).
int ensure_device_impl(mcdp_propagator &P, DeviceCopy *&out) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MCDP_E_DEVICE, "no HIP device available: the Monte Carlo engine runs on the GPU only");
    int device = 0;
    HIP_TRY(hipGetDevice(&device), "hipGetDevice");
    for (auto &d : P.dev)
        if (d.device == device) {
            out = &d;
            return MCDP_OK;
        }
    const size_t a = 256;
    size_t off = 0;
    auto reserve = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, a);
        return o;
    };
    const size_t o_step = reserve(P.step_off.size() * sizeof(int32_t));
    const size_t o_recs = reserve(P.recs.size() * sizeof(Record));
    const size_t o_evoff = reserve(P.ev_off.size() * sizeof(int32_t));
    const size_t o_slots = reserve(P.slot_list.size() * sizeof(int32_t));
    // each (step, wave)'s slots as one row of k, -1 past its events
    std::vector<int32_t> step_slots((size_t)P.n_steps * P.waves * P.k, -1);
    for (size_t g = 0; g + 1 < P.ev_off.size(); ++g)
        for (int32_t i = P.ev_off[g]; i < P.ev_off[g + 1]; ++i)
            step_slots[g * (size_t)P.k + (size_t)(i - P.ev_off[g])] = P.slot_list[(size_t)i];
    const size_t o_sslots = reserve(step_slots.size() * sizeof(int32_t));
    const size_t o_orph = reserve(P.orph.size() * sizeof(Record));
    const size_t o_dist = reserve(P.dist.size() * sizeof(DistParams));
    const size_t o_alias = reserve(P.alias.size() * sizeof(AliasRec));
    const size_t o_earl = reserve(P.earliest.size() * sizeof(double));
    const size_t o_evrec = reserve(P.ev_rec.size() * sizeof(int32_t));
    const size_t total = std::max<size_t>(off, a);
    std::vector<unsigned char> host(total, 0);
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (bytes) std::memcpy(host.data() + o, src, bytes);
    };
    put(o_step, P.step_off.data(), P.step_off.size() * sizeof(int32_t));
    put(o_recs, P.recs.data(), P.recs.size() * sizeof(Record));
    put(o_evoff, P.ev_off.data(), P.ev_off.size() * sizeof(int32_t));
    put(o_slots, P.slot_list.data(), P.slot_list.size() * sizeof(int32_t));
    put(o_sslots, step_slots.data(), step_slots.size() * sizeof(int32_t));
    put(o_orph, P.orph.data(), P.orph.size() * sizeof(Record));
    put(o_dist, P.dist.data(), P.dist.size() * sizeof(DistParams));
    put(o_alias, P.alias.data(), P.alias.size() * sizeof(AliasRec));
    put(o_earl, P.earliest.data(), P.earliest.size() * sizeof(double));
    put(o_evrec, P.ev_rec.data(), P.ev_rec.size() * sizeof(int32_t));
    void *blob = nullptr;
    hipError_t e = hipMalloc(&blob, total);
    if (e != hipSuccess) return fail(MCDP_E_NOMEM, std::string("hipMalloc (plan): ") + hipGetErrorString(e));
    e = hipMemcpy(blob, host.data(), total, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(blob);
        return hip_fail(e, "hipMemcpy (plan)");
    }
    DeviceCopy d;
    d.device = device;
    d.blob = blob;
    // the side stream of wide steps and its fork / join events, made here so
    // that a launch creates nothing (it may be under stream capture)
    e = hipStreamCreateWithFlags(&d.side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&d.fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&d.join, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (d.side) (void)hipStreamDestroy(d.side);
        if (d.fork) (void)hipEventDestroy(d.fork);
        (void)hipFree(blob);
        return hip_fail(e, "side stream");
    }
    auto *b = static_cast<unsigned char *>(blob);
    d.base.step_off = reinterpret_cast<const int32_t *>(b + o_step);
    d.base.recs = reinterpret_cast<const Record *>(b + o_recs);
    d.base.ev_off = reinterpret_cast<const int32_t *>(b + o_evoff);
    d.base.slot_list = reinterpret_cast<const int32_t *>(b + o_slots);
    d.base.step_slots = reinterpret_cast<const int32_t *>(b + o_sslots);
    d.base.orph = reinterpret_cast<const Record *>(b + o_orph);
    d.base.dists = reinterpret_cast<const DistParams *>(b + o_dist);
    d.base.alias = reinterpret_cast<const AliasRec *>(b + o_alias);
    d.base.earliest = reinterpret_cast<const double *>(b + o_earl);
    d.base.ev_rec = reinterpret_cast<const int32_t *>(b + o_evrec);
    d.base.n_steps = P.n_steps;
    d.base.n_orphans = (int32_t)P.orph.size();
    d.base.waves = P.waves;
    d.base.n_alias = (int32_t)P.alias.size();
    d.base.slots = P.slots;
    P.dev.push_back(d);
    out = &P.dev.back();
    return MCDP_OK;
}
// The fast kernel variant needs every read predecessor in LDS (or stale) and
// the alias tables small enough to stage in LDS.
WalkVariant walk_variant(const mcdp_propagator &P, bool injected) {
    const bool alias_fits = P.alias.size() * sizeof(AliasRec) <= kAliasLdsMaxBytes;
    const bool fast = P.n_global == 0 && (injected || alias_fits);
    return WalkVariant{injected, fast ? kPathsFast : kPathsGeneral, P.k, P.spl};
}

size_t walk_lds_bytes(const mcdp_propagator &P, WalkVariant v) {
    size_t bytes = (size_t)P.slots * kLanes * (size_t)v.spl * sizeof(double);
    if (!v.injected && v.paths == kPathsFast) bytes += P.alias.size() * sizeof(AliasRec);
    return bytes;
}

// A schedule shape: waves per workgroup, events per wave and step, samples
// per lane.
struct Shape {
    int32_t waves, k, spl;
};

bool shape_valid(int32_t waves, int32_t k, int32_t spl) {
    if (waves < 1 || waves > kMaxWaves) return false;
    if (spl != 1 && spl != 2) return false;
    bool k_ok = false;
    for (int c : kEventsPerWaveChoices) k_ok |= c == k;
    return k_ok && (spl == 1 || k <= kMaxEventsPerWaveSpl2);
}

// Shapes worth trying: every valid shape whose step holds the widest level
// (or that is the largest shape available).
std::vector<Shape> shape_candidates(const mcdp_propagator &P) {
    size_t max_width = 1;
    for (const auto &v : P.levels) max_width = std::max(max_width, v.size());
    static const int32_t kWaveChoices[] = {1, 2, 4, 6, 8, 10, 12, 14, 16};
    std::vector<Shape> out;
    for (int spl : kSplChoices) {
        const int kmax = spl == 1 ? kMaxEventsPerWave : kMaxEventsPerWaveSpl2;
        const size_t need = std::min<size_t>(max_width, (size_t)kMaxWaves * (size_t)kmax);
        for (int k : kEventsPerWaveChoices)
            for (int32_t w : kWaveChoices)
                if (shape_valid(w, k, spl) && (size_t)w * (size_t)k >= need) out.push_back(Shape{w, k, spl});
    }
    return out;
}

// Picks a shape for the current device without running anything: the most
// samples resident per CU (workgroups per CU x 64 x samples per lane), then
// two samples per lane (less scalar work per sample), then the least work on
// the busiest wave of each step (events x samples per lane, summed over
// steps), then more waves.  Candidate schedules are built on the host; P is
// left with the winner.
int choose_shape(mcdp_propagator &P) {
    Shape best{P.waves, P.k, 1};
    int64_t best_res = -1, best_load = 0;
    for (const Shape &c : shape_candidates(P)) {
        int rc = build_schedule(P, c.waves, c.k);
        if (rc) return rc;
        P.spl = c.spl;
        const WalkVariant v = walk_variant(P, false);
        int bpc = 0;
        if (walk_occupancy(v, P.waves, walk_lds_bytes(P, v), &bpc) != hipSuccess || bpc < 1) continue;
        const int64_t res = (int64_t)bpc * kLanes * c.spl;
        int64_t load = 0;
        for (int32_t t = 0; t < P.n_steps; ++t) {
            int32_t m = 0;
            for (int32_t w = 0; w < P.waves; ++w)
                m = std::max(m, P.ev_off[(size_t)t * P.waves + w + 1] - P.ev_off[(size_t)t * P.waves + w]);
            load += (int64_t)m * c.spl;
        }
        const bool better =
            res > best_res ||
            (res == best_res &&
             (c.spl > best.spl ||
              (c.spl == best.spl && (load < best_load || (load == best_load && c.waves > best.waves)))));
        if (better) {
            best_res = res;
            best_load = load;
            best = c;
        }
    }
    P.spl = best.spl;
    return build_schedule(P, best.waves, best.k);
}

// cause_pos (sampled walks, no cause): the winning entry's position per
// sample instead of the cause event.
int launch(mcdp_propagator &P, bool injected, const int32_t *seeds, const double *dur_in, int64_t n, int64_t ld,
           double *realized, double *durations, int32_t *cause, hipStream_t stream, uint8_t *cause_pos = nullptr) {
    if (n < 0) return fail(MCDP_E_INVALID, "n must be non-negative");
    if (cause_pos && (injected || cause)) return fail(MCDP_E_INVALID, "internal error: positions replace sampled causes");
    if (n > 0 && ld < n) return fail(MCDP_E_INVALID, "ld must be >= n");
    if (n > kMaxSamplesPerLaunch || ld > std::numeric_limits<int32_t>::max())
        return fail(MCDP_E_INVALID, "n and ld must be below 2^31 - 256 per call");
    if (injected && !dur_in && P.L > 0 && n > 0) return fail(MCDP_E_INVALID, "durations input is null");
    if (!injected && !seeds && n > 0) return fail(MCDP_E_INVALID, "seeds is null");
    if (!realized && P.E > 0 && n > 0) return fail(MCDP_E_INVALID, "realized must not be null");
    std::lock_guard<std::mutex> lock(P.mu);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MCDP_E_DEVICE, "no HIP device available: the Monte Carlo engine runs on the GPU only");
    if (!P.shape_fixed) {  // first launch: fit the schedule to this device
        int rc = choose_shape(P);
        if (rc) return rc;
        P.shape_fixed = true;
    }
    DeviceCopy *d = nullptr;
    int rc = ensure_device_impl(P, d);
    if (rc) return rc;
    if (n == 0 || (P.E == 0 && (injected || P.orph.empty()))) return MCDP_OK;
    WalkParams p = d->base;
    p.seeds = seeds;
    p.dur_in = dur_in;
    p.ld = (int32_t)ld;
    p.realized = realized;
    p.durations = injected ? nullptr : durations;
    p.cause = cause;
    p.cause_pos = cause_pos;
    WalkVariant v = walk_variant(P, injected);
    v.pos = cause_pos != nullptr;
    const size_t lds = walk_lds_bytes(P, v);
    int cus = 0, bpc = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->device), "hipDeviceGetAttribute");
    HIP_TRY(walk_occupancy(v, P.waves, lds, &bpc), "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (bpc < 1) return fail(MCDP_E_DEVICE, "walk kernel cannot be resident with " + std::to_string(lds) + " bytes of LDS");
    const int64_t resident = (int64_t)cus * bpc;
    const int64_t tile = (int64_t)kLanes * v.spl;

    // A tile's run time hardly depends on how many of its lanes are live, so a
    // last round of only a few tiles costs almost a full round.  The samples
    // past the last full round take one of two cheaper routes:
    //   * wide steps: one launch per step, each event of the step a wave of
    //     its own for every group of 64 samples, predecessors read from the
    //     output rows; the tail's latency is then about the steps' launch gaps
    //     plus their costliest events, instead of a whole tile's walk;
    //   * with two samples per lane, the one-sample-per-lane kernel of the
    //     same schedule, whose tiles finish sooner.
    // The wide steps run beside the full rounds' tiles (below), so they cost
    // little as long as they end before the tiles do: they are taken when
    // their estimated latency (from the schedule) is below that of the full
    // rounds.
    // Results do not depend on the split or the route.
    int64_t n_main = n, n_tail = 0;
    bool wide = false;
    WalkVariant v1 = v;
    v1.spl = 1;
    const size_t lds1 = walk_lds_bytes(P, v1);
    int64_t resident1 = 0;
    {
        const int64_t full_rounds = ((n + tile - 1) / tile) / resident;
        const int64_t rest = n - full_rounds * resident * tile;
        const int64_t groups = (rest + kLanes - 1) / kLanes;
        if (full_rounds == 0 && rest > 0) {
            // less than one round: the wide steps alone against one round of
            // tiles that hardly fills the device
            if (P.tail_mode == 2 || (P.tail_mode == 0 && groups <= kWideMaxGroups && P.wide_cost < P.tile_cost)) {
                wide = true;
                n_main = 0;
                n_tail = n;
            }
        } else if (full_rounds >= 1 && rest > 0) {
            if (P.tail_mode == 2 ||
                (P.tail_mode == 0 && groups <= kWideMaxGroups && P.wide_cost < full_rounds * P.tile_cost)) {
                wide = true;
                n_main = n - rest;
                n_tail = rest;
            } else if (v.spl == 2) {
                int bpc1 = 0;
                HIP_TRY(walk_occupancy(v1, P.waves, lds1, &bpc1), "hipOccupancyMaxActiveBlocksPerMultiprocessor");
                resident1 = (int64_t)cus * bpc1;
                if (bpc1 >= 1 && groups <= resident1) {
                    n_main = n - rest;
                    n_tail = rest;
                }
            }
        }
    }
    p.n = (int32_t)n_main;
    p.n_tiles = (int32_t)((n_main + tile - 1) / tile);
    // The wide steps need no LDS and fewer registers than a tile leaves free
    // on a SIMD, so they run on a second stream beside the tiles, forked from
    // the caller's stream before the tiles and joined after.
    hipStream_t tail_stream = stream;
    if (wide && n_main > 0) {
        HIP_TRY(hipEventRecord(d->fork, stream), "hipEventRecord");
        HIP_TRY(hipStreamWaitEvent(d->side, d->fork, 0), "hipStreamWaitEvent");
        tail_stream = d->side;
    }
    if (n_main > 0)
        HIP_TRY(launch_walk(p, v, (int)std::min<int64_t>(p.n_tiles, resident), lds, stream), "walk kernel launch");
    if (n_tail > 0) {
        WalkParams q = p;
        q.seeds = seeds ? seeds + n_main : nullptr;
        q.dur_in = dur_in ? dur_in + n_main : nullptr;
        q.realized = realized + n_main;
        q.durations = p.durations ? p.durations + n_main : nullptr;
        q.cause = cause ? cause + n_main : nullptr;
        q.cause_pos = cause_pos ? cause_pos + n_main : nullptr;
        q.n = (int32_t)n_tail;
        q.n_tiles = (int32_t)((n_tail + kLanes - 1) / kLanes);
        if (wide) {
            const int32_t groups = q.n_tiles;
            const int32_t W = P.waves;
            for (int32_t t = 0; t < P.n_steps; ++t) {
                const int32_t e0 = P.ev_off[(size_t)t * W], e1 = P.ev_off[(size_t)(t + 1) * W];
                HIP_TRY(launch_walk_wide(q, injected, e0, e1 - e0, groups, tail_stream), "wide step launch");
            }
            if (!injected) HIP_TRY(launch_walk_wide_orphans(q, groups, tail_stream), "wide orphans launch");
            if (tail_stream != stream) {
                HIP_TRY(hipEventRecord(d->join, tail_stream), "hipEventRecord");
                HIP_TRY(hipStreamWaitEvent(stream, d->join, 0), "hipStreamWaitEvent");
            }
        } else {
            HIP_TRY(launch_walk(q, v1, (int)std::min<int64_t>(q.n_tiles, resident1), lds1, stream), "walk kernel launch");
        }
    }
    return MCDP_OK;
}

}  // namespace

int mcdp::host_workspace(mcdp_propagator &P, size_t need, void *&buf, hipStream_t &stream) {
    int device = 0;
    HIP_TRY(hipGetDevice(&device), "hipGetDevice");
    HostWork *hw = nullptr;
    for (auto &h : P.host_work)
        if (h.device == device) hw = &h;
    if (!hw) {
        HostWork h;
        h.device = device;
        HIP_TRY(hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking), "hipStreamCreate");
        P.host_work.push_back(h);
        hw = &P.host_work.back();
    }
    if (hw->bytes < need) {
        if (hw->buf) {
            (void)hipStreamSynchronize(hw->stream);  // idle already: every call ends by waiting on it
            hw->drop_graph();
            (void)hipFree(hw->buf);
            hw->buf = nullptr;
            hw->bytes = 0;
        }
        hipError_t e = hipMalloc(&hw->buf, std::max<size_t>(need, 256));
        if (e != hipSuccess) {
            hw->buf = nullptr;
            return fail(MCDP_E_NOMEM, std::string("hipMalloc (host-call workspace): ") + hipGetErrorString(e));
        }
        hw->bytes = std::max<size_t>(need, 256);
    }
    buf = hw->buf;
    stream = hw->stream;
    return MCDP_OK;
}

void mcdp::host_workspace_trim(mcdp_propagator &P) {
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return;
    for (auto &h : P.host_work)
        if (h.device == device && h.buf && h.bytes > kHostWorkKeepBytes) {
            (void)hipStreamSynchronize(h.stream);
            h.drop_graph();
            (void)hipFree(h.buf);
            h.buf = nullptr;
            h.bytes = 0;
        }
}

namespace {

// The caller's output arrays, faulted in by several threads while the device
// works, for calls whose results copy directly (no staging buffers): a page's
// first touch costs a fault, which those copies would otherwise take one page at
// a time on one thread (for the network DAG's 2,048 samples, 1.3 GB: about 50 of
// 130 ms).  The pages get zeros, which the copies then overwrite; regions under
// 64 MB are left alone.
class Prefault {
   public:
    void start(const std::vector<std::pair<char *, size_t>> &regions) {
        size_t total = 0;
        for (const auto &r : regions) total += r.second;
        if (total < ((size_t)64 << 20)) return;
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const size_t nt = std::min<size_t>(16, hw);
        const size_t share = (total + nt - 1) / nt;
        std::vector<std::pair<char *, size_t>> pieces;  // (start, bytes) of every thread, in order
        size_t taken = 0;
        for (const auto &r : regions)
            for (size_t o = 0; o < r.second;) {
                const size_t room = share - taken % share;
                const size_t len = std::min(room, r.second - o);
                pieces.emplace_back(r.first + o, len);
                o += len;
                taken += len;
            }
        try {
            threads_.reserve(nt);
            for (size_t t = 0; t < nt; ++t) spawn(pieces, t, share);
        } catch (...) {  // no threads to spare: the copies fault the pages in themselves
            join();
        }
    }
    void join() {
        for (auto &t : threads_) t.join();
        threads_.clear();
    }
    ~Prefault() { join(); }

   private:
    void spawn(const std::vector<std::pair<char *, size_t>> &pieces, size_t t, size_t share) {
        threads_.emplace_back([pieces, t, share] {
            size_t at = 0;
            for (const auto &pc : pieces) {  // the pieces whose bytes fall in this thread's share
                const size_t lo = at, hi = at + pc.second;
                at = hi;
                const size_t a = std::max(lo, t * share), b = std::min(hi, (t + 1) * share);
                if (a < b) {
                    char *p = pc.first + (a - lo);
#ifdef MADV_POPULATE_WRITE
                    // whole pages inside the range in one call; the edges by hand
                    const uintptr_t pg = (uintptr_t)sysconf(_SC_PAGESIZE), x = (uintptr_t)p, y = x + (b - a);
                    const uintptr_t x1 = (x + pg - 1) / pg * pg, y1 = y / pg * pg;
                    if (y1 > x1 && madvise((void *)x1, y1 - x1, MADV_POPULATE_WRITE) == 0) {
                        std::memset(p, 0, x1 - x);
                        std::memset((void *)y1, 0, y - y1);
                        continue;
                    }
#endif
                    std::memset(p, 0, b - a);
                }
            }
        });
    }
    std::vector<std::thread> threads_;
};

// Large copies between the device and the caller's (pageable) arrays through
// pinned staging: kStageThreads threads, each with its own stream and pinned
// buffer of kStageBytes, take pieces of the arrays in turn and move each
// through their buffer.  A copy straight to or from pageable memory goes
// through one staging path in the runtime whatever the number of threads
// issuing it; this way several transfers run at once, and the host-side copies
// overlap them.  The buffers and streams exist once per device (64 MB of pinned
// memory, held until the process ends); calls take turns on them.
constexpr int kStageThreads = 8;
constexpr size_t kStageBytes = (size_t)8 << 20;
constexpr size_t kStageMinBytes = (size_t)16 << 20;  // smaller arrays copy directly

struct StageSet {
    int device = -1;
    void *buf[kStageThreads] = {};
    hipStream_t stream[kStageThreads] = {};
};
std::mutex g_stage_mu;
std::vector<StageSet> g_stage;

// One array of a chunk: rows of `width` bytes, `dev_pitch` apart on the device
// and `host_pitch` apart in the caller's array.
struct CopyPlane {
    char *host;
    char *dev;
    size_t rows, width, dev_pitch, host_pitch;
};

// A piece: rows [r0, r0 + k) and bytes [b0, b0 + len) of each of them.  Its
// device bytes are the contiguous span from row r0's byte b0 to row r0+k-1's
// byte b0+len, at most kStageBytes; an upload also writes the span's bytes
// between rows (columns past `width`, which the chunk does not use).
struct CopyPiece {
    const CopyPlane *plane;
    size_t r0, k, b0, len;
};

// The current device's staging set, created on first use; nullptr when its
// buffers cannot be had.  Call with g_stage_mu held.
StageSet *stage_set() {
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return nullptr;
    for (auto &g : g_stage)
        if (g.device == device) return &g;
    StageSet g;
    g.device = device;
    for (int t = 0; t < kStageThreads; ++t) {
        hipError_t e = hipHostMalloc(&g.buf[t], kStageBytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&g.stream[t], hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (int u = 0; u <= t; ++u) {
                if (g.buf[u]) (void)hipHostFree(g.buf[u]);
                if (g.stream[u]) (void)hipStreamDestroy(g.stream[u]);
            }
            return nullptr;
        }
    }
    g_stage.push_back(g);
    return &g_stage.back();
}

// True when large arrays of this device can go through the staging buffers.
bool staging_available() {
    std::lock_guard<std::mutex> lock(g_stage_mu);
    return stage_set() != nullptr;
}

// Moves every plane device -> host (to_host) or host -> device, synchronously.
// The device side must be ready (no kernel still writing or reading it).
// Returns an mcdp_status, or -1 when the staging buffers cannot be had (then
// nothing was copied).
int staged_copy(const std::vector<CopyPlane> &planes, bool to_host) {
    std::lock_guard<std::mutex> lock(g_stage_mu);
    StageSet *set = stage_set();
    if (!set) return -1;
    std::vector<CopyPiece> pieces;
    for (const auto &pl : planes) {
        if (pl.rows == 0 || pl.width == 0) continue;
        if (pl.width == pl.dev_pitch && pl.width == pl.host_pitch) {
            // rows back to back on both sides: one long row
            const size_t total = pl.rows * pl.width;
            for (size_t b = 0; b < total; b += kStageBytes)
                pieces.push_back({&pl, 0, 1, b, std::min(kStageBytes, total - b)});
        } else if (pl.dev_pitch <= kStageBytes) {
            const size_t per = kStageBytes / pl.dev_pitch;
            for (size_t r = 0; r < pl.rows; r += per) pieces.push_back({&pl, r, std::min(per, pl.rows - r), 0, pl.width});
        } else {
            for (size_t r = 0; r < pl.rows; ++r)
                for (size_t b = 0; b < pl.width; b += kStageBytes)
                    pieces.push_back({&pl, r, 1, b, std::min(kStageBytes, pl.width - b)});
        }
    }
    std::atomic<size_t> next{0};
    std::atomic<int> failed{0};
    hipError_t first_error = hipSuccess;
    std::mutex err_mu;
    auto work = [&](int t) {
        if (t > 0 && hipSetDevice(set->device) != hipSuccess) return;  // a new thread starts on device 0
        char *stage = static_cast<char *>(set->buf[t]);
        for (size_t i; !failed.load(std::memory_order_relaxed) && (i = next.fetch_add(1)) < pieces.size();) {
            const CopyPiece &pc = pieces[i];
            const CopyPlane &pl = *pc.plane;
            const bool flat = pl.width == pl.dev_pitch && pl.width == pl.host_pitch;
            const size_t pitch = flat ? 0 : pl.dev_pitch;  // of the rows in the staging buffer, as on the device
            const size_t host_pitch = flat ? 0 : pl.host_pitch;
            const size_t span = (pc.k - 1) * pitch + pc.len;
            char *dev = pl.dev + pc.r0 * pitch + pc.b0;
            if (!to_host)
                for (size_t j = 0; j < pc.k; ++j)
                    std::memcpy(stage + j * pitch, pl.host + (pc.r0 + j) * host_pitch + pc.b0, pc.len);
            hipError_t e = to_host ? hipMemcpyAsync(stage, dev, span, hipMemcpyDeviceToHost, set->stream[t])
                                   : hipMemcpyAsync(dev, stage, span, hipMemcpyHostToDevice, set->stream[t]);
            if (e == hipSuccess) e = hipStreamSynchronize(set->stream[t]);
            if (e != hipSuccess) {
                std::lock_guard<std::mutex> g(err_mu);
                if (first_error == hipSuccess) first_error = e;
                failed.store(1);
                return;
            }
            if (to_host)
                for (size_t j = 0; j < pc.k; ++j)
                    std::memcpy(pl.host + (pc.r0 + j) * host_pitch + pc.b0, stage + j * pitch, pc.len);
        }
    };
    std::vector<std::thread> threads;
    try {
        for (int t = 1; t < kStageThreads; ++t) threads.emplace_back(work, t);
    } catch (...) {  // fewer threads: the ones running take every piece
    }
    work(0);
    for (auto &th : threads) th.join();
    if (failed.load()) return hip_fail(first_error, to_host ? "hipMemcpy (outputs)" : "hipMemcpy (inputs)");
    return MCDP_OK;
}

// Short runs through the host path replay a captured graph of their launches
// (HostWork): for up to this many samples, which stay below one round of tiles
// on any device, so the launches are one stream's sequence; and for plans of at
// least kGraphMinSteps steps.  A replay spends about 1.7 us per dependent launch
// against about 3 us for launches one by one, but a graph of a single kernel
// (the tiles' route) is slower to launch than the kernel alone.
constexpr int64_t kGraphMaxSamples = 1024;
constexpr int32_t kGraphMinSteps = 32;

// The chunk's launches (launch()).  A run of one chunk of at most
// kGraphMaxSamples samples whose key (ShortRunKey) matches the previous call's
// is captured into a graph and launched as one; while the key stays, later
// calls replay that graph.  Any other call, or a capture that fails, launches
// directly.  The caller holds P.host_mu.
int launch_host_chunk(mcdp_propagator &P, HostWork *hw, bool injected, const int32_t *d_seeds, double *d_dur,
                      int64_t m, int64_t ld, double *d_real, int32_t *d_cause, hipStream_t st) {
    auto direct = [&] {
        return launch(P, injected, d_seeds, d_dur, m, ld, d_real, injected ? nullptr : d_dur, d_cause, st);
    };
    if (!hw || m != ld || m > kGraphMaxSamples || P.n_steps < kGraphMinSteps) return direct();
    ShortRunKey key;
    key.buf = hw->buf;
    key.n = m;
    key.ld = ld;
    key.injected = injected;
    key.generation = P.generation.load();
    key.tail_mode = P.tail_mode;
    if (hw->graph && hw->captured == key) {
        HIP_TRY(hipGraphLaunch(hw->graph, st), "hipGraphLaunch");
        return MCDP_OK;
    }
    if (!(hw->seen == key)) {  // first sight: the shape and the plan's upload happen here, not in a capture
        hw->seen = key;
        return direct();
    }
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return direct();
    }
    const int rc = direct();
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(st, &g);
    hipGraphExec_t exec = nullptr;
    const bool ok = rc == MCDP_OK && ee == hipSuccess && g && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) (void)hipGraphDestroy(g);
    if (!ok) {  // launch directly; this key is not captured again until it is seen twice anew
        (void)hipGetLastError();
        hw->seen = ShortRunKey{};
        return direct();
    }
    if (hw->graph) (void)hipGraphExecDestroy(hw->graph);
    hw->graph = exec;
    hw->captured = key;
    HIP_TRY(hipGraphLaunch(exec, st), "hipGraphLaunch");
    return MCDP_OK;
}

// The host-path state of the current device (after host_workspace).
HostWork *current_host_work(mcdp_propagator &P) {
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return nullptr;
    for (auto &h : P.host_work)
        if (h.device == device) return &h;
    return nullptr;
}

// Host-buffer convenience: runs in chunks of samples so device memory stays
// bounded, and copies each chunk into the caller's event-major arrays.  The
// device buffers and the stream are the propagator's (HostWork), kept from
// call to call; the call waits on that stream only, so other streams of the
// process (a server's, a caller's own) keep running.
int run_host(mcdp_propagator &P, bool injected, const int32_t *seeds, const double *dur_in, int64_t n,
             double *realized, double *durations, int32_t *cause) {
    if (n < 0) return fail(MCDP_E_INVALID, "n must be non-negative");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MCDP_E_DEVICE, "no HIP device available: the Monte Carlo engine runs on the GPU only");
    if (n == 0) return MCDP_OK;
    const int64_t E = P.E, L = P.L;
    const int64_t per_sample = 12 * E + 8 * L + 8;
    int64_t chunk = std::max<int64_t>(kLanes, ((int64_t)1 << 31) / std::max<int64_t>(per_sample, 1));
    chunk = std::min<int64_t>(n, chunk / kLanes * kLanes);
    const size_t o_cause = align_up(sizeof(double) * E * chunk, 256);
    const size_t o_dur = o_cause + align_up(sizeof(int32_t) * E * chunk, 256);
    const size_t o_seeds = o_dur + align_up(sizeof(double) * L * chunk, 256);
    const size_t need = o_seeds + align_up(sizeof(int32_t) * chunk, 256);

    std::lock_guard<std::mutex> host_lock(P.host_mu);
    void *work = nullptr;
    hipStream_t st = nullptr;
    int rc = host_workspace(P, need, work, st);
    if (rc) return rc;
    HostWork *const hw = current_host_work(P);
    auto *base = static_cast<unsigned char *>(work);
    auto *d_real = reinterpret_cast<double *>(base);
    auto *d_cause = reinterpret_cast<int32_t *>(base + o_cause);
    auto *d_dur = reinterpret_cast<double *>(base + o_dur);
    auto *d_seeds = reinterpret_cast<int32_t *>(base + o_seeds);
    // With staging, the copying threads fault the caller's pages in themselves, in parallel and while
    // other pieces are in flight; without it the copies would take them one at a time.
    const size_t total_out = (size_t)n * ((realized ? 8 * E : 0) + (cause ? 4 * E : 0) +
                                          (durations && !injected ? 8 * L : 0));
    const bool staged = total_out >= kStageMinBytes && staging_available();
    Prefault prefault;
    if (!staged)
        prefault.start({{reinterpret_cast<char *>(realized), realized ? sizeof(double) * (size_t)(E * n) : 0},
                        {reinterpret_cast<char *>(cause), cause ? sizeof(int32_t) * (size_t)(E * n) : 0},
                        {reinterpret_cast<char *>(durations),
                         (durations && !injected) ? sizeof(double) * (size_t)(L * n) : 0}});
    for (int64_t c0 = 0; c0 < n && rc == MCDP_OK; c0 += chunk) {
        const int64_t m = std::min(chunk, n - c0);
        hipError_t e = hipSuccess;
        // one chunk of all samples: host and device rows have the same pitch, so every array moves as
        // one contiguous copy (a 2-D copy of many short rows, e.g. one sample, costs a transfer per row)
        const bool whole = chunk == n;
        if (injected) {
            // the workspace is idle here: every chunk ends by waiting on the stream
            const size_t in_bytes = sizeof(double) * (size_t)L * (size_t)m;
            int sc = -1;
            if (L > 0 && in_bytes >= kStageMinBytes)
                sc = staged_copy({{const_cast<char *>(reinterpret_cast<const char *>(dur_in + c0)),
                                   reinterpret_cast<char *>(d_dur), (size_t)L, sizeof(double) * m,
                                   sizeof(double) * chunk, sizeof(double) * n}},
                                 false);
            if (sc > 0) {
                rc = sc;
                break;
            }
            if (L > 0 && sc < 0)
                e = whole ? hipMemcpyAsync(d_dur, dur_in, sizeof(double) * L * n, hipMemcpyHostToDevice, st)
                          : hipMemcpy2DAsync(d_dur, sizeof(double) * chunk, dur_in + c0, sizeof(double) * n,
                                             sizeof(double) * m, (size_t)L, hipMemcpyHostToDevice, st);
        } else {
            e = hipMemcpyAsync(d_seeds, seeds + c0, sizeof(int32_t) * m, hipMemcpyHostToDevice, st);
        }
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipMemcpy (inputs)");
            break;
        }
        rc = launch_host_chunk(P, hw, injected, d_seeds, d_dur, m, chunk, d_real, d_cause, st);
        if (rc) break;
        prefault.join();  // the pages are ready before the first copy writes them
        const size_t out_bytes = (size_t)m * ((realized ? 8 * E : 0) + (cause ? 4 * E : 0) +
                                              (durations && !injected ? 8 * L : 0));
        if (staged && out_bytes >= kStageMinBytes) {
            if ((e = hipStreamSynchronize(st)) != hipSuccess) {
                rc = hip_fail(e, "walk kernel");
                break;
            }
            std::vector<CopyPlane> planes;
            auto plane = [&](void *host, void *dev, size_t elem, int64_t rows) {
                if (host && rows > 0)
                    planes.push_back({static_cast<char *>(host) + elem * c0, static_cast<char *>(dev), (size_t)rows,
                                      elem * m, elem * chunk, elem * n});
            };
            plane(realized, d_real, sizeof(double), E);
            plane(cause, d_cause, sizeof(int32_t), E);
            if (!injected) plane(durations, d_dur, sizeof(double), L);
            const int sc = staged_copy(planes, true);
            if (sc >= 0) {
                rc = sc;
                continue;
            }
        }
        auto copy_out = [&](void *dst, const void *src, size_t elem, int64_t rows) -> hipError_t {
            if (!dst || rows == 0) return hipSuccess;
            if (whole) return hipMemcpyAsync(dst, src, elem * (size_t)n * (size_t)rows, hipMemcpyDeviceToHost, st);
            return hipMemcpy2DAsync(static_cast<char *>(dst) + elem * c0, elem * n, src, elem * chunk, elem * m,
                                    (size_t)rows, hipMemcpyDeviceToHost, st);
        };
        if ((e = copy_out(realized, d_real, sizeof(double), E)) != hipSuccess ||
            (e = copy_out(cause, d_cause, sizeof(int32_t), E)) != hipSuccess ||
            (!injected && (e = copy_out(durations, d_dur, sizeof(double), L)) != hipSuccess)) {
            rc = hip_fail(e, "hipMemcpy (outputs)");
            break;
        }
        if ((e = hipStreamSynchronize(st)) != hipSuccess) rc = hip_fail(e, "walk kernel");
    }
    if (rc) (void)hipStreamSynchronize(st);  // nothing of this call may still use the buffers
    host_workspace_trim(P);
    return rc;
}

}  // namespace

int mcdp::ensure_device(mcdp_propagator &P, DeviceCopy *&out) {
    // The uploaded plan must be the one the walk kernel will run: fix the
    // schedule shape first, as the first launch would.
    if (!P.shape_fixed) {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
            return fail(MCDP_E_DEVICE, "no HIP device available: the Monte Carlo engine runs on the GPU only");
        int rc = choose_shape(P);
        if (rc) return rc;
        P.shape_fixed = true;
    }
    return ensure_device_impl(P, out);
}
mcdp::WalkVariant mcdp::plan_variant(const mcdp_propagator &P, bool injected) { return walk_variant(P, injected); }
size_t mcdp::plan_lds_bytes(const mcdp_propagator &P, WalkVariant v) { return walk_lds_bytes(P, v); }
int mcdp::configure_shape(mcdp_propagator *prop, int32_t waves, int32_t k, int32_t spl) {
    return mcdp_propagator_configure(prop, waves, k, spl);
}
mcdp::Record mcdp::link_record(const mcdp_propagator &P, int32_t l) { return edge_record(P, l); }

int32_t mcdp::max_entries(const mcdp_propagator &P) {
    int64_t m = 0;
    for (int32_t e = 0; e < P.E; ++e) m = std::max(m, P.pred_off[(size_t)e + 1] - P.pred_off[(size_t)e]);
    return (int32_t)std::min<int64_t>(m, std::numeric_limits<int32_t>::max());
}

int mcdp::run_many_positions(mcdp_propagator &P, const int32_t *seeds_dev, int64_t n, int64_t ld, double *realized,
                             uint8_t *cause_pos, hipStream_t s) {
    if (!cause_pos) return fail(MCDP_E_INVALID, "positions output is null");
    if (max_entries(P) > kMaxPositionEntries)
        return fail(MCDP_E_INVALID, "winner positions need events of at most 254 predecessor entries");
    if (n > 0 && !seeds_dev) return fail(MCDP_E_INVALID, "seeds is null");
    return launch(P, false, seeds_dev, nullptr, n, ld, realized, nullptr, nullptr, s, cause_pos);
}

// The range of every duration the sampler can draw for link l, as the device
// rounds it (mcdp_sampler.hpp: base + extra, extra = X * d, one rounding
// each): X in [x0, x1] for the multiplier kinds (exponential and gamma X >= 0,
// at most max_scale), the reachable alias-table values for the empirical
// kinds.  Rounding is monotone, so the two ends bound every draw.  A link
// that can draw NaN gets [-inf, +inf]: a NaN candidate never wins the max, so
// it acts as the smallest candidate.
void mcdp::link_duration_range(const mcdp_propagator &P, int32_t l, double &lo, double &hi) {
    constexpr double inf = std::numeric_limits<double>::infinity();
    const double d = P.dmin[(size_t)l];
    lo = -inf;
    hi = inf;
    if (std::isnan(d)) return;
    const uint32_t di = P.ldist[(size_t)l];
    if (di == kNoDistIndex) {
        lo = hi = d;
        return;
    }
    const DistRec &D = P.dist[di];
    auto rel = [&](double x) {  // d + x d, rounded as the device does
        const double extra = x * d;
        return d + extra;
    };
    auto ends = [&](double a, double b) {
        if (std::isnan(a) || std::isnan(b)) return;
        lo = std::min(a, b);
        hi = std::max(a, b);
    };
    // multiplier X in [x0, x1] (x1 may be +inf while every draw is finite)
    auto mult = [&](double x0, double x1) {
        if (!std::isfinite(d)) return;
        if (d == 0.0) {  // a finite X times a zero is a zero
            lo = hi = rel(0.0);
            return;
        }
        ends(rel(x0), std::isinf(x1) ? (d > 0.0 ? inf : -inf) : rel(x1));
    };
    switch (D.kind) {
        case kNone: lo = hi = d; return;
        case kConstant:
        case kEmpiricalRel1: ends(rel(D.p[0]), rel(D.p[0])); return;
        case kEmpiricalAbs1: ends(d + D.p[0], d + D.p[0]); return;
        case kExpInf: mult(0.0, inf); return;
        case kExpTrunc: mult(0.0, D.p[2]); return;
        case kGammaErlang2:
        case kGammaHalf:
        case kGammaMT:
        case kGammaPow: mult(0.0, D.p[3]); return;
        case kEmpiricalAbs:
        case kEmpiricalRel: {
            uint64_t packed;
            std::memcpy(&packed, &D.p[0], sizeof(double));
            const size_t n = (uint32_t)packed, off = (size_t)(uint32_t)(packed >> 32);
            double vmin = inf, vmax = -inf;
            bool nan = false;
            for (size_t b = 0; b < n; ++b) {  // a bin's own value needs prob > 0; its alias is always reachable
                const AliasRec &a = P.alias[off + b];
                for (int side = 0; side < 2; ++side) {
                    if (side == 0 && !(a.prob > 0.0)) continue;
                    const double v = side == 0 ? a.v_self : a.v_alias;
                    if (std::isnan(v)) nan = true;
                    vmin = std::min(vmin, v);
                    vmax = std::max(vmax, v);
                }
            }
            if (nan) return;
            if (D.kind == kEmpiricalAbs) {
                if (std::isfinite(d)) ends(d + vmin, d + vmax);
            } else if (std::isfinite(d)) {
                if (d == 0.0) {
                    if (std::isfinite(vmin) && std::isfinite(vmax)) lo = hi = rel(0.0);
                } else {
                    ends(rel(vmin), rel(vmax));
                }
            }
            return;
        }
        default: return;
    }
}

// Every event's realized time lies in [ev_lo[e], ev_hi[e]] for every seed:
// the walk's recurrence (upper = earliest + max_delay; per predecessor in
// input order t = min(upper, r_src + duration), best = t when t >= best;
// realized = min(upper, best); a stale predecessor reads its earliest time)
// over the per-link duration ranges, which it preserves because each step is
// monotone.  False when an earliest time or max_delay is not finite.
bool mcdp::event_ranges(const mcdp_propagator &P, std::vector<double> &ev_lo, std::vector<double> &ev_hi) {
    if (!std::isfinite(P.max_delay)) return false;
    for (double e : P.earliest)
        if (!std::isfinite(e)) return false;
    std::vector<double> dlo((size_t)P.L), dhi((size_t)P.L);
    for (int32_t l = 0; l < P.L; ++l) link_duration_range(P, l, dlo[(size_t)l], dhi[(size_t)l]);
    ev_lo.assign((size_t)P.E, 0.0);
    ev_hi.assign((size_t)P.E, 0.0);
    for (const auto &level : P.levels)
        for (int32_t e : level) {
            const double upper = P.earliest[(size_t)e] + P.max_delay;
            double blo = P.earliest[(size_t)e], bhi = blo;
            for (int64_t k = P.pred_off[(size_t)e]; k < P.pred_off[(size_t)e + 1]; ++k) {
                const PredEntry &pe = P.preds[(size_t)k];
                const bool valid = pe.edge >= 0 && pe.edge < P.L;
                const double rlo = pe.stale ? P.earliest[(size_t)pe.src] : ev_lo[(size_t)pe.src];
                const double rhi = pe.stale ? P.earliest[(size_t)pe.src] : ev_hi[(size_t)pe.src];
                double tlo = rlo + (valid ? dlo[(size_t)pe.edge] : 0.0);
                double thi = rhi + (valid ? dhi[(size_t)pe.edge] : 0.0);
                tlo = (upper < tlo) ? upper : tlo;
                thi = (upper < thi) ? upper : thi;
                if (tlo >= blo) blo = tlo;
                if (thi >= bhi) bhi = thi;
            }
            ev_lo[(size_t)e] = (upper < blo) ? upper : blo;
            ev_hi[(size_t)e] = (upper < bhi) ? upper : bhi;
        }
    return true;
}

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char *mcdp_last_error(void) { return g_last_error.c_str(); }
int mcdp_abi_version(void) { return MCDP_ABI_VERSION; }

int mcdp_generator_create(mcdp_generator **out) {
    if (!out) return fail(MCDP_E_INVALID, "out is null");
    *out = new (std::nothrow) mcdp_generator();
    return *out ? MCDP_OK : fail(MCDP_E_NOMEM, "out of host memory");
}
void mcdp_generator_destroy(mcdp_generator *gen) { delete gen; }
int mcdp_generator_set_seed(mcdp_generator *gen, int32_t) {
    return gen ? MCDP_OK : fail(MCDP_E_INVALID, "generator is null");
}
int mcdp_generator_add_constant(mcdp_generator *gen, int32_t t, double factor) {
    if (!gen) return fail(MCDP_E_INVALID, "generator is null");
    DistSpec s;
    s.kind = kApiConstant;
    s.a = factor;
    gen->dists[t] = s;
    return MCDP_OK;
}
int mcdp_generator_add_exponential(mcdp_generator *gen, int32_t t, double lambda_, double max_scale) {
    if (!gen) return fail(MCDP_E_INVALID, "generator is null");
    DistSpec s;
    s.kind = kApiExponential;
    s.a = lambda_;
    s.b = max_scale;
    gen->dists[t] = s;
    return MCDP_OK;
}
int mcdp_generator_add_gamma(mcdp_generator *gen, int32_t t, double shape, double scale, double max_scale) {
    if (!gen) return fail(MCDP_E_INVALID, "generator is null");
    DistSpec s;
    s.kind = kApiGamma;
    s.a = shape;
    s.b = scale;
    s.c = max_scale;
    gen->dists[t] = s;
    return MCDP_OK;
}
static int add_empirical(mcdp_generator *gen, int32_t t, const double *v, const double *w, int64_t n, ApiKind k) {
    if (!gen) return fail(MCDP_E_INVALID, "generator is null");
    if (n < 0 || (n > 0 && (!v || !w))) return fail(MCDP_E_INVALID, "invalid empirical table");
    DistSpec s;
    s.kind = k;
    s.values.assign(v, v + n);
    s.weights.assign(w, w + n);
    gen->dists[t] = std::move(s);
    return MCDP_OK;
}
int mcdp_generator_add_empirical_absolute(mcdp_generator *gen, int32_t t, const double *values, const double *weights,
                                          int64_t n) {
    return add_empirical(gen, t, values, weights, n, kApiEmpAbs);
}
int mcdp_generator_add_empirical_relative(mcdp_generator *gen, int32_t t, const double *factors,
                                          const double *weights, int64_t n) {
    return add_empirical(gen, t, factors, weights, n, kApiEmpRel);
}

int mcdp_propagator_create(int32_t n_events, const double *earliest, int32_t n_activities,
                           const int32_t *activity_idx, const double *minimal_duration,
                           const int32_t *activity_type, int32_t n_entries, const int32_t *entry_target,
                           const int64_t *entry_offsets, const int32_t *pred_event, const int32_t *pred_activity,
                           double max_delay, const mcdp_generator *gen, mcdp_propagator **out) {
    if (!out) return fail(MCDP_E_INVALID, "out is null");
    *out = nullptr;
    if (!gen) return fail(MCDP_E_INVALID, "generator is null");
    if ((n_events > 0 && !earliest) || (n_activities > 0 && (!activity_idx || !minimal_duration || !activity_type)) ||
        (n_entries > 0 && (!entry_target || !entry_offsets)) ||
        (n_entries > 0 && entry_offsets[n_entries] > 0 && (!pred_event || !pred_activity)))
        return fail(MCDP_E_INVALID, "null input array");
    try {
        std::unique_ptr<mcdp_propagator> P(new mcdp_propagator());
        int rc = compile_propagator(*P, n_events, earliest, n_activities, activity_idx, minimal_duration,
                                    activity_type, n_entries, entry_target, entry_offsets, pred_event,
                                    pred_activity, max_delay, *gen);
        if (rc) return rc;
        *out = P.release();
        return MCDP_OK;
    } catch (const std::bad_alloc &) {
        return fail(MCDP_E_NOMEM, "out of host memory");
    }
}
void mcdp_propagator_destroy(mcdp_propagator *prop) { delete prop; }
int32_t mcdp_propagator_node_count(const mcdp_propagator *prop) { return prop ? prop->E : -1; }
int32_t mcdp_propagator_activity_count(const mcdp_propagator *prop) { return prop ? prop->L : -1; }

int mcdp_run_many_device(mcdp_propagator *prop, const int32_t *seeds_dev, int64_t n, int64_t ld,
                         double *realized_dev, double *durations_dev, int32_t *cause_dev, void *stream) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    return launch(*prop, false, seeds_dev, nullptr, n, ld, realized_dev, durations_dev, cause_dev,
                  static_cast<hipStream_t>(stream));
}
int mcdp_propagate_device(mcdp_propagator *prop, const double *durations_dev, int64_t n, int64_t ld,
                          double *realized_dev, int32_t *cause_dev, void *stream) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    return launch(*prop, true, nullptr, durations_dev, n, ld, realized_dev, nullptr, cause_dev,
                  static_cast<hipStream_t>(stream));
}
int mcdp_run_many_host(mcdp_propagator *prop, const int32_t *seeds, int64_t n, double *realized, double *durations,
                       int32_t *cause) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (n > 0 && !seeds) return fail(MCDP_E_INVALID, "seeds is null");
    return run_host(*prop, false, seeds, nullptr, n, realized, durations, cause);
}
int mcdp_propagate_host(mcdp_propagator *prop, const double *durations, int64_t n, double *realized,
                        int32_t *cause) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (n > 0 && prop->L > 0 && !durations) return fail(MCDP_E_INVALID, "durations is null");
    return run_host(*prop, true, nullptr, durations, n, realized, nullptr, cause);
}

int mcdp_propagator_plan_info(const mcdp_propagator *prop, int64_t *out, int n_out) {
    if (!prop || !out) return fail(MCDP_E_INVALID, "null argument");
    const int64_t v[12] = {prop->n_steps,
                          prop->waves,
                          prop->slots,
                          (int64_t)prop->slots * kLanes * prop->spl * (int64_t)sizeof(double),
                          prop->n_global,
                          prop->n_stale,
                          (int64_t)prop->orph.size(),
                          prop->max_per_wave,
                          prop->k,
                          prop->spl,
                          prop->tile_cost,
                          prop->wide_cost};
    for (int i = 0; i < n_out && i < 12; ++i) out[i] = v[i];
    return MCDP_OK;
}

int mcdp_propagator_ranges(const mcdp_propagator *prop, double *link_lo, double *link_hi, double *event_lo,
                           double *event_hi) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    for (int32_t l = 0; l < prop->L; ++l) {
        double lo, hi;
        link_duration_range(*prop, l, lo, hi);
        if (link_lo) link_lo[l] = lo;
        if (link_hi) link_hi[l] = hi;
    }
    if (!event_lo && !event_hi) return MCDP_OK;
    std::vector<double> elo, ehi;
    if (!event_ranges(*prop, elo, ehi))
        return fail(MCDP_E_INVALID, "event ranges need finite earliest times and a finite max_delay");
    if (event_lo) std::copy(elo.begin(), elo.end(), event_lo);
    if (event_hi) std::copy(ehi.begin(), ehi.end(), event_hi);
    return MCDP_OK;
}

int mcdp_propagator_configure(mcdp_propagator *prop, int32_t waves, int32_t events_per_wave,
                              int32_t samples_per_lane) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (!shape_valid(waves, events_per_wave, samples_per_lane))
        return fail(MCDP_E_INVALID,
                    "invalid shape: waves in [1, 16], events_per_wave in {4, 8, 12, 16}, samples_per_lane 1 or 2 "
                    "(with events_per_wave <= 8)");
    std::lock_guard<std::mutex> lock(prop->mu);
    int rc = build_schedule(*prop, waves, events_per_wave);
    if (rc) return rc;
    prop->spl = samples_per_lane;
    prop->drop_device_copies();
    prop->shape_fixed = true;
    return MCDP_OK;
}

int mcdp_propagator_autotune(mcdp_propagator *prop, int64_t n_samples, int32_t *waves, int32_t *events_per_wave,
                             int32_t *samples_per_lane, float *ms) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (n_samples < 1) return fail(MCDP_E_INVALID, "n_samples must be positive");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(MCDP_E_DEVICE, "no HIP device available");
    const int64_t n = n_samples, E = prop->E, L = prop->L;
    std::vector<int32_t> hseeds((size_t)n);
    for (int64_t i = 0; i < n; ++i) hseeds[(size_t)i] = (int32_t)i;
    void *buf = nullptr;
    const size_t bytes = align_up((size_t)n * 4, 256) + (size_t)n * (size_t)(12 * E + 8 * L);
    hipError_t e = hipMalloc(&buf, bytes);
    if (e != hipSuccess) return fail(MCDP_E_NOMEM, std::string("hipMalloc (autotune): ") + hipGetErrorString(e));
    auto *b = static_cast<unsigned char *>(buf);
    int32_t *d_seeds = reinterpret_cast<int32_t *>(b);
    double *d_real = reinterpret_cast<double *>(b + align_up((size_t)n * 4, 256));
    double *d_dur = d_real + (size_t)(E * n);
    int32_t *d_cause = reinterpret_cast<int32_t *>(d_dur + (size_t)(L * n));
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int rc = MCDP_OK;
    float best_ms = -1.0f;
    Shape best{prop->waves, prop->k, prop->spl};
    std::vector<Shape> cands;
    {
        std::lock_guard<std::mutex> lock(prop->mu);
        cands = shape_candidates(*prop);
    }
    e = hipMemcpy(d_seeds, hseeds.data(), (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    if (e != hipSuccess) rc = hip_fail(e, "autotune setup");
    for (size_t c = 0; rc == MCDP_OK && c < cands.size(); ++c) {
        rc = mcdp_propagator_configure(prop, cands[c].waves, cands[c].k, cands[c].spl);
        if (rc) break;
        {  // skip shapes that cannot be resident
            std::lock_guard<std::mutex> lock(prop->mu);
            const WalkVariant v = walk_variant(*prop, false);
            int bpc = 0;
            if (walk_occupancy(v, prop->waves, walk_lds_bytes(*prop, v), &bpc) != hipSuccess || bpc < 1) continue;
        }
        rc = launch(*prop, false, d_seeds, nullptr, n, n, d_real, d_dur, d_cause, nullptr);  // warm-up + upload
        if (rc) break;
        if ((e = hipEventRecord(t0, nullptr)) != hipSuccess) {
            rc = hip_fail(e, "hipEventRecord");
            break;
        }
        rc = launch(*prop, false, d_seeds, nullptr, n, n, d_real, d_dur, d_cause, nullptr);
        if (rc) break;
        float t = 0.0f;
        if ((e = hipEventRecord(t1, nullptr)) != hipSuccess || (e = hipEventSynchronize(t1)) != hipSuccess ||
            (e = hipEventElapsedTime(&t, t0, t1)) != hipSuccess) {
            rc = hip_fail(e, "autotune timing");
            break;
        }
        if (best_ms < 0.0f || t < best_ms) {
            best_ms = t;
            best = cands[c];
        }
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    (void)hipFree(buf);
    if (rc) return rc;
    rc = mcdp_propagator_configure(prop, best.waves, best.k, best.spl);
    if (rc) return rc;
    if (waves) *waves = best.waves;
    if (events_per_wave) *events_per_wave = best.k;
    if (samples_per_lane) *samples_per_lane = best.spl;
    if (ms) *ms = best_ms;
    return MCDP_OK;
}

// What a schedule shape's speed depends on: the precedence structure (events,
// links, every entry's source, edge and stale flag), which distribution each
// link draws from and each distribution's kind and flags (the code path of a
// draw), and the alias tables' size (LDS).  Times and distribution
// parameters are left out.  FNV-1a, 64 bits, over those values' bytes.
static uint64_t shape_fingerprint(const mcdp_propagator &P) {
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&](const void *data, size_t n) {
        const auto *b = static_cast<const unsigned char *>(data);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 0x100000001b3ull;
        }
    };
    auto mix64 = [&](int64_t v) { mix(&v, sizeof v); };
    mix("mcdp-shape-1", 12);
    mix64(P.E);
    mix64(P.L);
    for (int64_t o : P.pred_off) mix64(o);
    for (const PredEntry &pe : P.preds) {
        mix64(pe.src);
        mix64(pe.edge);
        mix64(pe.stale ? 1 : 0);
    }
    for (uint32_t d : P.ldist) mix64(d);
    for (const DistRec &d : P.dist) {
        mix64(d.kind);
        mix64(d.flags);
    }
    mix64((int64_t)P.alias.size());
    return h;
}

int mcdp_propagator_fingerprint(const mcdp_propagator *prop, uint64_t *fingerprint) {
    if (!prop || !fingerprint) return fail(MCDP_E_INVALID, "propagator or fingerprint is null");
    *fingerprint = shape_fingerprint(*prop);
    return MCDP_OK;
}

int mcdp_propagator_autotune_cached(mcdp_propagator *prop, int64_t n_samples, const char *path, int32_t *waves,
                                    int32_t *events_per_wave, int32_t *samples_per_lane, float *ms,
                                    int32_t *from_file) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (!path || !*path) return fail(MCDP_E_INVALID, "tuning file path is empty");
    if (n_samples < 1) return fail(MCDP_E_INVALID, "n_samples must be positive");
    int dev = 0;
    hipDeviceProp_t props{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&props, dev) != hipSuccess)
        return fail(MCDP_E_DEVICE, "no HIP device available");
    // the device kind: architecture name up to its first ':' or space, and compute units
    std::string arch(props.gcnArchName);
    arch = arch.substr(0, arch.find_first_of(": "));
    if (arch.empty()) arch = "unknown";
    int64_t bucket = 1;
    while (bucket < n_samples) bucket *= 2;
    if (arch.size() > 64) arch.resize(64);
    char key[160];
    std::snprintf(key, sizeof key, "mcdp-shape 1 %016" PRIx64 " %s %d %" PRId64, shape_fingerprint(*prop), arch.c_str(),
                  props.multiProcessorCount, bucket);
    const std::string prefix = std::string(key) + " ";

    // the last valid line for this key wins (a later tuning appends a newer choice)
    bool found = false;
    Shape got{0, 0, 0};
    float got_ms = 0.0f;
    {
        std::ifstream in(path);
        std::string line;
        while (in && std::getline(in, line)) {
            if (line.compare(0, prefix.size(), prefix) != 0) continue;
            std::istringstream rest(line.substr(prefix.size()));
            Shape c{0, 0, 0};
            float t = 0.0f;
            if (!(rest >> c.waves >> c.k >> c.spl >> t) || !shape_valid(c.waves, c.k, c.spl)) continue;
            found = true;
            got = c;
            got_ms = t;
        }
    }
    if (found) {
        int rc = mcdp_propagator_configure(prop, got.waves, got.k, got.spl);
        if (rc) return rc;
        bool resident = false;
        {
            std::lock_guard<std::mutex> lock(prop->mu);
            const WalkVariant v = walk_variant(*prop, false);
            int bpc = 0;
            resident = walk_occupancy(v, prop->waves, walk_lds_bytes(*prop, v), &bpc) == hipSuccess && bpc >= 1;
        }
        if (resident) {
            if (waves) *waves = got.waves;
            if (events_per_wave) *events_per_wave = got.k;
            if (samples_per_lane) *samples_per_lane = got.spl;
            if (ms) *ms = got_ms;
            if (from_file) *from_file = 1;
            return MCDP_OK;
        }
    }
    int32_t w = 0, k = 0, spl = 0;
    float t = 0.0f;
    int rc = mcdp_propagator_autotune(prop, n_samples, &w, &k, &spl, &t);
    if (rc) return rc;
    if (waves) *waves = w;
    if (events_per_wave) *events_per_wave = k;
    if (samples_per_lane) *samples_per_lane = spl;
    if (ms) *ms = t;
    if (from_file) *from_file = 0;
    // one write of the whole line to a file opened for appending: concurrent
    // tunings add whole lines
    char line[256];
    const int len = std::snprintf(line, sizeof line, "%s%d %d %d %.6f\n", prefix.c_str(), w, k, spl, (double)t);
    if (len <= 0 || (size_t)len >= sizeof line) return fail(MCDP_E_INVALID, "tuning file line too long");
    const int fd = ::open(path, O_WRONLY | O_APPEND | O_CREAT | O_CLOEXEC, 0644);
    if (fd < 0) return fail(MCDP_E_INVALID, std::string("cannot open tuning file ") + path + ": " + std::strerror(errno));
    const ssize_t wr = ::write(fd, line, (size_t)len);
    const int werr = errno;
    ::close(fd);
    if (wr != (ssize_t)len)
        return fail(MCDP_E_INVALID, std::string("cannot write tuning file ") + path + ": " + std::strerror(werr));
    return MCDP_OK;
}

int mcdp_debug_tail_mode(mcdp_propagator *prop, int32_t mode) {
    if (!prop) return fail(MCDP_E_INVALID, "propagator is null");
    if (mode < 0 || mode > 2) return fail(MCDP_E_INVALID, "tail mode must be 0, 1 or 2");
    std::lock_guard<std::mutex> lock(prop->mu);
    prop->tail_mode = mode;
    return MCDP_OK;
}

// ---- device plumbing --------------------------------------------------------
int mcdp_device_count(int *count) {
    if (!count) return fail(MCDP_E_INVALID, "count is null");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) *count = 0;
    return MCDP_OK;
}
int mcdp_malloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(MCDP_E_INVALID, "ptr is null");
    hipError_t e = hipMalloc(ptr, bytes);
    return e == hipSuccess ? MCDP_OK : fail(MCDP_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
}
int mcdp_free(void *ptr) {
    HIP_TRY(hipFree(ptr), "hipFree");
    return MCDP_OK;
}
int mcdp_memcpy_htod(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy HtoD");
    return MCDP_OK;
}
int mcdp_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy DtoH");
    return MCDP_OK;
}
int mcdp_memset(void *dst, int value, size_t bytes) {
    HIP_TRY(hipMemset(dst, value, bytes), "hipMemset");
    return MCDP_OK;
}
int mcdp_memcpy2d_dtoh(void *dst, size_t dst_pitch, const void *src, size_t src_pitch, size_t width,
                       size_t height) {
    HIP_TRY(hipMemcpy2D(dst, dst_pitch, src, src_pitch, width, height, hipMemcpyDeviceToHost), "hipMemcpy2D DtoH");
    return MCDP_OK;
}
int mcdp_graph_capture_begin(void *stream) {
    HIP_TRY(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeGlobal),
            "hipStreamBeginCapture");
    return MCDP_OK;
}
int mcdp_graph_capture_end(void *stream, void **graph_exec) {
    if (!graph_exec) return fail(MCDP_E_INVALID, "graph_exec is null");
    *graph_exec = nullptr;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g), "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *graph_exec = x;
    return MCDP_OK;
}
int mcdp_graph_launch(void *graph_exec, void *stream) {
    if (!graph_exec) return fail(MCDP_E_INVALID, "graph_exec is null");
    HIP_TRY(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), static_cast<hipStream_t>(stream)),
            "hipGraphLaunch");
    return MCDP_OK;
}
int mcdp_graph_destroy(void *graph_exec) {
    if (graph_exec) HIP_TRY(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)), "hipGraphExecDestroy");
    return MCDP_OK;
}
int mcdp_stream_create(void **stream) {
    if (!stream) return fail(MCDP_E_INVALID, "stream is null");
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreate(&s), "hipStreamCreate");
    *stream = s;
    return MCDP_OK;
}
int mcdp_stream_destroy(void *stream) {
    HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(stream)), "hipStreamDestroy");
    return MCDP_OK;
}
int mcdp_stream_synchronize(void *stream) {
    HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
    return MCDP_OK;
}
int mcdp_event_create(void **event) {
    if (!event) return fail(MCDP_E_INVALID, "event is null");
    hipEvent_t e = nullptr;
    HIP_TRY(hipEventCreate(&e), "hipEventCreate");
    *event = e;
    return MCDP_OK;
}
int mcdp_event_destroy(void *event) {
    HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
    return MCDP_OK;
}
int mcdp_event_record(void *event, void *stream) {
    HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)), "hipEventRecord");
    return MCDP_OK;
}
int mcdp_event_elapsed_ms(void *start, void *stop, float *ms) {
    if (!ms) return fail(MCDP_E_INVALID, "ms is null");
    HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)),
            "hipEventElapsedTime");
    return MCDP_OK;
}

int mcdp_debug_copy(double *dst, const double *src, int64_t n, void *stream) {
    if (n < 0 || (n > 0 && (!dst || !src))) return fail(MCDP_E_INVALID, "invalid arguments");
    HIP_TRY(launch_copy(dst, src, n, static_cast<hipStream_t>(stream)), "copy kernel launch");
    return MCDP_OK;
}

int mcdp_debug_philox(int variant, int64_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    if (variant != 2 && variant != 4) return fail(MCDP_E_INVALID, "variant must be 2 or 4");
    if (n < 0 || (n > 0 && (!ctr || !key || !out))) return fail(MCDP_E_INVALID, "invalid arguments");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(MCDP_E_DEVICE, "no HIP device available");
    if (n == 0) return MCDP_OK;
    const size_t cw = (size_t)variant, kw = (size_t)(variant / 2);
    uint32_t *d = nullptr;
    const size_t bytes = sizeof(uint32_t) * (size_t)n * (cw + kw + cw);
    HIP_TRY(hipMalloc(&d, bytes), "hipMalloc");
    uint32_t *dc = d, *dk = d + (size_t)n * cw, *dout = dk + (size_t)n * kw;
    hipError_t e = hipMemcpy(dc, ctr, sizeof(uint32_t) * n * cw, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dk, key, sizeof(uint32_t) * n * kw, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_philox_debug(variant, n, dc, dk, dout, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(uint32_t) * n * cw, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "philox debug");
    return MCDP_OK;
}

int mcdp_debug_math(int fn, int64_t n, const double *x, double *out) {
    if (fn < 0 || fn > 2) return fail(MCDP_E_INVALID, "fn must be 0 (log), 1 (cos2pi) or 2 (exp)");
    if (n < 0 || (n > 0 && (!x || !out))) return fail(MCDP_E_INVALID, "invalid arguments");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(MCDP_E_DEVICE, "no HIP device available");
    if (n == 0) return MCDP_OK;
    double *d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(double) * 2 * (size_t)n), "hipMalloc");
    hipError_t e = hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_math_debug(fn, n, d, d + n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, d + n, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "math debug");
    return MCDP_OK;
}

}  // extern "C"
