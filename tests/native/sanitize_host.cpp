// Host-side checks under AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Built by tests/native/Makefile from the library's host sources (the DAG
// compiler, schedule builder, reverse-plan builder, statistics and serving
// bookkeeping) and the C oracle, all instrumented; the device code is linked
// uninstrumented and never runs here (no GPU: every run entry point must fail
// cleanly with MCDP_E_DEVICE).  For random DAGs with the reference's unusual
// inputs -- repeated precedence targets, gaps in activity indices, edges
// outside the link table, unregistered and reserved types, cycles, events
// out of range -- it checks that library and oracle accept and reject the same
// contexts, builds every schedule shape and the reverse plan, round-trips the
// times, sizes serving requests, and runs the oracle's propagation (both
// layouts) and gradient on durations with NaN, infinities and signed zeros.
// Any out-of-bounds access, use after free or undefined operation aborts with
// a report; the last line is SANITIZE_OK.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <set>
#include <utility>
#include <vector>

#include "mcdp.h"

extern "C" {
typedef struct orc_ctx orc_ctx;
int orc_compile(int E, const double *earliest, int n_acts, const int *act_idx, const double *act_dur,
                const int *act_type, int n_prec, const int *prec_target, const int64_t *prec_off,
                const int *prec_src, const int *prec_edge, double max_delay, int n_reg, const int *reg_types,
                orc_ctx **out, char *err, int errlen);
void orc_free(orc_ctx *c);
int orc_event_count(const orc_ctx *c);
int orc_link_count(const orc_ctx *c);
void orc_propagate(const orc_ctx *c, const double *durations, int64_t ld_dur, int64_t n, double *realized,
                   int *cause);
void orc_propagate_columns(const orc_ctx *c, const double *durations, int64_t ld_dur, int64_t n, double *realized,
                           int *cause, int64_t ld_out);
int orc_gradient(const orc_ctx *c, const double *realized, const int *cause, const double *durations,
                 int64_t ld_dur, const double *mult1, int64_t n, const double *weights, const int *link_rel,
                 double *obj, double *grad_dmin, double *grad_earliest, char *err, int errlen);
}

namespace {

struct Rng {  // SplitMix64
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1p-53; }
    int below(int n) { return n <= 1 ? 0 : (int)(next() % (uint64_t)n); }
};

int failures = 0, accepted = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);            \
            std::fprintf(stderr, "\n");                   \
            ++failures;                                   \
        }                                                 \
    } while (0)

struct Dag {
    int E = 0;
    std::vector<double> earliest;
    std::vector<int> idx, type;
    std::vector<double> dur;
    std::vector<int> target, src, edge;
    std::vector<int64_t> off{0};
    double max_delay = 0.0;
};

// A random DAG over a random topological permutation (tests/dags.py's shape):
// options add repeated targets, index gaps, out-of-table edges, a cycle, an
// event out of range.
Dag random_dag(Rng &r, int n, bool duplicates, bool gaps, bool oob, bool cycle, bool out_of_range) {
    Dag d;
    d.E = n;
    std::vector<int> perm((size_t)n);
    for (int i = 0; i < n; ++i) perm[(size_t)i] = i;
    for (int i = n - 1; i > 0; --i) std::swap(perm[(size_t)i], perm[(size_t)r.below(i + 1)]);
    for (int i = 0; i < n; ++i) d.earliest.push_back(std::round(r.uniform() * 50.0));
    std::set<std::pair<int, int>> have;
    int next_idx = 0;
    std::vector<std::pair<int, std::vector<std::pair<int, int>>>> prec;
    for (int pos = 1; pos < n; ++pos) {
        const int e = perm[(size_t)pos];
        if (r.uniform() < 0.15) continue;
        const int k = 1 + r.below(4);
        std::vector<std::pair<int, int>> preds;
        for (int j = 0; j < k && j < pos; ++j) {
            const int s = perm[(size_t)r.below(pos)];
            if (have.count({s, e})) continue;
            have.insert({s, e});
            next_idx += (gaps && r.uniform() < 0.2) ? 2 + r.below(2) : 1;
            const int a = next_idx - 1;
            d.idx.push_back(a);
            d.dur.push_back(std::round(r.uniform() * 20.0));
            d.type.push_back(-1 + r.below(6));
            preds.push_back({s, a});
        }
        if (oob && !preds.empty() && r.uniform() < 0.1)
            preds.push_back({perm[0], r.uniform() < 0.5 ? next_idx + 50 : -3});
        prec.push_back({e, preds});
    }
    if (duplicates && !prec.empty())
        for (size_t c = 0; c < std::max<size_t>(1, prec.size() / 10); ++c) {
            const auto entry = prec[(size_t)r.below((int)prec.size())];
            std::vector<std::pair<int, int>> sub;
            for (const auto &p : entry.second)
                if (r.uniform() < 0.5) sub.push_back(p);
            prec.insert(prec.begin() + r.below((int)prec.size() + 1), {entry.first, sub});
        }
    if (cycle && n >= 2) prec.push_back({perm[0], {{perm[(size_t)n - 1], 0}}});
    if (out_of_range && !prec.empty()) prec[0].second.push_back({n + 3, 0});
    for (int i = (int)prec.size() - 1; i > 0; --i) std::swap(prec[(size_t)i], prec[(size_t)r.below(i + 1)]);
    for (const auto &entry : prec) {
        d.target.push_back(entry.first);
        for (const auto &p : entry.second) {
            d.src.push_back(p.first);
            d.edge.push_back(p.second);
        }
        d.off.push_back((int64_t)d.src.size());
    }
    const double delays[] = {0.0, 5.0, 30.0, 1e6};
    d.max_delay = delays[r.below(4)];
    return d;
}

// Adds distributions for types 0..4 (and, sometimes, the reserved type -1).
void add_dists(Rng &r, mcdp_generator *g, std::vector<int> &reg, bool reserved) {
    const double inf = std::numeric_limits<double>::infinity();
    for (int t = 0; t < 5; ++t) {
        int rc = 0;
        switch (r.below(7)) {
            case 0: rc = mcdp_generator_add_constant(g, t, 1.5); break;
            case 1: rc = mcdp_generator_add_exponential(g, t, 0.3, r.uniform() < 0.5 ? inf : 0.4); break;
            case 2: rc = mcdp_generator_add_gamma(g, t, 2.0, 0.2, inf); break;
            case 3: rc = mcdp_generator_add_gamma(g, t, 0.4, 1.0, r.uniform() < 0.5 ? inf : 0.05); break;
            case 4: {
                const double v[4] = {0.0, 0.5, 1.0, 2.0}, w[4] = {1.0, 2.0, 0.0, 1.0};
                rc = mcdp_generator_add_empirical_relative(g, t, v, w, 1 + r.below(4));
                break;
            }
            case 5: {
                const double v[3] = {0.0, 2.0, 5.0}, w[3] = {1.0, 1.0, 2.0};
                rc = mcdp_generator_add_empirical_absolute(g, t, v, w, 1 + r.below(3));
                break;
            }
            default: continue;  // type t stays unregistered
        }
        CHECK(rc == MCDP_OK, "add_* type %d: %s", t, mcdp_last_error());
        reg.push_back(t);
    }
    if (reserved) {
        CHECK(mcdp_generator_add_constant(g, -1, 1.0) == MCDP_OK, "reserved type is rejected at compile time");
        reg.push_back(-1);
    }
}

void check_dag(Rng &r, int trial) {
    const int n = 1 + r.below(90);
    const bool cycle = r.uniform() < 0.05, out_of_range = r.uniform() < 0.05;
    Dag d = random_dag(r, n, r.uniform() < 0.4, r.uniform() < 0.4, r.uniform() < 0.4, cycle, out_of_range);
    if (r.uniform() < 0.03) d.max_delay = -1.0;
    mcdp_generator *g = nullptr;
    CHECK(mcdp_generator_create(&g) == MCDP_OK, "generator");
    std::vector<int> reg;
    add_dists(r, g, reg, r.uniform() < 0.03);
    mcdp_propagator *p = nullptr;
    const int rc = mcdp_propagator_create(d.E, d.earliest.data(), (int32_t)d.idx.size(), d.idx.data(), d.dur.data(),
                                          d.type.data(), (int32_t)d.target.size(), d.target.data(), d.off.data(),
                                          d.src.data(), d.edge.data(), d.max_delay, g, &p);
    orc_ctx *o = nullptr;
    char err[256];
    const int orc = orc_compile(d.E, d.earliest.data(), (int)d.idx.size(), d.idx.data(), d.dur.data(), d.type.data(),
                                (int)d.target.size(), d.target.data(), d.off.data(), d.src.data(), d.edge.data(),
                                d.max_delay, (int)reg.size(), reg.data(), &o, err, sizeof err);
    CHECK((rc == MCDP_OK) == (orc == 0), "trial %d: library rc %d (%s), oracle rc %d (%s)", trial, rc,
          mcdp_last_error(), orc, err);
    if (rc == MCDP_OK && orc == 0) {
        ++accepted;
        const int32_t E = mcdp_propagator_node_count(p), L = mcdp_propagator_activity_count(p);
        CHECK(E == orc_event_count(o) && L == orc_link_count(o), "trial %d: counts", trial);
        int64_t info[12];
        CHECK(mcdp_propagator_plan_info(p, info, 12) == MCDP_OK, "plan_info");
        const int waves[] = {1, 2, 3, 8, 16}, ks[] = {4, 8, 12, 16};
        for (int w : waves)
            for (int k : ks)
                for (int spl = 1; spl <= 2; ++spl) {
                    const int c = mcdp_propagator_configure(p, w, k, spl);
                    CHECK(c == MCDP_OK || (spl == 2 && k > 8), "configure %d %d %d: %s", w, k, spl, mcdp_last_error());
                    if (c == MCDP_OK) CHECK(mcdp_propagator_plan_info(p, info, 12) == MCDP_OK, "plan_info");
                }
        int64_t ginfo[6];
        CHECK(mcdp_propagator_gradient_info(p, ginfo, 6) == MCDP_OK, "gradient_info: %s", mcdp_last_error());
        std::vector<int64_t> poff((size_t)E + 1);
        CHECK(mcdp_propagator_pred_offsets(p, poff.data()) == MCDP_OK, "pred_offsets");
        std::vector<double> dmin((size_t)L), e0((size_t)E), dmin2((size_t)L), e2((size_t)E);
        CHECK(mcdp_propagator_get_times(p, dmin.data(), e0.data()) == MCDP_OK, "get_times");
        for (auto &x : dmin) x *= 1.0 + r.uniform();
        CHECK(mcdp_propagator_set_times(p, dmin.data(), e0.data()) == MCDP_OK, "set_times");
        CHECK(mcdp_propagator_get_times(p, dmin2.data(), e2.data()) == MCDP_OK, "get_times");
        CHECK(dmin2 == dmin && e2 == e0, "times round trip");
        CHECK(mcdp_propagator_gradient_info(p, ginfo, 6) == MCDP_OK, "gradient_info after set_times");
        // bounds for every seed: the event ranges are the oracle's propagation of the link ranges' two ends
        {
            std::vector<double> llo((size_t)std::max(L, 1)), lhi((size_t)std::max(L, 1));
            std::vector<double> elo((size_t)std::max(E, 1)), ehi((size_t)std::max(E, 1));
            const int rrc = mcdp_propagator_ranges(p, llo.data(), lhi.data(), elo.data(), ehi.data());
            bool finite = std::isfinite(d.max_delay);
            for (double x : d.earliest) finite = finite && std::isfinite(x);
            CHECK((rrc == MCDP_OK) == finite, "trial %d: ranges rc %d", trial, rrc);
            if (rrc == MCDP_OK && E > 0) {
                std::vector<double> ends((size_t)std::max(L, 1) * 2), r2((size_t)E * 2);
                std::vector<int> c2((size_t)E * 2);
                for (int l = 0; l < L; ++l) {
                    CHECK(!(llo[(size_t)l] > lhi[(size_t)l]), "trial %d: link %d range", trial, l);
                    ends[(size_t)l] = llo[(size_t)l];
                    ends[(size_t)(L + l)] = lhi[(size_t)l];
                }
                orc_propagate(o, ends.data(), L, 2, r2.data(), c2.data());
                for (int e = 0; e < E; ++e)
                    CHECK(std::memcmp(&r2[(size_t)e], &elo[(size_t)e], 8) == 0 &&
                              std::memcmp(&r2[(size_t)(E + e)], &ehi[(size_t)e], 8) == 0,
                          "trial %d: event %d range", trial, e);
            }
        }
        // no device here: every run entry point fails cleanly
        int32_t seeds[3] = {1, 2, 3};
        std::vector<double> rr((size_t)E * 3), dd((size_t)L * 3);
        std::vector<int32_t> cc((size_t)E * 3);
        CHECK(mcdp_run_many_host(p, seeds, 3, rr.data(), dd.data(), cc.data()) == MCDP_E_DEVICE, "run without GPU");
        // the oracle on special durations, both layouts, and its gradient
        const int64_t m = 5;
        std::vector<double> dur((size_t)(m * std::max(L, 1)));
        for (auto &x : dur) {
            const double u = r.uniform();
            x = u < 0.03 ? std::nan("") : u < 0.05 ? INFINITY : u < 0.07 ? -INFINITY : u < 0.1 ? -0.0
                                                                                 : std::round(r.uniform() * 22 - 2);
        }
        std::vector<double> ro((size_t)(m * std::max(E, 1))), rc2((size_t)(m * std::max(E, 1)));
        std::vector<int> co((size_t)(m * std::max(E, 1))), cc2((size_t)(m * std::max(E, 1)));
        orc_propagate(o, dur.data(), L, m, ro.data(), co.data());
        std::vector<double> durT((size_t)(m * std::max(L, 1)));
        for (int64_t s = 0; s < m; ++s)
            for (int l = 0; l < L; ++l) durT[(size_t)(l * m + s)] = dur[(size_t)(s * L + l)];
        orc_propagate_columns(o, durT.data(), m, m, rc2.data(), cc2.data(), m);
        for (int64_t s = 0; s < m; ++s)
            for (int e = 0; e < E; ++e) {
                const double a = ro[(size_t)(s * E + e)], b = rc2[(size_t)(e * m + s)];
                CHECK(std::memcmp(&a, &b, 8) == 0 && co[(size_t)(s * E + e)] == cc2[(size_t)(e * m + s)],
                      "trial %d: layouts differ", trial);
            }
        for (auto &x : dur) x = std::round(r.uniform() * 20.0);  // the gradient needs finite durations
        orc_propagate(o, dur.data(), L, m, ro.data(), co.data());
        std::vector<double> w((size_t)std::max(E, 1), 1.0), gd((size_t)std::max(L, 1)), ge((size_t)std::max(E, 1));
        std::vector<int> rel((size_t)std::max(L, 1), 1);
        double obj = 0.0;
        std::vector<double> mult1((size_t)(m * std::max(L, 1)), 1.5);
        CHECK(orc_gradient(o, ro.data(), co.data(), dur.data(), L, mult1.data(), m, w.data(), rel.data(), &obj,
                           gd.data(), ge.data(), err, sizeof err) == 0,
              "trial %d: gradient: %s", trial, err);
    }
    if (p) mcdp_propagator_destroy(p);
    if (o) orc_free(o);
    mcdp_generator_destroy(g);
}

void check_arguments() {
    // null and invalid arguments fail without touching memory
    CHECK(mcdp_propagator_plan_info(nullptr, nullptr, 0) == MCDP_E_INVALID, "null plan_info");
    CHECK(mcdp_propagator_configure(nullptr, 4, 8, 1) == MCDP_E_INVALID, "null configure");
    // distributions the engine cannot sample are refused when a propagator is compiled with them
    const double v[2] = {1.0, 2.0}, w_bad[2] = {-1.0, 1.0}, w_zero[2] = {0.0, 0.0};
    const double one[1] = {1.0}, zero_idx_dur[1] = {2.0};
    const int32_t idx0[1] = {0}, type0[1] = {0}, tgt1[1] = {1}, src0[1] = {0}, edge0[1] = {0};
    const int64_t off1[2] = {0, 1};
    const double e2[2] = {0.0, 1.0};
    for (int bad = 0; bad < 4; ++bad) {
        mcdp_generator *gb = nullptr;
        CHECK(mcdp_generator_create(&gb) == MCDP_OK, "generator");
        if (bad == 0) (void)mcdp_generator_add_empirical_absolute(gb, 0, v, w_bad, 2);
        if (bad == 1) (void)mcdp_generator_add_empirical_absolute(gb, 0, v, w_zero, 2);
        if (bad == 2) (void)mcdp_generator_add_empirical_absolute(gb, 0, v, w_zero, 0);
        if (bad == 3) (void)mcdp_generator_add_exponential(gb, 0, 1.0, -1.0);
        mcdp_propagator *pb = nullptr;
        const int rc = mcdp_propagator_create(2, e2, 1, idx0, zero_idx_dur, type0, 1, tgt1, off1, src0, edge0, 10.0,
                                              gb, &pb);
        CHECK(rc != MCDP_OK && pb == nullptr, "unsamplable distribution %d accepted", bad);
        mcdp_propagator_destroy(pb);
        mcdp_generator_destroy(gb);
    }
    (void)one;
    mcdp_generator *g = nullptr;
    CHECK(mcdp_generator_create(&g) == MCDP_OK, "generator");
    // an empty DAG and a DAG of isolated events
    mcdp_propagator *p = nullptr;
    CHECK(mcdp_propagator_create(0, nullptr, 0, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr,
                                 10.0, g, &p) == MCDP_OK,
          "empty DAG: %s", mcdp_last_error());
    int64_t info[12];
    if (p) CHECK(mcdp_propagator_plan_info(p, info, 12) == MCDP_OK, "empty plan");
    mcdp_propagator_destroy(p);
    const double e3[3] = {1.0, 2.0, 3.0};
    const int64_t off0[1] = {0};
    p = nullptr;
    CHECK(mcdp_propagator_create(3, e3, 0, nullptr, nullptr, nullptr, 0, nullptr, off0, nullptr, nullptr, 10.0, g,
                                 &p) == MCDP_OK,
          "isolated events");
    int64_t ginfo[6];
    if (p) CHECK(mcdp_propagator_gradient_info(p, ginfo, 6) == MCDP_OK, "isolated gradient plan");
    mcdp_propagator_destroy(p);
    mcdp_generator_destroy(g);
    // serving's footprint rule over a grid of sizes
    for (int64_t E : {1, 1000, 200000})
        for (int64_t n : {1, 127, 128, 100000, 262144})
            for (int64_t cap : {(int64_t)1024, (int64_t)1 << 20, (int64_t)1 << 30, (int64_t)1 << 36})
                for (int32_t entries : {4, 300}) {
                    int64_t chunk = -1, bytes = -1;
                    const int rc = mcdp_debug_serve_chunk(n, E, 3 * E, entries, 3, cap, &chunk, &bytes);
                    CHECK(rc == MCDP_OK || rc == MCDP_E_NOMEM, "serve chunk rc %d", rc);
                    if (rc == MCDP_OK) CHECK(bytes <= cap && chunk >= 0 && chunk <= n, "serve chunk bounds");
                }
}

}  // namespace

int main(int argc, char **argv) {
    const int trials = argc > 1 ? std::atoi(argv[1]) : 400;
    Rng r{20260099};
    check_arguments();
    for (int t = 0; t < trials; ++t) check_dag(r, t);
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("SANITIZE_OK %d trials, %d compiled\n", trials, accepted);
    return 0;
}
