// Test-only harness around the reference Monte Carlo engine (oracle infrastructure).
//
// This translation unit compiles the reference's own engine source
// (src/mc_dagprop/monte_carlo/_core.cpp) in place, by include path, and exposes a
// flat C ABI so the test suite can drive the *real* reference `Simulator` from
// ctypes: same context compile (_core.cpp:193-307), same `run` (_core.cpp:312-353),
// same Xoshiro256++/libstdc++ streams.  Nothing here is part of the product; it
// is only ever loaded by tests/, the golden-fixture generator and the
// cpu_baseline leg of bench.py.
//
// `Simulator::run_many` (_core.cpp:355-361) releases the GIL through pybind11 and
// therefore cannot be entered from a plain C caller; `ref_run_many_timed`
// performs the same loop (results.emplace_back(run(s)) for every seed) without
// the GIL release, one Simulator copy per thread as in
// test/test_simulator.py:201-215.
#include "_core.cpp"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>

namespace {

void write_err(char *err, int errlen, const char *msg) {
    if (err && errlen > 0) {
        std::strncpy(err, msg, size_t(errlen) - 1);
        err[errlen - 1] = '\0';
    }
}

}  // namespace

extern "C" {

// Kinds of delay distribution, in the order of the reference's DistVar
// alternatives (_core.cpp:143).
enum { REF_CONSTANT = 0, REF_EXPONENTIAL = 1, REF_GAMMA = 2, REF_EMP_ABS = 3, REF_EMP_REL = 4 };

// Builds DagContext + GenericDelayGenerator exactly as the pybind11 casters
// would (_core.cpp:423-443, 519-542) and constructs a reference Simulator.
// Distributions are registered in the given order, so the last one per type
// wins (README.md:44-45).  Returns nullptr and fills `err` on any exception.
void *ref_sim_create(int n_events, const double *earliest, const double *latest, const double *actual,
                     int n_acts, const int *act_src, const int *act_dst, const int *act_idx,
                     const double *act_dur, const int *act_type, int n_prec, const int *prec_target,
                     const int64_t *prec_off, const int *prec_src, const int *prec_edge, double max_delay,
                     int n_dists, const int *dist_type, const int *dist_kind, const double *dist_params,
                     const int64_t *table_off, const double *table_values, const double *table_weights,
                     char *err, int errlen) {
    try {
        vector<Event> events;
        events.reserve(n_events);
        for (int i = 0; i < n_events; ++i)
            events.push_back(Event{std::to_string(i), EventTimestamp{earliest[i], latest[i], actual[i]}});
        unordered_map<pair<EventIndex, EventIndex>, Activity> amap;
        for (int i = 0; i < n_acts; ++i)
            amap[{act_src[i], act_dst[i]}] = Activity{act_idx[i], act_dur[i], act_type[i]};
        vector<pair<EventIndex, Preds>> plist;
        plist.reserve(n_prec);
        for (int i = 0; i < n_prec; ++i) {
            Preds p;
            for (int64_t k = prec_off[i]; k < prec_off[i + 1]; ++k) p.emplace_back(prec_src[k], prec_edge[k]);
            plist.emplace_back(prec_target[i], std::move(p));
        }
        DagContext ctx(std::move(events), std::move(amap), std::move(plist), max_delay);

        GenericDelayGenerator gen;
        for (int i = 0; i < n_dists; ++i) {
            const double *p = dist_params + 3 * i;
            const int t = dist_type[i];
            switch (dist_kind[i]) {
                case REF_CONSTANT: gen.add_constant(t, p[0]); break;
                case REF_EXPONENTIAL: gen.add_exponential(t, p[0], p[1]); break;
                case REF_GAMMA: gen.add_gamma(t, p[0], p[1], p[2]); break;
                case REF_EMP_ABS:
                case REF_EMP_REL: {
                    std::vector<double> v(table_values + table_off[i], table_values + table_off[i + 1]);
                    std::vector<double> w(table_weights + table_off[i], table_weights + table_off[i + 1]);
                    if (dist_kind[i] == REF_EMP_ABS)
                        gen.dist_map_[t] = EmpiricalAbsoluteDist{std::move(v), std::move(w)};
                    else
                        gen.dist_map_[t] = EmpiricalRelativeDist{std::move(v), std::move(w)};
                    break;
                }
                default: throw std::runtime_error("unknown distribution kind");
            }
        }
        return new Simulator(std::move(ctx), std::move(gen));
    } catch (const std::exception &e) {
        write_err(err, errlen, e.what());
        return nullptr;
    }
}

void ref_sim_destroy(void *sim) { delete static_cast<Simulator *>(sim); }
int ref_sim_node_count(void *sim) { return static_cast<Simulator *>(sim)->node_count(); }
int ref_sim_activity_count(void *sim) { return static_cast<Simulator *>(sim)->activity_count(); }

// Calls Simulator::run(seed) for every seed, in order, on the one Simulator
// (so the gamma normal cache carries across seeds exactly as it does inside
// run_many).  Outputs are sample-major: realized[S][E], durations[S][L],
// cause[S][E].
void ref_sim_run(void *sim_, const int *seeds, int64_t n, double *realized, double *durations, int *cause) {
    Simulator &sim = *static_cast<Simulator *>(sim_);
    const int E = sim.node_count(), L = sim.activity_count();
    for (int64_t i = 0; i < n; ++i) {
        SimResult r = sim.run(seeds[i]);
        std::memcpy(realized + i * E, r.realized.data(), sizeof(double) * E);
        std::memcpy(durations + i * L, r.durations.data(), sizeof(double) * L);
        std::memcpy(cause + i * E, r.cause_event.data(), sizeof(int) * E);
    }
}

// Wall time (seconds) of the run_many loop over `n` seeds split contiguously
// over `threads` workers, each with its own copy of the Simulator.  The
// SimResult vectors are materialised exactly as run_many materialises them.
double ref_run_many_timed(void *sim_, const int *seeds, int64_t n, int threads) {
    const Simulator &proto = *static_cast<Simulator *>(sim_);
    if (threads < 1) threads = 1;
    std::vector<Simulator> sims(threads, proto);
    std::vector<std::thread> pool;
    auto t0 = std::chrono::steady_clock::now();
    for (int w = 0; w < threads; ++w) {
        const int64_t lo = n * w / threads, hi = n * (w + 1) / threads;
        pool.emplace_back([&, w, lo, hi] {
            std::vector<SimResult> results;
            results.reserve(size_t(hi - lo));
            for (int64_t i = lo; i < hi; ++i) results.emplace_back(sims[w].run(seeds[i]));
        });
    }
    for (auto &t : pool) t.join();
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
