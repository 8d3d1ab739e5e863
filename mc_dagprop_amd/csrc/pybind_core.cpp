// Python module mc_dagprop_amd.monte_carlo._core: the reference's pybind11
// API (WonJayne/mc_dagprop, src/mc_dagprop/monte_carlo/_core.cpp:365-552,
// typed by _core.pyi) over the C ABI of libmcdp (include/mcdp.h).
//
// Classes, constructors, keyword names, attribute names, repr formats, the
// frozen-dataclass decoration and the exception types follow the reference;
// the Monte Carlo work itself runs in libmcdp's gfx950 kernels.
//
// MCDP_TYPES_MODULE names the module whose NewType aliases annotate the data
// classes.  Built as this package's module it is mc_dagprop_amd.types; built
// as the reference's own mc_dagprop.monte_carlo._core (a drop-in for the
// extension of _core.cpp) pass -DMCDP_TYPES_MODULE="mc_dagprop.types", the
// module the reference imports at _core.cpp:453.
#ifndef MCDP_TYPES_MODULE
#define MCDP_TYPES_MODULE "mc_dagprop_amd.types"
#endif
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <sys/mman.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/mcdp.h"

namespace py = pybind11;

namespace {

// Hash of an activity-map key.  Same combination as the reference
// (_core.cpp:28-33) so that DagContext.activities iterates in the same order.
struct PairHash {
    size_t operator()(std::pair<int, int> const &p) const noexcept {
        return std::hash<int>{}(p.first) ^ (std::hash<int>{}(p.second) << 1);
    }
};

using Preds = std::vector<std::pair<int, int>>;

struct EventTimestamp {
    double earliest, latest, actual;
};
struct Event {
    std::string event_id;
    EventTimestamp ts;
};
struct Activity {
    int idx;
    double duration;
    int activity_type;
};
struct DagContext {
    std::vector<Event> events;
    std::unordered_map<std::pair<int, int>, Activity, PairHash> activity_map;
    std::vector<std::pair<int, Preds>> precedence_list;
    double max_delay;
    DagContext(std::vector<Event> ev, std::unordered_map<std::pair<int, int>, Activity, PairHash> am,
               std::vector<std::pair<int, Preds>> pl, double md)
        : events(std::move(ev)), activity_map(std::move(am)), precedence_list(std::move(pl)), max_delay(md) {}
};

[[noreturn]] void throw_last() { throw std::runtime_error(mcdp_last_error()); }
void check(int rc) {
    if (rc != MCDP_OK) throw_last();
}

// An array whose elements start uninitialized: the results overwrite every
// one, so there is no zero-filling pass over hundreds of megabytes first.
// Arrays of 4 MB and more sit on 2 MB boundaries and ask for transparent huge
// pages, so their first touch takes one fault per 2 MB instead of per 4 KB.
template <typename T>
struct RawArray {
    struct Free {
        void operator()(T *q) const { std::free(q); }
    };
    std::unique_ptr<T, Free> p;
    void resize(size_t n) {
        constexpr size_t kHuge = (size_t)2 << 20;
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        void *q = nullptr;
        if (bytes >= 2 * kHuge) {
            const size_t whole = (bytes + kHuge - 1) / kHuge * kHuge;
            q = std::aligned_alloc(kHuge, whole);
            if (q) madvise(q, whole, MADV_HUGEPAGE);  // advice only: a refusal leaves small pages
        } else {
            q = std::malloc(bytes);
        }
        if (!q) throw std::bad_alloc();
        p.reset(static_cast<T *>(q));
    }
    T *data() { return p.get(); }
    const T *data() const { return p.get(); }
};

// Results of one run_many call, event-major with ld = n (column s is sample s).
struct Batch {
    int64_t n = 0, E = 0, L = 0;
    RawArray<double> realized, durations;
    RawArray<int32_t> cause;
};

struct SimResult {
    std::shared_ptr<Batch> batch;
    int64_t index = 0;
};

struct Generator {
    mcdp_generator *g = nullptr;
    Generator() { check(mcdp_generator_create(&g)); }
    ~Generator() { mcdp_generator_destroy(g); }
    Generator(const Generator &) = delete;
    Generator &operator=(const Generator &) = delete;
};

struct Propagator {
    mcdp_propagator *p = nullptr;
    Propagator(const DagContext &ctx, const Generator &gen) {
        const int32_t E = (int32_t)ctx.events.size();
        std::vector<double> earliest((size_t)E);
        for (int32_t i = 0; i < E; ++i) earliest[(size_t)i] = ctx.events[(size_t)i].ts.earliest;
        std::vector<int32_t> idx, type;
        std::vector<double> dur;
        idx.reserve(ctx.activity_map.size());
        for (const auto &kv : ctx.activity_map) {
            idx.push_back(kv.second.idx);
            dur.push_back(kv.second.duration);
            type.push_back(kv.second.activity_type);
        }
        std::vector<int32_t> target, psrc, pedge;
        std::vector<int64_t> off(1, 0);
        for (const auto &entry : ctx.precedence_list) {
            target.push_back(entry.first);
            for (const auto &pr : entry.second) {
                psrc.push_back(pr.first);
                pedge.push_back(pr.second);
            }
            off.push_back((int64_t)psrc.size());
        }
        check(mcdp_propagator_create(E, earliest.data(), (int32_t)idx.size(), idx.data(), dur.data(), type.data(),
                                     (int32_t)target.size(), target.data(), off.data(), psrc.data(), pedge.data(),
                                     ctx.max_delay, gen.g, &p));
    }
    ~Propagator() { mcdp_propagator_destroy(p); }
    Propagator(const Propagator &) = delete;
    Propagator &operator=(const Propagator &) = delete;

    int node_count() const { return mcdp_propagator_node_count(p); }
    int activity_count() const { return mcdp_propagator_activity_count(p); }

    std::shared_ptr<Batch> run_batch(const std::vector<int> &seeds) {
        auto b = std::make_shared<Batch>();
        b->n = (int64_t)seeds.size();
        b->E = node_count();
        b->L = activity_count();
        b->realized.resize((size_t)(b->E * b->n));
        b->durations.resize((size_t)(b->L * b->n));
        b->cause.resize((size_t)(b->E * b->n));
        std::vector<int32_t> s(seeds.begin(), seeds.end());
        int rc;
        {
            py::gil_scoped_release release;
            rc = mcdp_run_many_host(p, s.data(), b->n, b->realized.data(), b->durations.data(), b->cause.data());
        }
        check(rc);
        return b;
    }
};

// Column `index` of an event-major [rows][n] buffer as a 1-D view whose base
// keeps `owner` alive.
template <typename T>
py::array column_view(T *base, int64_t rows, int64_t n, int64_t index, py::handle owner) {
    return py::array(py::dtype::of<T>(), std::vector<py::ssize_t>{(py::ssize_t)rows},
                     std::vector<py::ssize_t>{(py::ssize_t)(n * (int64_t)sizeof(T))}, base + index, owner);
}

// Whole batch as an (n, rows) array (a transposed view of the [rows][n] buffer).
template <typename T>
py::array batch_view(RawArray<T> &v, int64_t rows, int64_t n, py::handle owner) {
    return py::array(py::dtype::of<T>(), std::vector<py::ssize_t>{(py::ssize_t)n, (py::ssize_t)rows},
                     std::vector<py::ssize_t>{(py::ssize_t)sizeof(T), (py::ssize_t)(n * (int64_t)sizeof(T))},
                     v.data(), owner);
}

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "Monte Carlo DAG propagation on MI355X (gfx950 HIP kernels behind the mc_dagprop API)";

    py::class_<EventTimestamp> ts_cls(m, "EventTimestamp",
                                     "Earliest and latest times of an event and its scheduled time.");
    ts_cls
        .def(py::init<double, double, double>(), py::arg("earliest"), py::arg("latest"), py::arg("actual"),
             "Create an event timestamp (earliest, latest, actual).")
        .def_readwrite("earliest", &EventTimestamp::earliest, "Earliest bound")
        .def_readwrite("latest", &EventTimestamp::latest, "Latest bound")
        .def_readwrite("actual", &EventTimestamp::actual, "Scheduled time")
        .def("__repr__", [](const EventTimestamp &ts) {
            return py::str("EventTimestamp(earliest={}, latest={}, actual={})").format(ts.earliest, ts.latest, ts.actual);
        });

    py::class_<Event> event_cls(m, "Event", "An event (node): identifier and timestamp window.");
    event_cls
        .def(py::init<std::string, EventTimestamp>(), py::arg("event_id"), py::arg("timestamp"),
             "An event node with its ID and timestamp")
        .def_readwrite("event_id", &Event::event_id, "Node identifier")
        .def_readwrite("timestamp", &Event::ts, "Event timing info")
        .def("__repr__", [](const Event &ev) {
            return py::str("Event(event_id={}, timestamp={})").format(py::repr(py::cast(ev.event_id)), py::repr(py::cast(ev.ts)));
        });

    py::class_<Activity> activity_cls(m, "Activity", "An activity (edge): index, minimal duration and delay type.");
    activity_cls
        .def(py::init<int, double, int>(), py::arg("idx"), py::arg("minimal_duration"), py::arg("activity_type"),
             "An activity (edge) with index, base duration and type")
        .def_readwrite("idx", &Activity::idx, "Index of the activity")
        .def_readwrite("minimal_duration", &Activity::duration, "Base duration")
        .def_readwrite("activity_type", &Activity::activity_type, "Type ID for delay dist.")
        .def("__repr__", [](const Activity &a) {
            return py::str("Activity(idx={}, minimal_duration={}, activity_type={})").format(a.idx, a.duration, a.activity_type);
        });

    py::class_<DagContext> ctx_cls(
        m, "DagContext",
        "The DAG: events, activities keyed by (src, dst), precedence list (any order) and max_delay.");
    ctx_cls
        .def(py::init<std::vector<Event>, std::unordered_map<std::pair<int, int>, Activity, PairHash>,
                      std::vector<std::pair<int, Preds>>, double>(),
             py::arg("events"), py::arg("activities"), py::arg("precedence_list"), py::arg("max_delay"),
             "Wraps a DAG: events, activity_map, precedence_list, max_delay")
        .def_readwrite("events", &DagContext::events)
        .def_readwrite("activities", &DagContext::activity_map)
        .def_readwrite("precedence_list", &DagContext::precedence_list)
        .def_readwrite("max_delay", &DagContext::max_delay)
        .def("__repr__", [](const DagContext &c) {
            return py::str("DagContext(events={}, activities={}, precedence_list={}, max_delay={})")
                .format(py::repr(py::cast(c.events)), py::repr(py::cast(c.activity_map)),
                        py::repr(py::cast(c.precedence_list)), c.max_delay);
        });

    // Same decoration as the reference (_core.cpp:445-488): annotations plus
    // dataclasses.dataclass(frozen=True, slots=True, init=False) applied to the
    // class (its return value is discarded, as there).  The classes carry
    // docstrings, so the decorator does not derive one from an inspect
    // signature (which CPython 3.10 cannot produce for these classes).
    {
        py::object dataclass = py::module_::import("dataclasses").attr("dataclass");
        py::dict opts;
        opts["frozen"] = true;
        opts["slots"] = true;
        opts["init"] = false;
        py::object deco = dataclass(**opts);
        py::module_ types_mod = py::module_::import(MCDP_TYPES_MODULE);
        py::object Second = types_mod.attr("Second");
        py::dict ts_ann;
        ts_ann["earliest"] = Second;
        ts_ann["latest"] = Second;
        ts_ann["actual"] = Second;
        ts_cls.attr("__annotations__") = ts_ann;
        deco(ts_cls);
        py::dict ev_ann;
        ev_ann["event_id"] = types_mod.attr("EventId");
        ev_ann["timestamp"] = ts_cls;
        event_cls.attr("__annotations__") = ev_ann;
        deco(event_cls);
        py::dict act_ann;
        act_ann["idx"] = types_mod.attr("ActivityIndex");
        act_ann["minimal_duration"] = Second;
        act_ann["activity_type"] = types_mod.attr("ActivityType");
        activity_cls.attr("__annotations__") = act_ann;
        deco(activity_cls);
        py::object typing = py::module_::import("typing");
        py::dict ctx_ann;
        ctx_ann["events"] = typing.attr("Sequence");
        ctx_ann["activities"] = typing.attr("Mapping");
        ctx_ann["precedence_list"] = typing.attr("Sequence");
        ctx_ann["max_delay"] = Second;
        ctx_cls.attr("__annotations__") = ctx_ann;
        deco(ctx_cls);
    }

    py::class_<SimResult>(m, "SimResult", py::buffer_protocol())
        .def_buffer([](SimResult &r) -> py::buffer_info {
            Batch &b = *r.batch;
            return py::buffer_info(b.realized.data() + r.index, sizeof(double), py::format_descriptor<double>::format(), 1,
                                   {(py::ssize_t)b.E}, {(py::ssize_t)(b.n * (int64_t)sizeof(double))});
        })
        .def_property_readonly(
            "realized",
            [](py::object self) {
                SimResult &r = self.cast<SimResult &>();
                return column_view(r.batch->realized.data(), r.batch->E, r.batch->n, r.index, self);
            },
            "Final event times as a NumPy array")
        .def_property_readonly(
            "durations",
            [](py::object self) {
                SimResult &r = self.cast<SimResult &>();
                return column_view(r.batch->durations.data(), r.batch->L, r.batch->n, r.index, self);
            },
            "Per-link durations (incl. extra) as a NumPy array")
        .def_property_readonly(
            "cause_event",
            [](py::object self) {
                SimResult &r = self.cast<SimResult &>();
                return column_view(r.batch->cause.data(), r.batch->E, r.batch->n, r.index, self);
            },
            "Index of predecessor causing each event as a NumPy array");

    py::class_<Generator>(m, "GenericDelayGenerator")
        .def(py::init<>(), "Create a new delay-generator")
        .def(
            "set_seed", [](Generator &g, int seed) { check(mcdp_generator_set_seed(g.g, seed)); }, py::arg("seed"),
            "Set RNG seed for reproducibility (runs are keyed by their own seed; this has no effect on results)")
        .def(
            "add_constant", [](Generator &g, int t, double f) { check(mcdp_generator_add_constant(g.g, t, f)); },
            py::arg("activity_type"), py::arg("factor"), "Constant: delay = factor * duration")
        .def(
            "add_exponential",
            [](Generator &g, int t, double lam, double mx) { check(mcdp_generator_add_exponential(g.g, t, lam, mx)); },
            py::arg("activity_type"), py::arg("lambda_"), py::arg("max_scale"), "Exponential(λ) truncated at max_scale")
        .def(
            "add_gamma",
            [](Generator &g, int t, double k, double s, double mx) { check(mcdp_generator_add_gamma(g.g, t, k, s, mx)); },
            py::arg("activity_type"), py::arg("shape"), py::arg("scale"),
            py::arg("max_scale") = std::numeric_limits<double>::infinity(), "Gamma(shape,scale) truncated at max_scale")
        .def(
            "add_empirical_absolute",
            [](Generator &g, int t, std::vector<double> values, std::vector<double> weights) {
                if (values.size() != weights.size())
                    throw std::runtime_error("EmpiricalAbsoluteDist: values and weights must have same length");
                check(mcdp_generator_add_empirical_absolute(g.g, t, values.data(), weights.data(), (int64_t)values.size()));
            },
            py::arg("activity_type"), py::arg("values"), py::arg("weights"),
            "Empirical absolute: draw one of your provided values, weighted by weights.")
        .def(
            "add_empirical_relative",
            [](Generator &g, int t, std::vector<double> factors, std::vector<double> weights) {
                if (factors.size() != weights.size())
                    throw std::runtime_error("EmpiricalRelativeDist: factors and weights must have same length");
                check(mcdp_generator_add_empirical_relative(g.g, t, factors.data(), weights.data(), (int64_t)factors.size()));
            },
            py::arg("activity_type"), py::arg("factors"), py::arg("weights"),
            "Empirical relative: draw a factor in [0,∞), then multiply by the activity duration.");

    py::class_<Propagator>(m, "MonteCarloPropagator")
        .def(py::init<const DagContext &, const Generator &>(), py::arg("context"), py::arg("generator"),
             "Construct simulator with context and delay-generator")
        .def("node_count", &Propagator::node_count, "Number of events")
        .def("activity_count", &Propagator::activity_count, "Number of links")
        .def(
            "run", [](Propagator &p, int seed) { return SimResult{p.run_batch(std::vector<int>{seed}), 0}; },
            py::arg("seed"), "Run single sim")
        .def(
            "run_many",
            [](Propagator &p, const std::vector<int> &seeds) {
                auto b = p.run_batch(seeds);
                std::vector<SimResult> out;
                out.reserve(seeds.size());
                for (int64_t i = 0; i < b->n; ++i) out.push_back(SimResult{b, i});
                return out;
            },
            py::arg("seeds"), "Run batch sims")
        .def(
            "run_many_arrays",
            [](Propagator &p, const std::vector<int> &seeds) {
                auto b = p.run_batch(seeds);
                py::capsule owner(new std::shared_ptr<Batch>(b),
                                  [](void *q) { delete static_cast<std::shared_ptr<Batch> *>(q); });
                return py::make_tuple(batch_view(b->realized, b->E, b->n, owner),
                                      batch_view(b->durations, b->L, b->n, owner),
                                      batch_view(b->cause, b->E, b->n, owner));
            },
            py::arg("seeds"),
            "run_many as three arrays of shape (len(seeds), rows): realized, durations, cause_event")
        .def(
            "propagate",
            [](Propagator &p, py::array_t<double, py::array::c_style | py::array::forcecast> durations) {
                const int64_t E = p.node_count(), L = p.activity_count();
                if (durations.ndim() != 2 || durations.shape(1) != L)
                    throw std::runtime_error("durations must have shape (n_samples, activity_count())");
                const int64_t n = durations.shape(0);
                auto b = std::make_shared<Batch>();
                b->n = n;
                b->E = E;
                b->L = L;
                b->realized.resize((size_t)(E * n));
                b->cause.resize((size_t)(E * n));
                std::vector<double> dT((size_t)(L * n));
                const double *src = durations.data();
                for (int64_t s = 0; s < n; ++s)
                    for (int64_t l = 0; l < L; ++l) dT[(size_t)(l * n + s)] = src[s * L + l];
                int rc;
                {
                    py::gil_scoped_release release;
                    rc = mcdp_propagate_host(p.p, dT.data(), n, b->realized.data(), b->cause.data());
                }
                check(rc);
                py::capsule owner(new std::shared_ptr<Batch>(b),
                                  [](void *q) { delete static_cast<std::shared_ptr<Batch> *>(q); });
                return py::make_tuple(batch_view(b->realized, E, n, owner), batch_view(b->cause, E, n, owner));
            },
            py::arg("durations"),
            "Propagation only, with injected durations of shape (n_samples, activity_count()): "
            "returns realized and cause_event of shape (n_samples, node_count())")
        .def(
            "plan_info",
            [](const Propagator &p) {
                int64_t v[10];
                check(mcdp_propagator_plan_info(p.p, v, 10));
                py::dict d;
                const char *names[10] = {"steps",   "waves",           "lds_slots",           "lds_bytes",
                                         "global_preds", "stale_preds", "orphans",             "max_events_per_wave",
                                         "events_per_wave", "samples_per_lane"};
                for (int i = 0; i < 10; ++i) d[names[i]] = v[i];
                return d;
            },
            "Facts about the compiled device plan")
        .def(
            "run_stats",
            [](Propagator &p, const std::vector<int> &seeds, const std::vector<double> &q) {
                const int64_t E = p.node_count();
                std::vector<int64_t> off((size_t)E + 1);
                check(mcdp_propagator_pred_offsets(p.p, off.data()));
                py::array_t<double> mean(E);
                py::array_t<double> quant(std::vector<py::ssize_t>{(py::ssize_t)E, (py::ssize_t)q.size()});
                py::array_t<int64_t> counts(off[(size_t)E] + E);
                py::array_t<int64_t> coff(E + 1);
                auto co = coff.mutable_unchecked<1>();
                for (int64_t e = 0; e <= E; ++e) co(e) = off[(size_t)e] + e;
                std::vector<int32_t> s(seeds.begin(), seeds.end());
                double *pm = mean.mutable_data(), *pq = quant.mutable_data();
                int64_t *pc = counts.mutable_data();
                int rc;
                {
                    py::gil_scoped_release nogil;
                    rc = mcdp_run_stats_host(p.p, s.data(), (int64_t)s.size(), (int32_t)q.size(), q.data(), pm, pq, pc);
                }
                check(rc);
                py::dict d;
                d["mean"] = mean;
                d["quantiles"] = quant;
                d["cause_counts"] = counts;
                d["cause_offsets"] = coff;
                return d;
            },
            py::arg("seeds"), py::arg("quantiles") = std::vector<double>{0.5, 0.9, 0.99},
            "run_many on the GPU reduced to per-event statistics: mean (n,), quantiles (n, len(quantiles)) "
            "as numpy.quantile's default method, and cause_counts, where event e's counters "
            "cause_counts[cause_offsets[e]:cause_offsets[e+1]] count cause -1 and then each predecessor "
            "position")
        .def(
            "set_times",
            [](Propagator &p, std::optional<std::vector<double>> dmin, std::optional<std::vector<double>> earliest) {
                if (dmin && (int64_t)dmin->size() != p.activity_count())
                    throw std::runtime_error("minimal_durations must have activity_count() entries");
                if (earliest && (int64_t)earliest->size() != p.node_count())
                    throw std::runtime_error("earliest must have node_count() entries");
                check(mcdp_propagator_set_times(p.p, dmin ? dmin->data() : nullptr,
                                                earliest ? earliest->data() : nullptr));
            },
            py::arg("minimal_durations") = py::none(), py::arg("earliest") = py::none(),
            "Replaces the minimal duration of every link (by Activity.idx) and/or the earliest time of every "
            "event; the compiled schedule is kept")
        .def(
            "get_times",
            [](const Propagator &p) {
                py::array_t<double> d(p.activity_count()), e(p.node_count());
                check(mcdp_propagator_get_times(p.p, d.mutable_data(), e.mutable_data()));
                return py::make_tuple(d, e);
            },
            "(minimal duration per link, earliest time per event)")
        .def(
            "gradient",
            [](Propagator &p, const std::vector<int> &seeds, const std::vector<double> &weights, int64_t chunk) {
                if ((int64_t)weights.size() != p.node_count())
                    throw std::runtime_error("weights must have node_count() entries");
                std::vector<int32_t> s(seeds.begin(), seeds.end());
                py::array_t<double> gd(p.activity_count()), ge(p.node_count());
                double obj = 0.0, *pd = gd.mutable_data(), *pe = ge.mutable_data();
                int rc;
                {
                    py::gil_scoped_release nogil;
                    rc = mcdp_gradient_host(p.p, s.data(), (int64_t)s.size(), weights.data(), chunk, &obj, pd, pe);
                }
                check(rc);
                return py::make_tuple(obj, gd, ge);
            },
            py::arg("seeds"), py::arg("weights"), py::arg("chunk") = 0,
            "Mean over the seeds of J = sum_e w_e max(0, realized_e - earliest_e) and its pathwise gradient: "
            "(J, dJ/d minimal_duration per link, dJ/d earliest per event)")
        .def(
            "configure",
            [](Propagator &p, int waves, int events_per_wave, int samples_per_lane) {
                check(mcdp_propagator_configure(p.p, waves, events_per_wave, samples_per_lane));
            },
            py::arg("waves"), py::arg("events_per_wave"), py::arg("samples_per_lane") = 1,
            "Fixes the device schedule shape (waves per workgroup, events per wave and step, samples per "
            "lane); results do not depend on it")
        .def(
            "autotune",
            [](Propagator &p, int64_t n_samples, std::optional<std::string> path) {
                int32_t w = 0, k = 0, spl = 0, hit = 0;
                float ms = 0.f;
                {
                    py::gil_scoped_release nogil;
                    if (path)
                        check(mcdp_propagator_autotune_cached(p.p, n_samples, path->c_str(), &w, &k, &spl, &ms, &hit));
                    else
                        check(mcdp_propagator_autotune(p.p, n_samples, &w, &k, &spl, &ms));
                }
                return py::make_tuple(w, k, spl, ms, hit != 0);
            },
            py::arg("n_samples"), py::arg("path") = py::none(),
            "Times each candidate schedule shape on n_samples samples and keeps the fastest; with a tuning "
            "file, a choice recorded there for this DAG structure and device is configured without timing, and "
            "a new choice is added to it.  Returns (waves, events_per_wave, samples_per_lane, ms, from_file)");
}
