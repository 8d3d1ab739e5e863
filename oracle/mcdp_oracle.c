/*
 * mcdp_oracle.c -- CPU oracle for the Monte Carlo DAG propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, what the
 * reference engine computes (WonJayne/mc_dagprop, src/mc_dagprop/monte_carlo/
 * _core.cpp) so that tests can check the GPU library against it.  It is
 * loaded only by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py.  The product (mc_dagprop_amd) never links or calls it.
 *
 * Contents and what pins each part:
 *   orc_compile     context compile, _core.cpp:193-307 (validation 196-201,
 *                   link table 213-229, Kahn order 232-264, CSR 266-285).
 *                   Pinned by the reference test KATs and by the compiled
 *                   reference harness (oracle/_ref) on random DAGs, including
 *                   duplicate precedence targets.
 *   orc_propagate   the propagation loop, _core.cpp:316-350, given injected
 *                   durations.  Pinned bit-exactly against oracle/_ref and
 *                   the golden fixtures it produced.
 *   orc_propagate_columns  the same loop over event-major arrays (samples
 *                   innermost), for full-size device outputs.  Pinned
 *                   against orc_propagate.
 *   orc_base_durations  durations that need no random numbers: no
 *                   distribution (_core.cpp:303-306) and constant
 *                   (_core.cpp:72-76 with :328).  Pinned by test KATs.
 *   orc_philox4x32 / orc_philox2x32  Philox-4x32-R and Philox-2x32-R
 *                   (Salmon, Moraes, Dror, Shaw, SC'11).  Pinned by the
 *                   Random123 known-answer vectors.
 *   orc_gradient    pathwise gradient of the expected weighted lateness
 *                   with respect to minimal durations and earliest times
 *                   (no counterpart in the reference: the derivative of
 *                   _core.cpp:332-350).  Pinned by central finite
 *                   differences of orc_propagate itself on random DAGs with
 *                   caps, stale reads, repeated sources and invalid edges,
 *                   and by closed forms on chains and diamonds.
 *
 * Behaviour the reference leaves undefined is given the definition the GPU
 * library uses (each one is listed where it applies):
 *   - precedence targets / sources outside [0, E): rejected (error 2);
 *   - negative Activity.idx: rejected (error 2);
 *   - an edge index outside [0, L) in a predecessor entry: that edge
 *     contributes duration 0.0 (the reference reads past its duration
 *     vector, _core.cpp:341; test_simulator.py:199 asserts the value a 0.0
 *     read gives).
 * Compile with -ffp-contract=off: every floating-point operation here must
 * be a single IEEE-754 binary64 operation, as in the reference.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_ctx {
    int E;             /* number of events */
    int L;             /* link_count = max(Activity.idx) + 1 */
    int64_t P;         /* total predecessor entries kept */
    double max_delay;
    double *earliest;  /* [E] */
    int *order;        /* [E]  evaluation order (Kahn, FIFO) */
    int64_t *off;      /* [E+1] CSR offsets, keyed by event id */
    int *src;          /* [P] predecessor event */
    int *edge;         /* [P] edge (activity) index */
    double *dmin;      /* [L] minimal duration (0.0 for gaps) */
    int *type;         /* [L] activity type (-1 for gaps) */
} orc_ctx;

static void set_err(char *err, int errlen, const char *msg) {
    if (err && errlen > 0) {
        strncpy(err, msg, (size_t)errlen - 1);
        err[errlen - 1] = '\0';
    }
}

void orc_free(orc_ctx *c) {
    if (!c) return;
    free(c->earliest);
    free(c->order);
    free(c->off);
    free(c->src);
    free(c->edge);
    free(c->dmin);
    free(c->type);
    free(c);
}

/*
 * Error codes: 0 ok, 1 reserved type -1 registered (_core.cpp:196-198),
 * 3 negative max_delay (:199-201), 4 cycle (:262-264), 2 index out of
 * range (undefined in the reference, rejected here), 5 out of memory.
 * `reg_types` lists the activity types that have a distribution.
 */
int orc_compile(int E, const double *earliest, int n_acts, const int *act_idx, const double *act_dur,
                const int *act_type, int n_prec, const int *prec_target, const int64_t *prec_off,
                const int *prec_src, const int *prec_edge, double max_delay, int n_reg, const int *reg_types,
                orc_ctx **out, char *err, int errlen) {
    *out = NULL;
    for (int i = 0; i < n_reg; ++i)
        if (reg_types[i] == -1) {
            set_err(err, errlen, "Activity type -1 is reserved for no delay");
            return 1;
        }
    if (max_delay < 0.0) {
        set_err(err, errlen, "max_delay must be non-negative");
        return 3;
    }
    int max_idx = -1;
    for (int i = 0; i < n_acts; ++i) {
        if (act_idx[i] < 0) {
            set_err(err, errlen, "activity index must be non-negative");
            return 2;
        }
        if (act_idx[i] > max_idx) max_idx = act_idx[i];
    }
    for (int i = 0; i < n_prec; ++i) {
        if (prec_target[i] < 0 || prec_target[i] >= E) {
            set_err(err, errlen, "precedence target out of range");
            return 2;
        }
        for (int64_t k = prec_off[i]; k < prec_off[i + 1]; ++k)
            if (prec_src[k] < 0 || prec_src[k] >= E) {
                set_err(err, errlen, "predecessor event out of range");
                return 2;
            }
    }

    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    if (!c) return 5;
    c->E = E;
    c->L = max_idx + 1;
    c->max_delay = max_delay;
    c->earliest = (double *)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    c->order = (int *)malloc(sizeof(int) * (size_t)(E > 0 ? E : 1));
    c->off = (int64_t *)calloc((size_t)E + 1, sizeof(int64_t));
    c->dmin = (double *)calloc((size_t)(c->L > 0 ? c->L : 1), sizeof(double));
    c->type = (int *)malloc(sizeof(int) * (size_t)(c->L > 0 ? c->L : 1));
    if (!c->earliest || !c->order || !c->off || !c->dmin || !c->type) {
        orc_free(c);
        return 5;
    }
    for (int i = 0; i < E; ++i) c->earliest[i] = earliest[i];
    /* _core.cpp:218-229: gaps keep {duration 0.0, type -1}. */
    for (int l = 0; l < c->L; ++l) c->type[l] = -1;
    for (int i = 0; i < n_acts; ++i) {
        c->dmin[act_idx[i]] = act_dur[i];
        c->type[act_idx[i]] = act_type[i];
    }

    /* _core.cpp:232-245: the last entry per target supplies its predecessor
     * list and indegree; every entry adds adjacency.                        */
    int *win = (int *)malloc(sizeof(int) * (size_t)(E > 0 ? E : 1)); /* winning entry per target */
    int *indeg = (int *)calloc((size_t)(E > 0 ? E : 1), sizeof(int));
    int64_t *adj_off = (int64_t *)calloc((size_t)E + 1, sizeof(int64_t));
    int64_t total_adj = prec_off[n_prec];
    int *adj = (int *)malloc(sizeof(int) * (size_t)(total_adj > 0 ? total_adj : 1));
    int *queue = (int *)malloc(sizeof(int) * (size_t)(E > 0 ? E : 1));
    if (!win || !indeg || !adj_off || !adj || !queue) {
        free(win); free(indeg); free(adj_off); free(adj); free(queue);
        orc_free(c);
        return 5;
    }
    for (int i = 0; i < E; ++i) win[i] = -1;
    for (int i = 0; i < n_prec; ++i) {
        win[prec_target[i]] = i;
        indeg[prec_target[i]] = (int)(prec_off[i + 1] - prec_off[i]);
        for (int64_t k = prec_off[i]; k < prec_off[i + 1]; ++k) adj_off[prec_src[k] + 1]++;
    }
    for (int i = 0; i < E; ++i) adj_off[i + 1] += adj_off[i];
    {
        int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(E > 0 ? E : 1));
        if (!pos) {
            free(win); free(indeg); free(adj_off); free(adj); free(queue);
            orc_free(c);
            return 5;
        }
        for (int i = 0; i < E; ++i) pos[i] = adj_off[i];
        /* adjacency[pr.first].push_back(tgt) in entry order, pred order */
        for (int i = 0; i < n_prec; ++i)
            for (int64_t k = prec_off[i]; k < prec_off[i + 1]; ++k) adj[pos[prec_src[k]]++] = prec_target[i];
        free(pos);
    }
    /* Kahn with a FIFO, seeded with indegree-0 events in id order
     * (_core.cpp:247-261). */
    int head = 0, tail = 0, n_out = 0;
    for (int i = 0; i < E; ++i)
        if (indeg[i] == 0) queue[tail++] = i;
    while (head < tail) {
        int n = queue[head++];
        c->order[n_out++] = n;
        for (int64_t k = adj_off[n]; k < adj_off[n + 1]; ++k)
            if (--indeg[adj[k]] == 0) queue[tail++] = adj[k];
    }
    free(indeg);
    free(adj_off);
    free(adj);
    free(queue);
    if (n_out != E) {
        free(win);
        orc_free(c);
        set_err(err, errlen, "Invalid DAG: cycle detected in precedence list");
        return 4;
    }

    /* CSR keyed by event id, predecessor order as given (_core.cpp:266-285). */
    for (int e = 0; e < E; ++e) {
        int w = win[e];
        c->off[e + 1] = c->off[e] + (w >= 0 ? (prec_off[w + 1] - prec_off[w]) : 0);
    }
    c->P = c->off[E];
    c->src = (int *)malloc(sizeof(int) * (size_t)(c->P > 0 ? c->P : 1));
    c->edge = (int *)malloc(sizeof(int) * (size_t)(c->P > 0 ? c->P : 1));
    if (!c->src || !c->edge) {
        free(win);
        orc_free(c);
        return 5;
    }
    for (int e = 0; e < E; ++e) {
        int w = win[e];
        if (w < 0) continue;
        int64_t o = c->off[e];
        for (int64_t k = prec_off[w]; k < prec_off[w + 1]; ++k, ++o) {
            c->src[o] = prec_src[k];
            c->edge[o] = prec_edge[k];
        }
    }
    free(win);
    *out = c;
    return 0;
}

int orc_event_count(const orc_ctx *c) { return c->E; }
int orc_link_count(const orc_ctx *c) { return c->L; }
void orc_get_order(const orc_ctx *c, int *order) { memcpy(order, c->order, sizeof(int) * (size_t)c->E); }

/*
 * Durations that need no random draw: links whose type has no distribution
 * keep minimal_duration (_core.cpp:303-306); constant links get
 * d + d*factor (_core.cpp:75, 328), multiply then add, two roundings.
 * `const_types`/`const_factor` list the constant distributions; any other
 * registered type is written as NaN (it needs a random draw).
 */
void orc_base_durations(const orc_ctx *c, int n_reg, const int *reg_types, const int *reg_is_const,
                        const double *const_factor, double *durations) {
    for (int l = 0; l < c->L; ++l) {
        const double d = c->dmin[l];
        int found = -1;
        for (int i = 0; i < n_reg; ++i)
            if (reg_types[i] == c->type[l]) found = i; /* last registration wins */
        if (found < 0) {
            durations[l] = d;
        } else if (reg_is_const[found]) {
            const double extra = d * const_factor[found];
            durations[l] = d + extra;
        } else {
            durations[l] = __builtin_nan("");
        }
    }
}

/*
 * Propagation given injected durations (_core.cpp:316-350), for n samples.
 * durations: [n][L] (row stride ld_dur), realized: [n][E], cause: [n][E].
 * std::min(a, b) is (b < a) ? b : a, so a NaN candidate is never selected
 * and a tie with the running maximum moves the cause to the later predecessor.
 */
void orc_propagate(const orc_ctx *c, const double *durations, int64_t ld_dur, int64_t n, double *realized,
                   int *cause) {
    const int E = c->E, L = c->L;
    for (int64_t s = 0; s < n; ++s) {
        const double *dur = durations + s * ld_dur;
        double *r = realized + s * (int64_t)E;
        int *ce = cause + s * (int64_t)E;
        for (int i = 0; i < E; ++i) {
            r[i] = c->earliest[i];
            ce[i] = -1;
        }
        for (int oi = 0; oi < E; ++oi) {
            const int e = c->order[oi];
            const double upper = c->earliest[e] + c->max_delay;
            double latest = r[e];
            int who = -1;
            for (int64_t k = c->off[e]; k < c->off[e + 1]; ++k) {
                const int src = c->src[k], edge = c->edge[k];
                const double d = (edge >= 0 && edge < L) ? dur[edge] : 0.0;
                double t = r[src] + d;
                t = (upper < t) ? upper : t;
                if (t >= latest) {
                    latest = t;
                    who = src;
                }
            }
            r[e] = (upper < latest) ? upper : latest;
            ce[e] = who;
        }
    }
}

/*
 * orc_propagate in the device's event-major layout: durations[l * ld_dur + s],
 * realized[e * ld_out + s], cause[e * ld_out + s] for samples s < n.  The same
 * operations in the same order per sample (the loop over samples is
 * innermost), so full-size device outputs can be checked without transposing
 * them.  Pinned against orc_propagate on random DAGs (stale reads, caps, NaN
 * and infinite durations, invalid edges).
 */
void orc_propagate_columns(const orc_ctx *c, const double *durations, int64_t ld_dur, int64_t n, double *realized,
                           int *cause, int64_t ld_out) {
    const int E = c->E, L = c->L;
    for (int i = 0; i < E; ++i) {
        double *r = realized + (int64_t)i * ld_out;
        int *ce = cause + (int64_t)i * ld_out;
        for (int64_t s = 0; s < n; ++s) {
            r[s] = c->earliest[i];
            ce[s] = -1;
        }
    }
    for (int oi = 0; oi < E; ++oi) {
        const int e = c->order[oi];
        const double upper = c->earliest[e] + c->max_delay;
        double *r = realized + (int64_t)e * ld_out; /* holds the running maximum until the cap below */
        int *ce = cause + (int64_t)e * ld_out;
        for (int64_t k = c->off[e]; k < c->off[e + 1]; ++k) {
            const int src = c->src[k], edge = c->edge[k];
            const double *rs = realized + (int64_t)src * ld_out;
            const double *d = (edge >= 0 && edge < L) ? durations + (int64_t)edge * ld_dur : NULL;
            for (int64_t s = 0; s < n; ++s) {
                double t = rs[s] + (d ? d[s] : 0.0);
                t = (upper < t) ? upper : t;
                if (t >= r[s]) {
                    r[s] = t;
                    ce[s] = src;
                }
            }
        }
        for (int64_t s = 0; s < n; ++s) r[s] = (upper < r[s]) ? upper : r[s];
    }
}

/* ---- Philox (Salmon et al. 2011), as published with Random123 ---------- */
static inline void philox4x32_round(uint32_t ctr[4], const uint32_t key[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
    const uint32_t y0 = (uint32_t)(p1 >> 32) ^ ctr[1] ^ key[0];
    const uint32_t y1 = (uint32_t)p1;
    const uint32_t y2 = (uint32_t)(p0 >> 32) ^ ctr[3] ^ key[1];
    const uint32_t y3 = (uint32_t)p0;
    ctr[0] = y0;
    ctr[1] = y1;
    ctr[2] = y2;
    ctr[3] = y3;
}

void orc_philox4x32(int rounds, const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t ctr[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
    uint32_t key[2] = {key_in[0], key_in[1]};
    for (int r = 0; r < rounds; ++r) {
        if (r) {
            key[0] += 0x9E3779B9u;
            key[1] += 0xBB67AE85u;
        }
        philox4x32_round(ctr, key);
    }
    memcpy(out, ctr, sizeof(ctr));
}

void orc_philox2x32(int rounds, const uint32_t ctr_in[2], uint32_t key, uint32_t out[2]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1];
    for (int r = 0; r < rounds; ++r) {
        if (r) key += 0x9E3779B9u;
        const uint64_t p = (uint64_t)0xD256D193u * c0;
        const uint32_t y0 = (uint32_t)(p >> 32) ^ key ^ c1;
        c1 = (uint32_t)p;
        c0 = y0;
    }
    out[0] = c0;
    out[1] = c1;
}

/* Vectorised helpers for tests: n independent counters. */
void orc_philox4x32_many(int rounds, int64_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    for (int64_t i = 0; i < n; ++i) orc_philox4x32(rounds, ctr + 4 * i, key + 2 * i, out + 4 * i);
}
void orc_philox2x32_many(int rounds, int64_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    for (int64_t i = 0; i < n; ++i) orc_philox2x32(rounds, ctr + 2 * i, key[i], out + 2 * i);
}

/* The link table as compiled: minimal duration and activity type per link. */
void orc_get_links(const orc_ctx *c, double *dmin, int *type) {
    memcpy(dmin, c->dmin, sizeof(double) * (size_t)c->L);
    memcpy(type, c->type, sizeof(int) * (size_t)c->L);
}

static int same_bits(double a, double b) { return memcmp(&a, &b, sizeof(double)) == 0; }

/*
 * Pathwise gradient of the expected weighted lateness, for n samples.
 *
 * There is no counterpart in the reference; this restates the derivative of
 * its propagation loop (_core.cpp:332-350) along the path that decided each
 * realized time.  Per sample s the objective is
 *     J_s = sum_e w[e] * max(0, realized[e] - earliest[e]),
 * and realized[e] is a max-plus expression of earliest times and durations.
 * Its derivative, where it exists, follows the recomputed winner of each
 * event (the last predecessor entry whose capped candidate reaches the
 * maximum, as the forward loop's `t >= latest` selects):
 *   - no entry beats earliest[e] (cause -1), or realized[e] equals the cap
 *     earliest[e] + max_delay: d realized[e] = d earliest[e];
 *   - otherwise realized[e] = base + duration, where base is realized[src]
 *     (or earliest[src] for an entry the forward loop read before src was
 *     evaluated: the reference's stale read) and d duration / d minimal
 *     duration is 1 for links whose type has no distribution or an empirical
 *     absolute one (duration = d + X), and 1 + X for the other types
 *     (duration = d + X*d), computed as duration * (1/d).  When d is 0 the
 *     multiplier comes from mult1 (the duration the same draw gives at d = 1).
 * Adjoints flow in reverse evaluation order.  The function also checks that
 * realized and cause are what the forward loop gives for these durations
 * (error 6 otherwise).
 *
 * Inputs are sample-major, as orc_propagate writes them: realized, cause
 * [n][E]; durations and mult1 [n][ld_dur].  link_rel[l] is 1 for the
 * duration = d + X*d types.  Outputs are SUMS over the n samples.
 */
int orc_gradient(const orc_ctx *c, const double *realized, const int *cause, const double *durations,
                 int64_t ld_dur, const double *mult1, int64_t n, const double *weights, const int *link_rel,
                 double *obj, double *grad_dmin, double *grad_earliest, char *err, int errlen) {
    const int E = c->E, L = c->L;
    int *pos = (int *)malloc(sizeof(int) * (size_t)(E > 0 ? E : 1));
    double *g = (double *)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    if (!pos || !g) {
        free(pos);
        free(g);
        set_err(err, errlen, "out of memory");
        return 5;
    }
    for (int i = 0; i < E; ++i) pos[c->order[i]] = i;
    *obj = 0.0;
    for (int l = 0; l < L; ++l) grad_dmin[l] = 0.0;
    for (int e = 0; e < E; ++e) grad_earliest[e] = 0.0;
    int rc = 0;
    for (int64_t s = 0; s < n && !rc; ++s) {
        const double *R = realized + s * (int64_t)E;
        const int *Cs = cause + s * (int64_t)E;
        const double *D = durations + s * ld_dur;
        for (int e = 0; e < E; ++e) g[e] = 0.0;
        for (int oi = E - 1; oi >= 0; --oi) {
            const int e = c->order[oi];
            const double r = R[e], ea = c->earliest[e];
            const double upper = ea + c->max_delay;
            const int active = r > ea;
            const double gt = g[e] + (active ? weights[e] : 0.0);
            if (active) {
                *obj += weights[e] * (r - ea);
                grad_earliest[e] -= weights[e];
            }
            /* the forward loop's winner, recomputed */
            double latest = ea;
            int64_t wk = -1;
            for (int64_t k = c->off[e]; k < c->off[e + 1]; ++k) {
                const int src = c->src[k], edge = c->edge[k];
                const double base = (pos[src] > pos[e]) ? c->earliest[src] : R[src];
                const double d = (edge >= 0 && edge < L) ? D[edge] : 0.0;
                double t = base + d;
                t = (upper < t) ? upper : t;
                if (t >= latest) {
                    latest = t;
                    wk = k;
                }
            }
            const double rr = (upper < latest) ? upper : latest;
            if ((wk >= 0 ? c->src[wk] : -1) != Cs[e] || !same_bits(rr, r)) {
                set_err(err, errlen, "realized/cause do not follow from these durations");
                rc = 6;
                break;
            }
            if (wk < 0 || r == upper) {
                grad_earliest[e] += gt;
                continue;
            }
            const int src = c->src[wk], edge = c->edge[wk];
            if (pos[src] > pos[e])
                grad_earliest[src] += gt;
            else
                g[src] += gt;
            if (edge >= 0 && edge < L) {
                double m = 1.0;
                if (link_rel[edge]) {
                    if (c->dmin[edge] != 0.0) {
                        const double inv = 1.0 / c->dmin[edge];
                        m = D[edge] * inv;
                    } else {
                        if (!mult1) {
                            set_err(err, errlen, "a relative link with minimal duration 0 needs mult1");
                            rc = 7;
                            break;
                        }
                        m = mult1[s * ld_dur + edge];
                    }
                }
                grad_dmin[edge] += gt * m;
            }
        }
    }
    free(pos);
    free(g);
    return rc;
}
