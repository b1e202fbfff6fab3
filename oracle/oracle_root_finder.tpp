// ORACLE — test infrastructure only (see oracle_interval.hpp).
// interval_root_finder, src/interval/interval_root_finder.cpp:97-203.
#pragma once

#include <atomic>
#include <stack>
#include <vector>

namespace oracle {

// Optional cap on inclusion-function evaluations per query (0 = unbounded,
// as in the reference, whose max_iterations is unused:
// interval_root_finder.cpp:127). When the cap is hit the search stops and
// budget_exceeded() reports it; tests then treat the query as unresolved.
inline long& eval_budget() { static thread_local long b = 0; return b; }
inline bool& budget_exceeded() { static thread_local bool e = false; return e; }

// n-dimensional version. f maps n intervals to n (or 3) intervals via out;
// is_domain_valid filters boxes; constraint_predicate accepts roots.
//
// shared_min (optional): the running minimum of the TOIs that other queries
// of the same earliest-TOI call have found. Boxes whose t.lo is at or above
// it are dropped, as the query's own earliest root drops them. A child's
// t.lo is never below its parent's, so the dropped subtrees hold no root
// below the bound: the minimum over the call's queries is the same, bit for
// bit, whatever the order in which they run (only queries whose own root
// lies at or above the bound report no root).
template <int N, typename F, typename V, typename P>
bool interval_root_finder_nd(F f, V is_domain_valid, P constraint_predicate,
                             const Interval (&x0)[N], double (&tol)[N], Interval (&x_out)[N],
                             long* n_evals = nullptr, const std::atomic<double>* shared_min = nullptr) {
    struct Box { Interval x[N]; };
    Box earliest_root;
    for (int i = 0; i < N; i++) earliest_root.x[i] = Interval(INFINITY);
    bool found_root = false;

    std::stack<Box> xs;
    Box b0;
    for (int i = 0; i < N; i++) b0.x[i] = x0[i];
    xs.push(b0);

    // If the start is a root then reduce the time tolerance (:117-125).
    {
        Box x_tol;
        for (int i = 0; i < N; i++) x_tol.x[i] = Interval(0, tol[i]);
        Interval y[N];
        f(x_tol.x, y);
        if (n_evals) (*n_evals)++;
        bool z = true;
        for (int i = 0; i < N; i++) z = z && zero_in(y[i]);
        if (z) tol[0] /= 1e2;
    }

    long local_evals = 0;
    budget_exceeded() = false;
    while (!xs.empty()) {
        if (eval_budget() > 0 && ++local_evals > eval_budget()) {
            budget_exceeded() = true;
            break;
        }
        Box x = xs.top();
        xs.pop();

        if (x.x[0].lo >= earliest_root.x[0].lo) continue;
        if (shared_min && x.x[0].lo >= shared_min->load(std::memory_order_relaxed)) continue;
        if (!is_domain_valid(x.x)) continue;

        Interval y[N];
        f(x.x, y);
        if (n_evals) (*n_evals)++;
        bool z = true;
        for (int i = 0; i < N; i++) z = z && zero_in(y[i]);
        if (!z) continue;

        double widths[N];
        bool all_tol_sat = true, all_widths_zero = true;
        for (int i = 0; i < N; i++) {
            widths[i] = width(x.x[i]);
            all_tol_sat = all_tol_sat && (widths[i] <= tol[i]);
            all_widths_zero = all_widths_zero && (widths[i] <= 1e-10);
        }
        if ((x.x[0].lo > 0 || all_widths_zero) && all_tol_sat) {
            if (constraint_predicate(x.x)) {
                earliest_root = x;
                found_root = true;
            }
            continue;
        }

        int split_i = -1;
        for (int i = 0; i < N; i++) {
            if ((all_tol_sat || widths[i] > tol[i])
                && (split_i == -1 || widths[i] * tol[split_i] > widths[split_i] * tol[i])) {
                split_i = i;
            }
        }

        const double m = median(x.x[split_i]);
        const Interval first = Interval::raw(x.x[split_i].lo, m);
        const Interval second = Interval::raw(m, x.x[split_i].hi);
        x.x[split_i] = second;
        xs.push(x);
        x.x[split_i] = first;
        xs.push(x);
    }
    for (int i = 0; i < N; i++) x_out[i] = earliest_root.x[i];
    return found_root;
}

template <typename F, typename P>
bool interval_root_finder_1d(F f, P constraint_predicate, const Interval& x0, double tol, Interval& x) {
    Interval x0v[1] = { x0 };
    double tolv[1] = { tol };
    Interval out[1];
    bool r = interval_root_finder_nd<1>(
        [&](const Interval* xi, Interval* yi) { yi[0] = f(xi[0]); },
        [](const Interval*) { return true; },
        [&](const Interval* xi) { return constraint_predicate(xi[0]); }, x0v, tolv, out);
    if (r) x = out[0];
    return r;
}

}  // namespace oracle
