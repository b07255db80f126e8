// Binomial checkpoint schedule for the reverse sweep of a T-step trajectory held in a pool of
// state slots (Griewank and Walther, "Algorithm 799: revolve", ACM TOMS 26 (2000)).  Slot 0
// holds the initial state; C further slots take intermediate states.  The schedule minimises
// the number of forward steps re-run during the reverse sweep:
//
//   cost(n, 0) = n (n - 1) / 2             (re-run from the stretch's first state every time)
//   cost(n, c) = min_{1 <= t < n} [ t + cost(n - t, c - 1) + cost(t, c) ]
//
// (advance t steps into a free slot, reverse the last n - t steps with one slot fewer, free
// it, reverse the first t).  For s = C + 1 slots and l steps this equals the paper's minimal
// count t l - binom(s + t, t - 1), with t the smallest integer such that binom(s + t, s) >= l.
//
// Plain host code (no HIP): the device solver drives it with callbacks that move states
// between its slots; the checks compile it with g++.
#pragma once
#include <algorithm>
#include <functional>
#include <utility>
#include <vector>

namespace zdm_binomial {

// Split table of the dynamic programme: split[n * (C + 1) + c] is the optimal first stretch
// for n steps and c free slots.  Empty when the table would be too large (T^2 (C + 1) above
// max_work), in which case the schedule splits stretches into equal parts.
inline std::vector<int> plan(int T, int C, std::vector<double>* cost_out = nullptr, double max_work = 3.0e8) {
    std::vector<int> split;
    if (T < 1 || C < 0 || (double)T * T * (C + 1) > max_work) return split;
    split.assign((size_t)(T + 1) * (C + 1), 1);
    std::vector<double> cost((size_t)(T + 1) * (C + 1), 0.0);
    auto at = [&](int n, int c) { return (size_t)n * (C + 1) + c; };
    for (int n = 2; n <= T; ++n) {
        cost[at(n, 0)] = 0.5 * n * (n - 1.0);
        for (int c = 1; c <= C; ++c) {
            double best = 1e300;
            int bt = 1;
            for (int t = 1; t < n; ++t) {
                const double v = t + cost[at(n - t, c - 1)] + cost[at(t, c)];
                if (v < best) {
                    best = v;
                    bt = t;
                }
            }
            cost[at(n, c)] = best;
            split[at(n, c)] = bt;
        }
    }
    if (cost_out) *cost_out = cost;
    return split;
}

// The schedule for T steps on slots 0 .. C, with its split table.
struct Schedule {
    int T = 0, C = 0;
    std::vector<int> split;

    Schedule() = default;
    Schedule(int T_, int C_, double max_work = 3.0e8) : T(T_), C(C_), split(plan(T_, C_, nullptr, max_work)) {}

    // first stretch for n steps with c free slots
    int stretch(int n, int c) const {
        const int want = split.empty() ? n / (c + 1) : split[(size_t)n * (C + 1) + c];
        return std::min(n - 1, std::max(1, want));
    }

    // The first sweep: every advance the schedule makes before its first adjoint runs forward
    // from state 0 without a break, through states the forward run computes anyway.  The
    // (state, slot) pairs it leaves (slot -1: the scratch state) can be stored by the forward
    // run itself; reverse(..., stored = true) then skips those advances, which saves exactly
    // T - 1 steps on every schedule.
    std::vector<std::pair<int, int>> first_sweep() const {
        std::vector<std::pair<int, int>> out{{0, 0}};
        int l = 0, c = C, slot = 1;
        while (T - l > 1 && c > 0) {
            l += stretch(T - l, c);
            out.push_back({l, slot++});
            --c;
        }
        if (T - l > 1) out.push_back({T - 1, -1});
        return out;
    }

    // Reverse sweep over steps T-1 .. 0 with slots 0 .. C (slot 0 holds state 0 on entry; with
    // stored, every pair of first_sweep() is in place too).
    //   advance(from, to, l, m): run steps l .. m-1 starting from the state in slot `from`
    //     (state l) and leave state m in slot `to`; to = -1 names the scratch state.
    //   adjoint(n, from): the adjoint of step n given its start state in slot `from` (-1: the
    //     scratch state).
    // Each callback returns 0 on success; the first failure ends the sweep and is returned.
    template <typename Advance, typename Adjoint>
    int reverse(Advance&& advance, Adjoint&& adjoint, bool stored = false) const {
        std::vector<int> freeslots;
        for (int k = C; k >= 1; --k) freeslots.push_back(k);
        // top: the stretch ends at T and its advances belong to the first sweep
        std::function<int(int, int, int, int, bool)> rev = [&](int l, int r, int sl, int c, bool top) -> int {
            const int n = r - l;
            if (n == 1) return adjoint(l, sl);
            if (c == 0 || freeslots.empty()) {
                for (int k = r - 1; k > l; --k) {
                    if (!(top && stored && k == r - 1))
                        if (int e = advance(sl, -1, l, k)) return e;
                    if (int e = adjoint(k, -1)) return e;
                }
                return adjoint(l, sl);
            }
            const int m = l + stretch(n, c);
            const int sm = freeslots.back();
            freeslots.pop_back();
            if (!(top && stored))
                if (int e = advance(sl, sm, l, m)) return e;
            if (int e = rev(m, r, sm, c - 1, top)) return e;
            freeslots.push_back(sm);
            return rev(l, m, sl, c, false);
        };
        return rev(0, T, 0, C, true);
    }
};

}  // namespace zdm_binomial
