// Binomial checkpoint schedule for the reverse sweep of a T-step trajectory held in a pool of
// state slots (Griewank and Walther, "Algorithm 799: revolve", ACM TOMS 26 (2000)).  Slot 0
// holds the initial state; C further slots take intermediate states.  The schedule minimises
// the number of forward steps re-run during the reverse sweep.
//
// Without a stretch tape (L = 0) every step's adjoint re-runs that step from its start state
// (its recording), and the programme counts the other re-runs:
//
//   cost(n, 0) = n (n - 1) / 2             (re-run from the stretch's first state every time)
//   cost(n, c) = min_{1 <= t < n} [ t + cost(n - t, c - 1) + cost(t, c) ]
//
// (advance t steps into a free slot, reverse the last n - t steps with one slot fewer, free
// it, reverse the first t).  For s = C + 1 slots and l steps this equals the paper's minimal
// count t l - binom(s + t, t - 1), with t the smallest integer such that binom(s + t, s) >= l.
//
// With a stretch tape of L >= 1 steps a stretch of at most L steps is re-run once, recording
// every step into the tape, and its adjoints then read the tape; the count includes these
// recordings:
//
//   cost(n, c) = n                                                  (n <= L)
//   cost(n, 0) = (n - L) + L + cost(n - L, 0)                       (n > L: the last L steps)
//   cost(n, c) = min_{1 <= t < n} [ t + cost(n - t, c - 1) + cost(t, c) ]   (n > L, c >= 1)
//
// L = 1 gives the schedule of L = 0 (each count is larger by n, the per-step recordings).
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
inline std::vector<int> plan(int T, int C, std::vector<double>* cost_out = nullptr, double max_work = 3.0e8,
                             int L = 0) {
    std::vector<int> split;
    if (T < 1 || C < 0 || (double)T * T * (C + 1) > max_work) return split;
    split.assign((size_t)(T + 1) * (C + 1), 1);
    std::vector<double> cost((size_t)(T + 1) * (C + 1), 0.0);
    auto at = [&](int n, int c) { return (size_t)n * (C + 1) + c; };
    for (int n = 1; n <= T; ++n) {
        if (L == 0)
            cost[at(n, 0)] = 0.5 * n * (n - 1.0);
        else
            cost[at(n, 0)] = n <= L ? n : n + cost[at(n - L, 0)];
        for (int c = 1; c <= C; ++c) {
            if (n <= std::max(L, 1)) {
                cost[at(n, c)] = L == 0 ? 0.0 : n;
                continue;
            }
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

// The schedule for T steps on slots 0 .. C and a stretch tape of L steps (0: none), with its
// split table.
struct Schedule {
    int T = 0, C = 0, L = 0;
    std::vector<int> split;

    Schedule() = default;
    Schedule(int T_, int C_, double max_work = 3.0e8, int L_ = 0)
        : T(T_), C(C_), L(L_), split(plan(T_, C_, nullptr, max_work, L_)) {}

    // first stretch for n steps with c free slots
    int stretch(int n, int c) const {
        const int want = split.empty() ? n / (c + 1) : split[(size_t)n * (C + 1) + c];
        return std::min(n - 1, std::max(1, want));
    }

    // The first sweep: every advance the schedule makes before its first adjoint runs forward
    // from state 0 without a break, through states the forward run computes anyway.  The
    // (state, slot) pairs it leaves (slot -1: the scratch state) can be stored by the forward
    // run itself; reverse(..., stored = true) then skips those advances, which saves exactly
    // T - 1 steps on every schedule without a stretch tape.  With one, the forward run also
    // records steps tape_begin() .. T-1, the first stretch the sweep records.
    std::vector<std::pair<int, int>> first_sweep() const {
        std::vector<std::pair<int, int>> out{{0, 0}};
        const int lim = std::max(L, 1);
        int l = 0, c = C, slot = 1;
        while (T - l > lim && c > 0) {
            l += stretch(T - l, c);
            out.push_back({l, slot++});
            --c;
        }
        if (L == 0 && T - l > 1) out.push_back({T - 1, -1});
        return out;
    }

    // the first step of the stretch a stored sweep takes as recorded by the forward run (L >= 1)
    int tape_begin() const {
        int l = 0, c = C;
        while (T - l > L && c > 0) {
            l += stretch(T - l, c);
            --c;
        }
        return T - l > L ? T - L : l;
    }

    // Reverse sweep over steps T-1 .. 0 with slots 0 .. C and no stretch tape (slot 0 holds
    // state 0 on entry; with stored, every pair of first_sweep() is in place too).
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

    // Reverse sweep with the stretch tape (L >= 1).
    //   advance(from, to, l, m): as above;
    //   record(from, l, m): run steps l .. m-1 (m - l <= L) from the state in slot `from`,
    //     recording step n at tape position n - l;
    //   adjoint(n, j): the adjoint of step n from tape position j.
    // With stored, the first sweep's states are in place and the forward run recorded steps
    // tape_begin() .. T-1 at positions 0 .. T-1-tape_begin().
    template <typename Advance, typename Record, typename Adjoint>
    int reverse_taped(Advance&& advance, Record&& record, Adjoint&& adjoint, bool stored = false) const {
        std::vector<int> freeslots;
        for (int k = C; k >= 1; --k) freeslots.push_back(k);
        auto taped = [&](int from, int l, int r, bool skip) -> int {
            if (!skip)
                if (int e = record(from, l, r)) return e;
            for (int k = r - 1; k >= l; --k)
                if (int e = adjoint(k, k - l)) return e;
            return 0;
        };
        std::function<int(int, int, int, int, bool)> rev = [&](int l, int r, int sl, int c, bool top) -> int {
            const int n = r - l;
            if (n <= L) return taped(sl, l, r, top && stored);
            if (c == 0 || freeslots.empty()) {
                // the last L steps from the scratch state, then the rest with the same slot
                for (int e = r; e > l;) {
                    const int b = std::max(l, e - L);
                    const bool first = top && stored && e == r;
                    if (b > l && !first)
                        if (int x = advance(sl, -1, l, b)) return x;
                    if (int x = taped(b > l ? -1 : sl, b, e, first)) return x;
                    e = b;
                }
                return 0;
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
