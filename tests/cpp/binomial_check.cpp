// Checks of the binomial checkpoint schedule (zoomy-dmplex_amd/csrc/binomial.h):
//  1. the dynamic programme's optimal re-run count equals Griewank and Walther's closed form
//     p(l, s) = t l - binom(s + t, t - 1), t the smallest integer with binom(s + t, s) >= l;
//  2. a simulated sweep (states as step indices) visits the adjoints in the order T-1 .. 0, each
//     from the state of its own step, never uses a slot that is held, never holds more than C
//     intermediate states, and re-runs exactly the programme's count of steps;
//  3. with the first sweep's states stored by the forward run, the same sweep re-runs exactly
//     T - 1 steps fewer;
//  4. the same for the equal-split schedule used when the table would be too large;
//  5. with a stretch tape of L steps: L = 1 splits as L = 0 and counts n more (the per-step
//     recordings); L < T <= (C + 1) L costs exactly 2 T - L (the advances to the last stretch, which
//     the forward records, and one recording of every step: T - L re-runs once the forward's
//     first sweep is stored), T <= L costs T; a
//     simulated sweep records stretches of at most L steps from their own start states and
//     reverses them in order from the tape, re-runs the programme's count of steps, and with the
//     first sweep stored by the forward run (its states, and steps tape_begin() .. T-1
//     recorded) exactly T fewer.
// Prints "cases N bad B" last.
#include <cstdio>
#include <vector>

#include "binomial.h"

static double binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    double r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

static double closed_form(int l, int s) {
    int t = 0;
    while (binom(s + t, s) < l) ++t;
    return t * (double)l - binom(s + t, t - 1);
}

struct Sim {
    int T, C;
    std::vector<int> slot;   // state index held in each slot (-1 free); scratch at C + 1
    std::vector<int> order;
    long long resteps = 0;
    int bad = 0;
    int at(int k) const { return k < 0 ? C + 1 : k; }
    int advance(int from, int to, int l, int m) {
        if (slot[at(from)] != l || m <= l || m > T) ++bad;
        if (to >= 0 && slot[at(to)] >= 0 && to != from) ++bad;   // overwrites a held slot
        resteps += m - l;
        slot[at(to)] = m;
        return 0;
    }
    int adjoint(int n, int from) {
        if (slot[at(from)] != n) ++bad;
        order.push_back(n);
        // a slot other than 0 is released once the sweep has gone below its state
        for (int k = 1; k <= C; ++k)
            if (slot[k] >= n) slot[k] = -1;
        return 0;
    }
};

// the sweep with a stretch tape: tape[0] .. tape[1] the steps the tape holds
struct SimTaped : Sim {
    int L = 1, tl = -1, tr = -1;
    int record(int from, int l, int m) {
        if (slot[at(from)] != l || m <= l || m > T || m - l > L) ++bad;
        resteps += m - l;
        tl = l;
        tr = m;
        return 0;
    }
    int adjoint_taped(int n, int j) {
        if (n < tl || n >= tr || j != n - tl) ++bad;
        order.push_back(n);
        for (int k = 1; k <= C; ++k)
            if (slot[k] >= n) slot[k] = -1;
        return 0;
    }
};

static int run_taped(int T, int C, int L, double max_work, long long want, bool stored) {
    const zdm_binomial::Schedule sch(T, C, max_work, L);
    SimTaped sm;
    sm.T = T;
    sm.C = C;
    sm.L = L;
    sm.slot.assign(C + 2, -1);
    sm.slot[0] = 0;
    int bad = 0;
    if (stored) {
        int last = -1;
        for (const auto& p : sch.first_sweep()) {
            if (p.first <= last || p.first >= T || p.second > C || p.second < 0 ||
                (sm.slot[p.second] >= 0 && p.first))
                ++bad;
            last = p.first;
            sm.slot[p.second] = p.first;
        }
        const int tb = sch.tape_begin();
        if (tb < last || T - tb > L || tb < 0) ++bad;
        sm.tl = tb;
        sm.tr = T;
    }
    sch.reverse_taped([&](int f, int t, int l, int m) { return sm.advance(f, t, l, m); },
                      [&](int f, int l, int m) { return sm.record(f, l, m); },
                      [&](int n, int j) { return sm.adjoint_taped(n, j); }, stored);
    bad += sm.bad;
    if ((int)sm.order.size() != T) ++bad;
    for (int i = 0; i < (int)sm.order.size(); ++i)
        if (sm.order[i] != T - 1 - i) {
            ++bad;
            break;
        }
    if (want >= 0 && sm.resteps != want) ++bad;
    return bad;
}

static int run(int T, int C, double max_work, long long want, bool stored) {
    const zdm_binomial::Schedule sch(T, C, max_work);
    Sim sm{T, C, std::vector<int>(C + 2, -1), {}, 0, 0};
    sm.slot[0] = 0;
    int bad = 0;
    if (stored) {
        // the forward run leaves the first sweep's states; it must name each slot at most once,
        // in increasing state order
        int last = -1;
        for (const auto& p : sch.first_sweep()) {
            if (p.first <= last || p.first >= T || p.second > C || (p.second >= 0 && sm.slot[p.second] >= 0 && p.first))
                ++bad;
            last = p.first;
            sm.slot[sm.at(p.second)] = p.first;
        }
    }
    sch.reverse([&](int f, int t, int l, int m) { return sm.advance(f, t, l, m); },
                [&](int n, int f) { return sm.adjoint(n, f); }, stored);
    bad += sm.bad;
    if ((int)sm.order.size() != T) ++bad;
    for (int i = 0; i < (int)sm.order.size(); ++i)
        if (sm.order[i] != T - 1 - i) {
            ++bad;
            break;
        }
    if (want >= 0 && sm.resteps != want) ++bad;
    return bad;
}

int main() {
    int cases = 0, bad = 0;
    for (int T = 1; T <= 120; ++T)
        for (int C = 0; C <= 7; ++C) {
            std::vector<double> cost;
            const std::vector<int> split = zdm_binomial::plan(T, C, &cost);
            const double dp = cost[(size_t)T * (C + 1) + C], cf = closed_form(T, C + 1);
            int b = 0;
            if (dp != cf) ++b;
            b += run(T, C, 3.0e8, (long long)dp, false);
            b += run(T, C, 3.0e8, (long long)dp - (T - 1), true);   // the forward's sweep is free
            b += run(T, C, 0.0, -1, false);
            b += run(T, C, 0.0, -1, true);
            ++cases;
            if (b) {
                ++bad;
                if (bad < 10) std::printf("T %d C %d dp %.0f closed %.0f bad %d\n", T, C, dp, cf, b);
            }
        }
    for (int T = 1; T <= 90; ++T)
        for (int C = 0; C <= 6; ++C) {
            std::vector<double> c0, c1;
            const std::vector<int> s0 = zdm_binomial::plan(T, C, &c0, 3.0e8, 0);
            const std::vector<int> s1 = zdm_binomial::plan(T, C, &c1, 3.0e8, 1);
            int b = 0;
            for (int n = 1; n <= T; ++n)
                for (int c = 0; c <= C; ++c) {
                    const size_t k = (size_t)n * (C + 1) + c;
                    if (c1[k] != c0[k] + n) ++b;
                    if (c > 0 && n > 1 && s0[k] != s1[k]) ++b;
                }
            for (int L : {1, 2, 3, 5, 8, 13}) {
                std::vector<double> cost;
                zdm_binomial::plan(T, C, &cost, 3.0e8, L);
                const double dp = cost[(size_t)T * (C + 1) + C];
                // every step recorded once, the forward's first sweep (T steps) included, and the
                // advances to the last stretch's start
                if (T <= L && dp != T) ++b;
                if (T > L && T <= (C + 1) * L && dp != 2 * T - L) ++b;
                b += run_taped(T, C, L, 3.0e8, (long long)dp, false);
                b += run_taped(T, C, L, 3.0e8, (long long)dp - T, true);
                b += run_taped(T, C, L, 0.0, -1, false);
                b += run_taped(T, C, L, 0.0, -1, true);
            }
            ++cases;
            if (b) {
                ++bad;
                if (bad < 10) std::printf("taped T %d C %d bad %d\n", T, C, b);
            }
        }
    std::printf("cases %d bad %d\n", cases, bad);
    return bad != 0;
}
