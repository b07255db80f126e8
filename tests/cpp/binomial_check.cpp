// Checks of the binomial checkpoint schedule (zoomy-dmplex_amd/csrc/binomial.h):
//  1. the dynamic programme's optimal re-run count equals Griewank and Walther's closed form
//     p(l, s) = t l - binom(s + t, t - 1), t the smallest integer with binom(s + t, s) >= l;
//  2. a simulated sweep (states as step indices) visits the adjoints in the order T-1 .. 0, each
//     from the state of its own step, never uses a slot that is held, never holds more than C
//     intermediate states, and re-runs exactly the programme's count of steps;
//  3. with the first sweep's states stored by the forward run, the same sweep re-runs exactly
//     T - 1 steps fewer;
//  4. the same for the equal-split schedule used when the table would be too large.
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
    std::printf("cases %d bad %d\n", cases, bad);
    return bad != 0;
}
