// Checks the reverse-mode hook Numerics<double>::numerical_flux_fluctuations_vjp against the
// transposed Jacobian of numerical_flux / numerical_fluctuations obtained with forward-mode
// duals through the same model functions.  Built per model directory:
//   g++ -std=c++17 -O1 -ffp-contract=off -I <model_dir> -I models flux_vjp_check.cpp
// Prints one line per case and "max_rel <x>" at the end; exits 1 on a mismatch above 1e-11.
#include "zoomy/UserFunctions.H"
#include "dual.H"
#include "Numerics.H"
#include <cstdio>
#include <cstdint>
#include <cmath>
#include <algorithm>

using M = Model<double>;
using NUMD = Numerics<double>;
constexpr int NQ = M::n_dof_q;
constexpr int NA = M::n_dof_qaux;
using DQ = Dual<NQ>;

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static double uni() {   // SplitMix64 -> [0, 1)
    uint64_t z = (rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// reference: g{L,R}[j] = sum_i bF_i dF_i/dq_j + sum_i bD_i dD_i/dq_j by duals seeded on one side
static void dual_vjp(const double* qL, const double* qR, const double* aL, const double* aR, const double* p,
                     const double* n, const double* bF, const double* bD, double* gL, double* gR) {
    for (int side = 0; side < 2; ++side) {
        DQ L[NQ], R[NQ], AL[NA > 0 ? NA : 1], AR[NA > 0 ? NA : 1], P[16], N[3];
        for (int i = 0; i < NQ; ++i) {
            L[i] = DQ(qL[i]);
            R[i] = DQ(qR[i]);
        }
        for (int i = 0; i < NQ; ++i) (side == 0 ? L : R)[i].d[i] = 1.0;
        for (int i = 0; i < NA; ++i) {
            AL[i] = DQ(aL[i]);
            AR[i] = DQ(aR[i]);
        }
        for (int i = 0; i < 16; ++i) P[i] = DQ(p[i]);
        for (int d = 0; d < 3; ++d) N[d] = DQ(n[d]);
        auto F = Numerics<DQ>::numerical_flux(L, R, AL, AR, P, N);
        auto D = Numerics<DQ>::numerical_fluctuations(L, R, AL, AR, P, N);
        double* g = side == 0 ? gL : gR;
        for (int j = 0; j < NQ; ++j) {
            double s = 0;
            for (int i = 0; i < NQ; ++i) s += bF[i] * F[i].d[j];
            for (int i = 0; i < 2 * NQ; ++i) s += bD[i] * D[i].d[j];
            g[j] = s;
        }
    }
}

int main() {
    const auto dp = M::default_parameters();
    double p[16] = {0};
    for (size_t i = 0; i < dp.size() && i < 16; ++i) p[i] = dp[i];
    double worst = 0;
    int bad = 0;
    const int ncase = 400;
    for (int c = 0; c < ncase; ++c) {
        double qL[NQ], qR[NQ], aL[NA > 0 ? NA : 1] = {0}, aR[NA > 0 ? NA : 1] = {0}, n[3] = {0, 0, 0};
        const double th = 2 * M_PI * uni();
        n[0] = std::cos(th);
        n[1] = M::dimension > 1 ? std::sin(th) : 0.0;
        if (M::dimension == 1) n[0] = uni() < 0.5 ? -1.0 : 1.0;
        const int kind = c % 8;
        for (int i = 0; i < NQ; ++i) {
            qL[i] = 2 * uni() - 1;
            qR[i] = 2 * uni() - 1;
        }
        qL[0] = 0.5 * uni();
        qR[0] = 0.5 * uni();
        qL[1] = 0.2 + 1.8 * uni();
        qR[1] = 0.2 + 1.8 * uni();
        if (kind == 1) qR[0] = qL[0];                         // bed tie: max takes the first
        if (kind == 2)                                        // lake at rest: zero momenta
            for (int i = 2; i < NQ; ++i) qL[i] = qR[i] = 0;
        if (kind == 3) qR[0] = qL[0] + qL[1] + 0.3;           // left side dry after reconstruction
        if (kind == 4) {                                      // shallow side (KP branch of hinv)
            qL[1] = 1e-7 * (1 + uni());
            for (int i = 2; i < NQ; ++i) qL[i] *= 1e-7;
        }
        if (kind == 5)                                        // identical states
            for (int i = 0; i < NQ; ++i) qR[i] = qL[i];
        for (int i = 0; i < NA; ++i) {
            aL[i] = M::update_aux_variables(qL, aL, p, 0.0, n)[i];
            aR[i] = M::update_aux_variables(qR, aR, p, 0.0, n)[i];
        }
        double bF[NQ], bD[2 * NQ];
        for (int i = 0; i < NQ; ++i) bF[i] = 2 * uni() - 1;
        for (int i = 0; i < 2 * NQ; ++i) bD[i] = 2 * uni() - 1;
        double rL[NQ], rR[NQ], hL[NQ] = {0}, hR[NQ] = {0};
        dual_vjp(qL, qR, aL, aR, p, n, bF, bD, rL, rR);
        NUMD::numerical_flux_fluctuations_vjp(qL, qR, p, n, bF, bD, hL, hR);
        double scale = 0, err = 0;
        for (int j = 0; j < NQ; ++j) {
            scale = std::max(scale, std::max(std::fabs(rL[j]), std::fabs(rR[j])));
            err = std::max(err, std::max(std::fabs(rL[j] - hL[j]), std::fabs(rR[j] - hR[j])));
        }
        const bool finite = std::isfinite(scale) && std::isfinite(err);
        if (kind == 3 && !finite) continue;   // sqrt at a dry face has no derivative either way
        const double rel = err / std::max(scale, 1e-300);
        worst = std::max(worst, rel);
        if (!(rel <= 1e-11)) {
            ++bad;
            if (bad <= 5) {
                std::printf("case %d kind %d rel %.3e\n", c, kind, rel);
                for (int j = 0; j < NQ; ++j)
                    std::printf("  j=%d dualL %.17g hookL %.17g dualR %.17g hookR %.17g\n", j, rL[j], hL[j], rR[j], hR[j]);
            }
        }
    }
    std::printf("cases %d bad %d max_rel %.3e\n", ncase, bad, worst);
    return bad ? 1 : 0;
}
