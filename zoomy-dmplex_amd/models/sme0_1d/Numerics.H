// SME level 0, 1-D: hydrostatic reconstruction + central flux + Rusanov fluctuations, the
// scheme of the reference's generated Numerics.H:29-109, restated from its equations:
//   h*_s = max(0, h_s + b_s - max(b_L, b_R)),  hu*_s = hu_s h*_s / max(1e-14, h_s)
//   numerical_flux  = 0.5 n [F(q*_L) + F(q*_R)],
//                     F = (0, hu*, (hu/h)^2 h*^2 kp(h*) + g/2 h*^2), kp = KP inverse, eps 1e-12
//   fluctuations    = right: (0, S dh*, n g/2 (h*_R^2 - h_R^2) + S dhu*)
//                     left:  (0, -S dh*, n g/2 (h_L^2 - h*_L^2) - S dhu*),
//                     S = 0.5 max |u* +- sqrt(g h*)| over both sides
//   local_max_abs_eigenvalue = max |n u +- sqrt(g h)|
#pragma once
#include "Model.H"

template <typename T>
struct Numerics {
    static constexpr int n_dof_q = 3;

    PORTABLE_FN static inline T kp_inv(T hs, T e) {
        const T em = std::max((T)9.9999999999999998e-13, e);
        const T a = hs * hs, b = em * em;
        return (T)1.4142135623730951 * hs / std::sqrt(a * a + b * b);
    }

    PORTABLE_FN static inline SimpleArray<T, 3> numerical_flux(const T* Q_minus, const T* Q_plus,
                                                               const T* Qaux_minus, const T* Qaux_plus,
                                                               const T* p, const T* n) {
        const T bstar = std::max(Q_minus[0], Q_plus[0]);
        const T eL = Q_minus[0] + Q_minus[1] - bstar, eR = Q_plus[0] + Q_plus[1] - bstar;
        const T hsL = std::max((T)0.0, eL), hsR = std::max((T)0.0, eR);
        const T dL = std::max((T)1.0e-14, Q_minus[1]), dR = std::max((T)1.0e-14, Q_plus[1]);
        const T uL = Q_minus[2] / dL, uR = Q_plus[2] / dR;
        const T half_n = (T)0.5 * n[0];
        const T g2 = (T)0.5 * p[0];
        SimpleArray<T, 3> r;
        r[0] = 0;
        r[1] = half_n * (Q_minus[2] * hsL / dL + Q_plus[2] * hsR / dR);
        const T mL = uL * uL * (hsL * hsL) * kp_inv(hsL, eL) + g2 * (hsL * hsL);
        const T mR = uR * uR * (hsR * hsR) * kp_inv(hsR, eR) + g2 * (hsR * hsR);
        r[2] = half_n * (mL + mR);
        return r;
    }

    PORTABLE_FN static inline SimpleArray<T, 6> numerical_fluctuations(const T* Q_minus, const T* Q_plus,
                                                                       const T* Qaux_minus, const T* Qaux_plus,
                                                                       const T* p, const T* n) {
        const T bstar = std::max(Q_minus[0], Q_plus[0]);
        const T eL = Q_minus[0] + Q_minus[1] - bstar, eR = Q_plus[0] + Q_plus[1] - bstar;
        const T hsL = std::max((T)0.0, eL), hsR = std::max((T)0.0, eR);
        const T fL = std::max((T)1.0e-14, eL), fR = std::max((T)1.0e-14, eR);
        const T sg = std::sqrt(p[0]) * n[0];
        const T cL = hsL * hsL * std::sqrt(hsL) * sg / (fL * fL);
        const T cR = hsR * hsR * std::sqrt(hsR) * sg / (fR * fR);
        const T huL = Q_minus[2] * hsL / std::max((T)1.0e-14, Q_minus[1]);
        const T huR = Q_plus[2] * hsR / std::max((T)1.0e-14, Q_plus[1]);
        const T unL = n[0] * huL / fL, unR = n[0] * huR / fR;
        const T S = (T)0.5 * std::max(std::max(std::fabs(cR + unR), std::fabs(cL + unL)),
                                      std::max(std::fabs(cR - unR), std::fabs(cL - unL)));
        const T g2 = (T)0.5 * p[0];
        SimpleArray<T, 6> r;
        r[0] = 0;
        r[1] = S * (hsR - hsL);
        r[2] = n[0] * (g2 * (hsR * hsR) - g2 * (Q_plus[1] * Q_plus[1])) + S * (huR - huL);
        r[3] = 0;
        r[4] = -r[1];
        r[5] = n[0] * (g2 * (Q_minus[1] * Q_minus[1]) - g2 * (hsL * hsL)) - S * (huR - huL);
        return r;
    }

    PORTABLE_FN static inline SimpleArray<T, 1> local_max_abs_eigenvalue(const T* Q, const T* Qaux,
                                                                          const T* p, const T* n) {
        const T hf = std::max((T)1.0e-14, Q[1]);
        const T h5 = Q[1] * Q[1] * Q[1] * Q[1] * Q[1];
        const T un = n[0] * Q[2] / hf;
        const T c = std::sqrt(p[0]) * n[0] * std::sqrt(std::max((T)0, h5)) / (hf * hf);
        SimpleArray<T, 1> r;
        r[0] = std::max(std::fabs(un + c), std::fabs(un - c));
        return r;
    }
};
