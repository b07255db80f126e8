// Shallow-water moment equations (SME) of level L on a 2-D horizontal mesh, written against
// the generated-model surface the reference solver consumes:
//   Model<T>    : Model.H:29-398   (constants, source + Jacobians, update_variables,
//                                   update_aux_variables, boundary_conditions, eigenvalues)
//   Numerics<T> : Numerics.H:29-109 (numerical_flux, numerical_fluctuations,
//                                   local_max_abs_eigenvalue)
// with the same signatures as the kernel-pointer typedefs of VirtualSolver.hpp:40-51, so the
// solver's kernels are instantiated on it exactly as on a generated header.
//
// State (n_dof_q = 2L + 4 + NTR):  [b, h, h*u_0 .. h*u_L, h*v_0 .. h*v_L, h*r_1 .. h*r_NTR]
//   u_0 / v_0 = depth-mean velocity, u_i / v_i (i >= 1) = Legendre moments alpha_i / beta_i,
//   r_k = NTR depth-integrated quantities advected with the mean velocity (VAM predictor:
//   the vertical-velocity moments).
// Aux (n_dof_qaux = 1):      [hinv], the Kurganov-Petrova regularised 1/h (Model.H:220 form).
// Parameters:                g = p[0], nu = p[PNU], lambda = p[PLAM] (gravity, kinematic
//                            viscosity, slip length); SME default layout [g, nu, lambda].
//
// Fluxes (Kowalski & Torrilhon 2019), normal component with un = u0.n, a_j = alpha_j.n:
//   h      : h un
//   h u_0  : h (u_0 un + sum_j a_j alpha_j/(2j+1)) + g h^2/2 n_x           (v analogous)
//   h u_i  : h (u_0 a_i + alpha_i un + sum_jk A_ijk alpha_j a_k)
// Non-conservative product (rows i >= 1), D_j = div(h alpha_j, h beta_j):
//   B(Q) dQ = -u_0 D_i + sum_jk B_ijk alpha_k D_j
// Source (Newtonian slip, u_b = u_0 + sum_j u_j the bottom velocity):
//   S_{h u_i} = -(2i+1) [ nu/lambda u_b + nu/h sum_j C_ij u_j ]
// Numerics: hydrostatic reconstruction (Audusse et al. 2004) of both face states, central
// flux of the reconstructed states, Rusanov dissipation 0.5 s (q*_R - q*_L) and the
// path-conservative non-conservative product on the straight path q*_L -> q*_R (3-point
// Gauss-Legendre), split half/half; wave speed bound s = |un| + sum_j |a_j| + sqrt(g h).
//
// Only +, -, *, /, sqrt, fabs and max/min are used, so host and device builds round alike.
#pragma once
#include <cmath>
#include <vector>
#include <string>
#include <algorithm>
#include "sme_coeffs.H"

#ifndef ZOOMY_SIMPLE_ARRAY
#define ZOOMY_SIMPLE_ARRAY
template <typename T, int N>
struct SimpleArray {
    T data[N];
    T& operator[](int i) { return data[i]; }
    const T& operator[](int i) const { return data[i]; }
    T* begin() { return data; }
    const T* begin() const { return data; }
    T* end() { return data + N; }
    const T* end() const { return data + N; }
};
#endif

#ifndef PORTABLE_FN
#define PORTABLE_FN
#endif

template <typename T, int L, int NTR = 0, int PNU = 1, int PLAM = 2, int NPAR = 3>
struct SmeModel2D {
    using K = SmeCoeffs<L>;
    static constexpr int level = L;
    static constexpr int NM = L + 1;                 // moments per direction incl. the mean
    static constexpr int n_tracers = NTR;
    static constexpr int IT = 2 + 2 * NM;            // first tracer slot
    static constexpr int n_dof_q = 2 * NM + 2 + NTR;
    static constexpr int n_dof_qaux = 1;
    static constexpr int n_parameters = NPAR;
    static constexpr int dimension = 2;
    static constexpr double dt_max = 5.0;
    static constexpr int n_dof_gradQ = 0;
    static constexpr bool has_diffusion = false;
    static constexpr bool has_free_surface = true;
    static constexpr int idx_b = 0;
    static constexpr int idx_h = 1;
    static constexpr int n_boundary_tags = 1;
    static constexpr int IU = 2;                     // first x-moment slot
    static constexpr int IV = 2 + NM;                // first y-moment slot
    static const std::vector<std::string> get_boundary_tags() { return {"default"}; }
    static const std::vector<std::string> parameter_names() { return {"g", "nu", "lambda"}; }
    static const std::vector<T> default_parameters() { return {9.81, 1.0e-3, 0.1}; }

    // Kurganov-Petrova regularised inverse depth: sqrt(2) max(0,h) / sqrt(h^4 + max(eps,h)^4),
    // which is 1/h wherever h >= eps (the wet case: one division instead of a square root and a
    // division)
    PORTABLE_FN static inline T hinv_kp(T h, T eps) {
        if (h >= eps) return (T)1 / h;
        const T h2 = h * h, e2 = eps * eps;
        return (T)1.4142135623730951 * std::max((T)0, h) / std::sqrt(h2 * h2 + e2 * e2);
    }

    PORTABLE_FN static inline SimpleArray<T, n_dof_q> source(const T* Q, const T* Qaux, const T* p,
                                                             const T time, const T dt, const T* X) {
        SimpleArray<T, n_dof_q> res;
        for (int i = 0; i < n_dof_q; ++i) res[i] = 0;
        const T hinv = Qaux[0];
        const T k_slip = p[PNU] / p[PLAM];
        const T k_visc = p[PNU] * hinv * hinv;
        T mu = 0, mv = 0;
        for (int j = 0; j < NM; ++j) { mu += Q[IU + j]; mv += Q[IV + j]; }
        const T ub = mu * hinv, vb = mv * hinv;
        for (int i = 0; i < NM; ++i) {
            T su = 0, sv = 0;
            for (int j = 1; j < NM; ++j) {
                su += (T)K::C[i * NM + j] * Q[IU + j];
                sv += (T)K::C[i * NM + j] * Q[IV + j];
            }
            const T w = (T)(2 * i + 1);
            res[IU + i] = -w * (k_slip * ub + k_visc * su);
            res[IV + i] = -w * (k_slip * vb + k_visc * sv);
        }
        return res;
    }

    // dS/dQ, row-major n x n, with d(hinv)/dh taken as -hinv^2 (as Model.H:169 does)
    PORTABLE_FN static inline SimpleArray<T, n_dof_q * n_dof_q> source_jacobian_wrt_variables(
        const T* Q, const T* Qaux, const T* p) {
        SimpleArray<T, n_dof_q * n_dof_q> J;
        for (int i = 0; i < n_dof_q * n_dof_q; ++i) J[i] = 0;
        const T hinv = Qaux[0];
        const T hinv2 = hinv * hinv;
        const T k_slip = p[PNU] / p[PLAM];
        const T k_visc = p[PNU] * hinv2;
        T mu = 0, mv = 0;
        for (int j = 0; j < NM; ++j) { mu += Q[IU + j]; mv += Q[IV + j]; }
        for (int i = 0; i < NM; ++i) {
            const T w = (T)(2 * i + 1);
            T su = 0, sv = 0;
            for (int j = 1; j < NM; ++j) {
                su += (T)K::C[i * NM + j] * Q[IU + j];
                sv += (T)K::C[i * NM + j] * Q[IV + j];
            }
            // d/dh: u_b = m hinv -> -m hinv^2 ; nu hinv^2 s -> -2 nu hinv^3 s
            J[(IU + i) * n_dof_q + 1] = w * (k_slip * mu * hinv2 + (T)2 * p[PNU] * hinv2 * hinv * su);
            J[(IV + i) * n_dof_q + 1] = w * (k_slip * mv * hinv2 + (T)2 * p[PNU] * hinv2 * hinv * sv);
            for (int j = 0; j < NM; ++j) {
                const T cij = (j >= 1) ? (T)K::C[i * NM + j] : (T)0;
                J[(IU + i) * n_dof_q + IU + j] = -w * (k_slip * hinv + k_visc * cij);
                J[(IV + i) * n_dof_q + IV + j] = -w * (k_slip * hinv + k_visc * cij);
            }
        }
        return J;
    }

    PORTABLE_FN static inline SimpleArray<T, n_dof_q * n_dof_qaux> source_jacobian_wrt_aux_variables(
        const T* Q, const T* Qaux, const T* p) {
        SimpleArray<T, n_dof_q * n_dof_qaux> J;
        for (int i = 0; i < n_dof_q * n_dof_qaux; ++i) J[i] = 0;
        return J;
    }

    // wet/dry clamp |h u_k| <= 30 max(0, h - 0.01) on every momentum slot (Model.H:195-207 form)
    PORTABLE_FN static inline SimpleArray<T, n_dof_q> update_variables(const T* Q, const T* Qaux,
                                                                       const T* p, const T dt) {
        const T t0 = (T)30.0 * std::max((T)0, Q[1] - (T)0.01);
        SimpleArray<T, n_dof_q> res;
        res[0] = Q[0];
        res[1] = Q[1];
        for (int i = 2; i < n_dof_q; ++i) res[i] = std::max(-t0, std::min(Q[i], t0));
        return res;
    }

    PORTABLE_FN static inline SimpleArray<T, n_dof_qaux> update_aux_variables(const T* Q, const T* Qaux,
                                                                              const T* p, const T time,
                                                                              const T* X) {
        SimpleArray<T, n_dof_qaux> res;
        res[0] = hinv_kp(Q[1], (T)1.0e-8);
        return res;
    }

    // wall: reflect the normal component of every (h u_k, h v_k) pair; Qaux may be null
    PORTABLE_FN static inline SimpleArray<T, n_dof_q> boundary_conditions(const int bc_idx, const T* Q,
                                                                          const T* Qaux, const T* p,
                                                                          const T* n, const T* X,
                                                                          const T time, const T dX) {
        SimpleArray<T, n_dof_q> res;
        if (bc_idx == 0) {
            res[0] = Q[0];
            res[1] = Q[1];
            for (int k = 0; k < NM; ++k) {
                const T qn = Q[IU + k] * n[0] + Q[IV + k] * n[1];
                res[IU + k] = Q[IU + k] - (T)2 * qn * n[0];
                res[IV + k] = Q[IV + k] - (T)2 * qn * n[1];
            }
            for (int k = 0; k < NTR; ++k) res[IT + k] = Q[IT + k];
            return res;
        }
        for (int i = 0; i < n_dof_q; ++i) res[i] = 0;
        return res;
    }

    // max |lambda| bound along n from the cell state and its aux: |un| + sum_j |a_j| + sqrt(g h)
    PORTABLE_FN static inline T wave_speed(const T* Q, T hinv, const T* p, const T* n) {
        const T un = (Q[IU] * n[0] + Q[IV] * n[1]) * hinv;
        T s = std::fabs(un);
        for (int j = 1; j < NM; ++j) s += std::fabs((Q[IU + j] * n[0] + Q[IV + j] * n[1]) * hinv);
        return s + std::sqrt(p[0] * std::max((T)0, Q[1]));
    }

    PORTABLE_FN static inline SimpleArray<T, n_dof_q> eigenvalues(const T* Q, const T* Qaux, const T* p,
                                                                  const T* n) {
        SimpleArray<T, n_dof_q> res;
        const T un = (Q[IU] * n[0] + Q[IV] * n[1]) * Qaux[0];
        const T c = wave_speed(Q, Qaux[0], p, n) - std::fabs(un);
        for (int i = 0; i < n_dof_q; ++i) res[i] = un;
        res[0] = 0;
        res[1] = un - c;
        res[n_dof_q - 1] = un + c;
        return res;
    }

    // Centred bathymetry term of the second-order hydrostatic reconstruction (Audusse et al.
    // 2004, sec. 4): per face side, with n the outward normal of the cell,
    //   + g/2 (eta_cell - b_face)^2 n   on the mean momenta,
    // which sums to -g h grad(b) over the cell and cancels the face pressure exactly at rest.
    PORTABLE_FN static inline SimpleArray<T, n_dof_q> hydrostatic_cell_source(const T* Qface, const T* Qcell,
                                                                              const T* p, const T* n) {
        SimpleArray<T, n_dof_q> res;
        for (int i = 0; i < n_dof_q; ++i) res[i] = 0;
        const T d = (Qcell[0] + Qcell[1]) - Qface[0];
        const T c = (T)0.5 * p[0] * (d * d);
        res[IU] = c * n[0];
        res[IV] = c * n[1];
        return res;
    }

    PORTABLE_FN static inline SimpleArray<T, n_dof_q> initial_condition(const T* X, const T* p) {
        SimpleArray<T, n_dof_q> res;
        for (int i = 0; i < n_dof_q; ++i) res[i] = 0;
        res[1] = 1.0;
        return res;
    }
};

template <typename T, int L, int NTR = 0, int PNU = 1, int PLAM = 2, int NPAR = 3>
struct SmeNumerics2D {
    using M = SmeModel2D<T, L, NTR, PNU, PLAM, NPAR>;
    static constexpr int IT = M::IT;
    using K = SmeCoeffs<L>;
    static constexpr int n_dof_q = M::n_dof_q;
    static constexpr int NM = M::NM;
    static constexpr int IU = M::IU;
    static constexpr int IV = M::IV;
    static constexpr T eps_flux = (T)1.0e-12;   // KP regularisation of the face inverse depth

    // Hydrostatically reconstructed state: h* = max(0, h + b - max(bL, bR)), momenta scaled by
    // h*/h; velocities of the reconstructed state through the KP inverse of h*.
    struct HR {
        T h;                    // h*
        T m[2 * NM + NTR];      // reconstructed momenta (x moments, y moments) and tracers
        T u[2 * NM];            // velocities            (x moments then y moments)
        T r, hi;                // h* / max(1e-14, h) and hinv_kp(h*), kept for the derivative
    };

    PORTABLE_FN static inline void reconstruct(const T* Q, T bstar, HR& s) {
        s.h = std::max((T)0.0, Q[0] + Q[1] - bstar);
        const T r = s.h / std::max((T)1.0e-14, Q[1]);
        const T hi = M::hinv_kp(s.h, eps_flux);
        for (int k = 0; k < 2 * NM; ++k) {
            s.m[k] = Q[IU + k] * r;
            s.u[k] = s.m[k] * hi;
        }
        for (int k = 0; k < NTR; ++k) s.m[2 * NM + k] = Q[IT + k] * r;
        s.r = r;
        s.hi = hi;
    }

    // normal flux F(q*) . n of a reconstructed state (b slot is zero)
    PORTABLE_FN static inline void phys_flux(const HR& s, const T* p, const T* n, T* F) {
        const T* u = s.u;       // u[k] = x moment k, u[NM + k] = y moment k
        const T* m = s.m;
        T an[NM];
        for (int k = 0; k < NM; ++k) an[k] = u[k] * n[0] + u[NM + k] * n[1];
        const T un = an[0];
        const T pr = (T)0.5 * p[0] * s.h * s.h;
        F[0] = 0;
        F[1] = m[0] * n[0] + m[NM] * n[1];
        // mean momenta
        T fx = m[0] * un, fy = m[NM] * un;
        for (int j = 1; j < NM; ++j) {
            fx += (T)K::A[j * NM + j] * m[j] * an[j];
            fy += (T)K::A[j * NM + j] * m[NM + j] * an[j];
        }
        F[IU] = fx + pr * n[0];
        F[IV] = fy + pr * n[1];
        // higher moments: h (u0 a_i + alpha_i un + sum_jk A_ijk alpha_j a_k)
        for (int i = 1; i < NM; ++i) {
            T gx = s.h * u[0] * an[i] + m[i] * un;
            T gy = s.h * u[NM] * an[i] + m[NM + i] * un;
            for (int j = 1; j < NM; ++j)
                for (int k = 1; k < NM; ++k) {
                    const T a = (T)K::A[(i * NM + j) * NM + k];
                    if (a != (T)0) {
                        gx += a * m[j] * an[k];
                        gy += a * m[NM + j] * an[k];
                    }
                }
            F[IU + i] = gx;
            F[IV + i] = gy;
        }
        // tracers ride the mean velocity: h r_k un
        for (int k = 0; k < NTR; ++k) F[IT + k] = m[2 * NM + k] * un;
    }

    PORTABLE_FN static inline T speed(const HR& s, const T* p, const T* n) {
        T v = std::fabs(s.u[0] * n[0] + s.u[NM] * n[1]);
        for (int j = 1; j < NM; ++j) v += std::fabs(s.u[j] * n[0] + s.u[NM + j] * n[1]);
        return v + std::sqrt(p[0] * s.h);
    }

    PORTABLE_FN static inline SimpleArray<T, n_dof_q> numerical_flux(const T* Q_minus, const T* Q_plus,
                                                                     const T* Qaux_minus, const T* Qaux_plus,
                                                                     const T* p, const T* n) {
        const T bstar = std::max(Q_minus[0], Q_plus[0]);
        HR sL, sR;
        reconstruct(Q_minus, bstar, sL);
        reconstruct(Q_plus, bstar, sR);
        T FL[n_dof_q], FR[n_dof_q];
        phys_flux(sL, p, n, FL);
        phys_flux(sR, p, n, FR);
        SimpleArray<T, n_dof_q> res;
        for (int i = 0; i < n_dof_q; ++i) res[i] = (T)0.5 * (FL[i] + FR[i]);
        return res;
    }

    // [0, n) = fluctuation into the right cell, [n, 2n) = into the left cell (Numerics.H:59-93
    // convention, applied as f_L -= D_left*A, f_R -= D_right*A by the transport step)
    PORTABLE_FN static inline SimpleArray<T, 2 * n_dof_q> numerical_fluctuations(
        const T* Q_minus, const T* Q_plus, const T* Qaux_minus, const T* Qaux_plus, const T* p,
        const T* n) {
        const T bstar = std::max(Q_minus[0], Q_plus[0]);
        HR sL, sR;
        reconstruct(Q_minus, bstar, sL);
        reconstruct(Q_plus, bstar, sR);
        const T s = (T)0.5 * std::max(speed(sL, p, n), speed(sR, p, n));
        SimpleArray<T, 2 * n_dof_q> res;
        T dq[n_dof_q];
        dq[0] = 0;
        dq[1] = sR.h - sL.h;
        for (int k = 0; k < 2 * NM + NTR; ++k) dq[IU + k] = sR.m[k] - sL.m[k];
        T nc[n_dof_q];
        for (int i = 0; i < n_dof_q; ++i) nc[i] = 0;
        if (NM > 1) {
            // path integral of B(Psi) dPsi along Psi(t) = q*_L + t (q*_R - q*_L)
            const T gt[3] = {(T)0.1127016653792583, (T)0.5, (T)0.8872983346207417};
            const T gw[3] = {(T)0.2777777777777778, (T)0.4444444444444444, (T)0.2777777777777778};
            T ubar[2 * NM];
            for (int k = 0; k < 2 * NM; ++k) ubar[k] = 0;
            for (int q = 0; q < 3; ++q) {
                const T h = sL.h + gt[q] * dq[1];
                const T hi = M::hinv_kp(h, eps_flux);
                for (int k = 0; k < 2 * NM; ++k) ubar[k] += gw[q] * ((sL.m[k] + gt[q] * dq[IU + k]) * hi);
            }
            T dD[NM];
            for (int j = 0; j < NM; ++j) dD[j] = dq[IU + j] * n[0] + dq[IV + j] * n[1];
            for (int i = 1; i < NM; ++i) {
                T bx = -ubar[0] * dD[i], by = -ubar[NM] * dD[i];
                for (int j = 1; j < NM; ++j)
                    for (int k = 1; k < NM; ++k) {
                        const T b = (T)K::B[(i * NM + j) * NM + k];
                        if (b != (T)0) {
                            bx += b * ubar[k] * dD[j];
                            by += b * ubar[NM + k] * dD[j];
                        }
                    }
                nc[IU + i] = bx;
                nc[IV + i] = by;
            }
        }
        const T gh = (T)0.5 * p[0];
        for (int i = 0; i < n_dof_q; ++i) {
            res[i] = (T)0.5 * nc[i] + s * dq[i];
            res[n_dof_q + i] = (T)0.5 * nc[i] - s * dq[i];
        }
        // hydrostatic-reconstruction pressure corrections on the mean momenta
        const T cR = gh * (sR.h * sR.h - Q_plus[1] * Q_plus[1]);
        const T cL = gh * (Q_minus[1] * Q_minus[1] - sL.h * sL.h);
        res[IU] += n[0] * cR;
        res[IV] += n[1] * cR;
        res[n_dof_q + IU] += n[0] * cL;
        res[n_dof_q + IV] += n[1] * cL;
        return res;
    }

    // Both of the above from one pair of reconstructed states (optional hook, detected by the
    // solver): the same operations in the same order as numerical_flux and
    // numerical_fluctuations, so the results are identical to calling them separately.
    PORTABLE_FN static inline void numerical_flux_fluctuations(const T* Q_minus, const T* Q_plus,
                                                               const T* Qaux_minus, const T* Qaux_plus,
                                                               const T* p, const T* n, T* Fout, T* Dout) {
        const T bstar = std::max(Q_minus[0], Q_plus[0]);
        HR sL, sR;
        reconstruct(Q_minus, bstar, sL);
        reconstruct(Q_plus, bstar, sR);
        {
            T FL[n_dof_q], FR[n_dof_q];
            phys_flux(sL, p, n, FL);
            phys_flux(sR, p, n, FR);
            for (int i = 0; i < n_dof_q; ++i) Fout[i] = (T)0.5 * (FL[i] + FR[i]);
        }
        const T s = (T)0.5 * std::max(speed(sL, p, n), speed(sR, p, n));
        T dq[n_dof_q];
        dq[0] = 0;
        dq[1] = sR.h - sL.h;
        for (int k = 0; k < 2 * NM + NTR; ++k) dq[IU + k] = sR.m[k] - sL.m[k];
        T nc[n_dof_q];
        for (int i = 0; i < n_dof_q; ++i) nc[i] = 0;
        if (NM > 1) {
            const T gt[3] = {(T)0.1127016653792583, (T)0.5, (T)0.8872983346207417};
            const T gw[3] = {(T)0.2777777777777778, (T)0.4444444444444444, (T)0.2777777777777778};
            T ubar[2 * NM];
            for (int k = 0; k < 2 * NM; ++k) ubar[k] = 0;
            for (int q = 0; q < 3; ++q) {
                const T h = sL.h + gt[q] * dq[1];
                const T hi = M::hinv_kp(h, eps_flux);
                for (int k = 0; k < 2 * NM; ++k) ubar[k] += gw[q] * ((sL.m[k] + gt[q] * dq[IU + k]) * hi);
            }
            T dD[NM];
            for (int j = 0; j < NM; ++j) dD[j] = dq[IU + j] * n[0] + dq[IV + j] * n[1];
            for (int i = 1; i < NM; ++i) {
                T bx = -ubar[0] * dD[i], by = -ubar[NM] * dD[i];
                for (int j = 1; j < NM; ++j)
                    for (int k = 1; k < NM; ++k) {
                        const T b = (T)K::B[(i * NM + j) * NM + k];
                        if (b != (T)0) {
                            bx += b * ubar[k] * dD[j];
                            by += b * ubar[NM + k] * dD[j];
                        }
                    }
                nc[IU + i] = bx;
                nc[IV + i] = by;
            }
        }
        const T gh = (T)0.5 * p[0];
        for (int i = 0; i < n_dof_q; ++i) {
            Dout[i] = (T)0.5 * nc[i] + s * dq[i];
            Dout[n_dof_q + i] = (T)0.5 * nc[i] - s * dq[i];
        }
        const T cR = gh * (sR.h * sR.h - Q_plus[1] * Q_plus[1]);
        const T cL = gh * (Q_minus[1] * Q_minus[1] - sL.h * sL.h);
        Dout[IU] += n[0] * cR;
        Dout[IV] += n[1] * cR;
        Dout[n_dof_q + IU] += n[0] * cL;
        Dout[n_dof_q + IV] += n[1] * cL;
    }

    // ---- reverse-mode derivative of numerical_flux_fluctuations (optional hook) ----
    // Branches of max / fabs are taken where the forward takes them: std::max(a, b) follows b
    // when a < b, fabs(a) has slope -1 when a < 0 (the choices forward-mode duals make through
    // the functions above), so the result is the exact transpose of their Jacobian.

    // d hinv_kp(h, eps) / dh times bres; hi = hinv_kp(h, eps)
    PORTABLE_FN static inline T hinv_kp_vjp(T h, T eps, T bres, T hi) {
        if (h >= eps) return -((bres * hi) * hi);   // d(1/h)/dh = -1/h^2
        const T h2 = h * h, e2 = eps * eps;
        const T den = std::sqrt(h2 * h2 + e2 * e2);
        const T c = (T)1.4142135623730951;
        const T bnum = c * bres / den;
        const T bden = -(c * std::max((T)0, h)) * bres / (den * den);
        const T barg = bden / ((T)2 * den);
        T bh = (T)2 * h * ((T)2 * h2 * barg);
        if ((T)0 < h) bh += bnum;
        return bh;
    }

    // adjoint of phys_flux: bF -> bh, bm, bu (accumulated)
    PORTABLE_FN static inline void phys_flux_vjp(const HR& s, const T* p, const T* n, const T* bF, T& bh, T* bm,
                                                 T* bu) {
        const T* u = s.u;
        const T* m = s.m;
        T an[NM], ban[NM];
        for (int k = 0; k < NM; ++k) {
            an[k] = u[k] * n[0] + u[NM + k] * n[1];
            ban[k] = 0;
        }
        const T un = an[0];
        T bun = m[0] * bF[IU] + m[NM] * bF[IV];
        bm[0] += n[0] * bF[1] + un * bF[IU];
        bm[NM] += n[1] * bF[1] + un * bF[IV];
        for (int j = 1; j < NM; ++j) {
            const T a = (T)K::A[j * NM + j];
            bm[j] += a * an[j] * bF[IU];
            bm[NM + j] += a * an[j] * bF[IV];
            ban[j] += a * (m[j] * bF[IU] + m[NM + j] * bF[IV]);
        }
        bh += p[0] * s.h * (n[0] * bF[IU] + n[1] * bF[IV]);   // pr = g h^2 / 2
        for (int i = 1; i < NM; ++i) {
            const T fx = bF[IU + i], fy = bF[IV + i];
            bh += an[i] * (u[0] * fx + u[NM] * fy);
            bu[0] += s.h * an[i] * fx;
            bu[NM] += s.h * an[i] * fy;
            ban[i] += s.h * (u[0] * fx + u[NM] * fy);
            bm[i] += un * fx;
            bm[NM + i] += un * fy;
            bun += m[i] * fx + m[NM + i] * fy;
            for (int j = 1; j < NM; ++j)
                for (int k = 1; k < NM; ++k) {
                    const T a = (T)K::A[(i * NM + j) * NM + k];
                    if (a != (T)0) {
                        bm[j] += a * an[k] * fx;
                        bm[NM + j] += a * an[k] * fy;
                        ban[k] += a * (m[j] * fx + m[NM + j] * fy);
                    }
                }
        }
        for (int k = 0; k < NTR; ++k) {
            bm[2 * NM + k] += un * bF[IT + k];
            bun += m[2 * NM + k] * bF[IT + k];
        }
        ban[0] += bun;
        for (int k = 0; k < NM; ++k) {
            bu[k] += n[0] * ban[k];
            bu[NM + k] += n[1] * ban[k];
        }
    }

    // adjoint of speed: bv -> bh, bu (accumulated)
    PORTABLE_FN static inline void speed_vjp(const HR& s, const T* p, const T* n, T bv, T& bh, T* bu) {
        for (int j = 0; j < NM; ++j) {
            const T a = s.u[j] * n[0] + s.u[NM + j] * n[1];
            const T sb = a < (T)0 ? -bv : bv;
            bu[j] += n[0] * sb;
            bu[NM + j] += n[1] * sb;
        }
        bh += p[0] * bv / ((T)2 * std::sqrt(p[0] * s.h));
    }

    // adjoint of reconstruct: (bh, bm, bu) -> g (accumulated); returns the adjoint of bstar
    PORTABLE_FN static inline T reconstruct_vjp(const T* Q, T bstar, const HR& s, T bh, T* bm, const T* bu,
                                                T* g) {
        const T den = std::max((T)1.0e-14, Q[1]);
        const T r = s.r, hi = s.hi;
        T bhi = 0;
        for (int k = 0; k < 2 * NM; ++k) {
            bm[k] += hi * bu[k];
            bhi += s.m[k] * bu[k];
        }
        bh += hinv_kp_vjp(s.h, eps_flux, bhi, hi);
        T br = 0;
        for (int k = 0; k < 2 * NM; ++k) {
            g[IU + k] += r * bm[k];
            br += Q[IU + k] * bm[k];
        }
        for (int k = 0; k < NTR; ++k) {
            g[IT + k] += r * bm[2 * NM + k];
            br += Q[IT + k] * bm[2 * NM + k];
        }
        const T brd = br / den;   // d r / d h* = 1 / den, d r / d h = -r / den
        bh += brd;
        if ((T)1.0e-14 < Q[1]) g[1] -= r * brd;
        if ((T)0 < Q[0] + Q[1] - bstar) {
            g[0] += bh;
            g[1] += bh;
            return -bh;
        }
        return (T)0;
    }

    // gL += (d Phi / d Q_minus)^T, gR += (d Phi / d Q_plus)^T for Phi = bF . F + bD . D, with
    // (F, D) = numerical_flux_fluctuations(Q_minus, Q_plus, ...).  These numerics read neither
    // aux argument, so the states are the only inputs.
    PORTABLE_FN static inline void numerical_flux_fluctuations_vjp(const T* Q_minus, const T* Q_plus, const T* p,
                                                                   const T* n, const T* bF, const T* bD, T* gL,
                                                                   T* gR) {
        constexpr int NMOM = 2 * NM + NTR;
        const T bstar = std::max(Q_minus[0], Q_plus[0]);
        HR sL, sR;
        reconstruct(Q_minus, bstar, sL);
        reconstruct(Q_plus, bstar, sR);
        T bhL = 0, bhR = 0, bmL[NMOM], bmR[NMOM], buL[2 * NM], buR[2 * NM];
        for (int k = 0; k < NMOM; ++k) bmL[k] = bmR[k] = 0;
        for (int k = 0; k < 2 * NM; ++k) buL[k] = buR[k] = 0;
        {   // F = (F(q*_L) + F(q*_R)) / 2
            T hb[n_dof_q];
            for (int i = 0; i < n_dof_q; ++i) hb[i] = (T)0.5 * bF[i];
            phys_flux_vjp(sL, p, n, hb, bhL, bmL, buL);
            phys_flux_vjp(sR, p, n, hb, bhR, bmR, buR);
        }
        const T spL = speed(sL, p, n), spR = speed(sR, p, n);
        const T sp = (T)0.5 * std::max(spL, spR);
        T dq[n_dof_q], bdq[n_dof_q];
        dq[0] = 0;
        dq[1] = sR.h - sL.h;
        for (int k = 0; k < NMOM; ++k) dq[IU + k] = sR.m[k] - sL.m[k];
        T bsp = 0;
        for (int i = 0; i < n_dof_q; ++i) {
            const T t = bD[i] - bD[n_dof_q + i];
            bsp += dq[i] * t;
            bdq[i] = sp * t;
        }
        if (NM > 1) {
            const T gt[3] = {(T)0.1127016653792583, (T)0.5, (T)0.8872983346207417};
            const T gw[3] = {(T)0.2777777777777778, (T)0.4444444444444444, (T)0.2777777777777778};
            T hq[3], hiq[3], ubar[2 * NM], bubar[2 * NM];
            for (int k = 0; k < 2 * NM; ++k) ubar[k] = bubar[k] = 0;
            for (int q = 0; q < 3; ++q) {
                hq[q] = sL.h + gt[q] * dq[1];
                hiq[q] = M::hinv_kp(hq[q], eps_flux);
                for (int k = 0; k < 2 * NM; ++k) ubar[k] += gw[q] * ((sL.m[k] + gt[q] * dq[IU + k]) * hiq[q]);
            }
            T dD[NM], bdD[NM];
            for (int j = 0; j < NM; ++j) {
                dD[j] = dq[IU + j] * n[0] + dq[IV + j] * n[1];
                bdD[j] = 0;
            }
            for (int i = 1; i < NM; ++i) {
                const T bx = (T)0.5 * (bD[IU + i] + bD[n_dof_q + IU + i]);
                const T by = (T)0.5 * (bD[IV + i] + bD[n_dof_q + IV + i]);
                bubar[0] -= dD[i] * bx;
                bubar[NM] -= dD[i] * by;
                bdD[i] -= ubar[0] * bx + ubar[NM] * by;
                for (int j = 1; j < NM; ++j)
                    for (int k = 1; k < NM; ++k) {
                        const T b = (T)K::B[(i * NM + j) * NM + k];
                        if (b != (T)0) {
                            bubar[k] += b * dD[j] * bx;
                            bubar[NM + k] += b * dD[j] * by;
                            bdD[j] += b * (ubar[k] * bx + ubar[NM + k] * by);
                        }
                    }
            }
            for (int j = 0; j < NM; ++j) {
                bdq[IU + j] += n[0] * bdD[j];
                bdq[IV + j] += n[1] * bdD[j];
            }
            for (int q = 0; q < 3; ++q) {
                T bhi = 0;
                for (int k = 0; k < 2 * NM; ++k) {
                    const T w = gw[q] * bubar[k];
                    bmL[k] += hiq[q] * w;
                    bdq[IU + k] += gt[q] * (hiq[q] * w);
                    bhi += (sL.m[k] + gt[q] * dq[IU + k]) * w;
                }
                const T bh = hinv_kp_vjp(hq[q], eps_flux, bhi, hiq[q]);
                bhL += bh;
                bdq[1] += gt[q] * bh;
            }
        }
        {   // hydrostatic-reconstruction pressure corrections
            const T gh = (T)0.5 * p[0];
            const T bcR = n[0] * bD[IU] + n[1] * bD[IV];
            const T bcL = n[0] * bD[n_dof_q + IU] + n[1] * bD[n_dof_q + IV];
            bhR += gh * ((T)2 * sR.h) * bcR;
            gR[1] -= gh * ((T)2 * Q_plus[1]) * bcR;
            gL[1] += gh * ((T)2 * Q_minus[1]) * bcL;
            bhL -= gh * ((T)2 * sL.h) * bcL;
        }
        if (spL < spR)
            speed_vjp(sR, p, n, (T)0.5 * bsp, bhR, buR);
        else
            speed_vjp(sL, p, n, (T)0.5 * bsp, bhL, buL);
        bhR += bdq[1];
        bhL -= bdq[1];
        for (int k = 0; k < NMOM; ++k) {
            bmR[k] += bdq[IU + k];
            bmL[k] -= bdq[IU + k];
        }
        const T bb = reconstruct_vjp(Q_minus, bstar, sL, bhL, bmL, buL, gL) +
                     reconstruct_vjp(Q_plus, bstar, sR, bhR, bmR, buR, gR);
        if (Q_minus[0] < Q_plus[0])
            gR[0] += bb;
        else
            gL[0] += bb;
    }

    PORTABLE_FN static inline SimpleArray<T, 1> local_max_abs_eigenvalue(const T* Q, const T* Qaux,
                                                                          const T* p, const T* n) {
        SimpleArray<T, 1> res;
        res[0] = M::wave_speed(Q, Qaux[0], p, n);
        return res;
    }
};
