// SME level 0 on a 1-D mesh with bathymetry in the state: the model the reference checks in
// (Model.H:29-398, generated by create_model.py from zoomy_core).  Restated from its
// equations, not copied; same state, aux, parameters, tags and semantics:
//   state  [b, h, hu]                      (Model.H:31-40)
//   aux    [dq0dx, dhdx, dbdx, hinv]       (slots 0-2 pass through the pointwise update,
//                                           Model.H:209-222, and are filled by the field-level
//                                           update_aux_variables, Model.H:250-262)
//   params [g, rho, nu, lambda_s, e_x] = [9.81, 1, 1, 0.5, 0]   (Model.H:43-44)
//   source S_hu = e_x g h - hinv lambda_s hu / rho               (Model.H:110-123)
//   clamp  |hu| <= 30 max(0, h - 0.01)                           (Model.H:195-207)
//   wall   hu -> -hu                                             (Model.H:356-375)
// pow() calls of the generated form are written as products and sqrt so that the host and
// device builds round identically.
#pragma once
#include <cmath>
#include <vector>
#include <string>
#include <algorithm>

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

template <typename T>
struct Model {
    static constexpr int n_dof_q = 3;
    static constexpr int n_dof_qaux = 4;
    static constexpr int n_parameters = 5;
    static constexpr int dimension = 1;
    static constexpr double dt_max = 5.0;
    static constexpr int n_dof_gradQ = 0;
    static constexpr bool has_diffusion = false;
    static constexpr bool has_free_surface = true;
    static constexpr int idx_b = 0;
    static constexpr int idx_h = 1;
    static constexpr int n_boundary_tags = 1;
    static const std::vector<std::string> get_boundary_tags() { return {"default"}; }
    static const std::vector<std::string> parameter_names() { return {"g", "rho", "nu", "lambda_s", "e_x"}; }
    static const std::vector<T> default_parameters() { return {9.81, 1.0, 1.0, 0.5, 0.0}; }

    PORTABLE_FN static inline SimpleArray<T, 3> source(const T* Q, const T* Qaux, const T* p,
                                                       const T time, const T dt, const T* X) {
        SimpleArray<T, 3> r;
        r[0] = 0;
        r[1] = 0;
        r[2] = p[4] * p[0] * Q[1] - Qaux[3] * p[3] * Q[2] / p[1];
        return r;
    }

    PORTABLE_FN static inline SimpleArray<T, 9> source_jacobian_wrt_variables(const T* Q, const T* Qaux,
                                                                              const T* p) {
        SimpleArray<T, 9> J;
        for (int i = 0; i < 9; ++i) J[i] = 0;
        const T k = p[3] / p[1];
        J[7] = p[4] * p[0] + Qaux[3] * Qaux[3] * Q[2] * k;   // dS_hu / dh
        J[8] = -Qaux[3] * k;                                  // dS_hu / dhu
        return J;
    }

    PORTABLE_FN static inline SimpleArray<T, 12> source_jacobian_wrt_aux_variables(const T* Q, const T* Qaux,
                                                                                   const T* p) {
        SimpleArray<T, 12> J;
        for (int i = 0; i < 12; ++i) J[i] = 0;
        return J;
    }

    PORTABLE_FN static inline SimpleArray<T, 3> update_variables(const T* Q, const T* Qaux, const T* p,
                                                                 const T dt) {
        const T lim = (T)30.0 * std::max((T)0, Q[1] - (T)0.01);
        SimpleArray<T, 3> r;
        r[0] = Q[0];
        r[1] = Q[1];
        r[2] = std::max(-lim, std::min(Q[2], lim));
        return r;
    }

    PORTABLE_FN static inline SimpleArray<T, 4> update_aux_variables(const T* Q, const T* Qaux, const T* p,
                                                                     const T time, const T* X) {
        SimpleArray<T, 4> r;
        r[0] = Qaux[0];
        r[1] = Qaux[1];
        r[2] = Qaux[2];
        const T h = Q[1], hm = std::max((T)1.0e-8, Q[1]);
        const T h2 = h * h, hm2 = hm * hm;
        r[3] = (T)1.4142135623730951 * std::max((T)0, h) / std::sqrt(h2 * h2 + hm2 * hm2);
        return r;
    }

    // Field-level, mesh-aware refresh of the derivative aux (Model.H:250-262): aux slots 0-2
    // are d(hu)/dx, dh/dx and db/dx of the per-cell fields, by compute_derivative
    // (ZoomyMesh and compute_derivative come from UserFunctions.H, included before this header)
    static inline void update_aux_variables(T* const* Q, T* const* Qaux, const T dt, const ZoomyMesh& mesh) {
        compute_derivative(Qaux[0], Q[2], 1, 0, 0, mesh);
        compute_derivative(Qaux[1], Q[1], 1, 0, 0, mesh);
        compute_derivative(Qaux[2], Q[0], 1, 0, 0, mesh);
    }

    PORTABLE_FN static inline SimpleArray<T, 3> boundary_conditions(const int bc_idx, const T* Q,
                                                                    const T* Qaux, const T* p, const T* n,
                                                                    const T* X, const T time, const T dX) {
        SimpleArray<T, 3> r;
        if (bc_idx == 0) {
            r[0] = Q[0];
            r[1] = Q[1];
            r[2] = -Q[2];
            return r;
        }
        r[0] = 0;
        r[1] = 0;
        r[2] = 0;
        return r;
    }

    PORTABLE_FN static inline SimpleArray<T, 3> eigenvalues(const T* Q, const T* Qaux, const T* p,
                                                            const T* n) {
        const T hf = std::max((T)1.0e-14, Q[1]);
        const T h5 = Q[1] * Q[1] * Q[1] * Q[1] * Q[1];
        const T c = std::sqrt(p[0]) * n[0] * std::sqrt(std::max((T)0, h5)) / (hf * hf);
        const T un = n[0] * Q[2] / hf;
        SimpleArray<T, 3> r;
        r[0] = 0;
        r[1] = un - c;
        r[2] = un + c;
        return r;
    }

    // Optional centred bathymetry term of the second-order hydrostatic reconstruction (not in
    // the generated header; used only when the solver's well_balanced extension is on).
    PORTABLE_FN static inline SimpleArray<T, 3> hydrostatic_cell_source(const T* Qface, const T* Qcell,
                                                                        const T* p, const T* n) {
        SimpleArray<T, 3> r;
        const T d = (Qcell[0] + Qcell[1]) - Qface[0];
        r[0] = 0;
        r[1] = 0;
        r[2] = (T)0.5 * p[0] * (d * d) * n[0];
        return r;
    }

    PORTABLE_FN static inline SimpleArray<T, 3> initial_condition(const T* X, const T* p) {
        SimpleArray<T, 3> r;
        r[0] = 1.0;
        r[1] = 0.0;
        r[2] = 0.0;
        return r;
    }
};
