// SME of level L with Manning bottom friction and a per-cell Manning coefficient, on the
// generated-model surface (Model.H:29-398).  Not emitted by the reference's generator: the
// friction law and its per-cell coefficient field are this build's calibration model.
//
//   aux (n_dof_qaux = 2): [hinv, n]   hinv = KP-regularised 1/h, recomputed from h;
//                                     n = Manning coefficient, passed through unchanged
//   parameters:           [g]
//   source on the mean momenta:  S_{hu_0} = -g n^2 |u| u h^{-1/3},  |u| = sqrt(u^2 + v^2 + eps^2)
//   (u = h u_0 hinv, eps = 1e-8 keeps |u| differentiable at rest); higher moments unforced.
// h^{-1/3} = cbrt(hinv) is computed with exact power-of-two scaling and a fixed number of
// Newton steps, so host and device builds round alike.  The Jacobians are the exact
// derivatives of the same source, taken with dual numbers.
#pragma once
#include "dual.H"   // first: the model templates below name std::sqrt / std::fabs on it
#include "sme2d.H"

// cube root of x >= 0: x = m 2^(3k) with m in [0.5, 4), Newton for m^(1/3) from 1
PORTABLE_FN inline double zdm_cbrt(double x) {
    if (!(x > 0.0)) return 0.0;
    int e = 0;
    const double m = std::frexp(x, &e);
    const int r = ((e % 3) + 3) % 3;
    const double mm = std::ldexp(m, r);
    double y = 1.0;
    for (int i = 0; i < 7; ++i) y = (2.0 * y + mm / (y * y)) / 3.0;
    return std::ldexp(y, (e - r) / 3);
}
template <int N>
PORTABLE_FN inline Dual<N> zdm_cbrt(const Dual<N>& x) {
    const double y = zdm_cbrt(x.v);
    Dual<N> r(y);
    for (int k = 0; k < N; ++k) r.d[k] = y > 0.0 ? x.d[k] / (3.0 * y * y) : 0.0;
    return r;
}

template <typename T, int L>
struct SmeManning2D : SmeModel2D<T, L> {
    using B = SmeModel2D<T, L>;
    using B::IU;
    using B::IV;
    static constexpr int n_dof_q = B::n_dof_q;
    static constexpr int n_dof_qaux = 2;
    static constexpr int n_parameters = 1;
    static const std::vector<std::string> parameter_names() { return {"g"}; }
    static const std::vector<T> default_parameters() { return {9.81}; }

    template <typename S>
    PORTABLE_FN static inline void source_of(const S* Q, const S* Qaux, double g, S* res) {
        for (int i = 0; i < n_dof_q; ++i) res[i] = 0;
        const S hinv = Qaux[0], n = Qaux[1];
        const S u = Q[IU] * hinv, v = Q[IV] * hinv;
        const S speed = std::sqrt(u * u + v * v + (S)1.0e-16);
        const S k = g * (n * n) * speed * zdm_cbrt(hinv);
        res[IU] = -(k * u);
        res[IV] = -(k * v);
    }

    PORTABLE_FN static inline SimpleArray<T, n_dof_q> source(const T* Q, const T* Qaux, const T* p,
                                                             const T time, const T dt, const T* X) {
        SimpleArray<T, n_dof_q> res;
        source_of<T>(Q, Qaux, zdm_value(p[0]), res.data);
        return res;
    }

    // d S / d Q at fixed aux, row-major (values only: the tangent of a Jacobian is not needed
    // by the Newton iteration that consumes it)
    PORTABLE_FN static inline SimpleArray<T, n_dof_q * n_dof_q> source_jacobian_wrt_variables(
        const T* Q, const T* Qaux, const T* p) {
        using D = Dual<n_dof_q>;
        D q[n_dof_q], a[2], s[n_dof_q];
        for (int i = 0; i < n_dof_q; ++i) q[i] = D::seed(zdm_value(Q[i]), i);
        a[0] = D(zdm_value(Qaux[0]));
        a[1] = D(zdm_value(Qaux[1]));
        source_of<D>(q, a, zdm_value(p[0]), s);
        SimpleArray<T, n_dof_q * n_dof_q> J;
        for (int i = 0; i < n_dof_q; ++i)
            for (int j = 0; j < n_dof_q; ++j) J[i * n_dof_q + j] = (T)s[i].d[j];
        return J;
    }

    PORTABLE_FN static inline SimpleArray<T, n_dof_q * n_dof_qaux> source_jacobian_wrt_aux_variables(
        const T* Q, const T* Qaux, const T* p) {
        using D = Dual<2>;
        D q[n_dof_q], a[2], s[n_dof_q];
        for (int i = 0; i < n_dof_q; ++i) q[i] = D(zdm_value(Q[i]));
        a[0] = D::seed(zdm_value(Qaux[0]), 0);
        a[1] = D::seed(zdm_value(Qaux[1]), 1);
        source_of<D>(q, a, zdm_value(p[0]), s);
        SimpleArray<T, n_dof_q * n_dof_qaux> J;
        for (int i = 0; i < n_dof_q; ++i)
            for (int j = 0; j < 2; ++j) J[i * 2 + j] = (T)s[i].d[j];
        return J;
    }

    PORTABLE_FN static inline SimpleArray<T, n_dof_qaux> update_aux_variables(const T* Q, const T* Qaux,
                                                                              const T* p, const T time,
                                                                              const T* X) {
        SimpleArray<T, n_dof_qaux> res;
        res[0] = B::hinv_kp(Q[1], (T)1.0e-8);
        res[1] = Qaux ? Qaux[1] : (T)0;
        return res;
    }
};
