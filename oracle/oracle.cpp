// ORACLE -- test infrastructure only.  CPU restatement of the reference's finite-volume time
// step (ZoomyLab/zoomy-dmplex), used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker.  Nothing in the product links or calls this file.
//
// It follows the reference's own loop structure (face loops that scatter into both cells,
// array-of-structs state x[c*n_q + i], gradient g[(c*n_q + i)*dim + d]) so that it is an
// independent restatement of the device path, which gathers per cell.  Each function cites
// the reference lines it restates.  It is instantiated on a model directory (Model.H +
// Numerics.H) given on the include path, exactly as the reference is built on its generated
// headers.
//
// Pinning: see tests/test_oracle_*.py (closed forms, invariants, the reference's KATs) and
// tests/test_reference_vectors.py: the model headers the oracle and the device share (the
// restated 1-D SME level 0 Model.H / Numerics.H and the VAM ChorinOps.H operators) against
// vectors evaluated from the reference's generated headers' own expression text.
// Parity unpinned: PETSc's TSSSP RKS2 / VecAXPBYPCZ rounding order (third-party, restated
// from its published algorithm), the DMPlex face numbering (the oracle uses the mesh
// service's ordering).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <vector>

#include <type_traits>
#include <utility>

#include "zoomy/UserFunctions.H"   // ZoomyMesh, compute_derivative, ...: before the model
#ifdef ORACLE_TANGENT
#include "dual.H"   // before the model: its templates name std::sqrt / std::fabs on duals
#endif
#include "Numerics.H"
#include "zdm_mesh.h"

template <typename TT, typename = void>
struct HasCellSource : std::false_type {};
template <typename TT>
struct HasCellSource<TT, std::void_t<decltype(Model<TT>::hydrostatic_cell_source(
                             std::declval<const TT*>(), std::declval<const TT*>(),
                             std::declval<const TT*>(), std::declval<const TT*>()))>>
    : std::true_type {};

// Model::update_variables (the wet/dry clamp) is optional in a generated model: without it the
// state is kept (HasUpdateVariables, TransportStep.hpp:23-28, 219)
template <typename TT, typename = void>
struct HasUpdateVariablesO : std::false_type {};
template <typename TT>
struct HasUpdateVariablesO<TT, std::void_t<decltype(Model<TT>::update_variables(
                                   std::declval<const TT*>(), std::declval<const TT*>(),
                                   std::declval<const TT*>(), std::declval<const TT>()))>> : std::true_type {};
template <typename S>
SimpleArray<S, Model<S>::n_dof_q> model_update_variables(const S* q, const S* a, const S* p, const S dt) {
    if constexpr (HasUpdateVariablesO<S>::value) {
        return Model<S>::update_variables(q, a, p, dt);
    } else {
        SimpleArray<S, Model<S>::n_dof_q> r;
        for (int i = 0; i < Model<S>::n_dof_q; ++i) r[i] = q[i];
        return r;
    }
}

// Model::diffusive_flux(q_face, a_face, gradQ_face, p), optional (TransportStep.hpp:10-18)
template <typename TT, typename = void>
struct HasDiffusiveFluxO : std::false_type {};
template <typename TT>
struct HasDiffusiveFluxO<TT, std::void_t<decltype(Model<TT>::diffusive_flux(
                                 std::declval<const TT*>(), std::declval<const TT*>(), std::declval<const TT*>(),
                                 std::declval<const TT*>()))>> : std::true_type {};

// the field-level, mesh-aware update_aux_variables(Q, Qaux, dt, mesh) (HasFieldAuxUpdate,
// TransportStep.hpp:30-38)
template <typename TT, typename = void>
struct HasFieldAuxUpdateO : std::false_type {};
template <typename TT>
struct HasFieldAuxUpdateO<TT, std::void_t<decltype(Model<TT>::update_aux_variables(
                                  std::declval<TT* const*>(), std::declval<TT* const*>(), std::declval<const TT>(),
                                  std::declval<const ZoomyMesh&>()))>> : std::true_type {};

// Model::initial_condition(X, p) is optional in a generated model (VirtualSolver.hpp:33-38)
template <typename MM, typename = void>
struct HasInitialConditionO : std::false_type {};
template <typename MM>
struct HasInitialConditionO<MM, std::void_t<decltype(MM::initial_condition(std::declval<const double*>(),
                                                                          std::declval<const double*>()))>>
    : std::true_type {};
template <typename MM>
void initial_condition_point(const double* X, const double* p, double* ic) {
    if constexpr (HasInitialConditionO<MM>::value) {
        auto v = MM::initial_condition(X, p);
        for (int i = 0; i < MM::n_dof_q; ++i) ic[i] = v[i];
    } else {
        for (int i = 0; i < MM::n_dof_q; ++i) ic[i] = 0;
    }
}

namespace {

#ifdef ORACLE_TANGENT
// tangent-linear build: every quantity carries one directional derivative
using T = Dual<1>;
#else
using T = double;
#endif
using M = Model<T>;
using NUM = Numerics<T>;
constexpr int NQ = M::n_dof_q;
constexpr int NA = M::n_dof_qaux;
constexpr int DIM = M::dimension;

// mesh coordinates (always double) as scalars of the build
struct P3 {
    T v[3];
    explicit P3(const double* p) {
        for (int d = 0; d < 3; ++d) v[d] = p[d];
    }
};
const T BIG = (T)std::numeric_limits<double>::max();
bool g_freeze_limiter = false;   // tangent build: hold each face's limiter value fixed
// solver.positivity at order 2: 0 none, 1 "mood" (eta-WB reconstruction, no theta, MOOD
// override), 2 "zhang_shu" (eta-WB reconstruction with the Zhang-Shu theta); wet_dry_eps
int g_positivity = 0;
double g_wet_eps = 1e-2;
long long g_troubled = 0;                    // MOOD: cells overridden so far
std::vector<uint8_t> g_last_mask;            // MOOD: the cells overridden in the last step that had any
const std::vector<uint8_t>* g_filter = nullptr;   // MOOD: faces touching these cells only

struct Opts {
    int order;            // 1 = PCM + forward Euler, 2 = MUSCL + SSP-RK
    int ssp_stages;       // s of SSP-RK(s,2)
    int use_limiters;     // Vacuum + TVD (ModularSolver.hpp:60, Reconstruction.hpp:170-175)
    int gradient;         // 0 = Green-Gauss (reference default), 1 = least squares
    int well_balanced;    // extension: Model::hydrostatic_cell_source per face side
    double cfl;
    double t_end;
    double min_dt;
};

// ---------------------------------------------------------------- TransportStep pieces

// TransportStep::UpdateState, TransportStep.hpp:196-230 (per cell clamp + local aux)
void update_state(int nc, const T* p, T* X, T* A) {
    const T X0[3] = {0, 0, 0};
    for (int c = 0; c < nc; ++c) {
        T* q = X + (size_t)c * NQ;
        T* a = A + (size_t)c * NA;
        auto rq = model_update_variables<T>(q, a, p, 0.0);
        for (int i = 0; i < NQ; ++i) q[i] = rq[i];
        auto ra = M::update_aux_variables(q, a, p, 0.0, X0);
        for (int i = 0; i < NA; ++i) a[i] = ra[i];
    }
}

// solver.refresh_derivative_aux (Settings.hpp:24, default true)
int g_refresh = 1;
int g_deriv_err = 0;

// compute_derivative on host fields (UserFunctions.H:60-112): the Green-Gauss face loop,
// 0.5 (s_L + s_R) N scattered into both cells in face order, a zero-Neumann face value s_L on
// boundary faces, then / volume (0 where the volume is not positive); order 0 copies
void host_field_derivative(void* ctx, double* result, const double* source, int dx, int dy, int dz) {
    const zdm_mesh* m = static_cast<const zdm_mesh*>(ctx);
    const int nc = m->ncells, total = dx + dy + dz;
    for (int c = 0; c < nc; ++c) result[c] = 0.0;
    if (total == 0) {
        for (int c = 0; c < nc; ++c) result[c] = source[c];
        return;
    }
    const int axis = dx == 1 ? 0 : (dy == 1 ? 1 : 2);
    if (total != 1 || axis >= DIM) {
        ++g_deriv_err;
        return;
    }
    std::vector<double> g(nc, 0.0);
    for (int f = 0; f < m->nfaces; ++f) {
        const double N = m->face_normal[3 * (size_t)f + axis];
        const int L = m->face_cells[2 * f], R = m->face_cells[2 * f + 1];
        if (R >= 0) {
            const double fv = 0.5 * (source[L] + source[R]);
            const double v = fv * N;
            g[L] += v;
            g[R] -= v;
        } else {
            g[L] += source[L] * N;
        }
    }
    for (int c = 0; c < nc; ++c)
        if (m->cell_volume[c] > 0.0) result[c] = g[c] / m->cell_volume[c];
}

// TransportStep::RefreshDerivativeAux (TransportStep.hpp:161-202): de-interleave X and A into
// per-field arrays, the model's field-level update_aux_variables(Q, Qaux, 0, mesh), scatter the
// aux back.  No-op for a model without it or with the setting off.
template <typename S>
void refresh_aux(const zdm_mesh* m, const S* X, S* A) {
    if constexpr (std::is_same<S, double>::value && HasFieldAuxUpdateO<S>::value) {
        if (!g_refresh) return;
        const int nc = m->ncells;
        std::vector<std::vector<S>> Qf(NQ, std::vector<S>(nc)), Af(NA, std::vector<S>(nc));
        for (int c = 0; c < nc; ++c) {
            for (int i = 0; i < NQ; ++i) Qf[i][c] = X[(size_t)c * NQ + i];
            for (int k = 0; k < NA; ++k) Af[k][c] = A[(size_t)c * NA + k];
        }
        std::vector<S*> Qp(NQ), Ap(NA);
        for (int i = 0; i < NQ; ++i) Qp[i] = Qf[i].data();
        for (int k = 0; k < NA; ++k) Ap[k] = Af[k].data();
        const ZoomyMesh mesh{const_cast<zdm_mesh*>(m), &host_field_derivative, DIM, 0, (long)nc};
        Model<S>::update_aux_variables(Qp.data(), Ap.data(), (S)0, mesh);
        for (int c = 0; c < nc; ++c)
            for (int k = 0; k < NA; ++k) A[(size_t)c * NA + k] = Af[k][c];
    }
    (void)m;
    (void)X;
    (void)A;
}

// TransportStep::UpdateNeighborBounds, TransportStep.hpp:122-154
void neighbor_bounds(const zdm_mesh* m, const T* X, T* vmin, T* vmax) {
    for (int c = 0; c < m->ncells; ++c) {
        const T* q = X + (size_t)c * NQ;
        T lo[NQ], hi[NQ];
        for (int i = 0; i < NQ; ++i) lo[i] = hi[i] = q[i];
        for (int k = 0; k < m->cell_nfaces[c]; ++k) {
            int n = m->cell_nbrs[(size_t)c * m->max_faces + k];
            if (n < 0) continue;
            const T* qn = X + (size_t)n * NQ;
            for (int i = 0; i < NQ; ++i) {
                if (qn[i] < lo[i]) lo[i] = qn[i];
                if (qn[i] > hi[i]) hi[i] = qn[i];
            }
        }
        for (int i = 0; i < NQ; ++i) {
            vmin[(size_t)c * NQ + i] = lo[i];
            vmax[(size_t)c * NQ + i] = hi[i];
        }
    }
}

// GreenGaussGradient::Compute, Gradient.hpp:46-190 (boundary face value = the model's BC
// ghost with Qaux = nullptr, Gradient.hpp:140-153 and ModularSolver.hpp:63-66)
void gradient_gg(const zdm_mesh* m, const int* face_bc, const T* p, const T* X, T* G) {
    std::memset(G, 0, sizeof(T) * (size_t)m->ncells * NQ * DIM);
    for (int f = 0; f < m->nfaces; ++f) {
        const double* N = m->face_normal + 3 * (size_t)f;
        const int cL = m->face_cells[2 * f], cR = m->face_cells[2 * f + 1];
        if (cR >= 0) {
            const T* qL = X + (size_t)cL * NQ;
            const T* qR = X + (size_t)cR * NQ;
            T* gL = G + (size_t)cL * NQ * DIM;
            T* gR = G + (size_t)cR * NQ * DIM;
            for (int i = 0; i < NQ; ++i) {
                T fv = 0.5 * (qL[i] + qR[i]);
                for (int d = 0; d < DIM; ++d) {
                    T v = fv * N[d];
                    gL[i * DIM + d] += v;
                    gR[i * DIM + d] -= v;
                }
            }
        } else if (face_bc[f] >= 0) {
            const T* qL = X + (size_t)cL * NQ;
            T area = 0;
            for (int d = 0; d < DIM; ++d) area += N[d] * N[d];
            area = std::sqrt(area);
            T nh[3] = {0, 0, 0};
            for (int d = 0; d < DIM; ++d) nh[d] = N[d] / area;
            auto qb = M::boundary_conditions(face_bc[f], qL, nullptr, p, nh,
                                             P3(m->face_centroid + 3 * (size_t)f).v, 0.0, 0.0);
            T* gL = G + (size_t)cL * NQ * DIM;
            for (int i = 0; i < NQ; ++i)
                for (int d = 0; d < DIM; ++d) gL[i * DIM + d] += qb[i] * N[d];
        }
    }
    for (int c = 0; c < m->ncells; ++c) {
        T vol = m->cell_volume[c];
        if (vol > 1e-15)
            for (int k = 0; k < NQ * DIM; ++k) G[(size_t)c * NQ * DIM + k] /= vol;
    }
}

// LeastSquaresGradient::SolveSmallSystemSPD, Gradient.hpp:205-217
bool chol_solve(int n, std::vector<T>& A, std::vector<T>& b) {
    std::vector<T> L(n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= i; j++) {
            T sum = 0;
            for (int k = 0; k < j; k++) sum += L[i * n + k] * L[j * n + k];
            if (i == j) {
                T val = A[i * n + i] - sum;
                if (val <= 0) return false;
                L[i * n + j] = std::sqrt(val);
            } else {
                L[i * n + j] = (A[i * n + j] - sum) / L[j * n + j];
            }
        }
    for (int i = 0; i < n; i++) {
        T sum = 0;
        for (int k = 0; k < i; k++) sum += L[i * n + k] * b[k];
        b[i] = (b[i] - sum) / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        T sum = 0;
        for (int k = i + 1; k < n; k++) sum += L[k * n + i] * b[k];
        b[i] = (b[i] - sum) / L[i * n + i];
    }
    return true;
}

// LeastSquaresGradient::Setup + Compute (order 1), Gradient.hpp:222-298, 300-359
void gradient_lsq(const zdm_mesh* m, const T* X, T* G) {
    std::memset(G, 0, sizeof(T) * (size_t)m->ncells * NQ * DIM);
    for (int c = 0; c < m->ncells; ++c) {
        const double* xc = m->cell_centroid + 3 * (size_t)c;
        std::vector<int> nb;
        std::vector<std::vector<T>> bv;
        for (int k = 0; k < m->cell_nfaces[c]; ++k) {
            int n = m->cell_nbrs[(size_t)c * m->max_faces + k];
            if (n < 0) continue;
            const double* xn = m->cell_centroid + 3 * (size_t)n;
            std::vector<T> b(DIM);
            for (int d = 0; d < DIM; ++d) b[d] = xn[d] - xc[d];
            nb.push_back(n);
            bv.push_back(b);
        }
        std::vector<T> ata(DIM * DIM, 0.0);
        for (const auto& b : bv)
            for (int i = 0; i < DIM; ++i)
                for (int j = 0; j < DIM; ++j) ata[i * DIM + j] += b[i] * b[j];
        const T* qc = X + (size_t)c * NQ;
        T* gc = G + (size_t)c * NQ * DIM;
        for (size_t k = 0; k < nb.size(); ++k) {
            std::vector<T> w = bv[k], sys = ata;
            if (!chol_solve(DIM, sys, w)) std::fill(w.begin(), w.end(), 0.0);
            const T* qn = X + (size_t)nb[k] * NQ;
            for (int i = 0; i < NQ; ++i) {
                T du = qn[i] - qc[i];
                for (int d = 0; d < DIM; ++d) gc[i * DIM + d] += w[d] * du;
            }
        }
    }
}

// LinearReconstructor::Reconstruct with VacuumLimiter + TVDLimiter,
// Reconstruction.hpp:29-35, 64-80, 203-243; PCMReconstructor 147-149 when grad == nullptr
void reconstruct(const T* q, const T* grad, const double* xc, const double* xf, const T* qmin,
                 const T* qmax, int use_limiters, T* out) {
    if (!grad) {
        for (int i = 0; i < NQ; ++i) out[i] = q[i];
        return;
    }
    T dx[3];
    for (int d = 0; d < DIM; ++d) dx[d] = xf[d] - xc[d];
    T unl[NQ], delta[NQ];
    for (int i = 0; i < NQ; ++i) {
        delta[i] = 0.0;
        for (int d = 0; d < DIM; ++d) delta[i] += grad[i * DIM + d] * dx[d];
        unl[i] = q[i] + delta[i];
    }
    T fin[NQ];
    for (int i = 0; i < NQ; ++i) fin[i] = 1.0;
    if (use_limiters) {
        // vacuum
        T tmp[NQ];
        for (int i = 0; i < NQ; ++i) tmp[i] = 1.0;
        if (NQ > 1 && qmin && qmin[1] < 1e-2)
            for (int i = 0; i < NQ; ++i) tmp[i] = 0.0;
        for (int i = 0; i < NQ; ++i)
            if (tmp[i] < fin[i]) fin[i] = tmp[i];
        // TVD (Barth-Jespersen, face-wise)
        for (int i = 0; i < NQ; ++i) tmp[i] = 1.0;
        if (qmin && qmax) {
            for (int i = 0; i < NQ; ++i) {
                T dq = unl[i] - q[i];
                if (std::fabs(dq) < 1e-12) continue;
                T a = 1.0;
                if (dq > 0) {
                    if (qmax[i] > q[i])
                        a = std::min((T)1.0, (qmax[i] - q[i]) / dq);
                    else
                        a = 0.0;
                } else {
                    if (qmin[i] < q[i])
                        a = std::min((T)1.0, (qmin[i] - q[i]) / dq);
                    else
                        a = 0.0;
                }
                tmp[i] = a;
            }
        }
        for (int i = 0; i < NQ; ++i)
            if (tmp[i] < fin[i]) fin[i] = tmp[i];
    }
    T alpha = 1.;
    for (int i = 0; i < NQ; ++i)
        if (alpha > fin[i]) alpha = fin[i];
#ifdef ORACLE_TANGENT
    if (g_freeze_limiter) alpha = T(alpha.v);   // frozen-limiter linearisation
#endif
    for (int i = 0; i < NQ; ++i) out[i] = q[i] + alpha * delta[i];
}

// TransportStep::LimitGradientsWB, TransportStep.hpp:241-323: the raw gradient G (of Q = [b,
// h, mom...]) becomes the effective W-gradients in place: slot b -> phi_b grad b, slot h ->
// phi_b grad b + theta sh, mom -> theta phi_mom grad mom; phi = Venkatakrishnan per W variable
// (W = b, eta = b + h, mom) over the face-neighbour extrema of the clamped states Xl, eps^2 =
// vol^(2/dim) (vol itself in 2-D; vol*vol in 1-D, where the reference's pow(vol, 2.0) may round
// differently in the last bit); theta = Zhang-Shu cap unless MOOD (skip_theta).
void limit_gradients_wb(const zdm_mesh* m, const T* Xl, T* G, bool skip_theta, double eps_wet) {
    constexpr int B = 0, H = 1;
    for (int c = 0; c < m->ncells; ++c) {
        T* gc = G + (size_t)c * NQ * DIM;
        const T* qc = Xl + (size_t)c * NQ;
        const double* cc = m->cell_centroid + 3 * (size_t)c;
        T g[NQ * DIM];
        for (int k = 0; k < NQ * DIM; ++k) g[k] = gc[k];
        T geta[DIM];
        for (int d = 0; d < DIM; ++d) geta[d] = g[H * DIM + d] + g[B * DIM + d];
        const T h_bar = qc[H], b_bar = qc[B];
        T Wc[NQ], Wmin[NQ], Wmax[NQ];
        for (int i = 0; i < NQ; ++i) Wc[i] = (i == H) ? (b_bar + h_bar) : qc[i];
        for (int i = 0; i < NQ; ++i) {
            Wmin[i] = Wc[i];
            Wmax[i] = Wc[i];
        }
        for (int k = 0; k < m->cell_nfaces[c]; ++k) {
            const int n = m->cell_nbrs[(size_t)c * m->max_faces + k];
            if (n < 0) continue;
            const T* qn = Xl + (size_t)n * NQ;
            for (int i = 0; i < NQ; ++i) {
                T w = (i == H) ? (qn[B] + qn[H]) : qn[i];
                if (w < Wmin[i]) Wmin[i] = w;
                if (w > Wmax[i]) Wmax[i] = w;
            }
        }
        const T vol = m->cell_volume[c];
        const T eps2 = DIM == 2 ? vol : vol * vol;
        T phi[NQ];
        for (int i = 0; i < NQ; ++i) phi[i] = 1.0;
        auto gW = [&](int i, int d) { return (i == H) ? geta[d] : g[i * DIM + d]; };
        for (int k = 0; k < m->cell_nfaces[c]; ++k) {
            const int f = m->cell_faces[(size_t)c * m->max_faces + k];
            const double* xf = m->face_centroid + 3 * (size_t)f;
            T r[3];
            for (int d = 0; d < DIM; ++d) r[d] = xf[d] - cc[d];
            for (int i = 0; i < NQ; ++i) {
                T delta = 0;
                for (int d = 0; d < DIM; ++d) delta += gW(i, d) * r[d];
                if (std::fabs(delta) < 1e-14) continue;
                T dm = (delta > 0) ? (Wmax[i] - Wc[i]) : (Wmin[i] - Wc[i]);
                T num = dm * dm + eps2 + 2 * delta * dm, den = dm * dm + 2 * delta * delta + delta * dm + eps2;
                T pf = (den > 1e-30) ? num / den : (T)1.0;
                pf = std::max((T)0, std::min((T)1, pf));
                if (pf < phi[i]) phi[i] = pf;
            }
        }
        if (h_bar < eps_wet)
            for (int i = 0; i < NQ; ++i) phi[i] = 0.0;
        phi[B] = std::min(phi[B], phi[H]);
        for (int i = 0; i < NQ; ++i)
            if (i != B && i != H) phi[i] = std::min(phi[i], phi[H]);
        T sh[DIM];
        for (int d = 0; d < DIM; ++d) sh[d] = phi[H] * geta[d] - phi[B] * g[B * DIM + d];
        T hmin = h_bar;
        for (int k = 0; k < m->cell_nfaces[c]; ++k) {
            const int f = m->cell_faces[(size_t)c * m->max_faces + k];
            const double* xf = m->face_centroid + 3 * (size_t)f;
            T hf = h_bar;
            for (int d = 0; d < DIM; ++d) hf += sh[d] * (xf[d] - cc[d]);
            if (hf < hmin) hmin = hf;
        }
        T theta = 1.0;
        if (!skip_theta && hmin < 0.0)
            theta = std::max((T)0, std::min((T)1, h_bar / std::max(h_bar - hmin, (T)1e-14)));
        for (int d = 0; d < DIM; ++d) {
            gc[B * DIM + d] = phi[B] * g[B * DIM + d];
            gc[H * DIM + d] = phi[B] * g[B * DIM + d] + theta * sh[d];
        }
        for (int i = 0; i < NQ; ++i)
            if (i != B && i != H)
                for (int d = 0; d < DIM; ++d) gc[i * DIM + d] = theta * phi[i] * g[i * DIM + d];
    }
}

// WBPositivityReconstructor::Reconstruct, Reconstruction.hpp:256-287: eta and b extrapolated
// with the effective gradients, h_f = max(eta_f - b_f, 0), momenta zeroed where h_f < eps_wet
void reconstruct_wb(const T* q, const T* grad, const double* xc, const double* xf, double eps_wet, T* out) {
    if (!grad) {
        for (int i = 0; i < NQ; ++i) out[i] = q[i];
        return;
    }
    T dx[3];
    for (int d = 0; d < DIM; ++d) dx[d] = xf[d] - xc[d];
    auto dot = [&](int slot) {
        T s = (T)0;
        for (int d = 0; d < DIM; ++d) s += grad[slot * DIM + d] * dx[d];
        return s;
    };
    const T b_bar = q[0], h_bar = q[1];
    const T b_f = b_bar + dot(0);
    const T eta_f = (b_bar + h_bar) + dot(1);
    const T h_f = std::max(eta_f - b_f, (T)0);
    for (int i = 0; i < NQ; ++i) {
        if (i == 0)
            out[i] = b_f;
        else if (i == 1)
            out[i] = h_f;
        else
            out[i] = (h_f < eps_wet) ? (T)0 : (q[i] + dot(i));
    }
}

// the face reconstruction the transport step is configured with (ModularSolver.hpp:46-61)
void face_state(const T* q, const T* grad, const double* xc, const double* xf, const T* qmin, const T* qmax,
                int use_limiters, T* out) {
    if (g_positivity != 0 && grad)
        reconstruct_wb(q, grad, xc, xf, g_wet_eps, out);
    else
        reconstruct(q, grad, xc, xf, qmin, qmax, use_limiters, out);
}

// TransportStep::ComputeFluxes, TransportStep.hpp:383-565 (owned = all cells on one rank)
// Extension (well_balanced): after a face's flux, each side adds
// area * Model::hydrostatic_cell_source(q_face, q_cell, p, outward normal).
template <typename S>
void cell_source(int wb, const S* qf, const S* qc, const S* p, const S* nh, double sign, S area, S* f) {
    if constexpr (HasCellSource<S>::value) {
        if (!wb) return;
        const S no[3] = {sign < 0 ? -nh[0] : nh[0], sign < 0 ? -nh[1] : nh[1], 0};
        auto Sv = Model<S>::hydrostatic_cell_source(qf, qc, p, no);
        for (int i = 0; i < NQ; ++i) f[i] += Sv[i] * area;
    }
}

// the diffusive face term (TransportStep.hpp:468-491): averaged state, aux and gradient of the
// two sides; fL += d area, fR -= d area
template <typename S>
void diffusive_face(const S* qL, const S* qR, const S* aL, const S* aR, const S* gL, const S* gR, const S* p,
                    S area, S* fL, S* fR) {
    if constexpr (HasDiffusiveFluxO<S>::value) {
        S gq[NQ * DIM], qf[NQ], af[NA];
        for (int k = 0; k < NQ * DIM; ++k) gq[k] = 0.5 * (gL[k] + gR[k]);
        for (int i = 0; i < NQ; ++i) qf[i] = 0.5 * (qL[i] + qR[i]);
        for (int i = 0; i < NA; ++i) af[i] = 0.5 * (aL[i] + aR[i]);
        auto d = Model<S>::diffusive_flux(qf, af, gq, p);
        for (int i = 0; i < NQ; ++i) {
            fL[i] += d[i] * area;
            fR[i] -= d[i] * area;
        }
    }
}
// ... at a boundary face (TransportStep.hpp:533-550): the interior cell's gradient, the state
// averaged with the ghost, the interior face aux
template <typename S>
void diffusive_boundary(const S* qL, const S* qR, const S* aL, const S* gL, const S* p, S area, S* fL) {
    if constexpr (HasDiffusiveFluxO<S>::value) {
        S qf[NQ], af[NA];
        for (int i = 0; i < NQ; ++i) qf[i] = 0.5 * (qL[i] + qR[i]);
        for (int i = 0; i < NA; ++i) af[i] = aL[i];
        auto d = Model<S>::diffusive_flux(qf, af, gL, p);
        for (int i = 0; i < NQ; ++i) fL[i] += d[i] * area;
    }
}

void compute_fluxes(const zdm_mesh* m, const int* face_bc, const T* p, T time, const T* X,
                    const T* A, const T* G, const T* vmin, const T* vmax, int use_limiters, int wb,
                    T* F) {
    for (int f = 0; f < m->nfaces; ++f) {
        const double* N = m->face_normal + 3 * (size_t)f;
        const double* xf = m->face_centroid + 3 * (size_t)f;
        const P3 xfT(xf);
        T area = 0;
        for (int d = 0; d < DIM; ++d) area += N[d] * N[d];
        area = std::sqrt(area);
        if (area <= 1e-15) continue;
        T nh[3] = {0, 0, 0};
        for (int d = 0; d < DIM; ++d) nh[d] = N[d] / area;
        const int cL = m->face_cells[2 * f], cR = m->face_cells[2 * f + 1];
        // local MOOD: only faces touching a troubled cell (TransportStep.hpp:413-418)
        if (g_filter && !(*g_filter)[cL] && !(cR >= 0 && (*g_filter)[cR])) continue;
        if (cR >= 0) {
            T qLf[NQ], qRf[NQ], aLf[NA], aRf[NA];
            face_state(X + (size_t)cL * NQ, G ? G + (size_t)cL * NQ * DIM : nullptr,
                        m->cell_centroid + 3 * (size_t)cL, xf, vmin ? vmin + (size_t)cL * NQ : nullptr,
                        vmax ? vmax + (size_t)cL * NQ : nullptr, use_limiters, qLf);
            face_state(X + (size_t)cR * NQ, G ? G + (size_t)cR * NQ * DIM : nullptr,
                        m->cell_centroid + 3 * (size_t)cR, xf, vmin ? vmin + (size_t)cR * NQ : nullptr,
                        vmax ? vmax + (size_t)cR * NQ : nullptr, use_limiters, qRf);
            for (int i = 0; i < NA; ++i) {
                aLf[i] = A[(size_t)cL * NA + i];
                aRf[i] = A[(size_t)cR * NA + i];
            }
            auto raL = M::update_aux_variables(qLf, aLf, p, 0.0, xfT.v);
            for (int i = 0; i < NA; ++i) aLf[i] = raL[i];
            auto raR = M::update_aux_variables(qRf, aRf, p, 0.0, xfT.v);
            for (int i = 0; i < NA; ++i) aRf[i] = raR[i];
            auto flux = NUM::numerical_flux(qLf, qRf, aLf, aRf, p, nh);
            T* fL = F + (size_t)cL * NQ;
            T* fR = F + (size_t)cR * NQ;
            auto nc = NUM::numerical_fluctuations(qLf, qRf, aLf, aRf, p, nh);
            for (int i = 0; i < NQ; ++i) {
                fL[i] -= nc[NQ + i] * area;
                fR[i] -= nc[i] * area;
            }
            for (int i = 0; i < NQ; ++i) {
                fL[i] -= flux[i] * area;
                fR[i] += flux[i] * area;
            }
            if (G) diffusive_face<T>(qLf, qRf, aLf, aRf, G + (size_t)cL * NQ * DIM, G + (size_t)cR * NQ * DIM, p, area, fL, fR);
            cell_source(wb, qLf, X + (size_t)cL * NQ, p, nh, 1.0, area, fL);
            cell_source(wb, qRf, X + (size_t)cR * NQ, p, nh, -1.0, area, fR);
        } else if (face_bc[f] >= 0) {
            T qLf[NQ], aLf[NA], aRf[NA];
            face_state(X + (size_t)cL * NQ, G ? G + (size_t)cL * NQ * DIM : nullptr,
                        m->cell_centroid + 3 * (size_t)cL, xf, vmin ? vmin + (size_t)cL * NQ : nullptr,
                        vmax ? vmax + (size_t)cL * NQ : nullptr, use_limiters, qLf);
            for (int i = 0; i < NA; ++i) aLf[i] = A[(size_t)cL * NA + i];
            auto raL = M::update_aux_variables(qLf, aLf, p, time, xfT.v);
            for (int i = 0; i < NA; ++i) aLf[i] = raL[i];
            auto qR = M::boundary_conditions(face_bc[f], qLf, aLf, p, nh, xfT.v, time, 0.0);
            auto raR = M::update_aux_variables(qR.data, aLf, p, time, xfT.v);
            for (int i = 0; i < NA; ++i) aRf[i] = raR[i];
            auto flux = NUM::numerical_flux(qLf, qR.data, aLf, aRf, p, nh);
            T* fL = F + (size_t)cL * NQ;
            auto nc = NUM::numerical_fluctuations(qLf, qR.data, aLf, aRf, p, nh);
            for (int i = 0; i < NQ; ++i) fL[i] -= nc[NQ + i] * area;
            for (int i = 0; i < NQ; ++i) fL[i] -= flux[i] * area;
            if (G) diffusive_boundary<T>(qLf, qR.data, aLf, G + (size_t)cL * NQ * DIM, p, area, fL);
            cell_source(wb, qLf, X + (size_t)cL * NQ, p, nh, 1.0, area, fL);
        } else if (wb) {
            T qLf[NQ];
            face_state(X + (size_t)cL * NQ, G ? G + (size_t)cL * NQ * DIM : nullptr,
                        m->cell_centroid + 3 * (size_t)cL, xf, vmin ? vmin + (size_t)cL * NQ : nullptr,
                        vmax ? vmax + (size_t)cL * NQ : nullptr, use_limiters, qLf);
            cell_source(wb, qLf, X + (size_t)cL * NQ, p, nh, 1.0, area, F + (size_t)cL * NQ);
        }
    }
    for (int c = 0; c < m->ncells; ++c) {
        T vol = m->cell_volume[c];
        if (vol > 1e-15)
            for (int i = 0; i < NQ; ++i) F[(size_t)c * NQ + i] /= vol;
    }
}

struct Work {
    std::vector<T> Xl, Al, G, vmin, vmax;
};

// TransportStep::FormRHS, TransportStep.hpp:325-357.  A_in stands for the scratch local aux
// vector the reference passes (its contents feed only pass-through aux slots).
void form_rhs(const zdm_mesh* m, const int* face_bc, const T* p, const Opts& o, T time, const T* X,
              const T* A_in, T* F, Work& w) {
    const size_t nc = m->ncells;
    std::fill(F, F + nc * NQ, 0.0);
    w.Xl.assign(X, X + nc * NQ);
    w.Al.assign(A_in, A_in + nc * NA);
    refresh_aux(m, w.Xl.data(), w.Al.data());   // UpdateState: the derivative aux of the unclamped X first
    update_state(m->ncells, p, w.Xl.data(), w.Al.data());
    const bool grad = o.order >= 2;
    if (grad) {
        w.vmin.resize(nc * NQ);
        w.vmax.resize(nc * NQ);
        neighbor_bounds(m, w.Xl.data(), w.vmin.data(), w.vmax.data());
        w.G.resize(nc * NQ * DIM);
        if (o.gradient == 1)
            gradient_lsq(m, X, w.G.data());
        else
            gradient_gg(m, face_bc, p, X, w.G.data());
        if (g_positivity != 0) limit_gradients_wb(m, w.Xl.data(), w.G.data(), g_positivity == 1, g_wet_eps);
    }
    compute_fluxes(m, face_bc, p, time, w.Xl.data(), w.Al.data(), grad ? w.G.data() : nullptr,
                   grad ? w.vmin.data() : nullptr, grad ? w.vmax.data() : nullptr, o.use_limiters,
                   o.well_balanced, F);
}

// VirtualSolver::ComputeTimeStep, VirtualSolver.hpp:285-357
T compute_time_step(const zdm_mesh* m, const T* p, const T* X, const T* A, T cfl, int* finite) {
    T local_min = BIG;
    for (int c = 0; c < m->ncells; ++c) {
        const T* qc = X + (size_t)c * NQ;
        const T* ac = A + (size_t)c * NA;
        T eig = 0.0;
        for (int d = 0; d < DIM; ++d) {
            T n[3] = {0, 0, 0};
            n[d] = 1.0;
            auto r = NUM::local_max_abs_eigenvalue(qc, ac, p, n);
            if (r[0] > eig) eig = r[0];
        }
        if (eig < 1e-14) continue;
        const double* xc = m->cell_centroid + 3 * (size_t)c;
        T rc = BIG;
        for (int k = 0; k < m->cell_nfaces[c]; ++k) {
            const double* xf = m->face_centroid + 3 * (size_t)m->cell_faces[(size_t)c * m->max_faces + k];
            T d2 = 0.0;
            for (int d = 0; d < DIM; ++d) {
                T dd = xc[d] - xf[d];
                d2 += dd * dd;
            }
            T dist = std::sqrt(d2);
            if (dist < rc) rc = dist;
        }
        if (rc < BIG) {
            T ratio = rc / eig;
            if (ratio < local_min) local_min = ratio;
        }
    }
    if (!std::isfinite(local_min) || local_min >= 0.5 * BIG) {
        if (finite) *finite = 0;
        return M::dt_max;
    }
    if (finite) *finite = 1;
    return cfl * local_min;
}

// SourceStep::SolveLinearSystem, SourceStep.hpp:211-242 (GEPP)
bool gepp(T* A, T* b, T* x, int n) {
    for (int k = 0; k < n; ++k) {
        int mr = k;
        T mv = std::fabs(A[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[i * n + k]) > mv) {
                mv = std::fabs(A[i * n + k]);
                mr = i;
            }
        if (mv < 1e-14) return false;
        if (mr != k) {
            for (int j = k; j < n; ++j) std::swap(A[k * n + j], A[mr * n + j]);
            std::swap(b[k], b[mr]);
        }
        for (int i = k + 1; i < n; ++i) {
            T fac = A[i * n + k] / A[k * n + k];
            for (int j = k; j < n; ++j) A[i * n + j] -= fac * A[k * n + j];
            b[i] -= fac * b[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        T sum = 0.0;
        for (int j = i + 1; j < n; ++j) sum += A[i * n + j] * x[j];
        x[i] = (b[i] - sum) / A[i * n + i];
    }
    return true;
}

// aux_jacobian_wrt_q (finite differences), SourceStep.hpp:26-53
void aux_jac_fd(const T* q, const T* aux, const T* p, T* J) {
    const T X0[3] = {0, 0, 0};
    for (int i = 0; i < NA * NQ; ++i) J[i] = 0;
    auto a0 = M::update_aux_variables(q, aux, p, 0.0, X0);
    T qp[NQ];
    for (int j = 0; j < NQ; ++j) {
        for (int k = 0; k < NQ; ++k) qp[k] = q[k];
        T h = (T)1e-7 * (std::fabs(q[j]) + (T)1e-7);
        qp[j] += h;
        auto a1 = M::update_aux_variables(qp, aux, p, 0.0, X0);
        for (int k = 0; k < NA; ++k) J[k * NQ + j] = (a1[k] - a0[k]) / h;
    }
}

// SourceStep::SolveLocalCell, SourceStep.hpp:125-209 (<= 15 Newton its, |R|_2 < 1e-8)
int solve_local_cell(T dt, T* q, const T* aux_init, const T* p) {
    const T X0[3] = {0, 0, 0};
    T qo[NQ], qc[NQ], ac[NA], r[NQ], dl[NQ], J[NQ * NQ], Jx[NA * NQ];
    for (int i = 0; i < NQ; ++i) qo[i] = qc[i] = q[i];
    for (int i = 0; i < NA; ++i) ac[i] = aux_init ? aux_init[i] : 0.0;
    int it = 0;
    for (it = 0; it < 15; ++it) {
        auto ua = M::update_aux_variables(qc, ac, p, 0.0, X0);
        for (int i = 0; i < NA; ++i) ac[i] = ua[i];
        auto S = M::source(qc, ac, p, (T)0, dt, X0);
        T rn = 0.0;
        for (int i = 0; i < NQ; ++i) {
            r[i] = qc[i] - qo[i] - dt * S[i];
            rn += r[i] * r[i];
        }
        if (std::sqrt(rn) < 1e-8) break;
        auto dSdQ = M::source_jacobian_wrt_variables(qc, ac, p);
        auto dSdA = M::source_jacobian_wrt_aux_variables(qc, ac, p);
        aux_jac_fd(qc, ac, p, Jx);
        for (int i = 0; i < NQ * NQ; ++i) J[i] = 0.0;
        for (int i = 0; i < NQ; ++i) J[i * NQ + i] = 1.0;
        for (int i = 0; i < NQ; ++i)
            for (int j = 0; j < NQ; ++j) {
                T term = dSdQ[i * NQ + j];
                for (int k = 0; k < NA; ++k) term += dSdA[i * NA + k] * Jx[k * NQ + j];
                J[i * NQ + j] -= dt * term;
            }
        for (int i = 0; i < NQ; ++i) r[i] = -r[i];
        if (gepp(J, r, dl, NQ)) {
            for (int i = 0; i < NQ; ++i) qc[i] += dl[i];
        } else {
            break;
        }
    }
    for (int i = 0; i < NQ; ++i) q[i] = qc[i];
    return it;
}

// SourceStep::Solve, SourceStep.hpp:72-122
void source_solve(int nc, const T* p, T dt, T* X, const T* A) {
    for (int c = 0; c < nc; ++c) solve_local_cell(dt, X + (size_t)c * NQ, A + (size_t)c * NA, p);
}

// MUSCLSolver::EnforcePhysicalConstraints, MUSCLSolver.hpp:234-254
void enforce(int nc, const T* p, T* X, T* A) { update_state(nc, p, X, A); }
// ... with the RefreshDerivativeAux of TransportStep::UpdateState (TransportStep.hpp:202)
void enforce(const zdm_mesh* m, const T* p, T* X, T* A) {
    refresh_aux(m, X, A);
    update_state(m->ncells, p, X, A);
}

// One TSStep: PETSc TSSSP "rks2" (s stages) for order >= 2, TSEULER otherwise
// (SolverStrategies.hpp:26-38).  SSP-RK(s,2):
//   u(i) = u(i-1) + dt/(s-1) L(u(i-1)),  i = 1..s-1          (VecAXPY)
//   u+   = (s-1)/s u(s-1) + dt/s L(u(s-1)) + 1/s u           (VecAXPBYPCZ)
// rhs(t, u, F) evaluates L; n = vector length.
template <typename RHS>
void ssp_or_euler(int order, int s, size_t n, T t0, T dt, T* X, RHS&& rhs) {
    std::vector<T> F(n);
    if (order < 2) {
        rhs(t0, X, F.data());
        for (size_t i = 0; i < n; ++i) X[i] += dt * F[i];
        return;
    }
    std::vector<T> w0(X, X + n);
    for (int i = 0; i < s - 1; ++i) {
        T st = t0 + dt * (i / (s - 1.));
        rhs(st, w0.data(), F.data());
        const T a = dt / (s - 1.);
        for (size_t k = 0; k < n; ++k) w0[k] += a * F[k];
    }
    rhs(t0 + dt, w0.data(), F.data());
    const T al = (s - 1.) / s, be = dt / s, ga = 1. / s;
    for (size_t k = 0; k < n; ++k) X[k] = al * w0[k] + be * F[k] + ga * X[k];
}

// IMEX step (IMEXStrategy, SolverStrategies.hpp:62-84: TSARKIMEX "2e", RHSFunction = FormRHS,
// IFunction = Xdot - source): the additive Runge-Kutta method ARK2 of Giraldo, Kelly and
// Constantinescu, SIAM J. Sci. Comput. 35(5), 2013, B1162-B1194 -- explicit first stage,
// L-stable SDIRK implicit part with diagonal g = 1 - 1/sqrt(2), weights b = (d, d, g),
// d = 1/(2 sqrt(2)), explicit a32 = (3 + 2 sqrt(2))/6, c = (0, 2g, 1).  PETSc itself is not
// available here: which member of the family its "2e" registers is unpinned.
//   Y1 = X;  Yi = X + dt sum_{j<i} (a_ij G_j + at_ij S_j) + dt g S(Yi);  X+ = X + dt sum b_i (G_i + S_i)
// rhs(t, U, F) = G; src(U, S) = S; solve(dg, Z, Y): Y - dg S(Y) = Z.  S_i = (Y_i - Z_i)/(dt g).
// Evaluated with running sums, each stage's terms added once they are known:
//   Z2 = X + dt (a21 G1 + at21 S1),  P3 = X + dt (a31 G1 + at31 S1 + at32 S2),
//   F = X + dt (b1 (G1 + S1) + b2 S2),  Z3 = P3 + dt a32 G2,  F += dt (b2 G2 + b3 S3),
//   X+ = F + dt b3 G3
namespace ark2o {
const double R2 = 1.41421356237309504880;
const double G = 1.0 - 1.0 / R2, D = 1.0 / (2.0 * R2);
const double A21 = 2.0 * G, A32 = (3.0 + 2.0 * R2) / 6.0, A31 = 1.0 - A32;
const double AT21 = G, AT31 = D, AT32 = D;
const double B1 = D, B2 = D, B3 = G, C2 = 2.0 * G;
}  // namespace ark2o

template <typename RHS, typename SRC, typename SOLVE>
void ark2_step(size_t n, T t0, T dt, T* X, RHS&& rhs, SRC&& src, SOLVE&& solve) {
    using namespace ark2o;
    std::vector<T> Gv(n), S1(n), Y(n), Z(n), P3(n), F(n);
    const T dg = dt * G;
    rhs(t0, X, Gv.data());
    src(X, S1.data());
    for (size_t k = 0; k < n; ++k) Z[k] = X[k] + dt * (A21 * Gv[k] + AT21 * S1[k]);
    solve(dg, Z.data(), Y.data());
    for (size_t k = 0; k < n; ++k) {
        const T s2 = (Y[k] - Z[k]) / dg;
        P3[k] = X[k] + dt * (A31 * Gv[k] + AT31 * S1[k] + AT32 * s2);
        F[k] = X[k] + dt * (B1 * (Gv[k] + S1[k]) + B2 * s2);
    }
    rhs(t0 + C2 * dt, Y.data(), Gv.data());
    for (size_t k = 0; k < n; ++k) Z[k] = P3[k] + dt * (A32 * Gv[k]);
    solve(dg, Z.data(), Y.data());
    for (size_t k = 0; k < n; ++k) {
        const T s3 = (Y[k] - Z[k]) / dg;
        F[k] = F[k] + dt * (B2 * Gv[k] + B3 * s3);
    }
    rhs(t0 + dt, Y.data(), Gv.data());
    for (size_t k = 0; k < n; ++k) X[k] = F[k] + dt * (B3 * Gv[k]);
}

// the IFunction's source (AddImplicitSourceToResidual, ModularSolver.hpp:82-124): per cell on
// the clamped state with refreshed aux (UpdateState of a local copy), zero where it is dry
// (h < 1e-6); source(Q, Qaux, p, 0, 0, 0)
void imex_source(int nc, const T* p, const T* U, const T* A, T* S) {
    const T X0[3] = {0, 0, 0};
    for (int c = 0; c < nc; ++c) {
        T q[NQ], a[NA];
        for (int i = 0; i < NQ; ++i) q[i] = U[(size_t)c * NQ + i];
        for (int i = 0; i < NA; ++i) a[i] = A[(size_t)c * NA + i];
        update_state(1, p, q, a);
        T* out = S + (size_t)c * NQ;
        if (q[1] < 1e-6) {
            for (int i = 0; i < NQ; ++i) out[i] = 0;
            continue;
        }
        auto sv = M::source(q, a, p, (T)0, (T)0, X0);
        for (int i = 0; i < NQ; ++i) out[i] = sv[i];
    }
}

int g_time_integration = 0;   // 0 splitting, 1 imex

void imex_step(const zdm_mesh* m, const int* face_bc, const T* p, const Opts& o, int source, T t0, T dt, T* X,
               const T* A, Work& w) {
    const int nc = m->ncells;
    const size_t n = (size_t)nc * NQ;
    ark2_step(
        n, t0, dt, X, [&](T t, const T* U, T* F) { form_rhs(m, face_bc, p, o, t, U, A, F, w); },
        [&](const T* U, T* S) {
            if (source)
                imex_source(nc, p, U, A, S);
            else
                std::fill(S, S + n, (T)0);
        },
        [&](T dg, const T* Z, T* Y) {
            // per cell: Y = Z, then the SourceStep Newton unless the clamped Z is dry
            for (size_t k = 0; k < n; ++k) Y[k] = Z[k];
            if (!source) return;
            for (int c = 0; c < nc; ++c) {
                T qc[NQ], ac[NA];
                for (int i = 0; i < NQ; ++i) qc[i] = Z[(size_t)c * NQ + i];
                for (int i = 0; i < NA; ++i) ac[i] = A[(size_t)c * NA + i];
                update_state(1, p, qc, ac);
                if (qc[1] >= 1e-6) solve_local_cell(dg, Y + (size_t)c * NQ, A + (size_t)c * NA, p);
            }
        });
}

#ifndef ORACLE_TANGENT
// Fully implicit step (FullyImplicitStrategy, SolverStrategies.hpp:88-195: TSBDF order 2 on the
// IFunction Xdot - source - FormRHS): the variable-step BDF2 equation
//   G(X) = (c0 X + c1 X_n + c2 X_{n-1} - S(X)) - R(X) = 0
// with c the derivative at t_{n+1} of the Lagrange interpolant through t_{n-1}, t_n, t_{n+1}
// (BDF1, c = (1/h, -1/h, 0), for the first step), R = FormRHS at t_{n+1}, S as in the IMEX
// step.  The oracle solves it to round-off by Newton with a dense central-difference
// Jacobian: it checks the discrete equation, not a Krylov path to it.  PETSc's own Newton
// stops at its SNES tolerances; that, and its TSBDF start-up, are parity unpinned.
void bdf_coeffs(double h, double h1, bool have_prev, double* c) {
    if (have_prev) {
        c[0] = 1.0 / h + 1.0 / (h + h1);
        c[1] = -(h + h1) / (h * h1);
        c[2] = h / (h1 * (h + h1));
    } else {
        c[0] = 1.0 / h;
        c[1] = -1.0 / h;
        c[2] = 0.0;
    }
}

std::vector<T> g_bdf_Xm;        // X_{n-1}
double g_bdf_dtp = 0;           // the previous dt
bool g_bdf_have = false;
double g_bdf_last_norm = 0;     // |G| at the end of the last Newton solve
int g_bdf_last_its = 0;

void bdf_residual(const zdm_mesh* m, const int* face_bc, const T* p, const Opts& o, T t1, const T* X,
                  const T* Xn, const T* Xm, const double* cf, const T* A, T* G, Work& w) {
    const int nc = m->ncells;
    const size_t n = (size_t)nc * NQ;
    std::vector<T> R(n), S(n);
    form_rhs(m, face_bc, p, o, t1, X, A, R.data(), w);
    imex_source(nc, p, X, A, S.data());
    for (size_t k = 0; k < n; ++k) {
        const T xd = cf[0] * X[k] + cf[1] * Xn[k] + cf[2] * Xm[k];
        G[k] = (xd - S[k]) - R[k];
    }
}

void bdf_step(const zdm_mesh* m, const int* face_bc, const T* p, const Opts& o, T t0, T dt, T* X, const T* A,
              Work& w) {
    const size_t n = (size_t)m->ncells * NQ;
    double cf[3];
    bdf_coeffs(dt, g_bdf_dtp, g_bdf_have, cf);
    const std::vector<T> Xn(X, X + n);
    const std::vector<T>& Xm = g_bdf_have ? g_bdf_Xm : Xn;
    const T t1 = t0 + dt;
    std::vector<T> G(n), Gp(n), Gm(n), J(n * n), rhs(n), dx(n), Xt(n);
    int it = 0;
    for (; it < 30; ++it) {
        bdf_residual(m, face_bc, p, o, t1, X, Xn.data(), Xm.data(), cf, A, G.data(), w);
        for (size_t j = 0; j < n; ++j) {
            const T hj = 1e-6 * (1.0 + std::fabs(X[j]));
            Xt.assign(X, X + n);
            Xt[j] = X[j] + hj;
            bdf_residual(m, face_bc, p, o, t1, Xt.data(), Xn.data(), Xm.data(), cf, A, Gp.data(), w);
            Xt[j] = X[j] - hj;
            bdf_residual(m, face_bc, p, o, t1, Xt.data(), Xn.data(), Xm.data(), cf, A, Gm.data(), w);
            for (size_t i = 0; i < n; ++i) J[i * n + j] = (Gp[i] - Gm[i]) / (2.0 * hj);
        }
        for (size_t i = 0; i < n; ++i) rhs[i] = -G[i];
        if (!gepp(J.data(), rhs.data(), dx.data(), (int)n)) break;
        T dmax = 0, xmax = 0;
        for (size_t i = 0; i < n; ++i) {
            X[i] += dx[i];
            dmax = std::max(dmax, (T)std::fabs(dx[i]));
            xmax = std::max(xmax, (T)std::fabs(X[i]));
        }
        if (dmax <= 1e-14 * (1.0 + xmax)) {
            ++it;
            break;
        }
    }
    bdf_residual(m, face_bc, p, o, t1, X, Xn.data(), Xm.data(), cf, A, G.data(), w);
    T gn = 0;
    for (size_t i = 0; i < n; ++i) gn += G[i] * G[i];
    g_bdf_last_norm = std::sqrt(gn);
    g_bdf_last_its = it;
    g_bdf_Xm = Xn;
    g_bdf_dtp = dt;
    g_bdf_have = true;
}

#endif  // ORACLE_TANGENT

void ts_step(const zdm_mesh* m, const int* face_bc, const T* p, const Opts& o, T t0, T dt, T* X,
             const T* A, Work& w) {
    const size_t n = (size_t)m->ncells * NQ;
    ssp_or_euler(o.order, o.ssp_stages, n, t0, dt, X,
                 [&](T t, const T* U, T* F) { form_rhs(m, face_bc, p, o, t, U, A, F, w); });
}

}  // namespace

namespace {

// Calibration loss with a frozen dt sequence (MUSCLSolver loop, MUSCLSolver.hpp:140-188):
// constraints once, then per step TSStep, constraints, source, constraints;
//   loss = sum over steps k*obs_every (k >= 1) and gauges g of (h_g - obs[k-1][g])^2.
// Run on double (orc_loss) and on one-direction dual numbers (orc_tl_loss, the
// tangent-linear reference for the discrete adjoint).
bool g_loss_source = true;   // the solver's `source` setting (false: no SourceStep)

T run_loss(const zdm_mesh* m, const int* face_bc, const T* P, int order, int ssp_stages,
           int use_limiters, int nsteps, const double* dts, T* X, T* A, int ngauge,
           const int* gauges, int obs_every, const double* obs) {
    const int nc = m->ncells;
    Opts o{order, ssp_stages, use_limiters, 0, 0, 0.5, 1e300, 0.0};
    Work w;
    enforce(nc, P, X, A);
    T L = 0;
    double t = 0;
    for (int n = 0; n < nsteps; ++n) {
        const T dt = T(dts[n]);
        ts_step(m, face_bc, P, o, T(t), dt, X, A, w);
        enforce(nc, P, X, A);
        if (g_loss_source) {
            source_solve(nc, P, dt, X, A);
            enforce(nc, P, X, A);
        }
        t += dts[n];
        if (obs_every > 0 && (n + 1) % obs_every == 0) {
            const int k = (n + 1) / obs_every - 1;
            for (int g = 0; g < ngauge; ++g) {
                const T r = X[(size_t)gauges[g] * NQ + 1] - T(obs[(size_t)k * ngauge + g]);
                L += r * r;
            }
        }
    }
    return L;
}

}  // namespace

#ifdef ZOOMY_MODEL_IS_VAM
#include "ChorinOps.H"
namespace {
#include "oracle_chorin.inc"
}  // namespace
#endif

#ifndef ORACLE_TANGENT
extern "C" {

int orc_model_info(int* nq, int* naux, int* npar, int* dim) {
    *nq = NQ;
    *naux = NA;
    *npar = M::n_parameters;
    *dim = DIM;
    return 0;
}

int orc_default_params(double* p) {
    auto v = M::default_parameters();
    for (size_t i = 0; i < v.size(); ++i) p[i] = v[i];
    return 0;
}

int orc_gradient(const zdm_mesh* m, const int* face_bc, const double* p, int method, const double* X,
                 double* G) {
    if (method == 1)
        gradient_lsq(m, X, G);
    else
        gradient_gg(m, face_bc, p, X, G);
    return 0;
}

int orc_bounds(const zdm_mesh* m, const double* X, double* vmin, double* vmax) {
    neighbor_bounds(m, X, vmin, vmax);
    return 0;
}

int orc_reconstruct(const double* q, const double* grad, const double* xc, const double* xf,
                    const double* qmin, const double* qmax, int use_limiters, double* out) {
    reconstruct(q, grad, xc, xf, qmin, qmax, use_limiters, out);
    return 0;
}

int orc_form_rhs(const zdm_mesh* m, const int* face_bc, const double* p, int order, int use_limiters,
                 int gradient, int well_balanced, double time, const double* X, const double* A,
                 double* F) {
    Opts o{order, 2, use_limiters, gradient, well_balanced, 0.5, 1.0, 1e-12};
    Work w;
    form_rhs(m, face_bc, p, o, time, X, A, F, w);
    return 0;
}

// eta-WB face states (positivity != 0 at order 2): clamp, Green-Gauss gradient of the raw X,
// LimitGradientsWB, WBPositivityReconstructor per face slot of every cell -> QF[c][k][i]
// (slots of cell_faces order, max_faces per cell); also the effective gradients G[c][i][d]
int orc_wb_faces(const zdm_mesh* m, const int* face_bc, const double* p, int skip_theta, double wet_eps,
                 const double* X, const double* A, double* QF, double* G) {
    const size_t nc = m->ncells;
    std::vector<T> Xl(X, X + nc * NQ), Al(A, A + nc * NA);
    update_state(m->ncells, p, Xl.data(), Al.data());
    gradient_gg(m, face_bc, p, X, G);
    limit_gradients_wb(m, Xl.data(), G, skip_theta != 0, wet_eps);
    for (int c = 0; c < m->ncells; ++c)
        for (int k = 0; k < m->cell_nfaces[c]; ++k) {
            const int f = m->cell_faces[(size_t)c * m->max_faces + k];
            reconstruct_wb(Xl.data() + (size_t)c * NQ, G + (size_t)c * NQ * DIM, m->cell_centroid + 3 * (size_t)c,
                           m->face_centroid + 3 * (size_t)f, wet_eps, QF + ((size_t)c * m->max_faces + k) * NQ);
        }
    return 0;
}

int orc_flux(const double* qL, const double* qR, const double* aL, const double* aR, const double* p,
             const double* n, double* flux, double* fluct) {
    auto f = NUM::numerical_flux(qL, qR, aL, aR, p, n);
    auto d = NUM::numerical_fluctuations(qL, qR, aL, aR, p, n);
    for (int i = 0; i < NQ; ++i) flux[i] = f[i];
    for (int i = 0; i < 2 * NQ; ++i) fluct[i] = d[i];
    return 0;
}

int orc_eigen(const double* q, const double* a, const double* p, const double* n, double* lam) {
    lam[0] = NUM::local_max_abs_eigenvalue(q, a, p, n)[0];
    return 0;
}

// The model header's per-point functions at one state (checked against vectors evaluated from
// the reference's generated Model.H): source, dS/dq (n x n), dS/da (n x n_aux), the clamp, the
// aux update, the eigenvalues along n, the ghost state of boundary condition bc and the
// initial condition at X.
int orc_model_point(const double* Q, const double* A, const double* p, const double* n, const double* X,
                    double time, double dt, int bc, double* src, double* dsdq, double* dsda, double* upd,
                    double* aux, double* eig, double* ghost, double* ic) {
    auto s = M::source(Q, A, p, time, dt, X);
    auto jq = M::source_jacobian_wrt_variables(Q, A, p);
    auto ja = M::source_jacobian_wrt_aux_variables(Q, A, p);
    auto u = model_update_variables<T>(Q, A, p, dt);
    auto a = M::update_aux_variables(Q, A, p, time, X);
    auto e = M::eigenvalues(Q, A, p, n);
    auto g = M::boundary_conditions(bc, Q, A, p, n, X, time, 0.0);
    for (int i = 0; i < NQ; ++i) {
        src[i] = s[i];
        upd[i] = u[i];
        eig[i] = e[i];
        ghost[i] = g[i];
    }
    for (int i = 0; i < NQ * NQ; ++i) dsdq[i] = jq[i];
    for (int i = 0; i < NQ * NA; ++i) dsda[i] = ja[i];
    for (int i = 0; i < NA; ++i) aux[i] = a[i];
    initial_condition_point<M>(X, p, ic);
    return 0;
}

#ifdef ZOOMY_MODEL_IS_VAM
// ChorinOps.H's three per-cell operators at one state: the 2x2 pressure blocks (VAM_N_PBLOCK of
// them, row-major), the pressure residual and the corrected slots (checked against vectors
// evaluated from the reference's generated ChorinOps.H)
int orc_vam_point(const double* Q, const double* A, const double* p, double dt, double* blocks, double* res,
                  double* corr) {
    vam_pressure_blocks(Q, A, p, dt, blocks);
    vam_pressure_residual(Q, A, p, dt, res);
    vam_corrector(Q, A, p, dt, corr);
    return 0;
}
#endif

int orc_update_state(int nc, const double* p, double* X, double* A) {
    update_state(nc, p, X, A);
    return 0;
}

// EnforcePhysicalConstraints on the mesh (RefreshDerivativeAux + clamp + aux)
int orc_enforce(const zdm_mesh* m, const double* p, double* X, double* A) {
    g_deriv_err = 0;
    enforce(m, p, X, A);
    return g_deriv_err ? -1 : 0;
}

int orc_set_refresh(int on) {
    g_refresh = on != 0;
    return 0;
}

double orc_compute_dt(const zdm_mesh* m, const double* p, const double* X, const double* A, double cfl,
                      int* finite) {
    return compute_time_step(m, p, X, A, cfl, finite);
}

int orc_source(int nc, const double* p, double dt, double* X, const double* A) {
    source_solve(nc, p, dt, X, A);
    return 0;
}

int orc_solve_cell(double dt, double* q, const double* aux, const double* p) {
    return solve_local_cell(dt, q, aux, p);
}

int orc_gepp(double* A, double* b, double* x, int n) { return gepp(A, b, x, n) ? 0 : -1; }

// The stage algebra on the scalar ODE u' = lam u (closed-form check of ssp_or_euler).
double orc_ssp_linear(int order, int s, double lam, double t0, double dt, double u0) {
    double u = u0;
    ssp_or_euler(order, s, 1, t0, dt, &u, [&](double, const double* U, double* F) { F[0] = lam * U[0]; });
    return u;
}

int orc_ts_step(const zdm_mesh* m, const int* face_bc, const double* p, int order, int ssp_stages,
                int use_limiters, int gradient, int well_balanced, double t0, double dt, double* X,
                const double* A) {
    Opts o{order, ssp_stages, use_limiters, gradient, well_balanced, 0.5, 1.0, 1e-12};
    Work w;
    ts_step(m, face_bc, p, o, t0, dt, X, A, w);
    return 0;
}

// solver.positivity / wet_dry_eps for the following calls (0 none, 1 mood, 2 zhang_shu); also
// resets the MOOD override count
// solver.time_integration of orc_run: 0 "splitting", 1 "imex", 2 "implicit"
int orc_set_time_integration(int ti) {
    g_time_integration = ti;
    return 0;
}

// the BDF coefficients of a step of size h after one of size h1 (have_prev = 0: BDF1)
int orc_bdf_coeffs(double h, double h1, int have_prev, double* c) {
    bdf_coeffs(h, h1, have_prev != 0, c);
    return 0;
}

// |G| after the last implicit step's Newton solve, and its iterations
double orc_bdf_last_residual(int* its) {
    if (its) *its = g_bdf_last_its;
    return g_bdf_last_norm;
}

// ARK2 on the scalar test equation y' = le y + li y (le explicit, li implicit): n steps of dt
// from y0, through the same stage code as the mesh step
double orc_ark2_linear(double le, double li, double dt, int n, double y0) {
    T y = y0;
    for (int k = 0; k < n; ++k)
        ark2_step(
            1, 0.0, dt, &y, [&](T, const T* U, T* F) { F[0] = le * U[0]; }, [&](const T* U, T* S) { S[0] = li * U[0]; },
            [&](T dg, const T* Z, T* Y) { Y[0] = Z[0] / (1.0 - dg * li); });
    return y;
}

int orc_set_positivity(int positivity, double wet_dry_eps) {
    g_positivity = positivity;
    g_wet_eps = wet_dry_eps;
    g_troubled = 0;
    g_last_mask.clear();
    return 0;
}
long long orc_troubled_count() { return g_troubled; }
// the cells the last MOOD override replaced (0 / 1 per cell; all 0 before any override)
int orc_last_troubled(int nc, int* mask) {
    for (int c = 0; c < nc; ++c) mask[c] = c < (int)g_last_mask.size() ? g_last_mask[c] : 0;
    return 0;
}

// MUSCLSolver::Run main loop, MUSCLSolver.hpp:54, 103, 140-188 (splitting):
//   constraints; dt = max(CFL, min_dt); while t < t_end and step < max_steps:
//     clamp dt to t_end; [save X;] TSStep; constraints; [MOOD;] source(dt); constraints;
//     t += dt; dt = CFL
// MOOD (positivity "mood", MUSCLSolver.hpp:130-171, 198-232): cells with h < -1e-10 after the
// constraints are overwritten by X_save + dt F, F = the first-order (PCM) right-hand side of
// the saved state over the faces touching them (FormRHSTroubledO1, TransportStep.hpp:363-381),
// then the constraints again.  zhang_shu at order >= 2 caps the CFL at 1/6 (MUSCLSolver.hpp:31-38).
// Returns the number of steps taken; *time and *dt_next are updated.
int orc_run(const zdm_mesh* m, const int* face_bc, const double* p, int order, int ssp_stages,
            int use_limiters, int gradient, int well_balanced, double cfl, double t_end,
            double min_dt, int max_steps, double* X, double* A, double* time, double* dt_io,
            int first) {
    if (g_positivity == 2 && order >= 2 && cfl > 1.0 / 6.0) cfl = 1.0 / 6.0;
    Opts o{order, ssp_stages, use_limiters, gradient, well_balanced, cfl, t_end, min_dt};
    Opts o1 = o;
    o1.order = 1;
    Work w;
    double t = *time;
    double dt = *dt_io;
    if (first) {
        enforce(m, p, X, A);
        dt = std::max(compute_time_step(m, p, X, A, cfl, nullptr), min_dt);
        g_bdf_have = false;
    }
    const size_t n = (size_t)m->ncells * NQ;
    std::vector<T> Xs, Fm;
    std::vector<uint8_t> mask;
    int steps = 0;
    while (t < t_end && steps < max_steps) {
        double t0 = t, dt0 = dt;
        if (t0 + dt0 > t_end) dt0 = t_end - t0;
        if (g_positivity == 1) Xs.assign(X, X + n);
        if (g_time_integration == 1)
            imex_step(m, face_bc, p, o, 1, t0, dt0, X, A, w);
        else if (g_time_integration == 2)
            bdf_step(m, face_bc, p, o, t0, dt0, X, A, w);
        else
            ts_step(m, face_bc, p, o, t0, dt0, X, A, w);
        enforce(m, p, X, A);
        if (g_positivity == 1) {
            mask.assign(m->ncells, 0);
            int nt = 0;
            for (int c = 0; c < m->ncells; ++c)
                if (X[(size_t)c * NQ + 1] < -1.0e-10) {
                    mask[c] = 1;
                    ++nt;
                }
            if (nt > 0) {
                Fm.assign(n, 0.0);
                g_filter = &mask;
                form_rhs(m, face_bc, p, o1, t0, Xs.data(), A, Fm.data(), w);
                g_filter = nullptr;
                for (int c = 0; c < m->ncells; ++c)
                    if (mask[c])
                        for (int i = 0; i < NQ; ++i)
                            X[(size_t)c * NQ + i] = Xs[(size_t)c * NQ + i] + dt0 * Fm[(size_t)c * NQ + i];
                enforce(m, p, X, A);
                g_troubled += nt;
                g_last_mask = mask;
            }
        }
        if (g_time_integration == 0) {
            source_solve(m->ncells, p, dt0, X, A);
            enforce(m, p, X, A);
        }
        ++steps;
        t = t0 + dt0;
        dt = compute_time_step(m, p, X, A, cfl, nullptr);
    }
    *time = t;
    *dt_io = dt;
    return steps;
}

}  // extern "C"

#endif  // ORACLE_TANGENT

// the loss runs with or without the SourceStep (both builds)
extern "C" void orc_set_loss_source(int on) { g_loss_source = on != 0; }

#ifdef ORACLE_TANGENT
// Tangent-linear run: the loss and its directional derivative along (dX0, dA0).
extern "C" int orc_tl_loss(const zdm_mesh* m, const int* face_bc, const double* p, int order,
                           int ssp_stages, int use_limiters, int nsteps, const double* dts,
                           const double* X0, const double* dX0, const double* A0,
                           const double* dA0, int ngauge, const int* gauges, int obs_every,
                           const double* obs, int freeze_limiter, double* loss, double* dloss) {
    const int nc = m->ncells;
    g_freeze_limiter = freeze_limiter != 0;
    std::vector<T> X((size_t)nc * NQ), A((size_t)nc * NA), P(M::n_parameters);
    for (size_t i = 0; i < X.size(); ++i) {
        X[i] = T(X0[i]);
        X[i].d[0] = dX0 ? dX0[i] : 0.0;
    }
    for (size_t i = 0; i < A.size(); ++i) {
        A[i] = T(A0[i]);
        A[i].d[0] = dA0 ? dA0[i] : 0.0;
    }
    for (int i = 0; i < M::n_parameters; ++i) P[i] = T(p[i]);
    const T L = run_loss(m, face_bc, P.data(), order, ssp_stages, use_limiters, nsteps, dts, X.data(),
                         A.data(), ngauge, gauges, obs_every, obs);
    *loss = L.v;
    *dloss = L.d[0];
    return 0;
}
#else
extern "C" double orc_loss(const zdm_mesh* m, const int* face_bc, const double* p, int order,
                           int ssp_stages, int use_limiters, int nsteps, const double* dts,
                           const double* X0, const double* A0, int ngauge, const int* gauges,
                           int obs_every, const double* obs) {
    std::vector<T> X(X0, X0 + (size_t)m->ncells * NQ), A(A0, A0 + (size_t)m->ncells * NA);
    return run_loss(m, face_bc, p, order, ssp_stages, use_limiters, nsteps, dts, X.data(), A.data(),
                    ngauge, gauges, obs_every, obs);
}
#endif
