// MI355X (gfx950) finite-volume time step for generated SME / VAM models.
//
// Built once per model directory (Model.H + Numerics.H on the include path), as the
// reference builds solver_cpu on its generated headers.  The headers are consumed unchanged:
// they are compiled for host and device through clang's force_cuda_host_device pragma.
//
// Device design (one GPU; FP64 throughout; results independent of scheduling: no floating-point
// atomics; the only atomics are integer counters, the MOOD troubled-cell count and the
// last-arrival counter of k_finalize_xcd, whose values do not depend on the order of arrival):
//   * structure-of-arrays state x[i*ld + c], cells renumbered by reverse Cuthill-McKee;
//   * k_prep  (per cell): wet/dry clamp + aux, face-neighbour bounds, Green-Gauss or
//     least-squares gradient, limited MUSCL reconstruction of every face of the cell into a
//     per-cell face-value buffer (so the gradient and bounds never go back to HBM);
//   * k_flux  (per cell): numerical flux + fluctuations of each face of the cell, gathered in
//     ascending face order (the reference's scatter order, TransportStep.hpp:403-466), divided
//     by the volume, with the SSP-RK stage update fused into the epilogue;
//   * k_post  (per cell): constraints -> implicit source Newton -> constraints -> local CFL
//     ratio, reduced per workgroup; k_finalize turns the partial minima into the next dt and
//     keeps time / dt on the device, so a time step needs no host synchronisation;
//   * k_flux_post: the final stage's k_flux and k_post in one pass, the new state in registers
//     between them (the default form of a step).
// Workgroups are remapped so that each XCD (own L2) works on a contiguous range of cells.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "zdm.h"

#pragma clang force_cuda_host_device begin
#include "zoomy/UserFunctions.H"   // ZoomyMesh, compute_derivative, eigen kernels: before the model
#include "dual.H"   // before the model: its templates name std::sqrt / std::fabs on duals
#include "Numerics.H"
#ifdef ZOOMY_MODEL_IS_VAM
#include "ChorinOps.H"
#endif
#pragma clang force_cuda_host_device end

#ifndef ZDM_MODEL_NAME
#define ZDM_MODEL_NAME "model"
#endif

// Model::initial_condition(X, p) is optional (VirtualSolver.hpp:33-38)
template <typename TT, typename = void>
struct HasInitialCondition : std::false_type {};
template <typename TT>
struct HasInitialCondition<TT, std::void_t<decltype(Model<TT>::initial_condition(
                                   std::declval<const TT*>(), std::declval<const TT*>()))>>
    : std::true_type {};

// Model::update_variables (the wet/dry state clamp) is optional: a model without it keeps its
// state (HasUpdateVariables, TransportStep.hpp:23-28, 219)
template <typename TT, typename = void>
struct HasUpdateVariables : std::false_type {};
template <typename TT>
struct HasUpdateVariables<TT, std::void_t<decltype(Model<TT>::update_variables(
                                  std::declval<const TT*>(), std::declval<const TT*>(),
                                  std::declval<const TT*>(), std::declval<const TT>()))>> : std::true_type {};

template <typename S>
__host__ __device__ __forceinline__ SimpleArray<S, Model<S>::n_dof_q> model_update_variables(const S* q, const S* a,
                                                                                              const S* p, const S dt) {
    if constexpr (HasUpdateVariables<S>::value) {
        return Model<S>::update_variables(q, a, p, dt);
    } else {
        SimpleArray<S, Model<S>::n_dof_q> r;
        for (int i = 0; i < Model<S>::n_dof_q; ++i) r[i] = q[i];
        return r;
    }
}

// Model::diffusive_flux(q_face, a_face, gradQ_face, p), optional (HasDiffusiveFlux,
// TransportStep.hpp:10-18): the face term added after the flux
template <typename TT, typename = void>
struct HasDiffusiveFlux : std::false_type {};
template <typename TT>
struct HasDiffusiveFlux<TT, std::void_t<decltype(Model<TT>::diffusive_flux(
                                std::declval<const TT*>(), std::declval<const TT*>(), std::declval<const TT*>(),
                                std::declval<const TT*>()))>> : std::true_type {};

// The field-level, mesh-aware update_aux_variables(Q, Qaux, dt, mesh) of a generated model
// (Model.H:250-262), which fills the spatial-derivative aux through compute_derivative
// (HasFieldAuxUpdate, TransportStep.hpp:30-38)
template <typename TT, typename = void>
struct HasFieldAuxUpdate : std::false_type {};
template <typename TT>
struct HasFieldAuxUpdate<TT, std::void_t<decltype(Model<TT>::update_aux_variables(
                                 std::declval<TT* const*>(), std::declval<TT* const*>(), std::declval<const TT>(),
                                 std::declval<const ZoomyMesh&>()))>> : std::true_type {};

// Optional numerics hook (not emitted by the reference generator): numerical_flux and
// numerical_fluctuations from one evaluation, with results identical to the two calls
template <typename NN, typename S, typename = void>
struct HasFusedFlux : std::false_type {};
template <typename NN, typename S>
struct HasFusedFlux<NN, S, std::void_t<decltype(NN::numerical_flux_fluctuations(
                               std::declval<const S*>(), std::declval<const S*>(), std::declval<const S*>(),
                               std::declval<const S*>(), std::declval<const S*>(), std::declval<const S*>(),
                               std::declval<S*>(), std::declval<S*>()))>> : std::true_type {};

// F = numerical_flux, D = numerical_fluctuations of one face (one evaluation when the numerics
// provide the fused hook)
template <typename NN, typename S>
__host__ __device__ __forceinline__ void face_flux(const S* qL, const S* qR, const S* aL, const S* aR,
                                                   const S* p, const S* n, S* F, S* D) {
    constexpr int nq = Model<S>::n_dof_q;
    if constexpr (HasFusedFlux<NN, S>::value) {
        NN::numerical_flux_fluctuations(qL, qR, aL, aR, p, n, F, D);
    } else {
        auto f = NN::numerical_flux(qL, qR, aL, aR, p, n);
        auto d = NN::numerical_fluctuations(qL, qR, aL, aR, p, n);
        for (int i = 0; i < nq; ++i) F[i] = f[i];
        for (int i = 0; i < 2 * nq; ++i) D[i] = d[i];
    }
}

// Optional numerics hook (not emitted by the reference generator): the reverse-mode derivative
// of numerical_flux / numerical_fluctuations, gL += J_L^T (bF, bD), gR += J_R^T (bF, bD), for
// numerics that read no aux; the adjoint uses it in place of forward-mode duals
template <typename NN, typename S, typename = void>
struct HasFluxVjp : std::false_type {};
template <typename NN, typename S>
struct HasFluxVjp<NN, S, std::void_t<decltype(NN::numerical_flux_fluctuations_vjp(
                             std::declval<const S*>(), std::declval<const S*>(), std::declval<const S*>(),
                             std::declval<const S*>(), std::declval<const S*>(), std::declval<const S*>(),
                             std::declval<S*>(), std::declval<S*>()))>> : std::true_type {};

// Optional model hook (not emitted by the reference generator): centred bathymetry term of
// the second-order hydrostatic reconstruction, added per face side when well_balanced is on.
template <typename TT, typename = void>
struct HasCellSource : std::false_type {};
template <typename TT>
struct HasCellSource<TT, std::void_t<decltype(Model<TT>::hydrostatic_cell_source(
                             std::declval<const TT*>(), std::declval<const TT*>(),
                             std::declval<const TT*>(), std::declval<const TT*>()))>>
    : std::true_type {};

namespace {

using T = double;
using M = Model<T>;
using NUM = Numerics<T>;
constexpr int NQ = M::n_dof_q;
constexpr int NA = M::n_dof_qaux;
// Face states QF: per face slot k, NQP / 2 planes of 16-byte pairs (components 2j, 2j+1 of cell
// c side by side), NQP = NQ rounded up to even: k_prep's stores and k_flux's own-state loads
// are 16-byte vector accesses contiguous across the wavefront, and a neighbour's face state is
// NQP / 2 vector loads.
constexpr int NQP = (NQ + 1) & ~1;
__host__ __device__ __forceinline__ size_t qf_idx(int ld, int k, int c, int i) {
    return (((size_t)k * (NQP / 2) + (i >> 1)) * ld + c) * 2 + (i & 1);
}
constexpr int DIM = M::dimension;
constexpr int NPAR = M::n_parameters;
constexpr int MAXPAR = 16;
constexpr int BLOCK = 256;
constexpr int NXCD = 8;
static_assert(NPAR <= MAXPAR, "too many model parameters");

thread_local std::string g_err;
int fail(const std::string& m) {
    g_err = m;
    return -1;
}

// Model::initial_condition(X, p) at one point (optional in a generated model,
// VirtualSolver.hpp:33-38), or an error for a model without it
template <typename S>
int model_initial_condition(const S* X, const S* p, S* Q) {
    if constexpr (HasInitialCondition<S>::value) {
        auto q = Model<S>::initial_condition(X, p);
        for (int i = 0; i < NQ; ++i) Q[i] = q[i];
        return 0;
    } else {
        return fail("No io.initial_condition_file set and the model emits no initial_condition(x,p)");
    }
}
#define HIPCHK(x)                                                                     \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) return fail(std::string(#x " : ") + hipGetErrorString(e_)); \
    } while (0)

struct Params {
    T p[MAXPAR];
};

// Device control block: time stepping state lives on the device between steps.
struct Ctrl {
    T t;          // simulation time at the start of the current step
    T dt;         // dt of the current step (clamped to t_end)
    T dt_cfl;     // unclamped CFL dt (ComputeTimeStep) for the current step
    T dt_query;   // result slot of a standalone ComputeTimeStep
    T a_stage;    // dt / (s-1)          (Euler: dt)
    T al, be, ga; // (s-1)/s, dt/s, 1/s
    T stage_t[16];
    int active;   // t < t_end
    int steps;
    int finite;
    int troubled; // MOOD overrides so far
};

struct DevMesh {
    int ncells, nfaces, ld;
    const T* vol;
    const T* cx;
    const T* cy;
    const T* rc;      // min centroid -> face-centroid distance
    const int* cface; // [MAXF][ld]
    const int* cnbr;  // [MAXF][ld]
    const int* cmeta; // [MAXF][ld]: bits 0-1 = slot of the face in the neighbour, bit 2 = this
                      // cell is support[1]
    const T* fNx;     // area-weighted normal
    const T* fNy;
    const T* fcx;     // face centroid
    const T* fcy;
    const T* farea;   // sqrt(N.N)
    const T* fnhx;    // N / area
    const T* fnhy;
    const int* fbc;   // model BC index of boundary faces, -1 = none
    // k_flux's packed copies: ctopo[2 (k ld + c)] = cface, [.. + 1] = cnbr | cmeta << 28
    // (-1 without a neighbour); fgeo[3 f] = (fnhx, fnhy, farea)
    const int* ctopo;
    const T* fgeo;
    // per face slot of each cell, two 16-byte records (sgeo_idx): the area-weighted normal
    // signed outward from the cell (-N where the cell is support[1]) and the face centroid
    // minus the cell centroid; zero for empty slots
    const double2* sgeo;
    const T* lsqw;    // [MAXF][DIM][ld] least-squares weights
    // ensemble of nmem member meshes in one launch (nmem = 0: a single mesh): each member owns
    // whole workgroups, blocks [blk_first[m], blk_first[m+1]) cover its cells
    // [cell_off[m], cell_off[m+1])
    const int* blk_member;
    const int* blk_first;
    const int* cell_off;
    int nmem;
    int cpt;          // cells per thread of the transport kernels: thread t of a workgroup takes
                      // cells first + t + j * blockDim, j < cpt
    int rev;          // 1: each XCD walks its range of workgroups from the last to the first
};

__host__ __device__ __forceinline__ size_t sgeo_idx(int ld, int k, int part, int c) {
    return ((size_t)k * 2 + part) * ld + c;
}

__device__ __forceinline__ int logical_block(int rev = 0) {
    // XCD-aware remap: hardware dispatches workgroup b to XCD b % 8; give each XCD a
    // contiguous range of workgroups so face neighbours share that XCD's L2.  rev: each XCD
    // takes its range from the last workgroup to the first.  The flux launches run against
    // k_prep's direction, so each XCD starts on the face states k_prep wrote last, which are
    // still in the 256 MB Infinity Cache (C2: k_flux 0.205 -> 0.193 ms, k_flux_post 0.234 ->
    // 0.222 ms); k_prep in turn starts where the flux launch ended.  Cells are independent and
    // the CFL reduction takes a minimum, so results do not depend on the direction.
    const int nb = gridDim.x;
    const int xcd = blockIdx.x % NXCD;
    const int q = nb / NXCD, r = nb % NXCD;
    const int cnt = q + (xcd < r ? 1 : 0);
    const int idx = rev ? cnt - 1 - (int)(blockIdx.x / NXCD) : (int)(blockIdx.x / NXCD);
    return xcd * q + (xcd < r ? xcd : r) + idx;
}
__device__ __forceinline__ int cell_index() { return logical_block() * blockDim.x + threadIdx.x; }

// first cell, logical block and ensemble member of this thread (c = ncells past the member's
// end).  MC (cells per thread > 1): cell(j) = its j-th cell, j < cpt, or ncells past the end;
// without MC only cell(0) = c exists and the index math is that of one cell per thread.
template <bool MC>
struct Loc {
    int c, lb, m, base, end, nc;
    __device__ __forceinline__ int cell(int j) const {
        if constexpr (!MC) return c;
        const int cj = base + j * (int)blockDim.x;
        return cj < end ? cj : nc;
    }
};
template <bool MC = false>
__device__ __forceinline__ Loc<MC> locate(const DevMesh& m) {
    Loc<MC> L;
    L.lb = logical_block(m.rev);
    if constexpr (!MC) {
        if (m.nmem == 0) {
            L.m = 0;
            L.c = L.lb * blockDim.x + threadIdx.x;
            return L;
        }
        L.m = m.blk_member[L.lb];
        const int c = m.cell_off[L.m] + (L.lb - m.blk_first[L.m]) * blockDim.x + threadIdx.x;
        L.c = c < m.cell_off[L.m + 1] ? c : m.ncells;
        return L;
    } else {
        L.nc = m.ncells;
        const int span = (int)blockDim.x * m.cpt;
        if (m.nmem == 0) {
            L.m = 0;
            L.base = L.lb * span + threadIdx.x;
            L.end = m.ncells;
        } else {
            L.m = m.blk_member[L.lb];
            L.base = m.cell_off[L.m] + (L.lb - m.blk_first[L.m]) * span + threadIdx.x;
            L.end = m.cell_off[L.m + 1];
        }
        L.c = L.cell(0);
        return L;
    }
}

template <int N>
__device__ __forceinline__ void load_soa(const T* __restrict__ a, int ld, int c, T* v) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = a[(size_t)i * ld + c];
}
template <int N>
__device__ __forceinline__ void store_soa(T* __restrict__ a, int ld, int c, const T* v) {
#pragma unroll
    for (int i = 0; i < N; ++i) a[(size_t)i * ld + c] = v[i];
}
// The transport kernels' read-once and write-once streams, non-temporal when NT: in the flux
// kernels the stage state, Q^n, the output state and aux, the cell's own slot topology and
// volume; in the Green-Gauss k_prep the cell's own slot topology, slot geometry and volume.
// On a mesh whose flux pass is far larger than the last-level cache they then do not displace
// the data that neighbours read again in the same launch (C2: 551 -> 512 B/cell moved by the
// stage flux launch), nor the face states that the next flux launch reads from the Infinity
// Cache (C2: k_prep 0.166 -> 0.162 ms).  On a smaller mesh the same hints keep data out of the
// 256 MB Infinity Cache that the next kernel or step would have found there, and cost time, so
// the solver picks them by size (zdm_flux_nontemporal).
typedef double dvec2 __attribute__((ext_vector_type(2)));
template <bool NT, typename S>
__device__ __forceinline__ S ldx(const S* p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}
template <bool NT, typename S>
__device__ __forceinline__ void stx(S v, S* p) {
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}
template <bool NT, int N>
__device__ __forceinline__ void load_soa_x(const T* __restrict__ a, int ld, int c, T* v) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = ldx<NT>(a + (size_t)i * ld + c);
}
template <bool NT, int N>
__device__ __forceinline__ void store_soa_x(T* __restrict__ a, int ld, int c, const T* v) {
#pragma unroll
    for (int i = 0; i < N; ++i) stx<NT>(v[i], a + (size_t)i * ld + c);
}

// TransportStep::UpdateState per cell (TransportStep.hpp:210-227)
__device__ __forceinline__ void update_state_cell(const Params& P, T* q, T* a) {
    const T X0[3] = {0, 0, 0};
    auto rq = model_update_variables<T>(q, a, P.p, 0.0);
#pragma unroll
    for (int i = 0; i < NQ; ++i) q[i] = rq[i];
    auto ra = M::update_aux_variables(q, a, P.p, 0.0, X0);
#pragma unroll
    for (int i = 0; i < NA; ++i) a[i] = ra[i];
}

// ------------------------------------------------------------------------------ k_prep
// GRAD: 0 = none (PCM), 1 = Green-Gauss, 2 = least squares
// 3 waves / SIMD: the hoisted neighbour loads need ~170 VGPRs (a 20-byte spill at the cap of
// 168); measured faster than the 4-wave form with the loads issued inside the face loop, and
// than 2 waves without the spill (0.188 against 0.209 ms at 2 M cells).
// Clamp + aux of cell c (UpdateState), its gradient of the raw state (GRAD 1: Green-Gauss,
// 2: least squares, 3: Green-Gauss with the eta-WB limiting of positivity "mood" / "zhang_shu"),
// the neighbour bounds of the clamped states and the limited linear reconstruction at every face
// of the cell (Reconstruction.hpp:203-243; GRAD 3: WBPositivityReconstructor, 256-287).
// put(k, i, v) receives component i of face slot k's state; the clamped aux is returned in ac.
// wet_eps / zs_theta: GRAD 3 only (wet_dry_eps; 1 = apply the Zhang-Shu theta).
// AMODE (limited linear reconstruction, GRAD 1 / 2): 0 computes the face limiter values alpha,
// 1 computes and stores them in abuf[k ld + c], 2 takes them from abuf (a frozen limiter)
template <int MAXF, int GRAD, typename Put, int AMODE = 0, bool NT = false>
__device__ __forceinline__ void recon_cell(const DevMesh& m, const Params& P, const T* __restrict__ X,
                                           const T* __restrict__ Ag, int c, int use_limiters, T* ac, Put put,
                                           T wet_eps = 0, int zs_theta = 0, T* abuf = nullptr, T* gout = nullptr) {
    constexpr bool GG = GRAD == 1 || GRAD == 3;
    const int ld = m.ld;
    T q[NQ], qc[NQ];
    load_soa<NQ>(X, ld, c, q);
    load_soa<NA>(Ag, ld, c, ac);
#pragma unroll
    for (int i = 0; i < NQ; ++i) qc[i] = q[i];
    update_state_cell(P, qc, ac);

    if constexpr (GRAD == 0) {
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            if (m.cface[(size_t)k * ld + c] < 0) continue;
#pragma unroll
            for (int i = 0; i < NQ; ++i) put(k, i, qc[i]);
        }
        return;
    } else {
        // all independent loads of the cell first (indices, then neighbour states and face
        // geometry through in-bounds clamped indices), so they are in flight together
        // face geometry from the cell's own slot records (coalesced): Nk = the area-weighted
        // normal signed outward from this cell, dxk = face centroid - cell centroid
        int fk[MAXF], nbk[MAXF];
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            fk[k] = ldx<NT>(m.cface + (size_t)k * ld + c);
            nbk[k] = ldx<NT>(m.cnbr + (size_t)k * ld + c);
        }
        T qn[MAXF][NQ], an[MAXF][NA], Nk[MAXF][2], dxk[MAXF][2];
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            const int nbs = nbk[k] >= 0 ? nbk[k] : c;
            load_soa<NQ>(X, ld, nbs, qn[k]);
            load_soa<NA>(Ag, ld, nbs, an[k]);
            const dvec2 ns = ldx<NT>(reinterpret_cast<const dvec2*>(m.sgeo + sgeo_idx(ld, k, 0, c)));
            const dvec2 dx = ldx<NT>(reinterpret_cast<const dvec2*>(m.sgeo + sgeo_idx(ld, k, 1, c)));
            Nk[k][0] = ns.x;
            Nk[k][1] = ns.y;
            dxk[k][0] = dx.x;
            dxk[k][1] = dx.y;
        }
        T lo[NQ], hi[NQ], g[NQ * DIM];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            // GRAD 3: extrema of W = (b, eta = b + h, momenta), LimitGradientsWB
            const T w = (GRAD == 3 && i == 1) ? qc[0] + qc[1] : qc[i];
            lo[i] = w;
            hi[i] = w;
        }
#pragma unroll
        for (int i = 0; i < NQ * DIM; ++i) g[i] = 0;

#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            const int f = fk[k];
            if (f < 0) continue;
            const int nb = nbk[k];
            const T* N = Nk[k];
            if (nb >= 0) {
                auto rn = model_update_variables<T>(qn[k], an[k], P.p, 0.0);
#pragma unroll
                for (int i = 0; i < NQ; ++i) {
                    const T w = (GRAD == 3 && i == 1) ? rn[0] + rn[1] : rn[i];
                    if (w < lo[i]) lo[i] = w;
                    if (w > hi[i]) hi[i] = w;
                }
                if constexpr (GG) {
                    // g -= fv N on the face's support[1] side is g += fv (-N) bit for bit
#pragma unroll
                    for (int i = 0; i < NQ; ++i) {
                        const T fv = (T)0.5 * (q[i] + qn[k][i]);
#pragma unroll
                        for (int d = 0; d < DIM; ++d) g[i * DIM + d] += fv * N[d];
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < NQ; ++i) {
                        const T du = qn[k][i] - q[i];
#pragma unroll
                        for (int d = 0; d < DIM; ++d)
                            g[i * DIM + d] += m.lsqw[((size_t)k * DIM + d) * ld + c] * du;
                    }
                }
            } else if constexpr (GG) {
                const int bc = m.fbc[f];
                if (bc >= 0) {
                    const T nh[3] = {m.fnhx[f], DIM > 1 ? m.fnhy[f] : (T)0, 0};
                    const T xf[3] = {m.fcx[f], DIM > 1 ? m.fcy[f] : (T)0, 0};
                    auto qb = M::boundary_conditions(bc, q, nullptr, P.p, nh, xf, 0.0, 0.0);
#pragma unroll
                    for (int i = 0; i < NQ; ++i)
#pragma unroll
                        for (int d = 0; d < DIM; ++d) g[i * DIM + d] += qb[i] * N[d];
                }
            }
        }
        if constexpr (GG) {
            const T vol = ldx<NT>(m.vol + c);
            if (vol > 1e-15) {
#pragma unroll
                for (int i = 0; i < NQ * DIM; ++i) g[i] /= vol;
            }
        }
        if constexpr (HasDiffusiveFlux<T>::value && GRAD != 3) {
            // the cell gradient the diffusive face term reads (TransportStep.hpp:470-475)
            if (gout) {
#pragma unroll
                for (int i = 0; i < NQ * DIM; ++i) gout[(size_t)i * ld + c] = g[i];
            }
        }
        if constexpr (GRAD == 3) {
            // TransportStep::LimitGradientsWB (TransportStep.hpp:241-323) on the raw gradient g:
            // Venkatakrishnan phi per W variable over the cell's faces (eps^2 = vol^(2/dim)),
            // first order below wet_eps, b and momentum slopes tied to eta, Zhang-Shu theta on
            // sh = phi_eta grad eta - phi_b grad b; then WBPositivityReconstructor
            // (Reconstruction.hpp:256-287) at every face
            constexpr int B = 0, H = 1;
            const T vol = m.vol[c];
            const T eps2 = DIM == 2 ? vol : vol * vol;
            T geta[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) geta[d] = g[H * DIM + d] + g[B * DIM + d];
            const T h_bar = qc[H], b_bar = qc[B];
            T phi[NQ];
#pragma unroll
            for (int i = 0; i < NQ; ++i) phi[i] = 1.0;
#pragma unroll
            for (int k = 0; k < MAXF; ++k) {
                if (fk[k] < 0) continue;
                const T* r = dxk[k];
#pragma unroll
                for (int i = 0; i < NQ; ++i) {
                    T delta = 0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) delta += ((i == H) ? geta[d] : g[i * DIM + d]) * r[d];
                    if (std::fabs(delta) < 1e-14) continue;
                    const T wc = (i == H) ? b_bar + h_bar : qc[i];
                    const T dm = (delta > 0) ? (hi[i] - wc) : (lo[i] - wc);
                    const T num = dm * dm + eps2 + 2 * delta * dm, den = dm * dm + 2 * delta * delta + delta * dm + eps2;
                    T pf = (den > 1e-30) ? num / den : (T)1.0;
                    pf = std::max((T)0, std::min((T)1, pf));
                    if (pf < phi[i]) phi[i] = pf;
                }
            }
            if (h_bar < wet_eps) {
#pragma unroll
                for (int i = 0; i < NQ; ++i) phi[i] = 0.0;
            }
            phi[B] = std::min(phi[B], phi[H]);
#pragma unroll
            for (int i = 0; i < NQ; ++i)
                if (i != B && i != H) phi[i] = std::min(phi[i], phi[H]);
            T sh[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) sh[d] = phi[H] * geta[d] - phi[B] * g[B * DIM + d];
            T hmin = h_bar;
#pragma unroll
            for (int k = 0; k < MAXF; ++k) {
                if (fk[k] < 0) continue;
                const T* r = dxk[k];
                T hf = h_bar;
#pragma unroll
                for (int d = 0; d < DIM; ++d) hf += sh[d] * r[d];
                if (hf < hmin) hmin = hf;
            }
            T theta = 1.0;
            if (zs_theta && hmin < 0.0) theta = std::max((T)0, std::min((T)1, h_bar / std::max(h_bar - hmin, (T)1e-14)));
            // effective W-gradients
            T ge[NQ * DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                ge[B * DIM + d] = phi[B] * g[B * DIM + d];
                ge[H * DIM + d] = phi[B] * g[B * DIM + d] + theta * sh[d];
            }
#pragma unroll
            for (int i = 0; i < NQ; ++i)
                if (i != B && i != H)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) ge[i * DIM + d] = theta * phi[i] * g[i * DIM + d];
#pragma unroll
            for (int k = 0; k < MAXF; ++k) {
                if (fk[k] < 0) continue;
                const T* dx = dxk[k];
                T dot[NQ];
#pragma unroll
                for (int i = 0; i < NQ; ++i) {
                    T sd = 0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) sd += ge[i * DIM + d] * dx[d];
                    dot[i] = sd;
                }
                const T b_f = b_bar + dot[B];
                const T eta_f = (b_bar + h_bar) + dot[H];
                const T h_f = std::max(eta_f - b_f, (T)0);
                put(k, B, b_f);
                put(k, H, h_f);
#pragma unroll
                for (int i = 2; i < NQ; ++i) put(k, i, (h_f < wet_eps) ? (T)0 : (qc[i] + dot[i]));
            }
            return;
        }
        // limited linear reconstruction at every face of the cell (Reconstruction.hpp:203-243)
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            if (fk[k] < 0) continue;
            const T* dx = dxk[k];
            T delta[NQ], fin[NQ];
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                delta[i] = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) delta[i] += g[i * DIM + d] * dx[d];
                fin[i] = 1.0;
            }
            if (AMODE != 2 && use_limiters) {
                // vacuum limiter: first order next to (nearly) dry cells
                if (NQ > 1 && lo[1] < 1e-2) {
#pragma unroll
                    for (int i = 0; i < NQ; ++i) fin[i] = 0.0;
                }
                // TVD / Barth-Jespersen, component-wise
#pragma unroll
                for (int i = 0; i < NQ; ++i) {
                    const T unl = qc[i] + delta[i];
                    const T dq = unl - qc[i];
                    if (std::fabs(dq) < 1e-12) continue;
                    // min(1, num / dq) without the division where the quotient is >= 1
                    // (rounding is monotone, so the result is the same bit for bit)
                    T a = 1.0;
                    if (dq > 0) {
                        const T num = hi[i] - qc[i];
                        if (hi[i] > qc[i])
                            a = num < dq ? num / dq : (T)1.0;
                        else
                            a = 0.0;
                    } else {
                        const T num = lo[i] - qc[i];
                        if (lo[i] < qc[i])
                            a = num > dq ? num / dq : (T)1.0;
                        else
                            a = 0.0;
                    }
                    if (a < fin[i]) fin[i] = a;
                }
            }
            T alpha = 1.;
#pragma unroll
            for (int i = 0; i < NQ; ++i)
                if (alpha > fin[i]) alpha = fin[i];
            if constexpr (AMODE == 1) abuf[(size_t)k * ld + c] = alpha;
            if constexpr (AMODE == 2) alpha = abuf[(size_t)k * ld + c];
#pragma unroll
            for (int i = 0; i < NQ; ++i) put(k, i, qc[i] + alpha * delta[i]);
        }
    }
}

template <int MAXF, int GRAD, bool MC, bool NT = false>
__global__ __launch_bounds__(BLOCK, 3) void k_prep(DevMesh m, Params P, const Ctrl* __restrict__ ctrl,
                                                const T* __restrict__ X, const T* __restrict__ Ag,
                                                T* __restrict__ As, T* __restrict__ QF, int use_limiters,
                                                T wet_eps, int zs_theta, T* __restrict__ G) {
    const auto L = locate<MC>(m);
    ctrl += L.m;
    if (!ctrl->active) return;
    const int ld = m.ld;
    auto body = [&](int c) {
        T ac[NA];
        auto put = [&](int k, int i, T v) { QF[qf_idx(ld, k, c, i)] = v; };
        recon_cell<MAXF, GRAD, decltype(put), 0, NT>(m, P, X, Ag, c, use_limiters, ac, put, wet_eps, zs_theta, nullptr,
                                                      G);
        store_soa<NA>(As, ld, c, ac);
    };
    if constexpr (MC) {
        for (int j = 0; j < m.cpt; ++j) {
            const int c = L.cell(j);
            if (c >= m.ncells) return;
            body(c);
        }
    } else {
        if (L.c >= m.ncells) return;
        body(L.c);
    }
}

// k_prep (Green-Gauss, the launch shape of k_prep) that also stores each face's limiter value
// alpha in abuf[k ld + c]: the recording forward of the adjoint
template <int MAXF, bool MC>
__global__ __launch_bounds__(BLOCK, 3) void k_prep_rec(DevMesh m, Params P, const Ctrl* __restrict__ ctrl,
                                                    const T* __restrict__ X, const T* __restrict__ Ag,
                                                    T* __restrict__ As, T* __restrict__ QF, int use_limiters,
                                                    T* __restrict__ abuf) {
    const auto L = locate<MC>(m);
    ctrl += L.m;
    if (!ctrl->active) return;
    const int ld = m.ld;
    auto body = [&](int c) {
        T ac[NA];
        auto put = [&](int k, int i, T v) { QF[qf_idx(ld, k, c, i)] = v; };
        recon_cell<MAXF, 1, decltype(put), 1>(m, P, X, Ag, c, use_limiters, ac, put, (T)0, 0, abuf);
        store_soa<NA>(As, ld, c, ac);
    };
    if constexpr (MC) {
        for (int j = 0; j < m.cpt; ++j) {
            const int c = L.cell(j);
            if (c >= m.ncells) return;
            body(c);
        }
    } else {
        if (L.c >= m.ncells) return;
        body(L.c);
    }
}

// k_prep of one mesh (single-cell launch shape) with the limiter values stored (AMODE 1) or
// taken from abuf (AMODE 2), for the matrix-free Jacobian of the implicit step
template <int MAXF, int GRAD, int AMODE>
__global__ __launch_bounds__(BLOCK) void k_prep_alpha(DevMesh m, Params P, const T* __restrict__ X,
                                                      const T* __restrict__ Ag, T* __restrict__ As, T* __restrict__ QF,
                                                      int use_limiters, T* __restrict__ abuf) {
    const int c = cell_index();
    if (c >= m.ncells) return;
    const int ld = m.ld;
    T ac[NA];
    auto put = [&](int k, int i, T v) { QF[qf_idx(ld, k, c, i)] = v; };
    recon_cell<MAXF, GRAD, decltype(put), AMODE>(m, P, X, Ag, c, use_limiters, ac, put, (T)0, 0, abuf);
    store_soa<NA>(As, ld, c, ac);
}

// ------------------------------------------------------------------------------ k_flux
// ARK2 tableau of the IMEX step (imex.inc): Giraldo, Kelly and Constantinescu, SIAM J. Sci.
// Comput. 35(5), 2013
namespace ark2 {
constexpr double R2 = 1.41421356237309504880;   // sqrt(2)
constexpr double G = 1.0 - 1.0 / R2;            // gamma, the SDIRK diagonal
constexpr double D = 1.0 / (2.0 * R2);
constexpr double A21 = 2.0 * G;
constexpr double A32 = (3.0 + 2.0 * R2) / 6.0;
constexpr double A31 = 1.0 - A32;
constexpr double AT21 = G, AT31 = D, AT32 = D;
constexpr double B1 = D, B2 = D, B3 = G;
constexpr double C2 = 2.0 * G;                   // c = (0, 2 g, 1)
}  // namespace ark2

// MODE_IMEX3: the last term of the IMEX step, X+ = F + dt b3 G3 with F = X0 (imex.inc)
enum { MODE_RHS = 0, MODE_EULER = 1, MODE_STAGE = 2, MODE_FINAL = 3, MODE_IMEX3 = 4 };

// One face slot of the face loop of ComputeFluxes (TransportStep.hpp:383-565) for a slot with
// fc += area * Model::hydrostatic_cell_source(q_face, q_cell, p, n_out) (the well_balanced
// extension); nothing for a model without the hook
template <typename S>
__device__ __forceinline__ void add_cell_source(const S* qf, const S* qcell, const S* p, const S* no, S area, S* fc) {
    if constexpr (HasCellSource<S>::value) {
        auto Sv = Model<S>::hydrostatic_cell_source(qf, qcell, p, no);
#pragma unroll
        for (int i = 0; i < NQ; ++i) fc[i] += Sv[i] * area;
    }
}

// The diffusive face term of a model with Model::diffusive_flux (TransportStep.hpp:468-491):
// state, aux and gradient averaged over the two sides, the term times the area added to the
// face's support[0] cell and subtracted from its support[1] cell (right = this cell is the
// latter).  Nothing for a model without the hook.
template <typename S>
__device__ __forceinline__ void add_diffusive_face(const S* qL, const S* qR, const S* aL, const S* aR, const S* gL,
                                                   const S* gR, const S* p, S area, bool right, S* fc) {
    if constexpr (HasDiffusiveFlux<S>::value) {
        S gq[NQ * DIM], qf[NQ], af[NA];
#pragma unroll
        for (int k = 0; k < NQ * DIM; ++k) gq[k] = 0.5 * (gL[k] + gR[k]);
#pragma unroll
        for (int i = 0; i < NQ; ++i) qf[i] = 0.5 * (qL[i] + qR[i]);
#pragma unroll
        for (int i = 0; i < NA; ++i) af[i] = 0.5 * (aL[i] + aR[i]);
        auto d = Model<S>::diffusive_flux(qf, af, gq, p);
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            if (right)
                fc[i] -= d[i] * area;
            else
                fc[i] += d[i] * area;
        }
    }
}
// ... at a boundary face (TransportStep.hpp:533-550): the cell's own gradient, the state
// averaged with the ghost, the cell side's face aux
template <typename S>
__device__ __forceinline__ void add_diffusive_boundary(const S* qL, const S* qR, const S* aL, const S* gL,
                                                       const S* p, S area, S* fc) {
    if constexpr (HasDiffusiveFlux<S>::value) {
        S qf[NQ], af[NA];
#pragma unroll
        for (int i = 0; i < NQ; ++i) qf[i] = 0.5 * (qL[i] + qR[i]);
#pragma unroll
        for (int i = 0; i < NA; ++i) af[i] = aL[i];
        auto d = Model<S>::diffusive_flux(qf, af, gL, p);
#pragma unroll
        for (int i = 0; i < NQ; ++i) fc[i] += d[i] * area;
    }
}

// f >= 0 and area > 1e-15: the face aux, flux and fluctuations of the (own qo, neighbour qn with
// clamped aux an) face states and the -D A -/+ F A accumulation into fc, or the boundary ghost
// (nb < 0), then the optional well-balanced cell term.  meta = the slot's cmeta entry (read only
// when nb >= 0); ac = the cell's clamped aux, qcell = its clamped state (well-balanced term only).
__device__ __forceinline__ void flux_face(const DevMesh& m, const Params& P, int f, int nb, int meta, T area,
                                          const T* nh, const T* xf, const T* qo, const T* qn, const T* an,
                                          const T* ac, T time, int wb, const T* qcell, T* fc,
                                          const T* go = nullptr, const T* gn = nullptr) {
    if (nb >= 0) {
        const bool right = (meta >> 2) & 1;
        const T* qL = right ? qn : qo;
        const T* qR = right ? qo : qn;
        const T* aLc = right ? an : ac;
        const T* aRc = right ? ac : an;
        auto aL = M::update_aux_variables(qL, aLc, P.p, 0.0, xf);
        auto aR = M::update_aux_variables(qR, aRc, P.p, 0.0, xf);
        T F[NQ], D[2 * NQ];
        face_flux<NUM, T>(qL, qR, aL.data, aR.data, P.p, nh, F, D);
        if (right) {
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                fc[i] -= D[i] * area;
                fc[i] += F[i] * area;
            }
        } else {
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                fc[i] -= D[NQ + i] * area;
                fc[i] -= F[i] * area;
            }
        }
        if (go) add_diffusive_face<T>(qL, qR, aL.data, aR.data, right ? gn : go, right ? go : gn, P.p, area, right, fc);
    } else {
        const int bc = m.fbc[f];
        if (bc >= 0) {
            auto aL = M::update_aux_variables(qo, ac, P.p, time, xf);
            auto qR = M::boundary_conditions(bc, qo, aL.data, P.p, nh, xf, time, 0.0);
            auto aR = M::update_aux_variables(qR.data, aL.data, P.p, time, xf);
            T F[NQ], D[2 * NQ];
            face_flux<NUM, T>(qo, qR.data, aL.data, aR.data, P.p, nh, F, D);
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                fc[i] -= D[NQ + i] * area;
                fc[i] -= F[i] * area;
            }
            if (go) add_diffusive_boundary<T>(qo, qR.data, aL.data, go, P.p, area, fc);
        }
    }
    if (wb) {
        const bool right = nb >= 0 && ((meta >> 2) & 1);
        const T no[3] = {right ? -nh[0] : nh[0], right ? -nh[1] : nh[1], 0};
        add_cell_source<T>(qo, qcell, P.p, no, area, fc);
    }
}

template <bool NT = false>
__device__ __forceinline__ void flux_divide_volume(const DevMesh& m, int c, T* fc) {
    const T vol = ldx<NT>(m.vol + c);
    if (vol > 1e-15) {
#pragma unroll
        for (int i = 0; i < NQ; ++i) fc[i] /= vol;
    }
}

// The face loop of ComputeFluxes for cell c, slot by slot in ascending order, then / vol.
// own(k, qo) loads face slot k's own state, nbr(k, nb, ks, qn, an) the neighbour's state and
// clamped aux.
template <int MAXF, typename Own, typename Nbr>
__device__ __forceinline__ void flux_cell(const DevMesh& m, const Params& P, int c, T time, const T* ac, int wb,
                                          const T* qcell, Own own, Nbr nbr, T* fc) {
    const int ld = m.ld;
#pragma unroll
    for (int i = 0; i < NQ; ++i) fc[i] = 0;
#pragma unroll
    for (int k = 0; k < MAXF; ++k) {
        const int f = m.cface[(size_t)k * ld + c];
        if (f < 0) continue;
        const T area = m.farea[f];
        if (area <= 1e-15) continue;
        const int nb = m.cnbr[(size_t)k * ld + c];
        const T nh[3] = {m.fnhx[f], DIM > 1 ? m.fnhy[f] : (T)0, 0};
        const T xf[3] = {m.fcx[f], DIM > 1 ? m.fcy[f] : (T)0, 0};
        T qo[NQ], qn[NQ], an[NA];
        own(k, qo);
        int meta = 0;
        if (nb >= 0) {
            meta = m.cmeta[(size_t)k * ld + c];
            nbr(k, nb, meta & 3, qn, an);
        }
        flux_face(m, P, f, nb, meta, area, nh, xf, qo, qn, an, ac, time, wb, qcell, fc);
    }
    flux_divide_volume(m, c, fc);
}

// Inputs of one face slot of cell c for the k_flux gather.  The topology of all slots is read
// first; every load of a slot then uses an in-bounds index whatever the slot holds (no face:
// face 0; boundary face: the cell itself, slot 0), so a slot's loads issue together: one memory
// round trip per slot in place of the dependent chain cface -> farea -> cnbr -> cmeta ->
// neighbour state.  (Prefetching the next slot during this one's flux needs 156 VGPRs, one wave
// per SIMD less, and measures slower.)
constexpr int NGD = HasDiffusiveFlux<T>::value ? NQ * DIM : 1;   // gradient of a diffusive model
struct FaceIn {
    T area, nh[3], xf[3];
    T qo[NQ], qn[NQ], an[NA];
    T gn[NGD];   // the neighbour's cell gradient (diffusive models only)
};

template <int MAXF, bool NT = false>
__device__ __forceinline__ void flux_cell_gather(const DevMesh& m, const Params& P, int c, T time, const T* ac,
                                                 int wb, const T* qcell, const T* __restrict__ QF,
                                                 const T* __restrict__ As, T* fc, const T* __restrict__ G = nullptr) {
    const int ld = m.ld;
    T go[NGD];   // this cell's gradient (diffusive models only)
    if constexpr (HasDiffusiveFlux<T>::value) {
        if (G) load_soa<NGD>(G, ld, c, go);
    }
    int fk[MAXF], nk[MAXF], mk[MAXF];
#pragma unroll
    for (int k = 0; k < MAXF; ++k) {
        const size_t e = 2 * ((size_t)k * ld + c);
        const int nm = ldx<NT>(m.ctopo + e + 1);
        fk[k] = ldx<NT>(m.ctopo + e);
        nk[k] = nm >= 0 ? (nm & 0x0FFFFFFF) : -1;
        mk[k] = nm >= 0 ? (nm >> 28) : 0;
    }
    auto load = [&](int k, FaceIn& in) {
        const int f = fk[k] >= 0 ? fk[k] : 0;
        const int nb = nk[k] >= 0 ? nk[k] : c;
        const int ks = nk[k] >= 0 ? (mk[k] & 3) : 0;
        const T* g = m.fgeo + 3 * (size_t)f;
        in.nh[0] = g[0];
        in.nh[1] = DIM > 1 ? g[1] : (T)0;
        in.area = g[2];
        in.nh[2] = 0;
        in.xf[0] = m.fcx[f];
        in.xf[1] = DIM > 1 ? m.fcy[f] : (T)0;
        in.xf[2] = 0;
#pragma unroll
        for (int i = 0; i < NQ; ++i) in.qo[i] = QF[qf_idx(ld, k, c, i)];
#pragma unroll
        for (int i = 0; i < NQ; ++i) in.qn[i] = QF[qf_idx(ld, ks, nb, i)];
        load_soa<NA>(As, ld, nb, in.an);
        if constexpr (HasDiffusiveFlux<T>::value) {
            if (G) load_soa<NGD>(G, ld, nb, in.gn);
        }
    };
#pragma unroll
    for (int i = 0; i < NQ; ++i) fc[i] = 0;
#pragma unroll
    for (int k = 0; k < MAXF; ++k) {
        FaceIn in;
        load(k, in);
        if (fk[k] >= 0 && in.area > 1e-15)
            flux_face(m, P, fk[k], nk[k], mk[k], in.area, in.nh, in.xf, in.qo, in.qn, in.an, ac, time, wb, qcell,
                      fc, HasDiffusiveFlux<T>::value && G ? go : nullptr, in.gn);
    }
    flux_divide_volume<NT>(m, c, fc);
}

// Out = the stage update of the cell from its right-hand side fc (MODE_RHS: fc itself)
template <int MODE, bool NT = false>
__device__ __forceinline__ void stage_update(const Ctrl* ctrl, int c, int ld, const T* fc, const T* In, T* Out,
                                             const T* X0) {
    if constexpr (MODE == MODE_RHS) {
        store_soa_x<NT, NQ>(Out, ld, c, fc);
    } else if constexpr (MODE == MODE_EULER || MODE == MODE_STAGE) {
        // VecAXPY(u, a, F): u += a*F   (Euler: a = dt; SSP stage: a = dt/(s-1))
        const T a = ctrl->a_stage;
#pragma unroll
        for (int i = 0; i < NQ; ++i) stx<NT>(ldx<NT>(In + (size_t)i * ld + c) + a * fc[i], Out + (size_t)i * ld + c);
    } else if constexpr (MODE == MODE_FINAL) {
        // VecAXPBYPCZ(X, (s-1)/s, dt/s, 1/s, W, F)
        const T al = ctrl->al, be = ctrl->be, ga = ctrl->ga;
#pragma unroll
        for (int i = 0; i < NQ; ++i)
            stx<NT>(al * ldx<NT>(In + (size_t)i * ld + c) + be * fc[i] + ga * ldx<NT>(X0 + (size_t)i * ld + c),
                    Out + (size_t)i * ld + c);
    } else {
        const T dt = ctrl->dt;
#pragma unroll
        for (int i = 0; i < NQ; ++i)
            stx<NT>(ldx<NT>(X0 + (size_t)i * ld + c) + dt * (ark2::B3 * fc[i]), Out + (size_t)i * ld + c);
    }
}

template <int MAXF, int MODE, bool MC, typename TT = T, bool NT = false>
__global__ __launch_bounds__(BLOCK) void k_flux(DevMesh m, Params P, const Ctrl* __restrict__ ctrl,
                                                int stage, const T* __restrict__ QF,
                                                const T* __restrict__ As, const T* In, T* Out,
                                                const T* __restrict__ Ag, int wb, const T* X0,
                                                const T* __restrict__ G) {
    const auto L = locate<MC>(m);
    ctrl += L.m;
    if (!ctrl->active) return;
    const int ld = m.ld;
    auto body = [&](int c) {
        const T time = ctrl->stage_t[stage];
        T fc[NQ], ac[NA];
        load_soa<NA>(As, ld, c, ac);
        T qcell[NQ];   // clamped cell state (only read by the well-balanced term)
        if constexpr (HasCellSource<TT>::value) {
            if (wb) {
                T a0[NA];
                load_soa<NQ>(In, ld, c, qcell);
                load_soa<NA>(Ag, ld, c, a0);
                auto rq = model_update_variables<T>(qcell, a0, P.p, 0.0);
    #pragma unroll
                for (int i = 0; i < NQ; ++i) qcell[i] = rq[i];
            }
        }
        flux_cell_gather<MAXF, NT>(m, P, c, time, ac, wb, qcell, QF, As, fc, G);
        stage_update<MODE, NT>(ctrl, c, ld, fc, In, Out, X0);
    };
    if constexpr (MC) {
        for (int j = 0; j < m.cpt; ++j) {
            const int c = L.cell(j);
            if (c >= m.ncells) return;
            body(c);
        }
    } else {
        if (L.c >= m.ncells) return;
        body(L.c);
    }
}

// ------------------------------------------------------------------------------ source
// SourceStep::SolveLinearSystem (GEPP) with register-resident rows: the pivot swap is a
// compile-time-indexed select so the n x n matrix never spills.
__device__ __forceinline__ bool gepp_reg(T* A, T* b, T* x) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        int mr = k;
        T mv = std::fabs(A[k * NQ + k]);
#pragma unroll
        for (int i = k + 1; i < NQ; ++i) {
            const T v = std::fabs(A[i * NQ + k]);
            if (v > mv) {
                mv = v;
                mr = i;
            }
        }
        if (mv < 1e-14) return false;
        // the selects of the swap run only when a lane of the wavefront pivots off the diagonal
        // (a diagonally dominant I - dt dS/dq rarely does)
        if (__any(mr != k))
#pragma unroll
        for (int i = k + 1; i < NQ; ++i) {
            const bool sw = (i == mr);
#pragma unroll
            for (int j = k; j < NQ; ++j) {
                const T u = A[k * NQ + j], v = A[i * NQ + j];
                A[k * NQ + j] = sw ? v : u;
                A[i * NQ + j] = sw ? u : v;
            }
            const T u = b[k], v = b[i];
            b[k] = sw ? v : u;
            b[i] = sw ? u : v;
        }
#pragma unroll
        for (int i = k + 1; i < NQ; ++i) {
            const T fac = A[i * NQ + k] / A[k * NQ + k];
#pragma unroll
            for (int j = k; j < NQ; ++j) A[i * NQ + j] -= fac * A[k * NQ + j];
            b[i] -= fac * b[k];
        }
    }
#pragma unroll
    for (int i = NQ - 1; i >= 0; --i) {
        T sum = 0.0;
#pragma unroll
        for (int j = i + 1; j < NQ; ++j) sum += A[i * NQ + j] * x[j];
        x[i] = (b[i] - sum) / A[i * NQ + i];
    }
    return true;
}

// The Newton residual R = q - qo - dt S(q, a) of SourceStep::SolveLocalCell (SourceStep.hpp:
// 151-158), its squared norm returned
__device__ __forceinline__ T newton_residual(const Params& P, T dt, const T* q, const T* qo, const T* a, T* r) {
    const T X0[3] = {0, 0, 0};
    auto S = M::source(q, a, P.p, (T)0, dt, X0);
    T rn = 0.0;
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        r[i] = q[i] - qo[i] - dt * S[i];
        rn += r[i] * r[i];
    }
    return rn;
}

// J = I - dt (dS/dq + dS/da da/dq) (SourceStep.hpp:160-197).  A zero term gives J = I's entry
// without the product: for finite dt that is the same value (0 - dt 0 = +0, 1 - dt 0 = 1), and
// entries the model's Jacobian fixes at zero become compile-time constants.
__device__ __forceinline__ void newton_jacobian(const Params& P, T dt, const T* q, const T* a, T* J) {
    const T X0[3] = {0, 0, 0};
    auto dSdQ = M::source_jacobian_wrt_variables(q, a, P.p);
    auto dSdA = M::source_jacobian_wrt_aux_variables(q, a, P.p);
    // The aux chain term dS/dA dA/dq only where dS/dA has a nonzero entry: for a model whose
    // dS/dA is identically zero the test folds at compile time and the finite-difference aux
    // Jacobian is not formed (with a finite dA/dq the term adds a signed zero, which leaves J
    // unchanged)
    bool aux_term = false;
#pragma unroll
    for (int i = 0; i < NQ * NA; ++i) aux_term |= dSdA[i] != (T)0;
    T Jx[NA * NQ];
#pragma unroll
    for (int i = 0; i < NA * NQ; ++i) Jx[i] = 0;
    if (aux_term) {
        auto a0 = M::update_aux_variables(q, a, P.p, 0.0, X0);
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            T qp[NQ];
#pragma unroll
            for (int l = 0; l < NQ; ++l) qp[l] = q[l];
            const T h = (T)1e-7 * (std::fabs(q[j]) + (T)1e-7);
            qp[j] += h;
            auto a1 = M::update_aux_variables(qp, a, P.p, 0.0, X0);
#pragma unroll
            for (int l = 0; l < NA; ++l) Jx[l * NQ + j] = (a1[l] - a0[l]) / h;
        }
    }
#pragma unroll
    for (int i = 0; i < NQ; ++i)
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            T term = dSdQ[i * NQ + j];
            if (aux_term) {
#pragma unroll
                for (int l = 0; l < NA; ++l) term += dSdA[i * NA + l] * Jx[l * NQ + j];
            }
            J[i * NQ + j] = (i == j ? (T)1.0 : (T)0.0) - (term == (T)0 ? (T)0 : dt * term);
        }
}

// Gaussian elimination of A x = b without row exchanges, valid (returns true) where GEPP
// (SourceStep::SolveLinearSystem, SourceStep.hpp:211-242) would keep every pivot on the diagonal
// (no |A[i][k]| > |A[k][k]| below it), no pivot is below 1e-14 or non-finite, every multiplier
// and every x is finite: then x is GEPP's bit for bit.  Zero entries are skipped: a zero below
// the pivot gives GEPP's multiplier 0 / pivot = copysign(0, pivot) and leaves the row of A as it
// is (A never holds -0 and the products are finite), a zero in the pivot row or in the back
// substitution adds a signed zero to a value that is not -0.  Entries the model fixes at zero
// are compile-time constants, so their registers and arithmetic fold away.
__device__ __forceinline__ bool ge_nopivot(T* A, T* b, T* x) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const T piv = A[k * NQ + k];
        const T mv = std::fabs(piv);
#pragma unroll
        for (int i = k + 1; i < NQ; ++i) ok &= !(std::fabs(A[i * NQ + k]) > mv);
        ok &= mv >= (T)1e-14 && mv <= DBL_MAX;
#pragma unroll
        for (int i = k + 1; i < NQ; ++i) {
            const T aik = A[i * NQ + k];
            const T fac = aik == (T)0 ? std::copysign((T)0, piv) : aik / piv;
            ok &= std::fabs(fac) <= DBL_MAX;
#pragma unroll
            for (int j = k + 1; j < NQ; ++j) {
                const T akj = A[k * NQ + j];
                A[i * NQ + j] = akj == (T)0 ? A[i * NQ + j] : A[i * NQ + j] - fac * akj;
            }
            b[i] -= fac * b[k];
        }
    }
#pragma unroll
    for (int i = NQ - 1; i >= 0; --i) {
        T sum = 0.0;
#pragma unroll
        for (int j = i + 1; j < NQ; ++j) sum = A[i * NQ + j] == (T)0 ? sum : sum + A[i * NQ + j] * x[j];
        x[i] = (b[i] - sum) / A[i * NQ + i];
        ok &= std::fabs(x[i]) <= DBL_MAX;
    }
    return ok;
}

// The Newton update of one iteration by the full GEPP with row exchanges (SourceStep::
// SolveLinearSystem, SourceStep.hpp:211-242, the operations of gepp_reg in the same order) for
// the lanes ge_nopivot declined, which are rare.  Those lanes take turns on one LDS workspace per
// wavefront (the matrix, right-hand side and solution, 576 bytes at n = 8), so the matrix holds
// neither registers nor scratch memory.  Each recomputes its residual and J exactly as the caller
// did.  Workgroups of at most BLOCK threads.
constexpr int EXCH_WAVES = BLOCK / 64;
constexpr int EXCH_WS = NQ * NQ + 2 * NQ;

__device__ __forceinline__ bool newton_update_exchange(const Params& P, T dt, const T* q, const T* qo, const T* a,
                                                       T* dl) {
    __shared__ T ws_all[EXCH_WAVES][EXCH_WS];
    T* A = ws_all[threadIdx.x >> 6];
    T* b = A + NQ * NQ;
    T* x = b + NQ;
    const int lane = threadIdx.x & 63;
    bool res = true;
    unsigned long long pend = __ballot(1);   // the lanes in here, served one at a time
    while (pend) {
        const int L = __ffsll((long long)pend) - 1;
        pend &= pend - 1;
        if (lane != L) continue;
        {
            // copies behind an empty asm: the compiler must not merge this residual and Jacobian
            // with the caller's (that would keep the caller's J live in registers)
            T qq[NQ], qqo[NQ], aa[NA];
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                qq[i] = q[i];
                qqo[i] = qo[i];
                asm volatile("" : "+v"(qq[i]), "+v"(qqo[i]));
            }
#pragma unroll
            for (int i = 0; i < NA; ++i) {
                aa[i] = a[i];
                asm volatile("" : "+v"(aa[i]));
            }
            T r[NQ], J[NQ * NQ];
            newton_residual(P, dt, qq, qqo, aa, r);
            newton_jacobian(P, dt, qq, aa, J);
#pragma unroll
            for (int i = 0; i < NQ * NQ; ++i) A[i] = J[i];
#pragma unroll
            for (int i = 0; i < NQ; ++i) b[i] = -r[i];
        }
        bool ok = true;
        #pragma unroll 1
        for (int k = 0; k < NQ && ok; ++k) {
            int mr = k;
            T mv = std::fabs(A[k * NQ + k]);
            #pragma unroll 1
            for (int i = k + 1; i < NQ; ++i) {
                const T v = std::fabs(A[i * NQ + k]);
                if (v > mv) {
                    mv = v;
                    mr = i;
                }
            }
            if (mv < 1e-14) {
                ok = false;
                break;
            }
            if (mr != k) {
                #pragma unroll 1
                for (int j = k; j < NQ; ++j) {
                    const T t = A[k * NQ + j];
                    A[k * NQ + j] = A[mr * NQ + j];
                    A[mr * NQ + j] = t;
                }
                const T t = b[k];
                b[k] = b[mr];
                b[mr] = t;
            }
            #pragma unroll 1
            for (int i = k + 1; i < NQ; ++i) {
                const T fac = A[i * NQ + k] / A[k * NQ + k];
                #pragma unroll 1
                for (int j = k; j < NQ; ++j) A[i * NQ + j] -= fac * A[k * NQ + j];
                b[i] -= fac * b[k];
            }
        }
        if (ok) {
            #pragma unroll 1
            for (int i = NQ - 1; i >= 0; --i) {
                T sum = 0.0;
                #pragma unroll 1
                for (int j = i + 1; j < NQ; ++j) sum += A[i * NQ + j] * x[j];
                x[i] = (b[i] - sum) / A[i * NQ + i];
            }
#pragma unroll
            for (int i = 0; i < NQ; ++i) dl[i] = x[i];
        }
        res = ok;
    }
    return res;
}

// SourceStep::SolveLocalCell (SourceStep.hpp:125-209): implicit Euler Newton, <= 15
// iterations, |R|_2 < 1e-8, aux Jacobian by forward differences (SourceStep.hpp:40-50)
__device__ __forceinline__ void newton_cell(const Params& P, T dt, T* q, const T* aux_init) {
    const T X0[3] = {0, 0, 0};
    T qo[NQ], a[NA];
#pragma unroll
    for (int i = 0; i < NQ; ++i) qo[i] = q[i];
#pragma unroll
    for (int i = 0; i < NA; ++i) a[i] = aux_init[i];
    for (int it = 0; it < 15; ++it) {
        auto ua = M::update_aux_variables(q, a, P.p, 0.0, X0);
#pragma unroll
        for (int i = 0; i < NA; ++i) a[i] = ua[i];
        T r[NQ];
        const T rn = newton_residual(P, dt, q, qo, a, r);
        if (std::sqrt(rn) < 1e-8) break;
        T J[NQ * NQ];
        newton_jacobian(P, dt, q, a, J);
#pragma unroll
        for (int i = 0; i < NQ; ++i) r[i] = -r[i];
        T dl[NQ];
        if (!ge_nopivot(J, r, dl) && !newton_update_exchange(P, dt, q, qo, a, dl)) break;
#pragma unroll
        for (int i = 0; i < NQ; ++i) q[i] += dl[i];
    }
}

// local CFL ratio r_c / lambda_c (VirtualSolver.hpp:316-345); DBL_MAX when wave-free
__device__ __forceinline__ T cfl_ratio(const Params& P, const T* q, const T* a, T rc) {
    T eig = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        T n[3] = {0, 0, 0};
        n[d] = 1.0;
        auto r = NUM::local_max_abs_eigenvalue(q, a, P.p, n);
        if (r[0] > eig) eig = r[0];
    }
    if (eig < 1e-14) return DBL_MAX;
    if (rc < DBL_MAX) return rc / eig;
    return DBL_MAX;
}

__device__ __forceinline__ T block_min(T v) {
    __shared__ T red[BLOCK / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) v = red[i] < v ? red[i] : v;
    }
    return v;
}

// POST: 0 = CFL only, 1 = constraints + CFL, 2 = constraints + source + constraints + CFL,
//       3 = source only (no CFL), 4 = constraints only (no CFL),
//       5 = source with the step's dt + constraints (no CFL; the Chorin predictor),
//       6 = source with the step's dt (no constraints, no CFL)
template <int POST, bool MC>
__global__ __launch_bounds__(BLOCK) void k_post(DevMesh m, Params P, const Ctrl* __restrict__ ctrl, T* X,
                                                T* A, T* __restrict__ partials, T dt_override) {
    const auto L = locate<MC>(m);
    ctrl += L.m;
    if (!ctrl->active) return;
    const int ld = m.ld;
    T ratio = DBL_MAX;
    auto body = [&](int c) -> T {
        T q[NQ], a[NA];
        load_soa<NQ>(X, ld, c, q);
        load_soa<NA>(A, ld, c, a);
        if constexpr (POST == 1 || POST == 2 || POST == 4) update_state_cell(P, q, a);
        if constexpr (POST == 2 || POST == 3 || POST == 5 || POST == 6) {
            const T dt = POST == 3 ? dt_override : ctrl->dt;
            newton_cell(P, dt, q, a);
        }
        if constexpr (POST == 2 || POST == 5) update_state_cell(P, q, a);
        if constexpr (POST != 0) {
            store_soa<NQ>(X, ld, c, q);
            store_soa<NA>(A, ld, c, a);
        }
        if constexpr (POST <= 2) return cfl_ratio(P, q, a, m.rc[c]);
        return (T)DBL_MAX;
    };
    if constexpr (MC) {
        for (int j = 0; j < m.cpt; ++j) {
            const int c = L.cell(j);
            if (c >= m.ncells) break;
            const T r = body(c);
            ratio = r < ratio ? r : ratio;
        }
    } else {
        if (L.c < m.ncells) ratio = body(L.c);
    }
    if constexpr (POST <= 2) {
        ratio = block_min(ratio);
        if (threadIdx.x == 0) partials[L.lb] = ratio;
    }
}

// The final SSP stage of a step and the post-stage of k_post<POST> (POST 1: constraints + CFL,
// 2: constraints, source, constraints + CFL) in one pass: each cell's new state
// X1 = al W + be F(W) + ga X0 stays in registers for the post-stage instead of a store, a load
// and a second store through HBM; the post-stage's result goes to Xout (X0 = Xout in place for a
// plain step).  The same operations on the same values as k_flux<MODE_FINAL> followed by
// k_post<POST>, so the results are bit for bit theirs.  REC: the recording step of the adjoint,
// which also keeps X1, and for POST 2 the aux before the source (A3) and the source's solution
// (X3), as k_post_record does, and the step's dt in dt_log.  Not for the well-balanced cell term.
template <int MAXF, int POST, bool MC, bool REC, bool NT = false>
__global__ __launch_bounds__(BLOCK) void k_flux_post(DevMesh m, Params P, const Ctrl* __restrict__ ctrl, int stage,
                                                     const T* __restrict__ QF, const T* __restrict__ As,
                                                     const T* __restrict__ W, const T* X0, T* Xout, T* A,
                                                     T* __restrict__ partials, const T* __restrict__ G,
                                                     T* __restrict__ X1, T* __restrict__ X3, T* __restrict__ A3,
                                                     T* __restrict__ dt_log) {
    static_assert(POST == 1 || POST == 2, "k_flux_post: POST 1 or 2");
    const auto L = locate<MC>(m);
    ctrl += L.m;
    if (!ctrl->active) return;   // uniform over the workgroup (members own whole workgroups)
    if (REC && dt_log && L.lb == 0 && threadIdx.x == 0) *dt_log = ctrl->dt;   // the step's dt
    const int ld = m.ld;
    T ratio = DBL_MAX;
    auto body = [&](int c) -> T {
        T q[NQ], a[NA];
        {
            const T time = ctrl->stage_t[stage];
            T fc[NQ], ac[NA];
            load_soa<NA>(As, ld, c, ac);
            flux_cell_gather<MAXF, NT>(m, P, c, time, ac, 0, nullptr, QF, As, fc, G);
            // VecAXPBYPCZ(X, (s-1)/s, dt/s, 1/s, W, F) as stage_update<MODE_FINAL>
            const T al = ctrl->al, be = ctrl->be, ga = ctrl->ga;
#pragma unroll
            for (int i = 0; i < NQ; ++i)
                q[i] = al * ldx<NT>(W + (size_t)i * ld + c) + be * fc[i] + ga * ldx<NT>(X0 + (size_t)i * ld + c);
        }
        if constexpr (REC) store_soa_x<NT, NQ>(X1, ld, c, q);
        load_soa_x<NT, NA>(A, ld, c, a);
        update_state_cell(P, q, a);
        if constexpr (POST == 2) {
            if constexpr (REC) store_soa_x<NT, NA>(A3, ld, c, a);
            newton_cell(P, ctrl->dt, q, a);
            if constexpr (REC) store_soa_x<NT, NQ>(X3, ld, c, q);
            update_state_cell(P, q, a);
        }
        store_soa_x<NT, NQ>(Xout, ld, c, q);
        store_soa_x<NT, NA>(A, ld, c, a);
        return cfl_ratio(P, q, a, ldx<NT>(m.rc + c));
    };
    if constexpr (MC) {
        for (int j = 0; j < m.cpt; ++j) {
            const int c = L.cell(j);
            if (c >= m.ncells) break;
            const T r = body(c);
            ratio = r < ratio ? r : ratio;
        }
    } else {
        if (L.c < m.ncells) ratio = body(L.c);
    }
    ratio = block_min(ratio);
    if (threadIdx.x == 0) partials[L.lb] = ratio;
}

// Post-stage of a step with the local MOOD override (MUSCLSolver.hpp:150-182): constraints on the
// transport result X; a cell left with h < -1e-10 is replaced by its first-order forward-Euler
// update from the saved start-of-step state Xs, X_save + dt F (F = k_flux's face loop over the
// clamped saved states of the cell and its neighbours, FormRHSTroubledO1, TransportStep.hpp:
// 363-381, at the step's start time), then constraints again; then [source, constraints,] CFL as
// k_post.  Cells are independent: Xs and As are read-only here, X and A are written per cell.
// Untroubled cells get one more clamp in the reference, a no-op (the clamp is idempotent).
template <int MAXF, int SRC, bool MC>
__global__ __launch_bounds__(BLOCK) void k_post_mood(DevMesh m, Params P, Ctrl* __restrict__ ctrl, T* X,
                                                     T* A, const T* __restrict__ Xs,
                                                     const T* __restrict__ As, T* __restrict__ partials) {
    const auto L = locate<MC>(m);
    ctrl += L.m;
    if (!ctrl->active) return;
    const int ld = m.ld;
    T ratio = DBL_MAX;
    auto body = [&](int c) -> T {
        T q[NQ], a[NA];
        load_soa<NQ>(X, ld, c, q);
        load_soa<NA>(A, ld, c, a);
        update_state_cell(P, q, a);
        if (q[1] < (T)-1.0e-10) {
            T qs[NQ], as[NA], qsc[NQ], fc[NQ];
            load_soa<NQ>(Xs, ld, c, qs);
            load_soa<NA>(As, ld, c, as);
#pragma unroll
            for (int i = 0; i < NQ; ++i) qsc[i] = qs[i];
            update_state_cell(P, qsc, as);
            flux_cell<MAXF>(
                m, P, c, ctrl->t, as, 0, nullptr,
                [&](int k, T* qo) {
#pragma unroll
                    for (int i = 0; i < NQ; ++i) qo[i] = qsc[i];
                },
                [&](int k, int nb, int ks, T* qn, T* an) {
                    load_soa<NQ>(Xs, ld, nb, qn);
                    load_soa<NA>(As, ld, nb, an);
                    update_state_cell(P, qn, an);
                },
                fc);
            const T dt = ctrl->dt;
#pragma unroll
            for (int i = 0; i < NQ; ++i) q[i] = qs[i] + dt * fc[i];
            update_state_cell(P, q, a);
            atomicAdd(&ctrl->troubled, 1);
        }
        if constexpr (SRC) {
            newton_cell(P, ctrl->dt, q, a);
            update_state_cell(P, q, a);
        }
        store_soa<NQ>(X, ld, c, q);
        store_soa<NA>(A, ld, c, a);
        return cfl_ratio(P, q, a, m.rc[c]);
    };
    if constexpr (MC) {
        for (int j = 0; j < m.cpt; ++j) {
            const int c = L.cell(j);
            if (c >= m.ncells) break;
            const T r = body(c);
            ratio = r < ratio ? r : ratio;
        }
    } else {
        if (L.c < m.ncells) ratio = body(L.c);
    }
    ratio = block_min(ratio);
    if (threadIdx.x == 0) partials[L.lb] = ratio;
}

// k_post<POST> (POST 1: constraints; POST 2: constraints, source, constraints; POST 5: source,
// constraints, the Chorin predictor's) of a step whose intermediate states are kept: reads the
// final stage's output X1, writes the source's solution X3 and the aux A3 it saw (POST 2, 5),
// the end-of-step state Xout and its aux A, and the CFL partial minima; the same operations as
// k_post.  dt_log (may be null) receives the step's dt.
template <int POST, bool MC>
__global__ __launch_bounds__(BLOCK) void k_post_record(DevMesh m, Params P, const Ctrl* __restrict__ ctrl,
                                                       const T* __restrict__ X1, T* __restrict__ Xout, T* A,
                                                       T* __restrict__ X3, T* __restrict__ A3,
                                                       T* __restrict__ partials, T* __restrict__ dt_log) {
    const auto L = locate<MC>(m);
    ctrl += L.m;
    if (!ctrl->active) return;
    if (dt_log && L.lb == 0 && threadIdx.x == 0) *dt_log = ctrl->dt;   // the step's dt
    const int ld = m.ld;
    T ratio = DBL_MAX;
    auto body = [&](int c) -> T {
        T q[NQ], a[NA];
        load_soa<NQ>(X1, ld, c, q);
        load_soa<NA>(A, ld, c, a);
        if constexpr (POST != 5) update_state_cell(P, q, a);
        if constexpr (POST == 2 || POST == 5) {
            store_soa<NA>(A3, ld, c, a);
            newton_cell(P, ctrl->dt, q, a);
            store_soa<NQ>(X3, ld, c, q);
            update_state_cell(P, q, a);
        }
        store_soa<NQ>(Xout, ld, c, q);
        store_soa<NA>(A, ld, c, a);
        return cfl_ratio(P, q, a, m.rc[c]);
    };
    if constexpr (MC) {
        for (int j = 0; j < m.cpt; ++j) {
            const int c = L.cell(j);
            if (c >= m.ncells) break;
            const T r = body(c);
            ratio = r < ratio ? r : ratio;
        }
    } else {
        if (L.c < m.ncells) ratio = body(L.c);
    }
    ratio = block_min(ratio);
    if (threadIdx.x == 0) partials[L.lb] = ratio;
}

// what: 0 = standalone ComputeTimeStep (-> dt_query), 1 = first dt, 2 = end of a step,
//       3 = end of a Chorin step (next dt = max(CFL, min_dt), ChorinVAM.hpp:103)
// The minimum of the partial minima in [b0, b1) with FACC independent minima per thread of a
// FBLOCK-thread workgroup (a 2 M-cell step's ~31 000 partials are a few rounds of loads in
// flight); the minimum is exact in any order.  Returns it in thread 0.
constexpr int FBLOCK = 1024;   // k_finalize's workgroup
constexpr int FACC = 16;

__device__ __forceinline__ T range_min(const T* __restrict__ partials, int b0, int b1) {
    T v[FACC];
#pragma unroll
    for (int a = 0; a < FACC; ++a) v[a] = DBL_MAX;
    const int st = blockDim.x;
    int i = b0 + threadIdx.x;
    for (; i + (FACC - 1) * st < b1; i += FACC * st) {
        T p[FACC];
#pragma unroll
        for (int a = 0; a < FACC; ++a) p[a] = partials[i + a * st];
#pragma unroll
        for (int a = 0; a < FACC; ++a) v[a] = p[a] < v[a] ? p[a] : v[a];
    }
    for (; i < b1; i += st) v[0] = partials[i] < v[0] ? partials[i] : v[0];
#pragma unroll
    for (int a = 1; a < FACC; ++a) v[0] = v[a] < v[0] ? v[a] : v[0];
    __shared__ T red[FBLOCK / 64];
    T w = v[0];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T x = __shfl_xor(w, o, 64);
        w = x < w ? x : w;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) w = red[k] < w ? red[k] : w;
    return w;
}

// The CFL dt from the minimum ratio vmin and the control block of the next step
// (MUSCLSolver.hpp:140-149); one thread.
__device__ __forceinline__ void finish_step(Ctrl* ctrl, T vmin, int what, int order, int s, T cfl, T min_dt,
                                            T t_end, T dt_max) {
    const bool fin = !(!isfinite(vmin) || vmin >= 0.5 * DBL_MAX);
    const T dtc = fin ? cfl * vmin : dt_max;
    if (what == 0) {
        ctrl->dt_query = dtc;
        ctrl->finite = fin;
        return;
    }
    if (what == 1) {
        ctrl->dt_cfl = (dtc < min_dt) ? min_dt : dtc;   // std::max(ComputeTimeStep(), min_dt)
    } else {
        ctrl->t = ctrl->t + ctrl->dt;
        ctrl->steps = ctrl->steps + 1;
        ctrl->dt_cfl = (what == 3 && dtc < min_dt) ? min_dt : dtc;
    }
    ctrl->finite = fin;
    // prepare the next step (MUSCLSolver.hpp:140-149)
    const T t0 = ctrl->t;
    ctrl->active = t0 < t_end;
    T dt0 = ctrl->dt_cfl;
    if (t0 + dt0 > t_end) dt0 = t_end - t0;
    ctrl->dt = dt0;
    if (order < 2) {
        ctrl->a_stage = dt0;
        ctrl->stage_t[0] = t0;
    } else {
        ctrl->a_stage = dt0 / (s - 1.);
        for (int i = 0; i < s - 1; ++i) ctrl->stage_t[i] = t0 + dt0 * (i / (s - 1.));
        ctrl->stage_t[s - 1] = t0 + dt0;
        ctrl->al = (s - 1.) / s;
        ctrl->be = dt0 / s;
        ctrl->ga = 1. / s;
    }
}

// One workgroup per ensemble member: the member's partial minima are those of its own
// workgroups.
__global__ __launch_bounds__(FBLOCK) void k_finalize(Ctrl* ctrl, const T* __restrict__ partials, int np,
                                                     const int* __restrict__ blk_first, int what,
                                                     int order, int s, T cfl, T min_dt, T t_end,
                                                     T dt_max) {
    ctrl += blockIdx.x;
    if ((what == 2 || what == 3) && !ctrl->active) return;
    const int b0 = blk_first ? blk_first[blockIdx.x] : 0;
    const int b1 = blk_first ? blk_first[blockIdx.x + 1] : np;
    const T vmin = range_min(partials, b0, b1);
    if (threadIdx.x == 0) finish_step(ctrl, vmin, what, order, s, cfl, min_dt, t_end, dt_max);
}

// A single run: NXCD workgroups, workgroup x over the partials the transport kernels' XCD x
// wrote (logical_block gives each XCD one contiguous range, so the reads hit that XCD's L2).
// Each writes its minimum to xmin[x]; the last to finish (an arrival counter, reset by it for
// the next launch) takes the minimum of the NXCD values and updates the control block.  The
// result does not depend on the arrival order.
__global__ __launch_bounds__(FBLOCK) void k_finalize_xcd(Ctrl* ctrl, const T* __restrict__ partials, int np,
                                                         T* xmin, unsigned* arrived, int what, int order,
                                                         int s, T cfl, T min_dt, T t_end, T dt_max,
                                                         T* dt_log) {
    if ((what == 2 || what == 3) && !ctrl->active) return;
    const int x = blockIdx.x, q = np / NXCD, r = np % NXCD;
    const int b0 = x * q + (x < r ? x : r), b1 = b0 + q + (x < r ? 1 : 0);
    const T w = range_min(partials, b0, b1);
    if (threadIdx.x != 0) return;
    xmin[x] = w;
    __threadfence();
    if (atomicAdd(arrived, 1u) != (unsigned)(NXCD - 1)) return;
    __threadfence();
    T vmin = DBL_MAX;
    for (int k = 0; k < NXCD; ++k) {
        const T v = __hip_atomic_load(xmin + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        vmin = v < vmin ? v : vmin;
    }
    *arrived = 0u;
    finish_step(ctrl, vmin, what, order, s, cfl, min_dt, t_end, dt_max);
    if (dt_log) *dt_log = ctrl->dt;   // the dt of the next step (optional record)
}

__global__ void k_aos_to_soa(const T* __restrict__ aos, T* __restrict__ soa, int n, int nc, int ld) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int i = 0; i < nc; ++i) soa[(size_t)i * ld + c] = aos[(size_t)c * nc + i];
}
__global__ void k_soa_to_aos(const T* __restrict__ soa, T* __restrict__ aos, int n, int nc, int ld) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int i = 0; i < nc; ++i) aos[(size_t)c * nc + i] = soa[(size_t)i * ld + c];
}
// STREAM-style copy used to measure the achievable HBM bandwidth and to calibrate the
// FETCH_SIZE / WRITE_SIZE counters against a kernel of known traffic.
__global__ __launch_bounds__(1024) void k_stream_copy(const double* __restrict__ a,
                                                      double* __restrict__ b, size_t n2) {
    // streaming (non-temporal) 2 x 8 bytes per element, coalesced across the workgroup
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) {
        const double x = __builtin_nontemporal_load(a + 2 * i);
        const double y = __builtin_nontemporal_load(a + 2 * i + 1);
        __builtin_nontemporal_store(x, b + 2 * i);
        __builtin_nontemporal_store(y, b + 2 * i + 1);
    }
}

__global__ void k_fill(T* __restrict__ a, size_t n, T v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Green-Gauss derivative along `axis` at cell c of a field given by get(cell), with a
// zero-Neumann boundary (face value = cell value) and ascending face order, then / vol
// (compute_derivative, UserFunctions.H:60-112)
template <int MAXF, typename GET>
__device__ __forceinline__ T gg_deriv(const DevMesh& m, int c, int axis, GET&& get) {
    const int ld = m.ld;
    const T fc = get(c);
    T g = 0;
#pragma unroll
    for (int k = 0; k < MAXF; ++k) {
        const int f = m.cface[(size_t)k * ld + c];
        if (f < 0) continue;
        const int nb = m.cnbr[(size_t)k * ld + c];
        const T N = axis == 0 ? m.fNx[f] : m.fNy[f];
        if (nb >= 0) {
            const T fv = (T)0.5 * (fc + get(nb));
            const T v = fv * N;
            if ((m.cmeta[(size_t)k * ld + c] >> 2) & 1)
                g -= v;
            else
                g += v;
        } else {
            g += fc * N;
        }
    }
    const T vol = m.vol[c];
    return vol > (T)0 ? g / vol : (T)0;
}

// out = the derivative along `axis` of the per-cell field src (the device side of
// compute_derivative for the field-level update_aux_variables)
template <int MAXF>
__global__ __launch_bounds__(BLOCK) void k_deriv1(DevMesh m, const T* __restrict__ src, int axis,
                                                  T* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.ncells) return;
    out[c] = gg_deriv<MAXF>(m, c, axis, [&](int cc) { return src[cc]; });
}

#include "krylov.inc"
#include "imex.inc"
#include "implicit.inc"
#include "binomial.h"
#include "adjoint.inc"
#ifdef ZOOMY_MODEL_IS_VAM
#include "chorin.inc"
#include "chorin_adjoint.inc"
#endif
struct ChorinHost;
struct AdjointHost;
struct ImplicitHost;
void adjoint_free(AdjointHost* ad);
void implicit_free(ImplicitHost* im);

}  // namespace

// ================================================================================ host side

struct zdm_solver {
    zdm_options opt;
    Params P;
    bool refresh = false;             // RefreshDerivativeAux before every UpdateState (the model has
                                      // the field-level update_aux_variables and the setting is on)
    T* Aref = nullptr;                // the aux a transport stage reads after the refresh
    T* G = nullptr;                   // cell gradients [NQ DIM][ld] of a model with diffusive_flux
    int deriv_err = 0;                // compute_derivative failures of the last refresh
    std::string deriv_msg;
    int ncells, nfaces, ld, maxf, nblocks;
    int bs;                           // workgroup size of the transport kernels (64, 128, 256)
    int ncb;                          // workgroups of 256 over all cells (per-cell kernels via cell_index)
    int maxblocks;                    // block tables / CFL partials capacity (bs = 64, cpt = 1)
    std::vector<int> coff_h;          // ensemble cell offsets (host copy)
    std::vector<double> tune_ms;      // last autotune: ms per step of each candidate
    hipStream_t stream;
    char* arena;
    size_t arena_bytes;
    DevMesh dm;
    T *X, *W, *F, *A, *As, *Atmp, *QF, *partials, *stage_aos;
    T* Xs;                            // MOOD: the state at the start of the step
    T* Im;                            // IMEX: stage right-hand side, stage state, running sums P3, F
    double cfl_in;                    // solver.cfl as given (opt.cfl: capped for zhang_shu)
    Ctrl *ctrl, *ctrl_rhs;
    T* dt_log = nullptr;              // single run: finalize also stores the next dt here (adjoint forward)
    bool timing;                      // record kernel events on the step being enqueued
    int timing_stride;                // zdm_enable_timing: events on every timing_stride-th step
    long long timing_step;
    std::vector<hipEvent_t> evpool;   // recorded in order; pending until resolved
    std::vector<int> evkind;          // kernel id that ends at event i (-1 = segment start)
    size_t evused;
    double kt_ms[4];
    int kt_n[4];
    bool fuse_post = true;            // zdm_fuse_post: the final stage and the post-step as k_flux_post
    int flux_nt_mode = -1;            // zdm_flux_nontemporal: -1 by size, 0 off, 1 on
    bool flux_nt = false;             // the stage / final flux launches use their non-temporal form
    hipEvent_t ev_caller = nullptr;   // orders the device-pointer entry points on a caller's stream
    int alpha_mode = 0;               // prep: 0 as configured, 1 store every face's limiter value at
    T* alpha_io = nullptr;            // alpha_io, 2 take them from it (advanced one stage per call)
    ChorinHost* ch;                   // pressure split state (VAM models)
    int nmem;                         // ensemble members (0: one mesh)
    int nctrl;                        // control blocks: max(1, nmem)
    hipGraph_t graph;                 // one captured time step (zdm_step_graph)
    hipGraphExec_t graph_exec;
    AdjointHost* adj;                 // calibration adjoint state (zdm_adjoint_*)
    ImplicitHost* imp;                // time_integration "implicit": Newton-Krylov state
    int step_err;                     // a host-driven step failed (the message is in zdm_last_error)
};

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// the gradient buffer k_prep writes and k_flux reads for a model with diffusive_flux at order 2
// (the reference computes the gradient only with a reconstruction, TransportStep.hpp:336-345)
T* grad_buffer(const zdm_solver* s) {
    return HasDiffusiveFlux<T>::value && s->opt.order >= 2 ? s->G : nullptr;
}

// ZoomyMesh::derivative of the solver: compute_derivative(result, source, dx, dy, dz) of the
// model's field-level update_aux_variables on device fields, ordered on the solver stream.
// Order 0 copies; a first derivative is the Green-Gauss kernel; other orders (which the
// reference does not support either, UserFunctions.H:62-65) are recorded as an error.
void field_derivative(void* ctx, double* result, const double* source, int dx, int dy, int dz) {
    zdm_solver* s = static_cast<zdm_solver*>(ctx);
    const int total = dx + dy + dz;
    if (total == 0) {
        if (hipMemcpyAsync(result, source, sizeof(T) * s->ncells, hipMemcpyDeviceToDevice, s->stream) != hipSuccess) {
            ++s->deriv_err;
            s->deriv_msg = "compute_derivative: device copy failed";
        }
        return;
    }
    const int axis = dx == 1 ? 0 : (dy == 1 ? 1 : 2);
    if (dx < 0 || dy < 0 || dz < 0 || total != 1 || axis >= DIM) {
        ++s->deriv_err;
        s->deriv_msg = "compute_derivative: only first derivatives along the mesh's axes are supported (D^(" +
                       std::to_string(dx) + "," + std::to_string(dy) + "," + std::to_string(dz) + ") requested)";
        return;
    }
    const unsigned g = (unsigned)((s->ncells + 255) / 256);
    if (s->maxf <= 2)
        k_deriv1<2><<<g, 256, 0, s->stream>>>(s->dm, source, axis, result);
    else if (s->maxf == 3)
        k_deriv1<3><<<g, 256, 0, s->stream>>>(s->dm, source, axis, result);
    else
        k_deriv1<4><<<g, 256, 0, s->stream>>>(s->dm, source, axis, result);
    if (hipGetLastError() != hipSuccess) {
        ++s->deriv_err;
        s->deriv_msg = "compute_derivative: kernel launch failed";
    }
}

// TransportStep::RefreshDerivativeAux (TransportStep.hpp:161-202): the model's field-level
// update_aux_variables(Q, Qaux, 0, mesh) over the per-field planes of X and A.  The solver's
// structure-of-arrays planes are the per-field arrays the overload expects, so nothing is
// de-interleaved; every compute_derivative is a kernel on the solver stream.
template <typename S>
int refresh_aux(zdm_solver* s, const S* X, S* A) {
    if constexpr (HasFieldAuxUpdate<S>::value) {
        S* Qp[NQ];
        S* Ap[NA];
        for (int i = 0; i < NQ; ++i) Qp[i] = const_cast<S*>(X) + (size_t)i * s->ld;
        for (int k = 0; k < NA; ++k) Ap[k] = A + (size_t)k * s->ld;
        const ZoomyMesh mesh{s, &field_derivative, DIM, 0, (long)s->ncells};
        s->deriv_err = 0;
        Model<S>::update_aux_variables(Qp, Ap, (S)0, mesh);
        if (s->deriv_err) return fail(s->deriv_msg);
    }
    (void)s;
    (void)X;
    (void)A;
    return 0;
}

// Workgroups of the transport kernels for workgroup size bs and cpt cells per thread: one
// range for a single mesh, whole workgroups per member for an ensemble (member tables).
// Returns the workgroup count.
int launch_tables(int nc, const std::vector<int>& coff, int bs, int cpt, std::vector<int>& blk_member,
                  std::vector<int>& blk_first) {
    const int span = bs * cpt;
    blk_member.clear();
    blk_first.clear();
    if (coff.empty()) return (nc + span - 1) / span;
    const int nmem = (int)coff.size() - 1;
    blk_first.assign(nmem + 1, 0);
    for (int i = 0; i < nmem; ++i) {
        const int cnt = coff[i + 1] - coff[i];
        blk_first[i + 1] = blk_first[i] + (cnt + span - 1) / span;
        for (int b = blk_first[i]; b < blk_first[i + 1]; ++b) blk_member.push_back(i);
    }
    return blk_first[nmem];
}

// Least-squares weights of LeastSquaresGradient::Setup (Gradient.hpp:222-298), order 1
bool chol_solve_host(int n, std::vector<T>& A, std::vector<T>& b) {
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

template <int MAXF, bool MC>
void launch_prep_t(zdm_solver* s, const Ctrl* ctrl, const T* Xin, const T* Ag, T* As, T* QF) {
    dim3 g(s->nblocks), b(s->bs);
    const int lim = s->opt.use_limiters, pos = s->opt.positivity;
    const T eps = s->opt.wet_dry_eps;
    if (s->opt.order < 2)
        k_prep<MAXF, 0, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, ctrl, Xin, Ag, As, QF, lim, eps, 0,
                                                           grad_buffer(s));
    else if (pos != 0)
        k_prep<MAXF, 3, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, ctrl, Xin, Ag, As, QF, lim, eps, pos == 2,
                                                           grad_buffer(s));
    else if (s->opt.gradient == 1)
        k_prep<MAXF, 2, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, ctrl, Xin, Ag, As, QF, lim, eps, 0,
                                                           grad_buffer(s));
    else if (!MC && s->flux_nt)
        k_prep<MAXF, 1, false, true><<<g, b, 0, s->stream>>>(s->dm, s->P, ctrl, Xin, Ag, As, QF, lim, eps, 0,
                                                             grad_buffer(s));
    else
        k_prep<MAXF, 1, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, ctrl, Xin, Ag, As, QF, lim, eps, 0,
                                                           grad_buffer(s));
}
template <int MAXF>
void launch_prep(zdm_solver* s, const Ctrl* ctrl, const T* Xin, const T* Ag, T* As, T* QF) {
    if (s->dm.cpt > 1)
        launch_prep_t<MAXF, true>(s, ctrl, Xin, Ag, As, QF);
    else
        launch_prep_t<MAXF, false>(s, ctrl, Xin, Ag, As, QF);
}

// MODE_FINAL: Out = al In + be F(In) + ga X0 (X0 = Out in place)
template <int MAXF, bool MC>
void launch_flux_t(zdm_solver* s, const Ctrl* ctrl, int mode, int stage, const T* In, T* Out, const T* Ag,
                 const T* QF, const T* As, const T* X0) {
    dim3 g(s->nblocks), b(s->bs);
    DevMesh dmr = s->dm;
    dmr.rev = 1;   // against k_prep's direction: start on the face states it wrote last
    const int wb = s->opt.well_balanced;
    switch (mode) {
        case MODE_RHS:
            if constexpr (!MC) {
                if (s->flux_nt) {
                    k_flux<MAXF, MODE_RHS, false, T, true><<<g, b, 0, s->stream>>>(
                        dmr, s->P, ctrl, stage, QF, As, In, Out, Ag, wb, X0, grad_buffer(s));
                    break;
                }
            }
            k_flux<MAXF, MODE_RHS, MC><<<g, b, 0, s->stream>>>(dmr, s->P, ctrl, stage, QF, As, In, Out, Ag, wb, X0,
                                                                grad_buffer(s));
            break;
        case MODE_EULER:
            k_flux<MAXF, MODE_EULER, MC><<<g, b, 0, s->stream>>>(dmr, s->P, ctrl, stage, QF, As, In, Out, Ag, wb, X0,
                                                                grad_buffer(s));
            break;
        case MODE_STAGE:
            if constexpr (!MC) {
                if (s->flux_nt) {
                    k_flux<MAXF, MODE_STAGE, false, T, true><<<g, b, 0, s->stream>>>(
                        dmr, s->P, ctrl, stage, QF, As, In, Out, Ag, wb, X0, grad_buffer(s));
                    break;
                }
            }
            k_flux<MAXF, MODE_STAGE, MC><<<g, b, 0, s->stream>>>(dmr, s->P, ctrl, stage, QF, As, In, Out, Ag, wb, X0,
                                                                grad_buffer(s));
            break;
        case MODE_IMEX3:
            if constexpr (!MC) {
                if (s->flux_nt) {
                    k_flux<MAXF, MODE_IMEX3, false, T, true><<<g, b, 0, s->stream>>>(
                        dmr, s->P, ctrl, stage, QF, As, In, Out, Ag, wb, X0, grad_buffer(s));
                    break;
                }
            }
            k_flux<MAXF, MODE_IMEX3, MC><<<g, b, 0, s->stream>>>(dmr, s->P, ctrl, stage, QF, As, In, Out, Ag, wb, X0,
                                                                grad_buffer(s));
            break;
        default:
            if constexpr (!MC) {
                if (s->flux_nt) {
                    k_flux<MAXF, MODE_FINAL, false, T, true><<<g, b, 0, s->stream>>>(
                        dmr, s->P, ctrl, stage, QF, As, In, Out, Ag, wb, X0, grad_buffer(s));
                    break;
                }
            }
            k_flux<MAXF, MODE_FINAL, MC><<<g, b, 0, s->stream>>>(dmr, s->P, ctrl, stage, QF, As, In, Out, Ag, wb, X0,
                                                                grad_buffer(s));
            break;
    }
}
template <int MAXF>
void launch_flux(zdm_solver* s, const Ctrl* ctrl, int mode, int stage, const T* In, T* Out, const T* Ag,
                 const T* QF, const T* As, const T* X0) {
    if (s->dm.cpt > 1)
        launch_flux_t<MAXF, true>(s, ctrl, mode, stage, In, Out, Ag, QF, As, X0);
    else
        launch_flux_t<MAXF, false>(s, ctrl, mode, stage, In, Out, Ag, QF, As, X0);
}

// prep (Green-Gauss) with the limiter values stored in abuf: the adjoint's recording forward
void prep_record(zdm_solver* s, const Ctrl* ctrl, const T* Xin, const T* Ag, T* QF, T* As, T* abuf) {
    dim3 g(s->nblocks), b(s->bs);
    const int lim = s->opt.use_limiters;
#define ZDM_PREC(MF)                                                                                      \
    if (s->dm.cpt > 1)                                                                                   \
        k_prep_rec<MF, true><<<g, b, 0, s->stream>>>(s->dm, s->P, ctrl, Xin, Ag, As, QF, lim, abuf);     \
    else                                                                                                 \
        k_prep_rec<MF, false><<<g, b, 0, s->stream>>>(s->dm, s->P, ctrl, Xin, Ag, As, QF, lim, abuf);
    if (s->maxf <= 2) {
        ZDM_PREC(2)
    } else if (s->maxf == 3) {
        ZDM_PREC(3)
    } else {
        ZDM_PREC(4)
    }
#undef ZDM_PREC
}

// face states / clamped aux into (QF, As): the solver's own buffers unless given
// k_prep with the face limiter values stored in (alpha_mode 1) or taken from (2) s->alpha_io:
// one mesh, order 2, the configured gradient
template <int MAXF>
void prep_alpha_io(zdm_solver* s, const T* Xin, const T* Ag, T* As, T* QF) {
    dim3 g((s->ncells + BLOCK - 1) / BLOCK), b(BLOCK);
    const int lim = s->opt.use_limiters;
    if (s->opt.gradient == 1) {
        if (s->alpha_mode == 1)
            k_prep_alpha<MAXF, 2, 1><<<g, b, 0, s->stream>>>(s->dm, s->P, Xin, Ag, As, QF, lim, s->alpha_io);
        else
            k_prep_alpha<MAXF, 2, 2><<<g, b, 0, s->stream>>>(s->dm, s->P, Xin, Ag, As, QF, lim, s->alpha_io);
    } else {
        if (s->alpha_mode == 1)
            k_prep_alpha<MAXF, 1, 1><<<g, b, 0, s->stream>>>(s->dm, s->P, Xin, Ag, As, QF, lim, s->alpha_io);
        else
            k_prep_alpha<MAXF, 1, 2><<<g, b, 0, s->stream>>>(s->dm, s->P, Xin, Ag, As, QF, lim, s->alpha_io);
    }
}

void prep(zdm_solver* s, const Ctrl* ctrl, const T* Xin, const T* Ag, T* QF = nullptr, T* As = nullptr) {
    QF = QF ? QF : s->QF;
    As = As ? As : s->As;
    if (s->refresh) {
        // FormRHS's UpdateState: the derivative aux of this stage's state first (from the
        // unclamped state, TransportStep.hpp:202), then the per-cell clamp + aux in k_prep
        if (hipMemcpyAsync(s->Aref, Ag, sizeof(T) * (size_t)NA * s->ld, hipMemcpyDeviceToDevice, s->stream) !=
                hipSuccess ||
            refresh_aux(s, Xin, s->Aref)) {
            s->step_err = 1;
            return;
        }
        Ag = s->Aref;
    }
    if (s->alpha_mode) {
        // the limiter values of this stage stored in (1) or taken from (2) alpha_io
        if (s->maxf <= 2)
            prep_alpha_io<2>(s, Xin, Ag, As, QF);
        else if (s->maxf == 3)
            prep_alpha_io<3>(s, Xin, Ag, As, QF);
        else
            prep_alpha_io<4>(s, Xin, Ag, As, QF);
        s->alpha_io += (size_t)(s->maxf <= 2 ? 2 : s->maxf) * s->ld;
        return;
    }
    if (s->maxf <= 2)
        launch_prep<2>(s, ctrl, Xin, Ag, As, QF);
    else if (s->maxf == 3)
        launch_prep<3>(s, ctrl, Xin, Ag, As, QF);
    else
        launch_prep<4>(s, ctrl, Xin, Ag, As, QF);
}

void flux(zdm_solver* s, const Ctrl* ctrl, int mode, int stage, const T* In, T* Out, const T* Ag,
          const T* QF = nullptr, const T* As = nullptr, const T* X0 = nullptr) {
    QF = QF ? QF : s->QF;
    As = As ? As : s->As;
    X0 = X0 ? X0 : Out;
    if (s->maxf <= 2)
        launch_flux<2>(s, ctrl, mode, stage, In, Out, Ag, QF, As, X0);
    else if (s->maxf == 3)
        launch_flux<3>(s, ctrl, mode, stage, In, Out, Ag, QF, As, X0);
    else
        launch_flux<4>(s, ctrl, mode, stage, In, Out, Ag, QF, As, X0);
}

// post-stage with the MOOD override (positivity "mood"): reads the saved state s->Xs
void post_mood(zdm_solver* s) {
    dim3 g(s->nblocks), b(s->bs);
    const bool src = s->opt.source && s->opt.time_integration == 0;   // IMEX: no split source
#define ZDM_MOOD1(MF, MC)                                                                                            \
    if (src)                                                                                                         \
        k_post_mood<MF, 1, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl, s->X, s->A, s->Xs, s->As, s->partials); \
    else                                                                                                             \
        k_post_mood<MF, 0, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl, s->X, s->A, s->Xs, s->As, s->partials);
#define ZDM_MOOD(MF)                                                                                                 \
    if (s->dm.cpt > 1) {                                                                                             \
        ZDM_MOOD1(MF, true)                                                                                          \
    } else {                                                                                                         \
        ZDM_MOOD1(MF, false)                                                                                         \
    }
    if (s->maxf <= 2) {
        ZDM_MOOD(2)
    } else if (s->maxf == 3) {
        ZDM_MOOD(3)
    } else {
        ZDM_MOOD(4)
    }
#undef ZDM_MOOD
#undef ZDM_MOOD1
}

template <bool MC>
void post_t(zdm_solver* s, int mode, T dt_override) {
    dim3 g(s->nblocks), b(s->bs);
    switch (mode) {
        case 5: k_post<5, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl, s->X, s->A, s->partials, dt_override); return;
        case 6: k_post<4, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl, s->X, s->A, s->partials, dt_override); return;
        case 7: k_post<6, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl, s->X, s->A, s->partials, dt_override); return;
        case 0: k_post<0, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl, s->X, s->A, s->partials, dt_override); break;
        case 1: k_post<1, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl, s->X, s->A, s->partials, dt_override); break;
        case 2: k_post<2, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl, s->X, s->A, s->partials, dt_override); break;
        case 3: k_post<3, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl_rhs, s->X, s->A, s->partials, dt_override); break;
        default: k_post<4, MC><<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl_rhs, s->X, s->A, s->partials, dt_override); break;
    }
}
void post(zdm_solver* s, int mode, T dt_override = 0) {
    if (s->dm.cpt > 1)
        post_t<true>(s, mode, dt_override);
    else
        post_t<false>(s, mode, dt_override);
}

// The final SSP stage fused with the post-stage of mode 1 (constraints + CFL) or 2 (constraints,
// source, constraints + CFL): k_flux_post from the stage state W (face states QF, clamped aux As)
// and the step's start state X0 into Xout, A.  rec (the adjoint's recording step, else null):
// X1, X3, A3 and dt_log as k_post_record keeps them.
struct FluxPostRec {
    T *X1, *X3, *A3, *dt_log;
};
template <int MAXF, bool MC>
void flux_post_t(zdm_solver* s, const Ctrl* ctrl, int mode, const T* QF, const T* As, const T* W, const T* X0,
                 T* Xout, T* A, const FluxPostRec* rec) {
    dim3 g(s->nblocks), b(s->bs);
    DevMesh dmr = s->dm;
    dmr.rev = 1;   // against k_prep's direction: start on the face states it wrote last
    const int stage = s->opt.ssp_stages - 1;
    const T* G = grad_buffer(s);
#define ZDM_FP(PO, RE, NT)                                                                                    \
    k_flux_post<MAXF, PO, MC, RE, NT><<<g, b, 0, s->stream>>>(dmr, s->P, ctrl, stage, QF, As, W, X0, Xout, A,    \
                                                              s->partials, G, rec ? rec->X1 : nullptr,         \
                                                              rec ? rec->X3 : nullptr, rec ? rec->A3 : nullptr, \
                                                              rec ? rec->dt_log : nullptr)
    // the non-temporal form for the one-cell-per-thread shape only (what a large mesh autotunes to)
    const bool nt = !MC && s->flux_nt;
    if (mode == 2) {
        if (rec) {
            if (nt)
                ZDM_FP(2, true, !MC);
            else
                ZDM_FP(2, true, false);
        } else {
            if (nt)
                ZDM_FP(2, false, !MC);
            else
                ZDM_FP(2, false, false);
        }
    } else {
        if (rec) {
            if (nt)
                ZDM_FP(1, true, !MC);
            else
                ZDM_FP(1, true, false);
        } else {
            if (nt)
                ZDM_FP(1, false, !MC);
            else
                ZDM_FP(1, false, false);
        }
    }
#undef ZDM_FP
}
template <int MAXF>
void flux_post_m(zdm_solver* s, const Ctrl* ctrl, int mode, const T* QF, const T* As, const T* W, const T* X0,
                 T* Xout, T* A, const FluxPostRec* rec) {
    if (s->dm.cpt > 1)
        flux_post_t<MAXF, true>(s, ctrl, mode, QF, As, W, X0, Xout, A, rec);
    else
        flux_post_t<MAXF, false>(s, ctrl, mode, QF, As, W, X0, Xout, A, rec);
}
void flux_post(zdm_solver* s, const Ctrl* ctrl, int mode, const T* QF, const T* As, const T* W, const T* X0,
               T* Xout, T* A, const FluxPostRec* rec = nullptr) {
    if (s->maxf <= 2)
        flux_post_m<2>(s, ctrl, mode, QF, As, W, X0, Xout, A, rec);
    else if (s->maxf == 3)
        flux_post_m<3>(s, ctrl, mode, QF, As, W, X0, Xout, A, rec);
    else
        flux_post_m<4>(s, ctrl, mode, QF, As, W, X0, Xout, A, rec);
}

void finalize(zdm_solver* s, int what) {
    if (s->dm.nmem == 0) {
        // xmin[NXCD] and the arrival counter follow the partials (zeroed with the arena)
        T* xmin = s->partials + s->maxblocks;
        k_finalize_xcd<<<NXCD, FBLOCK, 0, s->stream>>>(s->ctrl, s->partials, s->nblocks, xmin,
                                                       reinterpret_cast<unsigned*>(xmin + NXCD), what,
                                                       s->opt.order, s->opt.ssp_stages, s->opt.cfl,
                                                       s->opt.min_dt, s->opt.t_end, (T)M::dt_max, s->dt_log);
        return;
    }
    k_finalize<<<s->nctrl, FBLOCK, 0, s->stream>>>(s->ctrl, s->partials, s->nblocks, s->dm.blk_first,
                                                 what, s->opt.order, s->opt.ssp_stages, s->opt.cfl,
                                                 s->opt.min_dt, s->opt.t_end, (T)M::dt_max);
}

#ifdef ZOOMY_MODEL_IS_VAM
#include "chorin_host.inc"
#else
struct ChorinHost {};
void chorin_free(ChorinHost*) {}
#endif

int set_ctrl_rhs_time(zdm_solver* s, T t) {
    std::vector<Ctrl> cs(s->nctrl);
    for (Ctrl& c : cs) {
        std::memset(&c, 0, sizeof(c));
        c.active = 1;
        c.t = t;
        c.stage_t[0] = t;
    }
    HIPCHK(hipMemcpyAsync(s->ctrl_rhs, cs.data(), sizeof(Ctrl) * cs.size(), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

#include "implicit_host.inc"

// all control blocks (one per member) to the host
int read_ctrls(zdm_solver* s, std::vector<Ctrl>& cs) {
    cs.resize(s->nctrl);
    HIPCHK(hipMemcpyAsync(cs.data(), s->ctrl, sizeof(Ctrl) * cs.size(), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

}  // namespace

// Scratch of zdm_bsr2_gmres_noalloc: device memory the library holds on every GPU it runs on
// until the process ends, so that entry point allocates nothing.  Host code never uses the
// name `scratch` as an address (on the host it is only a stand-in for the device variable):
// scratch_on_current_device() resolves this GPU's copy once, and only that pointer reaches
// kernels and copies.
constexpr size_t ZDM_SCRATCH_BYTES = size_t(64) << 20;   // 64 MiB per GPU
static __device__ __align__(256) double scratch[ZDM_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int ZDM_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[ZDM_MAX_DEVICES] = {};

// this GPU's address of `scratch` (resolved on the first call with the device current)
int scratch_on_current_device(char** out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= ZDM_MAX_DEVICES) return fail("scratch: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail("scratch: no device address");
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// flag[0] = 1 unless rowptr[0] == 0 and rowptr never decreases
__global__ void k_bsr2_check_rows(int nb, const int* __restrict__ rowptr, int* __restrict__ flag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nb) return;
    const int lo = rowptr[c], hi = rowptr[c + 1];
    if (hi < lo || (c == 0 && lo != 0)) flag[0] = 1;
}

// flag[0] = 1 if a block column lies outside [0, nb)
__global__ void k_bsr2_check_cols(int nb, int nnzb, const int* __restrict__ col, int* __restrict__ flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnzb) return;
    const int j = col[k];
    if (j < 0 || j >= nb) flag[0] = 1;
}

// GMRES + point-block Jacobi on a 2x2-block CSR matrix whose arrays, b and x are on the device;
// Dinv (4 per block row) and kmem (gmres_bytes) are the caller's workspace.  The workspace is
// cleared by a fill kernel, so every call starts from the same state.
int bsr2_gmres_on_device(const Bsr2& A, const T* b, T* x, T* Dinv, char* kmem, const zdm_krylov_options& o,
                         hipStream_t st, const char* who, int* its, int* reason) {
    const size_t n = 2 * (size_t)A.nb;
    const size_t kd = gmres_bytes(o.restart, n) / sizeof(T);
    k_fill<<<(unsigned)((kd + 255) / 256), 256, 0, st>>>(reinterpret_cast<T*>(kmem), kd, 0.0);
    const unsigned cg = (unsigned)((A.nb + 255) / 256);
    k_pbj_from_bsr<<<cg, 256, 0, st>>>(A, Dinv);
    Gmres g;
    gmres_bind(g, kmem, o.restart, n);
    g.max_it = o.max_it;
    g.rtol = o.rtol;
    g.atol = o.atol;
    g.dtol = o.dtol;
    auto op = [&](const T* xx, T* y, const int* h) { k_bsr2_spmv<<<cg, KBLOCK, 0, st>>>(A, Dinv, xx, y, 0, h); };
    auto op_pc = [&](const T* xx, T* y, const int* h) { k_bsr2_spmv<<<cg, KBLOCK, 0, st>>>(A, Dinv, xx, y, 1, h); };
    auto pc = [&](const T* bb, const T* t, T* z) { k_pbj_residual<<<cg, 256, 0, st>>>(A.nb, Dinv, bb, t, z); };
    const int r = gmres_solve(g, st, op_pc, op, pc, b, x);
    if (r == -100 || hipGetLastError() != hipSuccess) return fail(std::string(who) + ": device error in the solve");
    if (its) *its = g.last_its;
    if (reason) *reason = g.last_reason;
    return 0;
}

}  // namespace

extern "C" {

const char* zdm_last_error(void) { return g_err.c_str(); }
const char* zdm_model_name(void) { return ZDM_MODEL_NAME; }

int zdm_model_info(int* n_q, int* n_aux, int* n_params, int* dim) {
    if (n_q) *n_q = NQ;
    if (n_aux) *n_aux = NA;
    if (n_params) *n_params = NPAR;
    if (dim) *dim = DIM;
    return 0;
}

int zdm_model_default_params(double* params) {
    auto v = M::default_parameters();
    for (size_t i = 0; i < v.size(); ++i) params[i] = v[i];
    return 0;
}

int zdm_model_boundary_tags(char* names64, int cap) {
    auto v = M::get_boundary_tags();
    for (int i = 0; i < (int)v.size() && i < cap; ++i) {
        std::memset(names64 + 64 * i, 0, 64);
        std::strncpy(names64 + 64 * i, v[i].c_str(), 63);
    }
    return (int)v.size();
}

int zdm_model_initial_condition(const double* X, const double* params, double* Q) {
    if (!X || !params || !Q) return fail("zdm_model_initial_condition: null argument");
    return model_initial_condition<T>(X, params, Q);
}

int zdm_model_parameter_names(char* names64, int cap) {
    auto v = M::parameter_names();
    for (int i = 0; i < (int)v.size() && i < cap; ++i) {
        std::memset(names64 + 64 * i, 0, 64);
        std::strncpy(names64 + 64 * i, v[i].c_str(), 63);
    }
    return (int)v.size();
}

// By size: the non-temporal flux form where one flux pass (48 n + 48 B per cell and 40 B per
// face) is larger than the 256 MB last-level cache.  Measured at C2 (2 M cells, 0.98 GB per
// pass): 2591 - 2597 -> 2616 - 2620 Mcell-steps/s; at C5 (250 k cells, 0.10 GB) the forward
// was 4 % slower with it.
static void choose_flux_nt(zdm_solver* s) {
    if (s->flux_nt_mode >= 0) {
        s->flux_nt = s->flux_nt_mode != 0;
        return;
    }
    const double pass = (double)s->ncells * (48.0 * NQ + 48.0) + 40.0 * (double)s->nfaces;
    s->flux_nt = pass > 256.0 * 1024.0 * 1024.0;
}

static int create_impl(const zdm_mesh* mesh, const int* face_bc, const double* params,
                       const zdm_options* opts, int nmem, const int* cell_off, zdm_solver** out) {
    if (!mesh || !face_bc || !opts || !out) return fail("zdm_solver_create: null argument");
    if (mesh->dim != DIM) return fail("mesh dimension does not match Model::dimension");
    if (mesh->max_faces > 4 || mesh->max_faces < 1) return fail("unsupported faces per cell");
    if (opts->order >= 2 && opts->ssp_stages < 2) return fail("ssp_stages must be >= 2");
    if (opts->ssp_stages > 16) return fail("ssp_stages must be <= 16");
    if (opts->block_size != 0 && opts->block_size != BLOCK) return fail("block_size must be 256");
    if (opts->well_balanced && !HasCellSource<T>::value)
        return fail("well_balanced needs Model::hydrostatic_cell_source");
    if (opts->positivity < 0 || opts->positivity > 2) return fail("positivity must be 0 (none), 1 (mood) or 2 (zhang_shu)");
    if (opts->positivity != 0 && opts->order >= 2 && opts->gradient != 0)
        return fail("positivity needs the Green-Gauss gradient (ModularSolver.hpp:62)");
    if (opts->positivity != 0 && opts->well_balanced)
        return fail("positivity and the well_balanced extension cannot be combined");
    if (opts->positivity != 0 && NQ < 2) return fail("positivity needs the [b, h, ...] state layout");
    if (opts->time_integration < 0 || opts->time_integration > 2)
        return fail("time_integration must be 0 (splitting), 1 (imex) or 2 (implicit)");
    if (opts->time_integration != 0 && nmem > 0)
        return fail("time_integration imex / implicit: not available for an ensemble");
    if (opts->time_integration != 0 && NQ < 2)
        return fail("time_integration imex / implicit needs the [b, h, ...] state layout");
    // RefreshDerivativeAux (solver.refresh_derivative_aux) runs for a model with the
    // field-level update_aux_variables in the splitting step (every stage's FormRHS and every
    // EnforcePhysicalConstraints); the fused MOOD override and the IMEX / implicit steps do
    // not carry it
    if (HasDiffusiveFlux<T>::value && opts->positivity != 0 && opts->order >= 2)
        return fail("a model with diffusive_flux is not available with the positivity-preserving (eta-WB) "
                    "reconstruction");
    const bool refresh = HasFieldAuxUpdate<T>::value && opts->refresh_derivative_aux != 0;
    if (refresh && opts->positivity == 1)
        return fail("refresh_derivative_aux with a field-level update_aux_variables is not available with "
                    "positivity \"mood\"; set solver.refresh_derivative_aux to false");
    if (refresh && opts->time_integration != 0)
        return fail("refresh_derivative_aux with a field-level update_aux_variables is available for the "
                    "splitting time integration only; set solver.refresh_derivative_aux to false");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device available");

    zdm_solver* s = new zdm_solver();
    std::memset(&s->P, 0, sizeof(Params));
    s->opt = *opts;
    s->cfl_in = opts->cfl;
    if (s->opt.wet_dry_eps <= 0) s->opt.wet_dry_eps = 1.0e-2;
    // Zhang-Shu positivity holds for CFL <= 1/6 (MUSCLSolver.hpp:31-38)
    if (s->opt.positivity == 2 && s->opt.order >= 2 && s->opt.cfl > 1.0 / 6.0) s->opt.cfl = 1.0 / 6.0;
    auto dp = M::default_parameters();
    for (int i = 0; i < NPAR; ++i) s->P.p[i] = params ? params[i] : dp[i];
    const int nc = mesh->ncells, nf = mesh->nfaces, MF = mesh->max_faces;
    s->ncells = nc;
    s->nfaces = nf;
    choose_flux_nt(s);
    s->maxf = MF;
    s->ld = (int)align_up((size_t)nc, 256);
    s->nmem = nmem;
    s->nctrl = nmem > 0 ? nmem : 1;
    s->ncb = (nc + BLOCK - 1) / BLOCK;
    s->bs = BLOCK;
    s->dm.cpt = 1;
    std::vector<int> blk_member, blk_first, coff;
    if (nmem > 0) coff.assign(cell_off, cell_off + nmem + 1);
    s->coff_h = coff;
    s->maxblocks = launch_tables(nc, coff, 64, 1, blk_member, blk_first);
    s->nblocks = launch_tables(nc, coff, BLOCK, 1, blk_member, blk_first);
    const int ld = s->ld;
    const int MFa = MF <= 2 ? 2 : MF;

    // ---- host-side geometry tables ----
    std::vector<T> vol(ld, 0), cx(ld, 0), cy(ld, 0), rc(ld, DBL_MAX);
    std::vector<int> cface((size_t)MFa * ld, -1), cnbr((size_t)MFa * ld, -1), cmeta((size_t)MFa * ld, 0);
    std::vector<T> fNx(nf), fNy(nf), fcx(nf), fcy(nf), farea(nf), fnhx(nf), fnhy(nf);
    std::vector<int> fbc(nf);
    std::vector<T> lsqw;
    if (nc >= (1 << 28)) {
        delete s;
        return fail("mesh too large: the packed neighbour index holds fewer than 2^28 cells");
    }
    std::vector<int> ctopo((size_t)2 * MFa * ld, -1);
    std::vector<T> fgeo((size_t)3 * std::max(nf, 1), 0.0);
    std::vector<double2> sgeo((size_t)MFa * 2 * ld, make_double2(0.0, 0.0));
    for (int f = 0; f < nf; ++f) {
        const double* N = mesh->face_normal + 3 * (size_t)f;
        fNx[f] = N[0];
        fNy[f] = DIM > 1 ? N[1] : 0;
        fcx[f] = mesh->face_centroid[3 * (size_t)f];
        fcy[f] = DIM > 1 ? mesh->face_centroid[3 * (size_t)f + 1] : 0;
        T area = 0;
        for (int d = 0; d < DIM; ++d) area += N[d] * N[d];
        area = std::sqrt(area);
        farea[f] = area;
        fnhx[f] = N[0] / area;
        fnhy[f] = DIM > 1 ? N[1] / area : 0;
        fbc[f] = mesh->face_cells[2 * f + 1] >= 0 ? -1 : face_bc[f];
        fgeo[3 * (size_t)f] = fnhx[f];
        fgeo[3 * (size_t)f + 1] = fnhy[f];
        fgeo[3 * (size_t)f + 2] = area;
    }
    for (int c = 0; c < nc; ++c) {
        vol[c] = mesh->cell_volume[c];
        cx[c] = mesh->cell_centroid[3 * (size_t)c];
        cy[c] = DIM > 1 ? mesh->cell_centroid[3 * (size_t)c + 1] : 0;
        T r = DBL_MAX;
        for (int k = 0; k < mesh->cell_nfaces[c]; ++k) {
            const int f = mesh->cell_faces[(size_t)c * MF + k];
            const int nb = mesh->cell_nbrs[(size_t)c * MF + k];
            cface[(size_t)k * ld + c] = f;
            cnbr[(size_t)k * ld + c] = nb;
            int meta = (mesh->face_cells[2 * f] == c) ? 0 : 4;
            if (nb >= 0) {
                int ks = -1;
                for (int j = 0; j < mesh->cell_nfaces[nb]; ++j)
                    if (mesh->cell_faces[(size_t)nb * MF + j] == f) ks = j;
                if (ks < 0) {
                    delete s;
                    return fail("inconsistent mesh: face missing from its neighbour");
                }
                meta |= ks;
            }
            cmeta[(size_t)k * ld + c] = meta;
            ctopo[2 * ((size_t)k * ld + c)] = f;
            ctopo[2 * ((size_t)k * ld + c) + 1] = nb >= 0 ? (nb | (meta << 28)) : -1;
            const T sgn = (meta & 4) ? (T)-1 : (T)1;
            sgeo[sgeo_idx(ld, k, 0, c)] = make_double2(sgn * fNx[f], sgn * fNy[f]);
            sgeo[sgeo_idx(ld, k, 1, c)] = make_double2(fcx[f] - cx[c], fcy[f] - cy[c]);
            const double* xf = mesh->face_centroid + 3 * (size_t)f;
            T d2 = 0.0;
            for (int d = 0; d < DIM; ++d) {
                T dd = mesh->cell_centroid[3 * (size_t)c + d] - xf[d];
                d2 += dd * dd;
            }
            T dist = std::sqrt(d2);
            if (dist < r) r = dist;
        }
        rc[c] = r;
    }
    if (opts->gradient == 1) {
        lsqw.assign((size_t)MFa * DIM * ld, 0.0);
        for (int c = 0; c < nc; ++c) {
            const double* xc = mesh->cell_centroid + 3 * (size_t)c;
            std::vector<int> slots;
            std::vector<std::vector<T>> bv;
            for (int k = 0; k < mesh->cell_nfaces[c]; ++k) {
                const int nb = mesh->cell_nbrs[(size_t)c * MF + k];
                if (nb < 0) continue;
                const double* xn = mesh->cell_centroid + 3 * (size_t)nb;
                std::vector<T> b(DIM);
                for (int d = 0; d < DIM; ++d) b[d] = xn[d] - xc[d];
                slots.push_back(k);
                bv.push_back(b);
            }
            std::vector<T> ata(DIM * DIM, 0.0);
            for (const auto& b : bv)
                for (int i = 0; i < DIM; ++i)
                    for (int j = 0; j < DIM; ++j) ata[i * DIM + j] += b[i] * b[j];
            for (size_t j = 0; j < slots.size(); ++j) {
                std::vector<T> w = bv[j], sys = ata;
                if (!chol_solve_host(DIM, sys, w)) std::fill(w.begin(), w.end(), 0.0);
                for (int d = 0; d < DIM; ++d) lsqw[((size_t)slots[j] * DIM + d) * ld + c] = w[d];
            }
        }
    }

    // ---- one device arena ----
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const size_t sQ = sizeof(T) * (size_t)NQ * ld, sA = sizeof(T) * (size_t)NA * ld;
    size_t oX = take(sQ), oW = take(sQ), oF = take(sQ), oA = take(sA), oAs = take(sA), oAt = take(sA);
    size_t oQF = take(sizeof(T) * (size_t)MFa * NQP * ld);
    size_t oXs = take(s->opt.positivity == 1 ? sQ : 8);   // MOOD: the saved start-of-step state
    size_t oAref = take(refresh ? sA : 8);                  // the refreshed aux of a transport stage
    size_t oG = take(HasDiffusiveFlux<T>::value ? sizeof(T) * (size_t)NQ * DIM * ld : 8);
    size_t oIm = take(s->opt.time_integration == 1 ? 4 * sQ : 8);   // IMEX: G, stage state, P3, F
    size_t oPart = take(sizeof(T) * ((size_t)s->maxblocks + NXCD + 1));   // + k_finalize_xcd's minima, counter
    size_t oStage = take(sizeof(T) * (size_t)std::max(NQ, NA) * ld);
    size_t oCtrl = take(sizeof(Ctrl) * s->nctrl), oCtrlR = take(sizeof(Ctrl) * s->nctrl);
    size_t oBm = take(sizeof(int) * std::max<size_t>(nmem > 0 ? (size_t)s->maxblocks : 0, 1)),
           oBf = take(sizeof(int) * std::max<size_t>(blk_first.size(), 1)),
           oCo = take(sizeof(int) * std::max<size_t>(coff.size(), 1));
    size_t oVol = take(sizeof(T) * ld), oCx = take(sizeof(T) * ld), oCy = take(sizeof(T) * ld), oRc = take(sizeof(T) * ld);
    size_t oCface = take(sizeof(int) * (size_t)MFa * ld), oCnbr = take(sizeof(int) * (size_t)MFa * ld),
           oCmeta = take(sizeof(int) * (size_t)MFa * ld);
    size_t oF7[7];
    for (int i = 0; i < 7; ++i) oF7[i] = take(sizeof(T) * (size_t)std::max(nf, 1));
    size_t oFbc = take(sizeof(int) * (size_t)std::max(nf, 1));
    size_t oCtopo = take(sizeof(int) * ctopo.size()), oFgeo = take(sizeof(T) * fgeo.size());
    size_t oSgeo = take(sizeof(double2) * sgeo.size());
    size_t oLsq = take(sizeof(T) * (lsqw.empty() ? 1 : lsqw.size()));
    s->arena_bytes = off;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return fail("hipStreamCreate failed");
    }
    if (hipMalloc(&s->arena, s->arena_bytes) != hipSuccess) {
        (void)hipStreamDestroy(s->stream);
        delete s;
        return fail("hipMalloc of the solver arena failed");
    }
    char* a = s->arena;
    s->X = (T*)(a + oX);
    s->Xs = (T*)(a + oXs);
    s->Aref = (T*)(a + oAref);
    s->G = (T*)(a + oG);
    s->refresh = refresh;
    s->Im = (T*)(a + oIm);
    s->W = (T*)(a + oW);
    s->F = (T*)(a + oF);
    s->A = (T*)(a + oA);
    s->As = (T*)(a + oAs);
    s->Atmp = (T*)(a + oAt);
    s->QF = (T*)(a + oQF);
    s->partials = (T*)(a + oPart);
    s->stage_aos = (T*)(a + oStage);
    s->ctrl = (Ctrl*)(a + oCtrl);
    s->ctrl_rhs = (Ctrl*)(a + oCtrlR);
    auto up = [&](size_t o, const void* src, size_t bytes) {
        return hipMemcpy(a + o, src, bytes, hipMemcpyHostToDevice);
    };
    hipError_t e = hipSuccess;
    e = e != hipSuccess ? e : hipMemset(a, 0, s->arena_bytes);
    e = e != hipSuccess ? e : up(oVol, vol.data(), sizeof(T) * ld);
    e = e != hipSuccess ? e : up(oCx, cx.data(), sizeof(T) * ld);
    e = e != hipSuccess ? e : up(oCy, cy.data(), sizeof(T) * ld);
    e = e != hipSuccess ? e : up(oRc, rc.data(), sizeof(T) * ld);
    e = e != hipSuccess ? e : up(oCface, cface.data(), sizeof(int) * cface.size());
    e = e != hipSuccess ? e : up(oCnbr, cnbr.data(), sizeof(int) * cnbr.size());
    e = e != hipSuccess ? e : up(oCmeta, cmeta.data(), sizeof(int) * cmeta.size());
    const std::vector<T>* fa[7] = {&fNx, &fNy, &fcx, &fcy, &farea, &fnhx, &fnhy};
    for (int i = 0; i < 7 && nf > 0; ++i) e = e != hipSuccess ? e : up(oF7[i], fa[i]->data(), sizeof(T) * nf);
    if (nf > 0) e = e != hipSuccess ? e : up(oFbc, fbc.data(), sizeof(int) * nf);
    e = e != hipSuccess ? e : up(oCtopo, ctopo.data(), sizeof(int) * ctopo.size());
    e = e != hipSuccess ? e : up(oFgeo, fgeo.data(), sizeof(T) * fgeo.size());
    e = e != hipSuccess ? e : up(oSgeo, sgeo.data(), sizeof(double2) * sgeo.size());
    if (!lsqw.empty()) e = e != hipSuccess ? e : up(oLsq, lsqw.data(), sizeof(T) * lsqw.size());
    if (nmem > 0) {
        e = e != hipSuccess ? e : up(oBm, blk_member.data(), sizeof(int) * blk_member.size());
        e = e != hipSuccess ? e : up(oBf, blk_first.data(), sizeof(int) * blk_first.size());
        e = e != hipSuccess ? e : up(oCo, coff.data(), sizeof(int) * coff.size());
    }
    if (e != hipSuccess) {
        (void)hipFree(s->arena);
        (void)hipStreamDestroy(s->stream);
        delete s;
        return fail(std::string("device upload failed: ") + hipGetErrorString(e));
    }
    DevMesh& d = s->dm;
    d.ncells = nc;
    d.nfaces = nf;
    d.ld = ld;
    d.vol = (T*)(a + oVol);
    d.cx = (T*)(a + oCx);
    d.cy = (T*)(a + oCy);
    d.rc = (T*)(a + oRc);
    d.cface = (int*)(a + oCface);
    d.cnbr = (int*)(a + oCnbr);
    d.cmeta = (int*)(a + oCmeta);
    d.fNx = (T*)(a + oF7[0]);
    d.fNy = (T*)(a + oF7[1]);
    d.fcx = (T*)(a + oF7[2]);
    d.fcy = (T*)(a + oF7[3]);
    d.farea = (T*)(a + oF7[4]);
    d.fnhx = (T*)(a + oF7[5]);
    d.fnhy = (T*)(a + oF7[6]);
    d.fbc = (int*)(a + oFbc);
    d.ctopo = (int*)(a + oCtopo);
    d.fgeo = (T*)(a + oFgeo);
    d.sgeo = (const double2*)(a + oSgeo);
    d.lsqw = (T*)(a + oLsq);
    d.blk_member = (int*)(a + oBm);
    d.blk_first = (int*)(a + oBf);
    d.cell_off = (int*)(a + oCo);
    d.nmem = nmem;
    d.rev = 0;
    s->timing = false;
    s->timing_stride = 0;
    s->timing_step = 0;
    s->evused = 0;
    for (int i = 0; i < 4; ++i) {
        s->kt_ms[i] = 0;
        s->kt_n[i] = 0;
    }
    std::vector<Ctrl> c0(s->nctrl);
    std::memset(c0.data(), 0, sizeof(Ctrl) * c0.size());
    if (hipMemcpy(s->ctrl, c0.data(), sizeof(Ctrl) * c0.size(), hipMemcpyHostToDevice) != hipSuccess) {
        zdm_solver_destroy(s);
        return fail("device upload of the control blocks failed");
    }
#ifdef ZOOMY_MODEL_IS_VAM
    if (nmem == 0 && chorin_build(s, mesh, krylov_defaults())) {
        const std::string msg = g_err;
        zdm_solver_destroy(s);
        return fail(msg);
    }
#endif
    *out = s;
    return 0;
}

int zdm_solver_create(const zdm_mesh* mesh, const int* face_bc, const double* params,
                      const zdm_options* opts, zdm_solver** out) {
    return create_impl(mesh, face_bc, params, opts, 0, nullptr, out);
}

int zdm_batch_create(const zdm_mesh* mesh, const int* face_bc, const double* params,
                     const zdm_options* opts, int nmembers, const int* cell_offsets, zdm_solver** out) {
    if (!mesh || nmembers < 1 || !cell_offsets) return fail("zdm_batch_create: bad arguments");
    if (cell_offsets[0] != 0 || cell_offsets[nmembers] != mesh->ncells)
        return fail("zdm_batch_create: cell offsets must run from 0 to ncells");
    for (int i = 0; i < nmembers; ++i)
        if (cell_offsets[i + 1] <= cell_offsets[i]) return fail("zdm_batch_create: every member needs cells");
    // members must not share faces: every neighbour of a member cell is in the member
    const int MF = mesh->max_faces;
    for (int i = 0; i < nmembers; ++i)
        for (int c = cell_offsets[i]; c < cell_offsets[i + 1]; ++c)
            for (int k = 0; k < mesh->cell_nfaces[c]; ++k) {
                const int nb = mesh->cell_nbrs[(size_t)c * MF + k];
                if (nb >= 0 && (nb < cell_offsets[i] || nb >= cell_offsets[i + 1]))
                    return fail("zdm_batch_create: members are coupled by a face");
            }
    return create_impl(mesh, face_bc, params, opts, nmembers, cell_offsets, out);
}

int zdm_solver_destroy(zdm_solver* s) {
    if (!s) return 0;
    // teardown: errors left by earlier work are not this call's to report
    (void)hipStreamSynchronize(s->stream);
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (hipEvent_t e : s->evpool) (void)hipEventDestroy(e);
    if (s->ev_caller) (void)hipEventDestroy(s->ev_caller);
    chorin_free(s->ch);
    adjoint_free(s->adj);
    implicit_free(s->imp);
    (void)hipFree(s->arena);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return 0;
}

static int upload_soa(zdm_solver* s, const double* host_aos, T* dst, int ncomp, bool from_device) {
    const size_t bytes = sizeof(T) * (size_t)s->ncells * ncomp;
    if (from_device)
        HIPCHK(hipMemcpyAsync(s->stage_aos, host_aos, bytes, hipMemcpyDeviceToDevice, s->stream));
    else
        HIPCHK(hipMemcpyAsync(s->stage_aos, host_aos, bytes, hipMemcpyHostToDevice, s->stream));
    k_aos_to_soa<<<(s->ncells + 255) / 256, 256, 0, s->stream>>>(s->stage_aos, dst, s->ncells, ncomp, s->ld);
    HIPCHK(hipGetLastError());
    if (!from_device) HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

static int download_soa(zdm_solver* s, const T* src, double* host_aos, int ncomp, bool to_device) {
    const size_t bytes = sizeof(T) * (size_t)s->ncells * ncomp;
    k_soa_to_aos<<<(s->ncells + 255) / 256, 256, 0, s->stream>>>(src, s->stage_aos, s->ncells, ncomp, s->ld);
    HIPCHK(hipGetLastError());
    if (to_device)
        HIPCHK(hipMemcpyAsync(host_aos, s->stage_aos, bytes, hipMemcpyDeviceToDevice, s->stream));
    else {
        HIPCHK(hipMemcpyAsync(host_aos, s->stage_aos, bytes, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return 0;
}

static int fill_zero(zdm_solver* s, T* dst, size_t n) {
    k_fill<<<(unsigned)((n + 255) / 256), 256, 0, s->stream>>>(dst, n, 0.0);
    HIPCHK(hipGetLastError());
    return 0;
}

int zdm_state_upload(zdm_solver* s, const double* X, const double* A) {
    if (!s || !X) return fail("zdm_state_upload: null argument");
    if (upload_soa(s, X, s->X, NQ, false)) return -1;
    if (A) return upload_soa(s, A, s->A, NA, false);
    if (fill_zero(s, s->A, (size_t)NA * s->ld)) return -1;
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int zdm_state_download(zdm_solver* s, double* X, double* A) {
    if (!s) return fail("zdm_state_download: null argument");
    if (X && download_soa(s, s->X, X, NQ, false)) return -1;
    if (A && download_soa(s, s->A, A, NA, false)) return -1;
    return 0;
}

int zdm_state_upload_device(zdm_solver* s, const double* dX, const double* dA) {
    if (!s || !dX) return fail("zdm_state_upload_device: null argument");
    if (upload_soa(s, dX, s->X, NQ, true)) return -1;
    if (dA) return upload_soa(s, dA, s->A, NA, true);
    return fill_zero(s, s->A, (size_t)NA * s->ld);
}

int zdm_state_download_device(zdm_solver* s, double* dX, double* dA) {
    if (!s) return fail("zdm_state_download_device: null argument");
    if (dX && download_soa(s, s->X, dX, NQ, true)) return -1;
    if (dA && download_soa(s, s->A, dA, NA, true)) return -1;
    return 0;
}

int zdm_initialize(zdm_solver* s) {
    if (!s) return fail("null solver");
    std::vector<Ctrl> c0(s->nctrl);
    std::memset(c0.data(), 0, sizeof(Ctrl) * c0.size());
    for (Ctrl& c : c0) c.active = 1;
    HIPCHK(hipMemcpyAsync(s->ctrl, c0.data(), sizeof(Ctrl) * c0.size(), hipMemcpyHostToDevice, s->stream));
    if (s->imp) s->imp->have_prev = false;   // BDF history restarts with BDF1
    if (s->refresh && refresh_aux(s, s->X, s->A)) return -1;
    post(s, 1);
    finalize(s, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

static void record(zdm_solver* s, int kind) {
    if (!s->timing) return;
    if (s->evused == s->evpool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        s->evpool.push_back(e);
        s->evkind.push_back(-1);
    }
    s->evkind[s->evused] = kind;
    (void)hipEventRecord(s->evpool[s->evused], s->stream);
    s->evused++;
}

// Fold the recorded events into per-kernel totals (synchronises the stream).
static int resolve_events(zdm_solver* s) {
    if (s->evused == 0) return 0;
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t i = 1; i < s->evused; ++i) {
        const int k = s->evkind[i];
        if (k < 0) continue;
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, s->evpool[i - 1], s->evpool[i]));
        s->kt_ms[k] += ms;
        s->kt_n[k] += 1;
    }
    s->evused = 0;
    return 0;
}

// the launches of one IMEX step (imex.inc): three transport right-hand sides, the per-cell
// implicit stages with the running sums, the last term; then constraints [with the MOOD
// override] and the CFL dt as after any step (no split source)
static void enqueue_imex_step(zdm_solver* s) {
    const size_t nQ = (size_t)NQ * s->ld;
    T* G = s->Im;
    T* Y = s->Im + nQ;
    T* P3 = s->Im + 2 * nQ;
    T* F = s->Im + 3 * nQ;
    const int src = s->opt.source;
    dim3 g(s->ncb), b(BLOCK);
    if (s->opt.positivity == 1)
        (void)hipMemcpyAsync(s->Xs, s->X, sizeof(T) * nQ, hipMemcpyDeviceToDevice, s->stream);
    record(s, -1);
    k_imex_times<<<1, 1, 0, s->stream>>>(s->ctrl);
    prep(s, s->ctrl, s->X, s->A);
    record(s, 0);
    flux(s, s->ctrl, MODE_RHS, 0, s->X, G, s->A);
    record(s, 1);
    k_imex_stage2<<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl, s->X, s->A, G, Y, P3, F, src);
    prep(s, s->ctrl, Y, s->A);
    record(s, 0);
    flux(s, s->ctrl, MODE_RHS, 1, Y, G, s->A);
    record(s, 1);
    k_imex_stage3<<<g, b, 0, s->stream>>>(s->dm, s->P, s->ctrl, s->A, G, P3, F, Y, src);
    prep(s, s->ctrl, Y, s->A);
    record(s, 0);
    flux(s, s->ctrl, MODE_IMEX3, 2, Y, s->X, s->A, nullptr, nullptr, F);   // X = F + dt b3 G3
    record(s, 1);
    if (s->opt.positivity == 1)
        post_mood(s);
    else
        post(s, 1);
    record(s, 2);
    finalize(s, 2);
}

// the launches of one MUSCLSolver time step
static void enqueue_step(zdm_solver* s) {
    if (s->opt.time_integration == 1) {
        enqueue_imex_step(s);
        return;
    }
    if (s->opt.time_integration == 2) {
        // host-driven Newton loop (synchronises): not capturable
        if (implicit_step(s)) s->step_err = 1;
        return;
    }
    const int order = s->opt.order, st = s->opt.ssp_stages;
    const bool mood = s->opt.positivity == 1;
    if (mood) (void)hipMemcpyAsync(s->Xs, s->X, sizeof(T) * (size_t)NQ * s->ld, hipMemcpyDeviceToDevice, s->stream);
    record(s, -1);
    if (order < 2) {
        prep(s, s->ctrl, s->X, s->A);
        record(s, 0);
        flux(s, s->ctrl, MODE_EULER, 0, s->X, s->X, s->A);
        record(s, 1);
    } else {
        for (int i = 0; i < st - 1; ++i) {
            const T* in = i == 0 ? s->X : s->W;
            prep(s, s->ctrl, in, s->A);
            record(s, 0);
            flux(s, s->ctrl, MODE_STAGE, i, in, s->W, s->A);
            record(s, 1);
        }
        prep(s, s->ctrl, s->W, s->A);
        record(s, 0);
        if (s->fuse_post && !mood && !s->refresh && !s->opt.well_balanced) {
            // final stage and post-step in one pass (k_flux_post): bitwise the separate kernels
            flux_post(s, s->ctrl, s->opt.source ? 2 : 1, s->QF, s->As, s->W, s->X, s->X, s->A);
            record(s, 3);
            finalize(s, 2);
            return;
        }
        flux(s, s->ctrl, MODE_FINAL, st - 1, s->W, s->X, s->A);
        record(s, 1);
    }
    if (mood) {
        post_mood(s);
    } else if (s->refresh) {
        // MUSCLSolver::Run after TSStep (MUSCLSolver.hpp:153-182) with the RefreshDerivativeAux
        // of each EnforcePhysicalConstraints: [refresh, constraints, source,] refresh,
        // constraints, CFL
        if (s->opt.source) {
            if (refresh_aux(s, s->X, s->A)) {
                s->step_err = 1;
                return;
            }
            post(s, 6);
            post(s, 7);
        }
        if (refresh_aux(s, s->X, s->A)) {
            s->step_err = 1;
            return;
        }
        post(s, 1);
    } else {
        post(s, s->opt.source ? 2 : 1);
    }
    record(s, 2);
    finalize(s, 2);
}

int zdm_step(zdm_solver* s, int n) {
    if (!s) return fail("null solver");
    for (int step = 0; step < n; ++step) {
        if (s->timing_stride > 0) s->timing = s->timing_step++ % s->timing_stride == 0;
        if (s->timing && s->evpool.size() > 4096 && s->evused + 64 > s->evpool.size())
            if (resolve_events(s)) return -1;
        enqueue_step(s);
        if (s->step_err) {
            s->step_err = 0;
            s->timing = s->timing_stride > 0;
            return -1;
        }
        HIPCHK(hipGetLastError());
    }
    s->timing = s->timing_stride > 0;
    return 0;
}

int zdm_step_graph(zdm_solver* s, int n) {
    if (!s) return fail("null solver");
    if (s->opt.time_integration == 2) return fail("zdm_step_graph: the implicit step synchronises (Newton) and is not captured");
    if (!s->graph_exec) {
        const bool was = s->timing;
        s->timing = false;
        HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        enqueue_step(s);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s->stream, &g);
        s->timing = was;
        if (e != hipSuccess) return fail(std::string("step capture failed: ") + hipGetErrorString(e));
        s->graph = g;
        HIPCHK(hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0));
    }
    for (int i = 0; i < n; ++i) HIPCHK(hipGraphLaunch(s->graph_exec, s->stream));
    return 0;
}

int zdm_enable_timing(zdm_solver* s, int on) {
    if (!s) return fail("null solver");
    if (resolve_events(s)) return -1;
    s->timing_stride = on == 0 ? 0 : on < 0 ? 1 : on;
    s->timing_step = 0;
    s->timing = on != 0;
    for (int i = 0; i < 4; ++i) {
        s->kt_ms[i] = 0;
        s->kt_n[i] = 0;
    }
    return 0;
}

int zdm_fuse_post(zdm_solver* s, int on) {
    if (!s) return fail("null solver");
    s->fuse_post = on != 0;
    return 0;
}

int zdm_flux_nontemporal(zdm_solver* s, int mode, int* active) {
    if (!s) return fail("null solver");
    if (mode < -1 || mode > 1) return fail("zdm_flux_nontemporal: mode must be -1, 0 or 1");
    s->flux_nt_mode = mode;
    choose_flux_nt(s);
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);   // a captured step holds the old kernels
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
    if (active) *active = s->flux_nt ? 1 : 0;
    return 0;
}

int zdm_kernel_time(zdm_solver* s, int kernel, double* avg_ms, int* launches) {
    if (!s || kernel < 0 || kernel > 3) return fail("bad kernel id");
    if (resolve_events(s)) return -1;
    *launches = s->kt_n[kernel];
    *avg_ms = s->kt_n[kernel] ? s->kt_ms[kernel] / s->kt_n[kernel] : 0.0;
    return 0;
}

int zdm_get_time(zdm_solver* s, double* time, double* dt, int* steps) {
    if (!s) return fail("null solver");
    Ctrl c;
    HIPCHK(hipMemcpyAsync(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (time) *time = c.t;
    if (dt) *dt = c.dt_cfl;
    if (steps) *steps = c.steps;
    return 0;
}

int zdm_run(zdm_solver* s, int max_steps, int check_every, int* steps_done) {
    if (!s) return fail("null solver");
    if (check_every < 1) check_every = 1;
    int done = 0, st0 = 0;
    if (zdm_get_time(s, nullptr, nullptr, &st0)) return -1;
    while (done < max_steps) {
        int n = std::min(check_every, max_steps - done);
        if (zdm_step(s, n)) return -1;
        done += n;
        std::vector<Ctrl> cs;
        if (read_ctrls(s, cs)) return -1;
        bool any = false;
        for (const Ctrl& c : cs) any = any || c.active;
        if (!any) break;
    }
    int st1 = 0;
    if (zdm_get_time(s, nullptr, nullptr, &st1)) return -1;
    if (steps_done) *steps_done = st1 - st0;
    return 0;
}

int zdm_batch_time(zdm_solver* s, int member, double* time, double* dt, int* steps) {
    if (!s || member < 0 || member >= s->nctrl) return fail("zdm_batch_time: bad member");
    std::vector<Ctrl> cs;
    if (read_ctrls(s, cs)) return -1;
    if (time) *time = cs[member].t;
    if (dt) *dt = cs[member].dt_cfl;
    if (steps) *steps = cs[member].steps;
    return 0;
}

int zdm_synchronize(zdm_solver* s) {
    if (!s) return fail("null solver");
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int zdm_transport_rhs(zdm_solver* s, double t, const double* X, const double* A, double* F) {
    if (!s || !X || !F) return fail("zdm_transport_rhs: null argument");
    if (upload_soa(s, X, s->W, NQ, false)) return -1;
    if (A) {
        if (upload_soa(s, A, s->Atmp, NA, false)) return -1;
    } else if (fill_zero(s, s->Atmp, (size_t)NA * s->ld)) {
        return -1;
    }
    if (set_ctrl_rhs_time(s, t)) return -1;
    prep(s, s->ctrl_rhs, s->W, s->Atmp);
    if (s->step_err) {
        s->step_err = 0;
        return -1;
    }
    flux(s, s->ctrl_rhs, MODE_RHS, 0, s->W, s->F, s->Atmp);
    HIPCHK(hipGetLastError());
    return download_soa(s, s->F, F, NQ, false);
}

// the right-hand side's control blocks at time t, on the stream (set_ctrl_rhs_time without the
// host copy and synchronisation)
__global__ void k_set_rhs_time(Ctrl* c, int n, T t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Ctrl z;
    z.t = t;
    z.dt = z.dt_cfl = z.dt_query = z.a_stage = z.al = z.be = z.ga = 0;
    for (int k = 0; k < 16; ++k) z.stage_t[k] = 0;
    z.stage_t[0] = t;
    z.active = 1;
    z.steps = z.finite = z.troubled = 0;
    c[i] = z;
}

int zdm_transport_rhs_device(zdm_solver* s, double t, const double* dX, const double* dA, double* dF,
                             void* stream) {
    if (!s || !dX || !dF) return fail("zdm_transport_rhs_device: null argument");
    if (s->refresh) return fail("zdm_transport_rhs_device: not for the derivative-aux refresh (it synchronises); "
                                "use zdm_transport_rhs");
    hipStream_t cs = (hipStream_t)stream, own = s->stream;
    const bool other = cs != own;
    if (other) {
        // the call's work starts after what the caller's stream holds
        if (!s->ev_caller) HIPCHK(hipEventCreateWithFlags(&s->ev_caller, hipEventDisableTiming));
        HIPCHK(hipEventRecord(s->ev_caller, cs));
        HIPCHK(hipStreamWaitEvent(own, s->ev_caller, 0));
    }
    const unsigned g = (unsigned)((s->ncells + 255) / 256);
    k_aos_to_soa<<<g, 256, 0, own>>>(dX, s->W, s->ncells, NQ, s->ld);
    if (dA)
        k_aos_to_soa<<<g, 256, 0, own>>>(dA, s->Atmp, s->ncells, NA, s->ld);
    else if (fill_zero(s, s->Atmp, (size_t)NA * s->ld))
        return -1;
    k_set_rhs_time<<<(s->nctrl + 63) / 64, 64, 0, own>>>(s->ctrl_rhs, s->nctrl, (T)t);
    prep(s, s->ctrl_rhs, s->W, s->Atmp);
    flux(s, s->ctrl_rhs, MODE_RHS, 0, s->W, s->F, s->Atmp);
    k_soa_to_aos<<<g, 256, 0, own>>>(s->F, dF, s->ncells, NQ, s->ld);
    HIPCHK(hipGetLastError());
    if (other) {
        // later work on the caller's stream sees F
        HIPCHK(hipEventRecord(s->ev_caller, own));
        HIPCHK(hipStreamWaitEvent(cs, s->ev_caller, 0));
    }
    return 0;
}

int zdm_source_step(zdm_solver* s, double dt) {
    if (!s) return fail("null solver");
    if (set_ctrl_rhs_time(s, 0.0)) return -1;
    post(s, 3, dt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int zdm_enforce_constraints(zdm_solver* s) {
    if (!s) return fail("null solver");
    if (set_ctrl_rhs_time(s, 0.0)) return -1;
    if (s->refresh && refresh_aux(s, s->X, s->A)) return -1;
    post(s, 4);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int zdm_stream_copy(size_t bytes, int iters, double* gbs, double* avg_ms) {
    if (!gbs || iters < 1 || bytes < 4096) return fail("zdm_stream_copy: bad arguments");
    const size_t n = bytes / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n * sizeof(double2)));
    if (hipMalloc(&b, n * sizeof(double2)) != hipSuccess) {
        (void)hipFree(a);
        return fail("hipMalloc failed in zdm_stream_copy");
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const unsigned grid = 65536;
    (void)hipMemset(a, 0, n * sizeof(double2));
    const double* ad = reinterpret_cast<const double*>(a);
    double* bd = reinterpret_cast<double*>(b);
    k_stream_copy<<<grid, 1024>>>(ad, bd, n);
    (void)hipEventRecord(e0, 0);
    for (int i = 0; i < iters; ++i) k_stream_copy<<<grid, 1024>>>(ad, bd, n);
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(std::string("zdm_stream_copy: ") + hipGetErrorString(e));
    const double per = ms / iters;
    *gbs = 2.0 * (double)(n * sizeof(double2)) / (per * 1e-3) / 1e9;
    if (avg_ms) *avg_ms = per;
    return 0;
}

int zdm_cfl_dt(zdm_solver* s, double* dt) {
    if (!s || !dt) return fail("null argument");
    if (s->nmem > 0) return fail("zdm_cfl_dt: use zdm_batch_time for an ensemble");
    // the standalone query must not disturb an inactive run: force the CFL pass on
    Ctrl c;
    HIPCHK(hipMemcpyAsync(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    const int was = c.active;
    c.active = 1;
    HIPCHK(hipMemcpyAsync(s->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, s->stream));
    post(s, 0);
    finalize(s, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    *dt = c.dt_query;
    c.active = was;
    HIPCHK(hipMemcpyAsync(s->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

// ------------------------------------------------------------------ Chorin pressure split

int zdm_bsr2_gmres(int nb, const int* rowptr, const int* col, const double* val, const double* b,
                   double* x, const zdm_krylov_options* opts, int* its, int* reason) {
    if (nb < 1 || !rowptr || !col || !val || !b || !x) return fail("zdm_bsr2_gmres: bad arguments");
    zdm_krylov_options o = opts ? *opts : krylov_defaults();
    if (check_krylov(o)) return -1;
    const int nnzb = rowptr[nb];
    if (nnzb < 0 || rowptr[0] != 0) return fail("zdm_bsr2_gmres: bad row pointers");
    for (int c = 0; c < nb; ++c)
        if (rowptr[c + 1] < rowptr[c]) return fail("zdm_bsr2_gmres: row pointers must not decrease");
    for (int k = 0; k < nnzb; ++k)
        if (col[k] < 0 || col[k] >= nb) return fail("zdm_bsr2_gmres: block column out of range");
    const size_t n = 2 * (size_t)nb;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t r = off;
        off = align_up(off + bytes, 256);
        return r;
    };
    const size_t oR = take(sizeof(int) * (nb + 1)), oC = take(sizeof(int) * std::max(nnzb, 1)),
                 oV = take(sizeof(T) * 4 * std::max(nnzb, 1)), oB = take(sizeof(T) * n), oX = take(sizeof(T) * n),
                 oD = take(sizeof(T) * 4 * (size_t)nb);
    char* mem = nullptr;
    char* kmem = nullptr;
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int rc = 0;
    if (hipMalloc(&mem, off) != hipSuccess || hipMalloc(&kmem, gmres_bytes(o.restart, n)) != hipSuccess) {
        rc = fail("zdm_bsr2_gmres: hipMalloc failed");
    } else {
        hipError_t e = hipMemcpyAsync(mem + oR, rowptr, sizeof(int) * (nb + 1), hipMemcpyHostToDevice, st);
        if (e == hipSuccess && nnzb) e = hipMemcpyAsync(mem + oC, col, sizeof(int) * nnzb, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && nnzb) e = hipMemcpyAsync(mem + oV, val, sizeof(T) * 4 * nnzb, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(mem + oB, b, sizeof(T) * n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(mem + oX, x, sizeof(T) * n, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) {
            rc = fail(std::string("zdm_bsr2_gmres: upload failed: ") + hipGetErrorString(e));
        } else {
            Bsr2 A;
            A.nb = nb;
            A.rowptr = (const int*)(mem + oR);
            A.col = (const int*)(mem + oC);
            A.val = (const T*)(mem + oV);
            rc = bsr2_gmres_on_device(A, (const T*)(mem + oB), (T*)(mem + oX), (T*)(mem + oD), kmem, o, st,
                                      "zdm_bsr2_gmres", its, reason);
            if (rc == 0 && (hipMemcpyAsync(x, mem + oX, sizeof(T) * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
                            hipStreamSynchronize(st) != hipSuccess))
                rc = fail("zdm_bsr2_gmres: download failed");
        }
    }
    (void)hipStreamSynchronize(st);
    if (mem) (void)hipFree(mem);
    if (kmem) (void)hipFree(kmem);
    (void)hipStreamDestroy(st);
    return rc;
}

int zdm_bsr2_gmres_noalloc(int nb, const int* rowptr, const int* col, const double* val, const double* b,
                           double* x, const zdm_krylov_options* opts, void* stream, int* its, int* reason) {
    if (nb < 1 || !rowptr || !col || !val || !b || !x) return fail("zdm_bsr2_gmres_noalloc: bad arguments");
    zdm_krylov_options o = opts ? *opts : krylov_defaults();
    if (check_krylov(o)) return -1;
    const size_t n = 2 * (size_t)nb;
    // scratch: [flags (256 B)] [Dinv, 4 per block] [GMRES workspace]
    const size_t oD = 256, oK = align_up(oD + sizeof(T) * 4 * (size_t)nb, 256);
    const size_t kbytes = gmres_bytes(o.restart, n), need = oK + kbytes;
    if (need > sizeof(scratch))
        return fail("zdm_bsr2_gmres_noalloc: " + std::to_string(nb) + " block rows with restart " +
                    std::to_string(o.restart) + " need " + std::to_string(need) +
                    " bytes of scratch, more than its " + std::to_string(sizeof(scratch)) + " bytes");
    char* sp = nullptr;
    if (scratch_on_current_device(&sp)) return -1;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int* flag = reinterpret_cast<int*>(sp);
    // the structure is checked on the device (the arrays never come to the host): first the row
    // pointers, and only when they start at 0 and never decrease, the columns they span
    k_fill<<<1, 64, 0, st>>>(reinterpret_cast<T*>(sp), 256 / sizeof(T), 0.0);
    const unsigned cg = (unsigned)((nb + 255) / 256);
    k_bsr2_check_rows<<<cg, 256, 0, st>>>(nb, rowptr, flag);
    HIPCHK(hipGetLastError());
    int hflag[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(hflag, flag, sizeof(hflag), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (hflag[0]) return fail("zdm_bsr2_gmres_noalloc: row pointers must start at 0 and not decrease");
    int nnzb = 0;
    HIPCHK(hipMemcpyAsync(&nnzb, rowptr + nb, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (nnzb > 0) {
        k_bsr2_check_cols<<<(unsigned)((nnzb + 255) / 256), 256, 0, st>>>(nb, nnzb, col, flag + 1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hflag, flag, sizeof(hflag), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (hflag[1]) return fail("zdm_bsr2_gmres_noalloc: block column out of range");
    }
    Bsr2 A;
    A.nb = nb;
    A.rowptr = rowptr;
    A.col = col;
    A.val = val;
    return bsr2_gmres_on_device(A, b, x, reinterpret_cast<T*>(sp + oD), sp + oK, o, st,
                                "zdm_bsr2_gmres_noalloc", its, reason);
}

#ifdef ZOOMY_MODEL_IS_VAM

int zdm_chorin_setup(zdm_solver* s, const zdm_krylov_options* opts) {
    if (!s || !s->ch) return fail("null solver");
    // ChorinVAMSolver::Run (ChorinVAM.hpp:86-129) has neither MOOD nor the zhang_shu CFL cap of
    // MUSCLSolver::Run; its predictor keeps the eta-WB reconstruction at order 2
    s->opt.cfl = s->cfl_in;
    const zdm_krylov_options o = opts ? *opts : krylov_defaults();
    if (check_krylov(o)) return -1;
    ChorinHost* ch = s->ch;
    if (o.restart != ch->g.m) {
        const size_t n = 2 * (size_t)s->ncells;
        char* km = nullptr;
        HIPCHK(hipStreamSynchronize(s->stream));
        HIPCHK(hipMalloc(&km, gmres_bytes(o.restart, n)));
        if (hipMemsetAsync(km, 0, gmres_bytes(o.restart, n), s->stream) != hipSuccess) {
            (void)hipFree(km);
            return fail("zdm_chorin_setup: clearing the GMRES workspace failed");
        }
        (void)hipFree(ch->kmem);
        ch->kmem = km;
        gmres_bind(ch->g, km, o.restart, n);
    }
    ch->opt = o;
    ch->g.max_it = o.max_it;
    ch->g.rtol = o.rtol;
    ch->g.atol = o.atol;
    ch->g.dtol = o.dtol;
    return 0;
}

int zdm_chorin_step(zdm_solver* s, int n, int* steps_done, int* its_total, int* last_reason) {
    if (!s) return fail("null solver");
    if (!s->ch) return fail("zdm_chorin_step: the pressure split runs on a single mesh, not an ensemble");
    int done = 0, its = 0, reason = 0;
    for (int i = 0; i < n; ++i) {
        const int r = chorin_one_step(s);
        if (r < 0) return -1;
        if (r == 1) break;
        ++done;
        its += s->ch->last_its;
        reason = s->ch->last_reason;
    }
    if (steps_done) *steps_done = done;
    if (its_total) *its_total = its;
    if (last_reason) *last_reason = reason;
    return 0;
}

int zdm_chorin_stats(zdm_solver* s, int* its, int* reason, double* rnorm, double* rnorm0) {
    if (!s || !s->ch) return fail("zdm_chorin_stats: no pressure solve on this solver");
    if (its) *its = s->ch->last_its;
    if (reason) *reason = s->ch->last_reason;
    if (rnorm) *rnorm = s->ch->last_rnorm;
    if (rnorm0) *rnorm0 = s->ch->last_rnorm0;
    return 0;
}

int zdm_chorin_pressure(zdm_solver* s, double* P) {
    if (!s || !s->ch || !P) return fail("zdm_chorin_pressure: null argument");
    HIPCHK(hipMemcpyAsync(P, s->ch->P, sizeof(T) * 2 * (size_t)s->ncells, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int zdm_vam_frozen_aux(zdm_solver* s, double* CA) {
    if (!s || !s->ch) return fail("null solver");
    vam_frozen(s);
    HIPCHK(hipGetLastError());
    if (CA) {
        std::vector<T> soa((size_t)VNCA * s->ld);
        HIPCHK(hipMemcpyAsync(soa.data(), s->ch->v.CA, sizeof(T) * soa.size(), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        for (int c = 0; c < s->ncells; ++c)
            for (int j = 0; j < VNCA; ++j) CA[(size_t)c * VNCA + j] = soa[(size_t)j * s->ld + c];
    }
    return 0;
}

int zdm_vam_residual(zdm_solver* s, double dt, const double* P, double* R) {
    if (!s || !s->ch || !P || !R) return fail("zdm_vam_residual: null argument");
    ChorinHost* ch = s->ch;
    const size_t bytes = sizeof(T) * 2 * (size_t)s->ncells;
    HIPCHK(hipMemcpyAsync(ch->P, P, bytes, hipMemcpyHostToDevice, s->stream));
    vam_apply(s, P2_RESIDUAL, dt, ch->P, ch->rhs);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(R, ch->rhs, bytes, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

int zdm_vam_pmat(zdm_solver* s, double dt, int cap, int* rowptr, int* col, double* val) {
    if (!s || !s->ch) return fail("null solver");
    ChorinHost* ch = s->ch;
    const int nnzb = ch->v.nnzb;
    if (cap < nnzb || !rowptr || !col || !val) return nnzb;
    vam_assemble(s, dt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(val, ch->v.bval, sizeof(T) * 4 * (size_t)nnzb, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    std::copy(ch->h_brow.begin(), ch->h_brow.end(), rowptr);
    std::copy(ch->h_bcol.begin(), ch->h_bcol.end(), col);
    return nnzb;
}

#else

static int not_vam() {
    return fail("solver.method=\"chorin\" needs a VAM model (ChorinOps.H and ZOOMY_MODEL_IS_VAM); "
                "this library was built for " ZDM_MODEL_NAME);
}
int zdm_chorin_setup(zdm_solver*, const zdm_krylov_options*) { return not_vam(); }
int zdm_chorin_step(zdm_solver*, int, int*, int*, int*) { return not_vam(); }
int zdm_chorin_pressure(zdm_solver*, double*) { return not_vam(); }
int zdm_chorin_stats(zdm_solver*, int*, int*, double*, double*) { return not_vam(); }
int zdm_vam_frozen_aux(zdm_solver*, double*) { return not_vam(); }
int zdm_vam_residual(zdm_solver*, double, const double*, double*) { return not_vam(); }
int zdm_vam_pmat(zdm_solver*, double, int, int*, int*, double*) { return not_vam(); }

#endif

}  // extern "C"

namespace {
#include "adjoint_host.inc"
}  // namespace

extern "C" {

int zdm_adjoint_setup(zdm_solver* s, int nsteps, int checkpoint_every, int obs_every, int ngauge,
                      const int* gauges, const double* obs, int exact_limiter) {
    if (!s) return fail("null solver");
    if (HasDiffusiveFlux<T>::value)
        return fail("the adjoint does not transpose a model's diffusive_flux");
    if (s->refresh)
        return fail("the adjoint does not transpose RefreshDerivativeAux: set solver.refresh_derivative_aux to false");
    return adjoint_alloc(s, nsteps, checkpoint_every, obs_every, ngauge, gauges, obs, 0, exact_limiter ? 0 : 1);
}

int zdm_adjoint_setup_chorin(zdm_solver* s, int nsteps, int checkpoint_every, int obs_every, int ngauge,
                             const int* gauges, const double* obs, int exact_limiter) {
    if (!s) return fail("null solver");
    if (HasDiffusiveFlux<T>::value)
        return fail("the adjoint does not transpose a model's diffusive_flux");
    if (s->refresh)
        return fail("the adjoint does not transpose RefreshDerivativeAux: set solver.refresh_derivative_aux to false");
    return adjoint_alloc(s, nsteps, checkpoint_every, obs_every, ngauge, gauges, obs, 1, exact_limiter ? 0 : 1);
}

int zdm_implicit_setup(zdm_solver* s, const zdm_newton_options* opts) {
    if (!s) return fail("null solver");
    if (s->opt.time_integration != 2) return fail("zdm_implicit_setup: the solver's time_integration is not implicit");
    const zdm_newton_options o = opts ? *opts : newton_defaults();
    const bool had = s->imp && s->imp->have_prev;
    const T dtp = s->imp ? s->imp->dt_prev : 0;
    std::vector<T> keep;
    if (had) {
        keep.resize((size_t)NQ * s->ld);
        HIPCHK(hipMemcpy(keep.data(), s->imp->Xm, sizeof(T) * keep.size(), hipMemcpyDeviceToHost));
    }
    if (implicit_build(s, o)) return -1;
    if (had) {
        // keep the BDF history across a change of settings
        HIPCHK(hipMemcpy(s->imp->Xm, keep.data(), sizeof(T) * keep.size(), hipMemcpyHostToDevice));
        s->imp->have_prev = true;
        s->imp->dt_prev = dtp;
    }
    return 0;
}

int zdm_implicit_stats(zdm_solver* s, long long* newton_its, long long* ksp_its, int* last_newton,
                       int* last_reason) {
    if (!s) return fail("null solver");
    const ImplicitHost* im = s->imp;
    if (newton_its) *newton_its = im ? im->newton_its : 0;
    if (ksp_its) *ksp_its = im ? im->ksp_its : 0;
    if (last_newton) *last_newton = im ? im->last_newton : 0;
    if (last_reason) *last_reason = im ? im->last_reason : 0;
    return 0;
}

int zdm_adjoint_krylov_its(zdm_solver* s, long long* its) {
    if (!s || !s->adj || !its) return fail("zdm_adjoint_krylov_its: bad arguments");
    *its = s->adj->its_t;
    return 0;
}

int zdm_set_launch(zdm_solver* s, int block_size, int cells_per_thread) {
    if (!s) return fail("null solver");
    if (block_size != 64 && block_size != 128 && block_size != 256)
        return fail("zdm_set_launch: block_size must be 64, 128 or 256");
    if (cells_per_thread < 1 || cells_per_thread > 8) return fail("zdm_set_launch: cells_per_thread must be 1..8");
    std::vector<int> bm, bf;
    const int nb = launch_tables(s->ncells, s->coff_h, block_size, cells_per_thread, bm, bf);
    if (nb > s->maxblocks) return fail("zdm_set_launch: too many workgroups");
    HIPCHK(hipStreamSynchronize(s->stream));
    if (!bm.empty()) {
        HIPCHK(hipMemcpy(const_cast<int*>(s->dm.blk_member), bm.data(), sizeof(int) * bm.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(const_cast<int*>(s->dm.blk_first), bf.data(), sizeof(int) * bf.size(), hipMemcpyHostToDevice));
    }
    s->bs = block_size;
    s->dm.cpt = cells_per_thread;
    s->nblocks = nb;
    // a captured step holds the old launch shapes
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
    return 0;
}

int zdm_get_launch(zdm_solver* s, int* block_size, int* cells_per_thread) {
    if (!s) return fail("null solver");
    if (block_size) *block_size = s->bs;
    if (cells_per_thread) *cells_per_thread = s->dm.cpt;
    return 0;
}

int zdm_autotune(zdm_solver* s, int steps, int* block_size, int* cells_per_thread, double* ms) {
    if (!s) return fail("null solver");
    if (steps < 1) return fail("zdm_autotune: steps must be >= 1");
    if (s->opt.time_integration == 2) return fail("zdm_autotune: not for the implicit step");
    static const int cand[9][2] = {{256, 1}, {128, 1}, {64, 1}, {256, 2}, {128, 2}, {64, 2}, {256, 4}, {128, 4}, {64, 4}};
    const size_t bX = sizeof(T) * (size_t)NQ * s->ld, bA = sizeof(T) * (size_t)NA * s->ld,
                 bC = sizeof(Ctrl) * (size_t)s->nctrl;
    char* save = nullptr;
    HIPCHK(hipMalloc(&save, bX + bA + bC));
    auto copy3 = [&](bool out) -> hipError_t {
        hipError_t e = hipMemcpyAsync(out ? save : (char*)s->X, out ? (char*)s->X : save, bX, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(out ? save + bX : (char*)s->A, out ? (char*)s->A : save + bX, bA, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(out ? save + bX + bA : (char*)s->ctrl, out ? (char*)s->ctrl : save + bX + bA, bC,
                               hipMemcpyDeviceToDevice, s->stream);
        return e;
    };
    const bool timing = s->timing;
    s->timing = false;
    int best = 0;
    double best_ms = 0;
    s->tune_ms.assign(9, 0.0);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (copy3(true) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail("zdm_autotune: setup failed");
    for (int k = 0; k < 9 && rc == 0; ++k) {
        if (zdm_set_launch(s, cand[k][0], cand[k][1])) {
            rc = -1;
            break;
        }
        float t = 0;
        hipError_t e = copy3(false);
        if (e == hipSuccess) {
            enqueue_step(s);   // warm-up
            e = copy3(false);
        }
        if (e == hipSuccess) e = hipEventRecord(e0, s->stream);
        for (int i = 0; i < steps && e == hipSuccess; ++i) enqueue_step(s);
        if (e == hipSuccess) e = hipEventRecord(e1, s->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) {
            rc = fail(std::string("zdm_autotune: ") + hipGetErrorString(e));
            break;
        }
        s->tune_ms[k] = (double)t / steps;
        if (k == 0 || s->tune_ms[k] < best_ms) {
            best_ms = s->tune_ms[k];
            best = k;
        }
    }
    if (rc == 0) rc = zdm_set_launch(s, cand[best][0], cand[best][1]);
    if (rc == 0 && (copy3(false) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess))
        rc = fail("zdm_autotune: restoring the state failed");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(save);
    s->timing = timing;
    if (rc) return rc;
    if (block_size) *block_size = cand[best][0];
    if (cells_per_thread) *cells_per_thread = cand[best][1];
    if (ms)
        for (int k = 0; k < 9; ++k) ms[k] = s->tune_ms[k];
    return 0;
}

int zdm_troubled_cells(zdm_solver* s, long long* count) {
    if (!s || !count) return fail("zdm_troubled_cells: null argument");
    std::vector<Ctrl> cs;
    if (read_ctrls(s, cs)) return -1;
    long long n = 0;
    for (const Ctrl& c : cs) n += c.troubled;
    *count = n;
    return 0;
}

int zdm_adjoint_layout(zdm_solver* s, int* checkpoint_every, int* tape) {
    if (!s || !s->adj) return fail("zdm_adjoint_layout: call zdm_adjoint_setup first");
    if (checkpoint_every) *checkpoint_every = s->adj->slots > 0 ? -2 - s->adj->slots : s->adj->ck;
    if (tape) *tape = s->adj->stretch ? 0 : s->adj->tape;
    return 0;
}

int zdm_adjoint_stretch_tape(zdm_solver* s, int steps, int buffers, int* steps_out, int* level_out,
                             int* buffers_out) {
    if (!s || !s->adj) return fail("zdm_adjoint_stretch_tape: call zdm_adjoint_setup first");
    AdjointHost* ad = s->adj;
    if (buffers < 0 || buffers > 2) return fail("zdm_adjoint_stretch_tape: buffers must be 0, 1 or 2");
    if (steps >= 0) {
        if (ad->tape && !ad->stretch) return fail("zdm_adjoint_stretch_tape: the layout is a full tape");
        if (stretch_setup(s, ad, steps, false, buffers)) return -1;
    }
    if (steps_out) *steps_out = ad->stretch ? ad->tlen : 0;
    if (level_out) *level_out = ad->stretch ? ad->tape : 0;
    if (buffers_out) *buffers_out = ad->stretch ? ad->tbufs : 0;
    return 0;
}

int zdm_adjoint_forward(zdm_solver* s, const double* X0, const double* A0, const double* dts, double* loss) {
    if (!s || !X0) return fail("zdm_adjoint_forward: null argument");
    if (!s->adj) return fail("zdm_adjoint_forward: call zdm_adjoint_setup first");
    return adjoint_forward(s, X0, A0, dts, loss);
}

int zdm_adjoint_forward_frozen_alpha(zdm_solver* s, const double* X0, const double* A0, const double* dts,
                                     int record, double* loss) {
    if (!s || !X0) return fail("zdm_adjoint_forward_frozen_alpha: null argument");
    if (!s->adj) return fail("zdm_adjoint_forward_frozen_alpha: call zdm_adjoint_setup first");
    if (s->dm.nmem > 0) return fail("zdm_adjoint_forward_frozen_alpha: not for an ensemble");
    return adjoint_forward_frozen_alpha(s, X0, A0, dts, record, loss);
}

int zdm_adjoint_gradient(zdm_solver* s, double* gX0, double* gA0) {
    if (!s) return fail("null solver");
    if (!s->adj) return fail("zdm_adjoint_gradient: call zdm_adjoint_setup first");
    return adjoint_reverse(s, gX0, gA0);
}

int zdm_adjoint_forward_device(zdm_solver* s, const double* dX0, const double* dA0, const double* dts,
                               double* loss) {
    if (!s || !dX0) return fail("zdm_adjoint_forward_device: null argument");
    if (!s->adj) return fail("zdm_adjoint_forward_device: call zdm_adjoint_setup first");
    return adjoint_forward(s, dX0, dA0, dts, loss, true);
}

int zdm_adjoint_gradient_device(zdm_solver* s, double* dgX0, double* dgA0) {
    if (!s) return fail("null solver");
    if (!s->adj) return fail("zdm_adjoint_gradient_device: call zdm_adjoint_setup first");
    return adjoint_reverse(s, dgX0, dgA0, true);
}

int zdm_solver_stream(zdm_solver* s, void** stream) {
    if (!s || !stream) return fail("zdm_solver_stream: null argument");
    *stream = (void*)s->stream;
    return 0;
}

int zdm_adjoint_dts(zdm_solver* s, double* dts) {
    if (!s || !s->adj || !dts) return fail("zdm_adjoint_dts: bad arguments");
    std::copy(s->adj->dts.begin(), s->adj->dts.end(), dts);
    return 0;
}

int zdm_adjoint_residuals(zdm_solver* s, double* r) {
    if (!s || !s->adj || !r) return fail("zdm_adjoint_residuals: bad arguments");
    std::copy(s->adj->res_h.begin(), s->adj->res_h.end(), r);
    return 0;
}

}  // extern "C"
