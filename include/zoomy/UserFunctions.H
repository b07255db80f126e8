// UserFunctions.H -- the functions a generated Model.H / Numerics.H calls by bare name and
// expects to find declared before it is included (the reference's UserFunctions.H:33-258):
//
//   ZoomyMesh           the mesh handle of the field-level update_aux_variables(Q, Qaux, dt,
//                       mesh) overload (Model.H:250-262)
//   compute_derivative  result = D^(dx,dy,dz) source for one per-cell field: the Green-Gauss
//                       cell gradient with a zero-Neumann boundary, first derivatives only
//                       (UserFunctions.H:60-112)
//   eigenvalues         idx-th eigenvalue (real part) of a row-major n x n matrix
//   eigensystem         idx-th entry of [eigenvalues (n), right eigenvectors R (n x n),
//                       L = R^-1 (n x n)]
//   solve               idx-th component of A^-1 b, arguments A (n x n) then b (n)
//
// No PETSc and no Eigen.  The field operations (compute_derivative) are carried out by the
// solver that owns the fields: ZoomyMesh names it and its derivative operator, so in the GPU
// solver the fields are device arrays and each derivative is a device kernel.  The dense
// kernels are fixed-size and allocation-free (no caches, no heap), so a generated Numerics.H
// may call them in device code: the matrix size n follows from the argument count at compile
// time.  Conventions kept from the reference (UserFunctions.H:7-26): a matrix with a
// non-finite entry, or one whose eigenvalue iteration does not converge, gives eigenvalues
// +inf and R = L = I, and the eigenvalue path never inverts the eigenbasis.  Each call
// recomputes the decomposition (the reference keeps a one-slot thread-local cache, which
// device code cannot; the values are the same whichever call computes them).
#pragma once
#include <cmath>
#include <type_traits>

// ---------------------------------------------------------------------------- ZoomyMesh
// Cells [cStart, cEnd) of the mesh whose per-cell fields are passed to the field-level
// update_aux_variables; derivative(ctx, result, source, dx, dy, dz) computes one field's
// derivative (the owner records an unsupported order or a device error and reports it after
// the model call returns).
struct ZoomyMesh {
    void* ctx;
    void (*derivative)(void* ctx, double* result, const double* source, int dx, int dy, int dz);
    int dim;
    long cStart, cEnd;
};

template <typename T>
inline void compute_derivative(T* result, const T* source, int dx, int dy, int dz, const ZoomyMesh& mesh) {
    static_assert(std::is_same<T, double>::value, "compute_derivative: fields are double");
    mesh.derivative(mesh.ctx, result, source, dx, dy, dz);
}

namespace zoomy_uf {

constexpr int side_of_square(int m) {
    int n = 0;
    while ((n + 1) * (n + 1) <= m) ++n;
    return n;
}
constexpr int side_of_system(int m) {   // m = n^2 + n
    int n = 0;
    while (n * n + n < m) ++n;
    return n;
}

inline bool all_finite(const double* a, int m) {
    for (int i = 0; i < m; ++i)
        if (!(std::fabs(a[i]) <= 1.7976931348623157e308)) return false;
    return true;
}

// x = A^-1 b by Gaussian elimination with partial pivoting on copies; false if a pivot is 0
template <int N>
inline bool gepp_solve(const double* A, const double* b, double* x) {
    double M[N * N], r[N];
    for (int i = 0; i < N * N; ++i) M[i] = A[i];
    for (int i = 0; i < N; ++i) r[i] = b[i];
    for (int k = 0; k < N; ++k) {
        int p = k;
        for (int i = k + 1; i < N; ++i)
            if (std::fabs(M[i * N + k]) > std::fabs(M[p * N + k])) p = i;
        if (M[p * N + k] == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < N; ++j) {
                const double t = M[k * N + j];
                M[k * N + j] = M[p * N + j];
                M[p * N + j] = t;
            }
            const double t = r[k];
            r[k] = r[p];
            r[p] = t;
        }
        for (int i = k + 1; i < N; ++i) {
            const double f = M[i * N + k] / M[k * N + k];
            for (int j = k; j < N; ++j) M[i * N + j] -= f * M[k * N + j];
            r[i] -= f * r[k];
        }
    }
    for (int i = N - 1; i >= 0; --i) {
        double s = r[i];
        for (int j = i + 1; j < N; ++j) s -= M[i * N + j] * x[j];
        x[i] = s / M[i * N + i];
    }
    return true;
}

// Real parts of the eigenvalues of the row-major N x N matrix A: orthogonal (Householder)
// reduction to upper Hessenberg form, then QR steps by Givens rotations with a Wilkinson shift
// on the active window, deflating 1 x 1 blocks and 2 x 2 blocks (a complex pair: real part
// twice).  False when the iteration has not converged after 60 N steps.
template <int N>
inline bool eig_real(const double* A, double* wr) {
    double H[N * N];
    for (int i = 0; i < N * N; ++i) H[i] = A[i];
    for (int k = 0; k + 2 < N; ++k) {   // Householder: zero H[k+2.., k]
        double nrm = 0.0;
        for (int i = k + 1; i < N; ++i) nrm += H[i * N + k] * H[i * N + k];
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) continue;
        const double alpha = H[(k + 1) * N + k] > 0 ? -nrm : nrm;
        double v[N];
        for (int i = 0; i < N; ++i) v[i] = 0.0;
        v[k + 1] = H[(k + 1) * N + k] - alpha;
        for (int i = k + 2; i < N; ++i) v[i] = H[i * N + k];
        double vv = 0.0;
        for (int i = k + 1; i < N; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        for (int j = 0; j < N; ++j) {   // H = (I - 2 v v^T / vv) H
            double s = 0.0;
            for (int i = k + 1; i < N; ++i) s += v[i] * H[i * N + j];
            s = 2.0 * s / vv;
            for (int i = k + 1; i < N; ++i) H[i * N + j] -= s * v[i];
        }
        for (int i = 0; i < N; ++i) {   // H = H (I - 2 v v^T / vv)
            double s = 0.0;
            for (int j = k + 1; j < N; ++j) s += H[i * N + j] * v[j];
            s = 2.0 * s / vv;
            for (int j = k + 1; j < N; ++j) H[i * N + j] -= s * v[j];
        }
        for (int i = k + 2; i < N; ++i) H[i * N + k] = 0.0;
    }
    const double eps = 2.220446049250313e-16;
    int hi = N - 1, its = 0;
    while (hi >= 0) {
        if (hi == 0) {
            wr[0] = H[0];
            break;
        }
        int l = hi;   // start of the active unreduced block
        while (l > 0) {
            const double s = std::fabs(H[(l - 1) * N + l - 1]) + std::fabs(H[l * N + l]);
            if (std::fabs(H[l * N + l - 1]) <= eps * (s > 0 ? s : 1.0)) {
                H[l * N + l - 1] = 0.0;
                break;
            }
            --l;
        }
        if (l == hi) {
            wr[hi] = H[hi * N + hi];
            --hi;
            its = 0;
            continue;
        }
        const double a = H[(hi - 1) * N + hi - 1], b = H[(hi - 1) * N + hi], c = H[hi * N + hi - 1],
                     d = H[hi * N + hi];
        const double tr = 0.5 * (a + d), det = a * d - b * c, disc = tr * tr - det;
        if (l == hi - 1) {   // a 2 x 2 block
            if (disc >= 0) {
                const double sq = std::sqrt(disc);
                wr[hi - 1] = tr + sq;
                wr[hi] = tr - sq;
            } else {
                wr[hi - 1] = tr;
                wr[hi] = tr;
            }
            hi -= 2;
            its = 0;
            continue;
        }
        if (++its > 60 * N) return false;
        // Wilkinson shift: the eigenvalue of the trailing 2 x 2 block nearer d (its real part
        // for a complex pair); an exceptional shift every 11th step
        double mu;
        if (disc >= 0) {
            const double sq = std::sqrt(disc);
            const double e1 = tr + sq, e2 = tr - sq;
            mu = std::fabs(e1 - d) < std::fabs(e2 - d) ? e1 : e2;
        } else {
            mu = tr;
        }
        if (its % 11 == 0) mu = d + 0.75 * std::fabs(c);
        // QR step on the window [l, hi]: H - mu I = Q R, H <- R Q + mu I (Givens)
        double cs[N], sn[N];
        for (int i = l; i <= hi; ++i) H[i * N + i] -= mu;
        for (int k = l; k < hi; ++k) {
            const double x = H[k * N + k], y = H[(k + 1) * N + k];
            const double r = std::sqrt(x * x + y * y);
            const double cc = r > 0 ? x / r : 1.0, ss = r > 0 ? y / r : 0.0;
            cs[k] = cc;
            sn[k] = ss;
            for (int j = k; j < N; ++j) {
                const double u = H[k * N + j], w = H[(k + 1) * N + j];
                H[k * N + j] = cc * u + ss * w;
                H[(k + 1) * N + j] = -ss * u + cc * w;
            }
        }
        for (int k = l; k < hi; ++k) {
            const double cc = cs[k], ss = sn[k];
            const int top = (k + 2 < hi ? k + 2 : hi);
            for (int i = 0; i <= top; ++i) {
                const double u = H[i * N + k], w = H[i * N + k + 1];
                H[i * N + k] = cc * u + ss * w;
                H[i * N + k + 1] = -ss * u + cc * w;
            }
        }
        for (int i = l; i <= hi; ++i) H[i * N + i] += mu;
    }
    return true;
}

// Eigenvector of A for the real eigenvalue lam by inverse iteration on A - (lam + delta) I,
// normalised to unit 2-norm; false if the shifted matrix is singular for every delta tried
template <int N>
inline bool eigvec(const double* A, double lam, double* v) {
    double scale = std::fabs(lam);
    for (int i = 0; i < N * N; ++i) scale = std::fabs(A[i]) > scale ? std::fabs(A[i]) : scale;
    if (scale == 0.0) scale = 1.0;
    for (double rel = 1e-10; rel < 1e-2; rel *= 1e3) {
        double S[N * N];
        for (int i = 0; i < N * N; ++i) S[i] = A[i];
        for (int i = 0; i < N; ++i) S[i * N + i] -= lam + rel * scale;
        double x[N];
        for (int i = 0; i < N; ++i) x[i] = 1.0 + 0.1 * i;
        bool ok = true;
        for (int it = 0; it < 4 && ok; ++it) {
            double y[N];
            ok = gepp_solve<N>(S, x, y) && all_finite(y, N);
            if (!ok) break;
            double nrm = 0.0;
            for (int i = 0; i < N; ++i) nrm += y[i] * y[i];
            nrm = std::sqrt(nrm);
            if (!(nrm > 0.0)) {
                ok = false;
                break;
            }
            for (int i = 0; i < N; ++i) x[i] = y[i] / nrm;
        }
        if (ok) {
            for (int i = 0; i < N; ++i) v[i] = x[i];
            return true;
        }
    }
    return false;
}

// Moore-Penrose pseudo-inverse of a singular N x N matrix V: cyclic Jacobi eigen-decomposition
// of V^T V = U diag(s) U^T, pinv = U diag(1/s, s > tol) U^T V^T
template <int N>
inline void pinv(const double* V, double* P) {
    double G[N * N], U[N * N];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double s = 0.0;
            for (int k = 0; k < N; ++k) s += V[k * N + i] * V[k * N + j];
            G[i * N + j] = s;
            U[i * N + j] = i == j ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < N; ++p)
            for (int q = p + 1; q < N; ++q) off += G[p * N + q] * G[p * N + q];
        if (off < 1e-30) break;
        for (int p = 0; p < N; ++p)
            for (int q = p + 1; q < N; ++q) {
                if (G[p * N + q] == 0.0) continue;
                const double th = (G[q * N + q] - G[p * N + p]) / (2.0 * G[p * N + q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < N; ++k) {   // G = J^T G J, U = U J
                    const double gp = G[k * N + p], gq = G[k * N + q];
                    G[k * N + p] = c * gp - s * gq;
                    G[k * N + q] = s * gp + c * gq;
                }
                for (int k = 0; k < N; ++k) {
                    const double gp = G[p * N + k], gq = G[q * N + k];
                    G[p * N + k] = c * gp - s * gq;
                    G[q * N + k] = s * gp + c * gq;
                }
                for (int k = 0; k < N; ++k) {
                    const double up = U[k * N + p], uq = U[k * N + q];
                    U[k * N + p] = c * up - s * uq;
                    U[k * N + q] = s * up + c * uq;
                }
            }
    }
    double smax = 0.0;
    for (int i = 0; i < N; ++i) smax = G[i * N + i] > smax ? G[i * N + i] : smax;
    const double tol = smax * 1e-24;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {   // P = U diag(1/s) U^T V^T
            double acc = 0.0;
            for (int k = 0; k < N; ++k) {
                const double sk = G[k * N + k];
                if (!(sk > tol)) continue;
                double uv = 0.0;   // (U^T V^T)[k][j] = sum_m U[m][k] V[j][m]
                for (int m = 0; m < N; ++m) uv += U[m * N + k] * V[j * N + m];
                acc += U[i * N + k] / sk * uv;
            }
            P[i * N + j] = acc;
        }
}

}  // namespace zoomy_uf

// ---------------------------------------------------------------------------- eigenvalues
template <typename I, typename... Ts>
inline double eigenvalues(I idx_, Ts... a_flat_args) {
    constexpr int M = sizeof...(a_flat_args);
    constexpr int N = zoomy_uf::side_of_square(M);
    static_assert(N * N == M && N > 0, "eigenvalues: a square matrix is expected");
    const double a[M] = {static_cast<double>(a_flat_args)...};
    const int idx = static_cast<int>(idx_);
    double w[N];
    if (!zoomy_uf::all_finite(a, M) || !zoomy_uf::eig_real<N>(a, w)) return __builtin_huge_val();
    return w[idx];
}

// ---------------------------------------------------------------------------- eigensystem
template <typename I, typename... Ts>
inline double eigensystem(I idx_, Ts... a_flat_args) {
    constexpr int M = sizeof...(a_flat_args);
    constexpr int N = zoomy_uf::side_of_square(M);
    static_assert(N * N == M && N > 0, "eigensystem: a square matrix is expected");
    const double a[M] = {static_cast<double>(a_flat_args)...};
    const int idx = static_cast<int>(idx_);
    double w[N], R[N * N], L[N * N];
    bool ok = zoomy_uf::all_finite(a, M) && zoomy_uf::eig_real<N>(a, w);
    for (int k = 0; ok && k < N; ++k) {   // column k of R: the eigenvector of w[k]
        double v[N];
        ok = zoomy_uf::eigvec<N>(a, w[k], v);
        for (int i = 0; ok && i < N; ++i) R[i * N + k] = v[i];
    }
    if (ok) {
        for (int j = 0; j < N; ++j) {   // L = R^-1 column by column, else the pseudo-inverse
            double e[N], x[N];
            for (int i = 0; i < N; ++i) e[i] = i == j ? 1.0 : 0.0;
            if (!zoomy_uf::gepp_solve<N>(R, e, x) || !zoomy_uf::all_finite(x, N)) {
                zoomy_uf::pinv<N>(R, L);
                break;
            }
            for (int i = 0; i < N; ++i) L[i * N + j] = x[i];
        }
    }
    if (!ok) {   // +inf eigenvalues, R = L = I (UserFunctions.H:150-166)
        if (idx < N) return __builtin_huge_val();
        const int e = (idx - N) % (N * N);
        return e / N == e % N ? 1.0 : 0.0;
    }
    if (idx < N) return w[idx];
    if (idx < N + N * N) return R[idx - N];
    return L[idx - N - N * N];
}

// ---------------------------------------------------------------------------- solve
template <typename I, typename... Ts>
inline double solve(I idx_, Ts... args) {
    constexpr int M = sizeof...(args);
    constexpr int N = zoomy_uf::side_of_system(M);
    static_assert(N * N + N == M && N > 0, "solve: a square matrix and a right-hand side are expected");
    const double v[M] = {static_cast<double>(args)...};
    double x[N];
    if (!zoomy_uf::gepp_solve<N>(v, v + N * N, x)) return __builtin_nan("");
    return x[static_cast<int>(idx_)];
}
