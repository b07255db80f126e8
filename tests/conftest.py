import ctypes
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "zoomy-dmplex_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")


def load_package():
    if "zoomy_dmplex_amd" in sys.modules:
        return sys.modules["zoomy_dmplex_amd"]
    spec = importlib.util.spec_from_file_location(
        "zoomy_dmplex_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["zoomy_dmplex_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


zdm = load_package()
import zoomy_dmplex_amd.cases as cases  # noqa: E402

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def dp(a):
    return None if a is None else a.ctypes.data_as(_dp)


def ip(a):
    return a.ctypes.data_as(_ip)


class Oracle:
    """ctypes view of oracle/liboracle_<model>.so (the CPU checker)."""

    def __init__(self, model):
        path = os.path.join(ORACLE_DIR, "liboracle_%s.so" % model)
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: %s (run oracle/build_oracle.py)" % path)
        L = ctypes.CDLL(path)
        M = ctypes.POINTER(zdm._ZdmMesh)
        L.orc_model_info.argtypes = [_ip] * 4
        L.orc_default_params.argtypes = [_dp]
        L.orc_gradient.argtypes = [M, _ip, _dp, ctypes.c_int, _dp, _dp]
        L.orc_bounds.argtypes = [M, _dp, _dp, _dp]
        L.orc_reconstruct.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.orc_form_rhs.argtypes = [M, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.orc_flux.argtypes = [_dp] * 8
        L.orc_eigen.argtypes = [_dp] * 5
        L.orc_update_state.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_enforce.argtypes = [M, _dp, _dp, _dp]
        L.orc_set_refresh.argtypes = [ctypes.c_int]
        L.orc_compute_dt.argtypes = [M, _dp, _dp, _dp, ctypes.c_double, _ip]
        L.orc_compute_dt.restype = ctypes.c_double
        L.orc_source.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp, _dp]
        L.orc_solve_cell.argtypes = [ctypes.c_double, _dp, _dp, _dp]
        L.orc_gepp.argtypes = [_dp, _dp, _dp, ctypes.c_int]
        L.orc_ssp_linear.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4
        L.orc_ssp_linear.restype = ctypes.c_double
        L.orc_ts_step.argtypes = [M, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  _dp, _dp]
        L.orc_set_positivity.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_last_troubled.argtypes = [ctypes.c_int, _ip]
        L.orc_wb_faces.argtypes = [M, _ip, _dp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_troubled_count.restype = ctypes.c_longlong
        L.orc_run.argtypes = [M, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        if hasattr(L, "orc_vam_info"):
            L.orc_vam_info.argtypes = [_ip] * 3
            L.orc_deriv1.argtypes = [M, _dp, ctypes.c_int, _dp]
            L.orc_fill_ca.argtypes = [M, ctypes.c_int, _dp, _dp, _dp]
            L.orc_pressure_residual.argtypes = [M, _dp, ctypes.c_double, _dp, _dp, _dp, _dp]
            L.orc_assemble_pmat.argtypes = [M, _dp, ctypes.c_double, _dp, _dp, ctypes.c_int,
                                            _ip, _ip, _dp]
            L.orc_csr_gmres.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip]
            L.orc_chorin_run.argtypes = [M, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp,
                                         _ip, _ip]
        self.L = L
        v = [ctypes.c_int() for _ in range(4)]
        L.orc_model_info(*[ctypes.byref(x) for x in v])
        self.n_q, self.n_aux, self.n_params, self.dim = [x.value for x in v]
        p = (ctypes.c_double * self.n_params)()
        L.orc_default_params(p)
        self.default_params = np.array(p[:])

    def flux(self, qL, qR, aL, aR, p, n):
        f = np.zeros(self.n_q)
        d = np.zeros(2 * self.n_q)
        args = [np.ascontiguousarray(x, dtype=np.float64) for x in (qL, qR, aL, aR, p, n)]
        self.L.orc_flux(*[dp(a) for a in args], dp(f), dp(d))
        return f, d

    def eigen(self, q, a, p, n):
        out = np.zeros(1)
        args = [np.ascontiguousarray(x, dtype=np.float64) for x in (q, a, p, n)]
        self.L.orc_eigen(*[dp(x) for x in args], dp(out))
        return out[0]

    def form_rhs(self, mesh, face_bc, p, X, A, order=2, use_limiters=1, gradient=0, t=0.0,
                 well_balanced=0, positivity="none", wet_dry_eps=1e-2, refresh_derivative_aux=True):
        F = np.zeros_like(X)
        self.L.orc_set_positivity(self.POSITIVITY[positivity], float(wet_dry_eps))
        self.L.orc_set_refresh(int(refresh_derivative_aux))
        try:
            self.L.orc_form_rhs(mesh.handle, ip(face_bc), dp(p), order, use_limiters, gradient,
                                int(well_balanced), t, dp(X), dp(A), dp(F))
        finally:
            self.L.orc_set_positivity(0, 1e-2)
            self.L.orc_set_refresh(1)
        return F

    def wb_faces(self, mesh, face_bc, p, X, A, skip_theta, wet_dry_eps=1e-2):
        """eta-WB face states [ncells, max_faces, n_q] and effective gradients."""
        QF = np.zeros((mesh.ncells, mesh.max_faces, self.n_q))
        G = np.zeros((mesh.ncells, self.n_q, self.dim))
        self.L.orc_wb_faces(mesh.handle, ip(face_bc), dp(p), int(skip_theta), float(wet_dry_eps),
                            dp(np.ascontiguousarray(X)), dp(np.ascontiguousarray(A)), dp(QF), dp(G))
        return QF, G

    def gradient(self, mesh, face_bc, p, X, method=0):
        G = np.zeros((mesh.ncells, self.n_q, self.dim))
        self.L.orc_gradient(mesh.handle, ip(face_bc), dp(p), method, dp(X), dp(G))
        return G

    def bounds(self, mesh, X):
        lo, hi = np.zeros_like(X), np.zeros_like(X)
        self.L.orc_bounds(mesh.handle, dp(X), dp(lo), dp(hi))
        return lo, hi

    def update_state(self, p, X, A):
        self.L.orc_update_state(X.shape[0], dp(p), dp(X), dp(A))

    def enforce(self, mesh, p, X, A, refresh_derivative_aux=True):
        """EnforcePhysicalConstraints: RefreshDerivativeAux (models with the field-level
        update_aux_variables) + clamp + aux."""
        self.L.orc_set_refresh(int(refresh_derivative_aux))
        try:
            if self.L.orc_enforce(mesh.handle, dp(p), dp(X), dp(A)) != 0:
                raise RuntimeError("oracle: compute_derivative order not supported")
        finally:
            self.L.orc_set_refresh(1)

    def source(self, p, dt, X, A):
        self.L.orc_source(X.shape[0], dp(p), float(dt), dp(X), dp(A))

    def compute_dt(self, mesh, p, X, A, cfl):
        fin = ctypes.c_int()
        return self.L.orc_compute_dt(mesh.handle, dp(p), dp(X), dp(A), cfl, ctypes.byref(fin))

    POSITIVITY = {"none": 0, "mood": 1, "zhang_shu": 2}

    def run(self, mesh, face_bc, p, X, A, order=2, ssp_stages=2, use_limiters=1, gradient=0,
            cfl=0.45, t_end=1e9, min_dt=1e-12, max_steps=10, time=0.0, dt=0.0, first=True,
            well_balanced=0, positivity="none", wet_dry_eps=1e-2, time_integration="splitting",
            refresh_derivative_aux=True):
        """MUSCLSolver loop; returns (steps, time, next dt).  self.troubled = MOOD overrides."""
        t = np.array([time])
        d = np.array([dt])
        self.L.orc_set_positivity(self.POSITIVITY[positivity], float(wet_dry_eps))
        self.L.orc_set_time_integration({"splitting": 0, "imex": 1, "implicit": 2}[time_integration])
        self.L.orc_set_refresh(int(refresh_derivative_aux))
        try:
            n = self.L.orc_run(mesh.handle, ip(face_bc), dp(p), order, ssp_stages, use_limiters,
                               gradient, int(well_balanced), cfl, t_end, min_dt, max_steps, dp(X),
                               dp(A), dp(t), dp(d), int(first))
            self.troubled = int(self.L.orc_troubled_count())
            mask = np.zeros(mesh.ncells, dtype=np.int32)
            self.L.orc_last_troubled(mesh.ncells, ip(mask))
            self.last_troubled = mask.astype(bool)
        finally:
            self.L.orc_set_positivity(0, 1e-2)
            self.L.orc_set_time_integration(0)
            self.L.orc_set_refresh(1)
        return n, t[0], d[0]


    # ---- Chorin VAM (liboracle_vam.so only)
    def vam_info(self):
        v = [ctypes.c_int() for _ in range(3)]
        self.L.orc_vam_info(*[ctypes.byref(x) for x in v])
        return [x.value for x in v]

    def deriv1(self, mesh, f, axis):
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros(mesh.ncells)
        self.L.orc_deriv1(mesh.handle, dp(f), axis, dp(out))
        return out

    def fill_ca(self, mesh, want_p, X, P, CA):
        self.L.orc_fill_ca(mesh.handle, int(want_p), dp(X), dp(P), dp(CA))

    def pressure_residual(self, mesh, p, dt, X, P, CA):
        out = np.zeros_like(P)
        self.L.orc_pressure_residual(mesh.handle, dp(p), dt, dp(X), dp(P), dp(CA), dp(out))
        return out

    def assemble_pmat(self, mesh, p, dt, X, CA):
        """scipy CSR of the assembled pressure matrix (scalar rows c*NP + k)."""
        import scipy.sparse as sp
        n = mesh.ncells * 2
        rp = np.zeros(n + 1, dtype=np.int32)
        z = np.zeros(1, dtype=np.int32)
        nnz = self.L.orc_assemble_pmat(mesh.handle, dp(p), dt, dp(X), dp(CA), 0, ip(rp), ip(z),
                                       dp(np.zeros(1)))
        ci = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz)
        self.L.orc_assemble_pmat(mesh.handle, dp(p), dt, dp(X), dp(CA), nnz, ip(rp), ip(ci), dp(v))
        return sp.csr_matrix((v, ci, rp), shape=(n, n))

    def csr_gmres(self, A, b, restart=30, max_it=400, rtol=1e-10, atol=1e-50, x0=None):
        rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        ci = np.ascontiguousarray(A.indices, dtype=np.int32)
        v = np.ascontiguousarray(A.data, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(A.shape[0]) if x0 is None else np.array(x0, dtype=np.float64)
        its = ctypes.c_int()
        reason = self.L.orc_csr_gmres(A.shape[0], ip(rp), ip(ci), dp(v), dp(b), dp(x), restart,
                                      max_it, rtol, atol, ctypes.byref(its))
        return x, reason, its.value

    def chorin_run(self, mesh, face_bc, p, X, A, P, order=1, use_limiters=1, cfl=0.45, t_end=1e9,
                   min_dt=1e-12, max_steps=1, restart=30, max_it=400, rtol=1e-10, atol=1e-50,
                   op_mode=0, time=0.0):
        t = np.array([time])
        its = ctypes.c_int()
        reason = ctypes.c_int()
        n = self.L.orc_chorin_run(mesh.handle, ip(face_bc), dp(p), order, use_limiters, cfl, t_end,
                                  min_dt, max_steps, restart, max_it, rtol, atol, op_mode, dp(X),
                                  dp(A), dp(P), dp(t), ctypes.byref(its), ctypes.byref(reason))
        return n, t[0], its.value, reason.value


    def bdf_coeffs(self, h, h1, have_prev):
        c = np.zeros(3)
        self.L.orc_bdf_coeffs(ctypes.c_double(h), ctypes.c_double(h1), int(have_prev), dp(c))
        return c

    def bdf_last_residual(self):
        """(|G| after the last implicit step's Newton solve, its Newton iterations)."""
        self.L.orc_bdf_last_residual.restype = ctypes.c_double
        its = ctypes.c_int()
        r = self.L.orc_bdf_last_residual(ctypes.byref(its))
        return r, its.value

    def ark2_linear(self, le, li, dt, n, y0=1.0):
        """The oracle's ARK2 step on y' = le y (explicit) + li y (implicit), n steps."""
        self.L.orc_ark2_linear.restype = ctypes.c_double
        self.L.orc_ark2_linear.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_double]
        return self.L.orc_ark2_linear(le, li, dt, n, y0)

    def chorin_loss(self, mesh, face_bc, p, X0, A0, dts, gauges, obs_every, obs, order=1, use_limiters=1,
                    restart=30, max_it=3000, rtol=1e-14, atol=1e-30):
        """Calibration loss over Chorin steps with frozen dts (assembled pressure operator)."""
        L = self.L
        L.orc_chorin_loss.restype = ctypes.c_double
        L.orc_chorin_loss.argtypes = [ctypes.POINTER(zdm._ZdmMesh), _ip, _dp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _ip, ctypes.c_int, _dp,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        g = np.ascontiguousarray(gauges, dtype=np.int32)
        c = lambda a: dp(np.ascontiguousarray(a, dtype=np.float64))
        return L.orc_chorin_loss(mesh.handle, ip(face_bc), dp(p), order, use_limiters, len(dts), c(dts), c(X0),
                                 c(A0), len(g), ip(g), obs_every, c(obs), restart, max_it, rtol, atol)

    # ---- calibration loss with frozen time steps (liboracle_<model>.so)
    def loss(self, mesh, face_bc, p, X0, A0, dts, gauges, obs_every, obs, order=2, ssp_stages=2,
             use_limiters=1):
        L = self.L
        L.orc_loss.restype = ctypes.c_double
        L.orc_loss.argtypes = [ctypes.POINTER(zdm._ZdmMesh), _ip, _dp, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _ip,
                               ctypes.c_int, _dp]
        g = np.ascontiguousarray(gauges, dtype=np.int32)
        return L.orc_loss(mesh.handle, ip(face_bc), dp(p), order, ssp_stages, use_limiters, len(dts),
                          dp(np.ascontiguousarray(dts, dtype=np.float64)), dp(np.ascontiguousarray(X0)),
                          dp(np.ascontiguousarray(A0)), len(g), ip(g), obs_every,
                          dp(np.ascontiguousarray(obs, dtype=np.float64)))


class TangentOracle:
    """liboracle_<model>_tl.so: the oracle on one-direction dual numbers (tangent-linear run)."""

    def __init__(self, model):
        path = os.path.join(ORACLE_DIR, "liboracle_%s_tl.so" % model)
        if not os.path.exists(path):
            raise RuntimeError("tangent oracle not built: %s" % path)
        L = ctypes.CDLL(path)
        L.orc_tl_loss.argtypes = [ctypes.POINTER(zdm._ZdmMesh), _ip, _dp, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                  _ip, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.orc_set_loss_source.argtypes = [ctypes.c_int]
        L.orc_set_loss_source.restype = None
        self.L = L

    def loss(self, mesh, face_bc, p, X0, dX0, A0, dA0, dts, gauges, obs_every, obs, order=2,
             ssp_stages=2, use_limiters=1, freeze_limiter=False, source=True):
        """(loss, directional derivative along (dX0, dA0)); freeze_limiter holds each face's
        limiter value fixed (the frozen-limiter linearisation); source=False drops the
        SourceStep (the solver's source setting)."""
        g = np.ascontiguousarray(gauges, dtype=np.int32)
        out, dout = np.zeros(1), np.zeros(1)
        c = lambda a: dp(np.ascontiguousarray(a, dtype=np.float64)) if a is not None else None
        self.L.orc_set_loss_source(int(bool(source)))
        try:
            self.L.orc_tl_loss(mesh.handle, ip(face_bc), dp(p), order, ssp_stages, use_limiters, len(dts),
                               c(dts), c(X0), c(dX0), c(A0), c(dA0), len(g), ip(g), obs_every, c(obs),
                               int(freeze_limiter), dp(out), dp(dout))
        finally:
            self.L.orc_set_loss_source(1)
        return out[0], dout[0]

    def chorin_loss(self, mesh, face_bc, p, X0, dX0, A0, dA0, dts, gauges, obs_every, obs, order=1,
                    use_limiters=1, freeze_limiter=False, restart=30, max_it=3000, rtol=1e-14, atol=1e-30):
        """(loss, directional derivative) over Chorin steps with frozen dts; the pressure's
        tangent is that of the exact solve."""
        L = self.L
        L.orc_tl_chorin_loss.argtypes = [ctypes.POINTER(zdm._ZdmMesh), _ip, _dp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _ip, ctypes.c_int,
                                         _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, _dp, _dp]
        g = np.ascontiguousarray(gauges, dtype=np.int32)
        out, dout = np.zeros(1), np.zeros(1)
        c = lambda a: dp(np.ascontiguousarray(a, dtype=np.float64)) if a is not None else None
        L.orc_tl_chorin_loss(mesh.handle, ip(face_bc), dp(p), order, use_limiters, len(dts), c(dts), c(X0), c(dX0),
                             c(A0), c(dA0), len(g), ip(g), obs_every, c(obs), int(freeze_limiter), restart,
                             max_it, rtol, atol, dp(out), dp(dout))
        return out[0], dout[0]


_oracles = {}


def oracle(model):
    if model not in _oracles:
        _oracles[model] = Oracle(model)
    return _oracles[model]


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pkg():
    return zdm


@pytest.fixture(scope="session")
def case_mod():
    return cases
