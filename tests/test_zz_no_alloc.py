"""zdm_bsr2_gmres_noalloc (device arrays, library-held scratch, no allocation) against
zdm_bsr2_gmres (host arrays, allocating) through the C ABI: the same kernels on the same system
must give bitwise the same solution, iteration count and converged reason, on two calls in a
row, the second on a smaller system, so the second call runs on scratch the first has left."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, cases, zdm
from test_gpu_chorin import csr_to_bsr
from test_oracle_vam import c3_small, read_petsc_mat

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


def _lib():
    lib = zdm.model_lib("vam")
    lib.zdm_bsr2_gmres_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, _ip, _ip]
    lib.zdm_bsr2_gmres_noalloc.restype = ctypes.c_int
    return lib


def _noalloc(lib, rp, col, val, b, x0, krylov):
    dev = torch.device("cuda:0")
    t_rp = torch.from_numpy(np.ascontiguousarray(rp, dtype=np.int32)).to(dev)
    t_col = torch.from_numpy(np.ascontiguousarray(col, dtype=np.int32)).to(dev)
    t_val = torch.from_numpy(np.ascontiguousarray(val, dtype=np.float64).reshape(-1)).to(dev)
    t_b = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float64)).to(dev)
    t_x = torch.from_numpy(np.array(x0, dtype=np.float64)).to(dev)
    stream = torch.cuda.current_stream(dev)
    torch.cuda.synchronize(dev)
    its, reason = ctypes.c_int(-1), ctypes.c_int(99)
    rc = lib.zdm_bsr2_gmres_noalloc(len(rp) - 1, t_rp.data_ptr(), t_col.data_ptr(), t_val.data_ptr(),
                                    t_b.data_ptr(), t_x.data_ptr(), ctypes.addressof(krylov),
                                    ctypes.c_void_p(stream.cuda_stream), ctypes.byref(its),
                                    ctypes.byref(reason))
    msg = lib.zdm_last_error().decode() if rc else ""
    torch.cuda.synchronize(dev)
    return rc, msg, t_x.cpu().numpy(), its.value, reason.value


def _pressure_system(n=40, dt=2e-3):
    """The assembled VAM pressure matrix of a jittered n x n grid (2 n^2 block rows)."""
    from conftest import oracle
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=3, renumber=True)
    X, _ = c3_small(mesh, oracle("vam"))
    s = zdm.Solver("vam", mesh, order=1, cfl=0.45, t_end=1e9)
    s.upload(X)
    s.vam_frozen_aux()
    rp, col, val = s.vam_pmat(dt)
    s.close()
    return rp, col, val


def _check_pair(lib, rp, col, val, b, krylov, label):
    x_ref, its_ref, reason_ref = zdm.bsr2_gmres(rp, col, val, b, krylov=krylov, model="vam")
    rc, msg, x, its, reason = _noalloc(lib, rp, col, val, b, np.zeros_like(b), krylov)
    print(label, "block rows", len(rp) - 1, "its", its_ref, its, "reason", reason_ref, reason, flush=True)
    assert rc == 0, msg
    assert reason == reason_ref and its == its_ref
    assert np.array_equal(x, x_ref)


def test_noalloc_gmres_matches_allocating_entry_point_twice():
    lib = _lib()
    kry = zdm.KrylovOptions(restart=30, max_it=400, rtol=1e-10, atol=1e-50)
    rng = np.random.default_rng(5)
    # first call: a 3 200-block-row pressure system
    rp, col, val = _pressure_system(40)
    b = rng.standard_normal(2 * (len(rp) - 1))
    _check_pair(lib, rp, col, val, b, kry, "pressure system")
    # second call, smaller: the reference's 60-block-row pmat.bin, on the scratch left behind
    A = read_petsc_mat(os.path.join(ROOT, "tests", "golden", "pmat.bin"))
    rp2, col2, val2 = csr_to_bsr(A)
    b2 = rng.standard_normal(120)
    _check_pair(lib, rp2, col2, val2, b2, kry, "pmat.bin")
    # and a capped solve (reason -3) on it
    kry3 = zdm.KrylovOptions(restart=30, max_it=7, rtol=1e-30, atol=1e-50)
    _check_pair(lib, rp2, col2, val2, b2, kry3, "pmat.bin capped")


def test_noalloc_gmres_rejects_what_does_not_fit_or_is_malformed():
    lib = _lib()
    A = read_petsc_mat(os.path.join(ROOT, "tests", "golden", "pmat.bin"))
    rp, col, val = csr_to_bsr(A)
    b = np.ones(120)
    # a system beyond the scratch size is refused on the host, before any launch
    big = zdm.KrylovOptions(restart=200, max_it=10, rtol=1e-10, atol=1e-50)
    dev = torch.device("cuda:0")
    t = torch.zeros(8, dtype=torch.float64, device=dev)
    its, reason = ctypes.c_int(), ctypes.c_int()
    rc = lib.zdm_bsr2_gmres_noalloc(1 << 20, t.data_ptr(), t.data_ptr(), t.data_ptr(), t.data_ptr(),
                                    t.data_ptr(), ctypes.addressof(big), None, ctypes.byref(its),
                                    ctypes.byref(reason))
    assert rc != 0 and b"scratch" in lib.zdm_last_error()
    # decreasing row pointers and an out-of-range column are found on the device
    kry = zdm.KrylovOptions(restart=30, max_it=10, rtol=1e-10, atol=1e-50)
    bad_rp = rp.copy()
    bad_rp[10] = bad_rp[11] + 1
    rc, msg, _, _, _ = _noalloc(lib, bad_rp, col, val, b, np.zeros(120), kry)
    assert rc != 0 and "row pointers" in msg
    bad_col = col.copy()
    bad_col[5] = 60
    rc, msg, _, _, _ = _noalloc(lib, rp, bad_col, val, b, np.zeros(120), kry)
    assert rc != 0 and "column" in msg
    # and the entry point still solves afterwards
    _check_pair(lib, rp, col, val, b, kry, "after rejections")
