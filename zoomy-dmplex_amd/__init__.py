"""zoomy-dmplex_amd: MI355X-native finite-volume time step for SME / VAM shallow-flow models.

Python host layer over the C ABI (include/zdm.h, include/zdm_mesh.h).  It mirrors the
reference's host objects:

  * Settings            -- Settings.hpp:13-138 (JSON keys and defaults)
  * Mesh                -- the DMPlex services the step uses (geometry, adjacency, tags)
  * boundary_face_bc    -- VirtualSolver.hpp:394-417, 500-511 (tag -> model BC index)
  * Solver              -- MUSCLSolver::Run (MUSCLSolver.hpp:22-194) on the device

The compute path is the HIP library libzdm_<model>.so built in-tree (see build.py); there
is no fallback: a missing library raises.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)


class ZdmError(RuntimeError):
    pass


_torch_first = False


def _bind_runtime():
    """Load PyTorch (when installed) before any HIP library of this package.

    PyTorch-ROCm ships its own HIP runtime with the same soname as the system's
    (libamdhip64.so.7).  Loaded first, it also serves this package's libraries, so one HIP /
    HSA runtime runs the process and torch tensors and these kernels share the device.  Loaded
    after a library that pulled in the system runtime, torch would bring a second runtime that
    finds no device.  ZDM_NO_TORCH=1 skips this (processes that never use torch)."""
    global _torch_first
    if _torch_first or os.environ.get("ZDM_NO_TORCH") == "1":
        return
    _torch_first = True
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def _load(name):
    # ZDM_LIB_DIR selects another build directory of the same libraries (e.g. a tuned build)
    _bind_runtime()
    path = os.path.join(os.environ.get("ZDM_LIB_DIR", HERE), name)
    if not os.path.exists(path):
        raise ZdmError("%s is not built (run zoomy-dmplex_amd/build.py or __graft_entry__.build())"
                       % path)
    return ctypes.CDLL(path)


# --------------------------------------------------------------------------- mesh service

class _ZdmMesh(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("ncells", ctypes.c_int), ("nfaces", ctypes.c_int),
        ("max_faces", ctypes.c_int), ("nverts", ctypes.c_int), ("max_cell_verts", ctypes.c_int),
        ("cell_centroid", _c_dbl_p), ("cell_volume", _c_dbl_p), ("cell_nfaces", _c_int_p),
        ("cell_faces", _c_int_p), ("cell_nbrs", _c_int_p), ("cell_perm", _c_int_p),
        ("cell_verts", _c_int_p), ("face_cells", _c_int_p), ("face_normal", _c_dbl_p),
        ("face_centroid", _c_dbl_p), ("face_tag", _c_int_p), ("verts", _c_dbl_p),
        ("nphysical", ctypes.c_int), ("physical_tag", _c_int_p), ("physical_dim", _c_int_p),
        ("physical_names", ctypes.c_void_p),
    ]


_mesh_lib = None


def mesh_lib():
    global _mesh_lib
    if _mesh_lib is None:
        lib = _load("libzdm_mesh.so")
        P = ctypes.POINTER(ctypes.POINTER(_ZdmMesh))
        lib.zdm_mesh_tri_grid.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_uint64, ctypes.c_int, P]
        lib.zdm_mesh_line.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, P]
        lib.zdm_mesh_read_msh.argtypes = [ctypes.c_char_p, ctypes.c_int, P]
        lib.zdm_mesh_from_cells.argtypes = [ctypes.c_int, ctypes.c_int, _c_dbl_p, ctypes.c_int,
                                            ctypes.c_int, _c_int_p, ctypes.c_int, _c_int_p,
                                            _c_int_p, ctypes.c_int, ctypes.c_int, P]
        lib.zdm_mesh_destroy.argtypes = [ctypes.POINTER(_ZdmMesh)]
        lib.zdm_mesh_concat.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.POINTER(_ZdmMesh)),
                                        _c_int_p, P]
        lib.zdm_mesh_last_error.restype = ctypes.c_char_p
        lib.zdm_splitmix64_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        lib.zdm_splitmix64_next.restype = ctypes.c_uint64
        lib.zdm_h5_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)]
        lib.zdm_h5_write.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_uint64)]
        _mesh_lib = lib
    return _mesh_lib


def _renum(r):
    if isinstance(r, str):
        return {"none": 0, "rcm": 1, "hilbert": 2}[r]
    return int(r)


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Mesh:
    """Finite-volume mesh in solver order (cells RCM-renumbered, faces by support)."""

    def __init__(self, ptr):
        self._ptr = ptr
        m = ptr.contents
        self.dim, self.ncells, self.nfaces = m.dim, m.ncells, m.nfaces
        self.max_faces = m.max_faces
        nc, nf, mf = self.ncells, self.nfaces, self.max_faces
        self.cell_centroid = _arr(m.cell_centroid, nc * 3, np.float64).reshape(nc, 3)
        self.cell_volume = _arr(m.cell_volume, nc, np.float64)
        self.cell_nfaces = _arr(m.cell_nfaces, nc, np.int32)
        self.cell_faces = _arr(m.cell_faces, nc * mf, np.int32).reshape(nc, mf)
        self.cell_nbrs = _arr(m.cell_nbrs, nc * mf, np.int32).reshape(nc, mf)
        self.cell_perm = _arr(m.cell_perm, nc, np.int32)
        self.face_cells = _arr(m.face_cells, nf * 2, np.int32).reshape(nf, 2)
        self.face_normal = _arr(m.face_normal, nf * 3, np.float64).reshape(nf, 3)
        self.face_centroid = _arr(m.face_centroid, nf * 3, np.float64).reshape(nf, 3)
        self.face_tag = _arr(m.face_tag, nf, np.int32)
        self.physical = []
        if m.nphysical > 0:
            raw = ctypes.string_at(m.physical_names, 64 * m.nphysical)
            tags = _arr(m.physical_tag, m.nphysical, np.int32)
            dims = _arr(m.physical_dim, m.nphysical, np.int32)
            for i in range(m.nphysical):
                name = raw[64 * i:64 * (i + 1)].split(b"\0", 1)[0].decode()
                self.physical.append((name, int(dims[i]), int(tags[i])))

    @property
    def handle(self):
        return self._ptr

    def __del__(self):
        try:
            if self._ptr:
                mesh_lib().zdm_mesh_destroy(self._ptr)
                self._ptr = None
        except Exception:
            pass

    @staticmethod
    def _make(fn, *args):
        out = ctypes.POINTER(_ZdmMesh)()
        rc = fn(*args, ctypes.byref(out))
        if rc != 0:
            raise ZdmError(mesh_lib().zdm_mesh_last_error().decode())
        return Mesh(out)

    @classmethod
    def tri_grid(cls, nx, ny=None, jitter=0.0, seed=0, renumber=True):
        """Unit square, nx x ny squares split SW->NE into right triangles; walls (tag 1).
        renumber: False/0 = none, True/1/"rcm" = reverse Cuthill-McKee, 2/"hilbert"."""
        return cls._make(mesh_lib().zdm_mesh_tri_grid, nx, ny or nx, float(jitter),
                         ctypes.c_uint64(seed), _renum(renumber))

    @classmethod
    def line(cls, n, x0=0.0, x1=1.0, renumber=False):
        return cls._make(mesh_lib().zdm_mesh_line, n, float(x0), float(x1), _renum(renumber))

    @classmethod
    def read_msh(cls, path, renumber=True):
        """Gmsh MSH 2.2 (the reference's mesh format, VirtualSolver.hpp:388)."""
        return cls._make(mesh_lib().zdm_mesh_read_msh, path.encode(), _renum(renumber))

    @classmethod
    def concat(cls, meshes):
        """Disjoint union (an ensemble); returns (mesh, cell_offsets[n+1])."""
        n = len(meshes)
        arr = (ctypes.POINTER(_ZdmMesh) * n)(*[m.handle for m in meshes])
        off = np.zeros(n + 1, dtype=np.int32)
        out = ctypes.POINTER(_ZdmMesh)()
        rc = mesh_lib().zdm_mesh_concat(n, arr, off.ctypes.data_as(_c_int_p), ctypes.byref(out))
        if rc != 0:
            raise ZdmError(mesh_lib().zdm_mesh_last_error().decode())
        return Mesh(out), off

    @classmethod
    def from_cells(cls, dim, xyz, cells, bfaces=None, btags=None, mark_all_tag=1, renumber=True):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        nbf = 0 if bfaces is None else len(btags)
        bf = np.ascontiguousarray(bfaces if bfaces is not None else np.zeros(1), dtype=np.int32)
        bt = np.ascontiguousarray(btags if btags is not None else np.zeros(1), dtype=np.int32)
        return cls._make(mesh_lib().zdm_mesh_from_cells, dim, xyz.shape[0],
                         xyz.ctypes.data_as(_c_dbl_p), cells.shape[0], cells.shape[1],
                         cells.ctypes.data_as(_c_int_p), nbf, bf.ctypes.data_as(_c_int_p),
                         bt.ctypes.data_as(_c_int_p), mark_all_tag, _renum(renumber))


def boundary_face_bc(mesh, model_tags, model_dim=None):
    """Model BC index per face (-1 = interior or no model BC for its tag).

    VirtualSolver.hpp:394-417: every model tag except "default" must be a physical name of
    the boundary dimension, else the run is refused.  VirtualSolver.hpp:500-511: a tag name
    found in the mesh maps its physical tag; "default" otherwise maps tag 1.
    """
    model_dim = mesh.dim if model_dim is None else model_dim
    mesh_map = {name: (d, t) for (name, d, t) in mesh.physical}
    for name in model_tags:
        if name == "default":
            continue
        if name not in mesh_map or mesh_map[name][0] != model_dim - 1:
            raise ZdmError("GMSH boundaries must exactly match model tags (excluding 'default').")
    bmap = {}
    for i, name in enumerate(model_tags):
        tag = -1
        if name in mesh_map:
            tag = mesh_map[name][1]
        elif name == "default":
            tag = 1
        if tag != -1:
            bmap[tag] = i
    out = np.full(mesh.nfaces, -1, dtype=np.int32)
    bnd = mesh.face_cells[:, 1] < 0
    for f in np.nonzero(bnd)[0]:
        out[f] = bmap.get(int(mesh.face_tag[f]), -1)
    return out


# ------------------------------------------------------------------------------- settings

class Settings:
    """JSON settings with the reference's keys and defaults (Settings.hpp:13-138).  Unknown
    keys are ignored; extension keys: solver.gradient ("green_gauss" | "lsq"),
    solver.well_balanced (bool)."""

    SOLVER_DEFAULTS = dict(t_end=1.0, cfl=0.5, use_deep_adjacency=True, reconstruction_order=1,
                           min_dt=1.0e-12, limiter="venkatakrishnan", time_integration="splitting",
                           method="muscl", refresh_derivative_aux=True, positivity="none",
                           wet_dry_eps=1.0e-2, ssp_stages=5, gradient="green_gauss",
                           well_balanced=False)
    IO_DEFAULTS = dict(directory="output", filename="sol", snapshots=10, snapshot_logic="snap",
                       clean_directory=False, mesh_path="", mesh_label="Face Sets",
                       initial_condition_file="", initial_condition_mask=[], write_3d=False,
                       n_layers_3d=10, output_3d_name="sol_3d")

    def __init__(self, d=None):
        d = d or {}
        self.name = d.get("name", "Simulation")
        self.io = dict(self.IO_DEFAULTS)
        self.io.update({k: v for k, v in d.get("io", {}).items() if k in self.IO_DEFAULTS})
        self.io["initial_condition_mask"] = [int(v) for v in self.io["initial_condition_mask"]
                                             if isinstance(v, int)]
        self.solver = dict(self.SOLVER_DEFAULTS)
        self.solver.update({k: v for k, v in d.get("solver", {}).items()
                            if k in self.SOLVER_DEFAULTS})
        params = d.get("model", {}).get("parameters", {})
        self.parameters = {k: float(v) for k, v in params.items()
                           if isinstance(v, (int, float)) and not isinstance(v, bool)}

    @classmethod
    def from_json(cls, path):
        if not os.path.exists(path):
            raise ZdmError("Could not open settings file: %s" % path)
        with open(path) as f:
            return cls(json.load(f))


# --------------------------------------------------------------------------------- solver

class _Options(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("ssp_stages", ctypes.c_int),
                ("use_limiters", ctypes.c_int), ("gradient", ctypes.c_int),
                ("source", ctypes.c_int), ("cfl", ctypes.c_double), ("t_end", ctypes.c_double),
                ("min_dt", ctypes.c_double), ("block_size", ctypes.c_int),
                ("well_balanced", ctypes.c_int), ("positivity", ctypes.c_int),
                ("wet_dry_eps", ctypes.c_double), ("time_integration", ctypes.c_int),
                ("refresh_derivative_aux", ctypes.c_int)]

POSITIVITY = {"none": 0, "mood": 1, "zhang_shu": 2}
TIME_INTEGRATION = {"splitting": 0, "imex": 1, "implicit": 2}


class NewtonOptions(ctypes.Structure):
    """zdm_newton_options: SNES / KSP settings of time_integration "implicit"."""
    _fields_ = [("snes_atol", ctypes.c_double), ("snes_rtol", ctypes.c_double), ("snes_max_it", ctypes.c_int),
                ("ksp_rtol", ctypes.c_double), ("ksp_restart", ctypes.c_int), ("ksp_max_it", ctypes.c_int)]

    def __init__(self, snes_atol=1e-4, snes_rtol=1e-3, snes_max_it=20, ksp_rtol=1e-5, ksp_restart=30,
                 ksp_max_it=10000):
        super().__init__(snes_atol, snes_rtol, snes_max_it, ksp_rtol, ksp_restart, ksp_max_it)


_model_libs = {}

SOLVER_SYMBOLS = [
    "zdm_model_info", "zdm_model_name", "zdm_model_default_params", "zdm_model_boundary_tags",
    "zdm_model_parameter_names",
    "zdm_solver_create", "zdm_solver_destroy", "zdm_state_upload", "zdm_state_download",
    "zdm_initialize", "zdm_step", "zdm_run", "zdm_get_time", "zdm_synchronize",
    "zdm_transport_rhs", "zdm_source_step", "zdm_enforce_constraints", "zdm_cfl_dt",
    "zdm_state_upload_device", "zdm_state_download_device", "zdm_enable_timing",
    "zdm_kernel_time", "zdm_stream_copy", "zdm_last_error",
    "zdm_chorin_setup", "zdm_chorin_step", "zdm_chorin_pressure", "zdm_chorin_stats", "zdm_vam_frozen_aux",
    "zdm_vam_residual", "zdm_vam_pmat", "zdm_bsr2_gmres", "zdm_batch_create", "zdm_batch_time",
    "zdm_step_graph", "zdm_adjoint_setup", "zdm_adjoint_forward", "zdm_adjoint_gradient",
    "zdm_adjoint_dts", "zdm_adjoint_residuals", "zdm_adjoint_layout", "zdm_troubled_cells",
    "zdm_set_launch", "zdm_get_launch", "zdm_autotune", "zdm_adjoint_setup_chorin",
    "zdm_adjoint_krylov_its", "zdm_implicit_setup", "zdm_implicit_stats", "zdm_adjoint_stretch_tape",
    "zdm_adjoint_forward_device", "zdm_adjoint_gradient_device", "zdm_solver_stream", "zdm_fuse_post",
    "zdm_transport_rhs_device", "zdm_adjoint_forward_frozen_alpha", "zdm_flux_nontemporal",
]


class KrylovOptions(ctypes.Structure):
    """zdm_krylov_options: GMRES(restart) + point-block Jacobi of the pressure solve.
    Defaults are the reference's KSP settings (ChorinVAM.hpp:403, 449)."""
    _fields_ = [("restart", ctypes.c_int), ("max_it", ctypes.c_int), ("rtol", ctypes.c_double),
                ("atol", ctypes.c_double), ("dtol", ctypes.c_double), ("op_mode", ctypes.c_int)]

    def __init__(self, restart=40, max_it=400, rtol=1e-9, atol=1e-12, dtol=1e5, op_mode=0):
        super().__init__(restart, max_it, rtol, atol, dtol, op_mode)


def model_lib(model):
    if model in _model_libs:
        return _model_libs[model]
    lib = _load("libzdm_%s.so" % model)
    V = ctypes.c_void_p
    lib.zdm_model_info.argtypes = [_c_int_p] * 4
    lib.zdm_model_name.restype = ctypes.c_char_p
    lib.zdm_model_default_params.argtypes = [_c_dbl_p]
    lib.zdm_model_boundary_tags.argtypes = [ctypes.c_char_p, ctypes.c_int]
    lib.zdm_model_parameter_names.argtypes = [ctypes.c_char_p, ctypes.c_int]
    lib.zdm_solver_create.argtypes = [ctypes.POINTER(_ZdmMesh), _c_int_p, _c_dbl_p,
                                      ctypes.POINTER(_Options), ctypes.POINTER(V)]
    lib.zdm_solver_destroy.argtypes = [V]
    lib.zdm_state_upload.argtypes = [V, _c_dbl_p, _c_dbl_p]
    lib.zdm_state_download.argtypes = [V, _c_dbl_p, _c_dbl_p]
    lib.zdm_state_upload_device.argtypes = [V, V, V]
    lib.zdm_state_download_device.argtypes = [V, V, V]
    lib.zdm_initialize.argtypes = [V]
    lib.zdm_step.argtypes = [V, ctypes.c_int]
    lib.zdm_run.argtypes = [V, ctypes.c_int, ctypes.c_int, _c_int_p]
    lib.zdm_get_time.argtypes = [V, _c_dbl_p, _c_dbl_p, _c_int_p]
    lib.zdm_synchronize.argtypes = [V]
    lib.zdm_transport_rhs.argtypes = [V, ctypes.c_double, _c_dbl_p, _c_dbl_p, _c_dbl_p]
    lib.zdm_source_step.argtypes = [V, ctypes.c_double]
    lib.zdm_enforce_constraints.argtypes = [V]
    lib.zdm_cfl_dt.argtypes = [V, _c_dbl_p]
    lib.zdm_enable_timing.argtypes = [V, ctypes.c_int]
    lib.zdm_kernel_time.argtypes = [V, ctypes.c_int, _c_dbl_p, _c_int_p]
    lib.zdm_fuse_post.argtypes = [V, ctypes.c_int]
    lib.zdm_flux_nontemporal.argtypes = [V, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.zdm_transport_rhs_device.argtypes = [V, ctypes.c_double, V, V, V, V]
    lib.zdm_adjoint_forward_frozen_alpha.argtypes = [V, _c_dbl_p, _c_dbl_p, _c_dbl_p, ctypes.c_int, _c_dbl_p]
    lib.zdm_stream_copy.argtypes = [ctypes.c_size_t, ctypes.c_int, _c_dbl_p, _c_dbl_p]
    lib.zdm_last_error.restype = ctypes.c_char_p
    lib.zdm_batch_create.argtypes = [ctypes.POINTER(_ZdmMesh), _c_int_p, _c_dbl_p,
                                     ctypes.POINTER(_Options), ctypes.c_int, _c_int_p,
                                     ctypes.POINTER(V)]
    lib.zdm_batch_time.argtypes = [V, ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_int_p]
    lib.zdm_step_graph.argtypes = [V, ctypes.c_int]
    lib.zdm_adjoint_setup.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      _c_int_p, _c_dbl_p, ctypes.c_int]
    lib.zdm_adjoint_setup_chorin.argtypes = lib.zdm_adjoint_setup.argtypes
    lib.zdm_adjoint_krylov_its.argtypes = [V, ctypes.POINTER(ctypes.c_longlong)]
    lib.zdm_implicit_setup.argtypes = [V, ctypes.POINTER(NewtonOptions)]
    lib.zdm_implicit_stats.argtypes = [V, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.zdm_adjoint_forward.argtypes = [V, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]
    lib.zdm_adjoint_gradient.argtypes = [V, _c_dbl_p, _c_dbl_p]
    lib.zdm_adjoint_dts.argtypes = [V, _c_dbl_p]
    lib.zdm_adjoint_residuals.argtypes = [V, _c_dbl_p]
    lib.zdm_adjoint_layout.argtypes = [V, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.zdm_adjoint_stretch_tape.argtypes = [V, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.zdm_adjoint_forward_device.argtypes = [V, V, V, _c_dbl_p, ctypes.POINTER(ctypes.c_double)]
    lib.zdm_adjoint_gradient_device.argtypes = [V, V, V]
    lib.zdm_solver_stream.argtypes = [V, ctypes.POINTER(ctypes.c_void_p)]
    lib.zdm_troubled_cells.argtypes = [V, ctypes.POINTER(ctypes.c_longlong)]
    lib.zdm_set_launch.argtypes = [V, ctypes.c_int, ctypes.c_int]
    lib.zdm_get_launch.argtypes = [V, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.zdm_autotune.argtypes = [V, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                 _c_dbl_p]
    K = ctypes.POINTER(KrylovOptions)
    lib.zdm_chorin_setup.argtypes = [V, K]
    lib.zdm_chorin_step.argtypes = [V, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p]
    lib.zdm_chorin_pressure.argtypes = [V, _c_dbl_p]
    lib.zdm_chorin_stats.argtypes = [V, _c_int_p, _c_int_p, _c_dbl_p, _c_dbl_p]
    lib.zdm_vam_frozen_aux.argtypes = [V, _c_dbl_p]
    lib.zdm_vam_residual.argtypes = [V, ctypes.c_double, _c_dbl_p, _c_dbl_p]
    lib.zdm_vam_pmat.argtypes = [V, ctypes.c_double, ctypes.c_int, _c_int_p, _c_int_p, _c_dbl_p]
    lib.zdm_bsr2_gmres.argtypes = [ctypes.c_int, _c_int_p, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p,
                                   K, _c_int_p, _c_int_p]
    _model_libs[model] = lib
    return lib


class ModelInfo:
    def __init__(self, model):
        lib = model_lib(model)
        v = [ctypes.c_int() for _ in range(4)]
        lib.zdm_model_info(*[ctypes.byref(x) for x in v])
        self.name = model
        self.n_q, self.n_aux, self.n_params, self.dim = [x.value for x in v]
        p = (ctypes.c_double * max(self.n_params, 1))()
        lib.zdm_model_default_params(p)
        self.default_params = np.array(p[:self.n_params])
        buf = ctypes.create_string_buffer(64 * 16)
        n = lib.zdm_model_boundary_tags(buf, 16)
        self.boundary_tags = [buf.raw[64 * i:64 * (i + 1)].split(b"\0", 1)[0].decode()
                              for i in range(n)]
        buf = ctypes.create_string_buffer(64 * 16)
        n = lib.zdm_model_parameter_names(buf, 16)
        self.parameter_names = [buf.raw[64 * i:64 * (i + 1)].split(b"\0", 1)[0].decode()
                                for i in range(n)]


def _dp(a):
    return a.ctypes.data_as(_c_dbl_p)


# ------------------------------------------------------------------- initial conditions

def read_h5(path, dataset="state"):
    """A float64 HDF5 dataset as a numpy array of its own shape (libhdf5 opened at run time,
    see zdm_mesh.h)."""
    lib = mesh_lib()
    nd = ctypes.c_int(0)
    dims = (ctypes.c_uint64 * 4)()
    if lib.zdm_h5_read(os.fsencode(path), dataset.encode(), None, 0, ctypes.byref(nd), dims):
        raise ZdmError(lib.zdm_mesh_last_error().decode())
    shape = tuple(int(dims[i]) for i in range(nd.value))
    out = np.empty(shape, dtype=np.float64)
    if lib.zdm_h5_read(os.fsencode(path), dataset.encode(), out.ctypes.data_as(ctypes.c_void_p), out.size,
                       ctypes.byref(nd), dims):
        raise ZdmError(lib.zdm_mesh_last_error().decode())
    return out


def write_h5(path, array, dataset="state"):
    """One float64 dataset (1 to 4 dimensions) into a new HDF5 file."""
    a = np.ascontiguousarray(array, dtype=np.float64)
    if not 1 <= a.ndim <= 4:
        raise ZdmError("write_h5: 1 to 4 dimensions")
    dims = (ctypes.c_uint64 * 4)(*a.shape)
    lib = mesh_lib()
    if lib.zdm_h5_write(os.fsencode(path), dataset.encode(), a.ctypes.data_as(ctypes.c_void_p), a.ndim, dims):
        raise ZdmError(lib.zdm_mesh_last_error().decode())


def initial_state(settings, mesh, model):
    """VirtualSolver::SetupInitialConditions (VirtualSolver.hpp:234-281): the state in the
    mesh's (solver) cell order, (ncells, n_q).  With io.initial_condition_file it is read as
    IOManager::LoadSolution reads it (IOManager.hpp:122-192): HDF5 dataset "state" of ncells *
    n_q values, cell-major in the mesh file's cell order (flat or (ncells, n_q)), mapped by
    cell_perm; io.initial_condition_mask selects components (the others stay 0).  A .npy of
    shape (ncells, k) is read the same way for k = n_q, and for k < n_q its columns are the
    mask's components in order.  Without a file, Model::initial_condition at the centroids."""
    info = model if isinstance(model, ModelInfo) else ModelInfo(model)
    st = settings if isinstance(settings, Settings) else Settings(settings)
    nc, nq = mesh.ncells, info.n_q
    X = np.zeros((nc, nq))
    path = st.io["initial_condition_file"]
    if path:
        if path.endswith(".npy"):
            d = np.load(path, allow_pickle=False)
            if d.ndim != 2 or d.shape[0] != nc:
                raise ZdmError("%s: expected shape (ncells, k)" % path)
        else:
            if not os.path.exists(path):
                raise ZdmError("initial condition file %s cannot be opened" % path)
            d = read_h5(path, "state")
            if d.size != nc * nq or d.ndim > 2 or (d.ndim == 2 and d.shape[1] != nq):
                raise ZdmError('%s: dataset "state" has %d values; the mesh and model need %d x %d'
                               % (path, d.size, nc, nq))
            d = d.reshape(nc, nq)
        mask = st.io["initial_condition_mask"]
        fields = [k for k in mask if 0 <= k < nq] if mask else list(range(nq))
        src = d[mesh.cell_perm]
        if d.shape[1] == nq:
            X[:, fields] = src[:, fields]
        else:
            if d.shape[1] < len(fields):
                raise ZdmError("%s: fewer columns than fields" % path)
            X[:, fields] = src[:, :len(fields)]
        return X
    lib = model_lib(info.name)
    p = np.array(info.default_params, dtype=np.float64)
    for k, v in st.parameters.items():
        if k in info.parameter_names:
            p[info.parameter_names.index(k)] = v
    cen = np.ascontiguousarray(mesh.cell_centroid, dtype=np.float64)
    row = np.zeros(nq)
    for c in range(nc):
        if lib.zdm_model_initial_condition(_dp(cen[c]), _dp(p), _dp(row)):
            raise ZdmError(lib.zdm_last_error().decode())
        X[c] = row
    return X


class Solver:
    """MUSCL / SSP-RK finite-volume solver on one GPU (MUSCLSolver, MUSCLSolver.hpp:14-194).

    Construction mirrors VirtualSolver::Initialize + InitializeComponents: parameters start
    from Model::default_parameters() and are overridden by name (VirtualSolver.hpp:199-217).
    """

    def __init__(self, model, mesh, settings=None, params=None, face_bc=None, source=True,
                 order=None, ssp_stages=None, cfl=None, t_end=None, gradient=None,
                 use_limiters=None, well_balanced=None, cell_offsets=None, positivity=None,
                 wet_dry_eps=None, time_integration=None, refresh_derivative_aux=None):
        self.model = model
        self.info = ModelInfo(model)
        self.lib = model_lib(model)
        self.mesh = mesh
        st = settings if settings is not None else Settings()
        sv = st.solver
        p = self.info.default_params.copy()
        names = []
        buf = params if params is not None else {}
        if isinstance(buf, dict):
            merged = dict(st.parameters)
            merged.update(buf)
            pn = self.info.parameter_names
            for k, v in merged.items():
                if k in pn:
                    p[pn.index(k)] = v
                else:
                    names.append(k)
        else:
            p = np.asarray(buf, dtype=np.float64)
        self.params = p
        self.unknown_params = names
        if face_bc is None:
            face_bc = boundary_face_bc(mesh, self.info.boundary_tags, self.info.dim)
        self.face_bc = np.ascontiguousarray(face_bc, dtype=np.int32)
        o = _Options()
        o.order = int(order if order is not None else sv["reconstruction_order"])
        o.ssp_stages = int(ssp_stages if ssp_stages is not None else sv["ssp_stages"])
        o.use_limiters = int(use_limiters if use_limiters is not None else sv["limiter"] != "none")
        g = gradient if gradient is not None else sv.get("gradient", "green_gauss")
        o.gradient = 1 if g in ("lsq", "least_squares", 1) else 0
        ti = time_integration if time_integration is not None else sv["time_integration"]
        if ti not in TIME_INTEGRATION:
            raise ZdmError("time_integration %r is not provided by this build (%s)"
                           % (ti, ", ".join(TIME_INTEGRATION)))
        o.time_integration = TIME_INTEGRATION[ti]
        o.source = int(bool(source))
        o.cfl = float(cfl if cfl is not None else sv["cfl"])
        o.t_end = float(t_end if t_end is not None else sv["t_end"])
        o.min_dt = float(sv["min_dt"])
        o.block_size = 0
        o.well_balanced = int(bool(well_balanced if well_balanced is not None
                                   else sv.get("well_balanced", False)))
        pos = positivity if positivity is not None else sv.get("positivity", "none")
        if pos not in POSITIVITY:
            raise ZdmError("positivity must be one of %s" % sorted(POSITIVITY))
        o.positivity = POSITIVITY[pos]
        o.wet_dry_eps = float(wet_dry_eps if wet_dry_eps is not None else sv.get("wet_dry_eps", 1e-2))
        o.refresh_derivative_aux = int(bool(refresh_derivative_aux if refresh_derivative_aux is not None
                                            else sv.get("refresh_derivative_aux", True)))
        self.options = o
        h = ctypes.c_void_p()
        if cell_offsets is None:
            rc = self.lib.zdm_solver_create(mesh.handle, self.face_bc.ctypes.data_as(_c_int_p),
                                            _dp(self.params), ctypes.byref(o), ctypes.byref(h))
        else:
            self.cell_offsets = np.ascontiguousarray(cell_offsets, dtype=np.int32)
            rc = self.lib.zdm_batch_create(mesh.handle, self.face_bc.ctypes.data_as(_c_int_p),
                                           _dp(self.params), ctypes.byref(o),
                                           len(self.cell_offsets) - 1,
                                           self.cell_offsets.ctypes.data_as(_c_int_p),
                                           ctypes.byref(h))
        self._check(rc)
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise ZdmError(self.lib.zdm_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            self.lib.zdm_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, X, A=None):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(self.mesh.ncells, self.info.n_q)
        if A is not None:
            A = np.ascontiguousarray(A, dtype=np.float64).reshape(self.mesh.ncells, self.info.n_aux)
        self._check(self.lib.zdm_state_upload(self._h, _dp(X), _dp(A) if A is not None else None))

    def download(self):
        X = np.empty((self.mesh.ncells, self.info.n_q))
        A = np.empty((self.mesh.ncells, self.info.n_aux))
        self._check(self.lib.zdm_state_download(self._h, _dp(X), _dp(A)))
        return X, A

    def initialize(self):
        self._check(self.lib.zdm_initialize(self._h))

    def step(self, n=1):
        self._check(self.lib.zdm_step(self._h, int(n)))

    def step_graph(self, n=1):
        """n steps by replaying one HIP-graph-captured step."""
        self._check(self.lib.zdm_step_graph(self._h, int(n)))

    def member_time(self, i):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(self.lib.zdm_batch_time(self._h, int(i), ctypes.byref(t), ctypes.byref(dt),
                                            ctypes.byref(n)))
        return t.value, dt.value, n.value

    def run(self, max_steps=1 << 30, check_every=16):
        n = ctypes.c_int()
        self._check(self.lib.zdm_run(self._h, int(max_steps), int(check_every), ctypes.byref(n)))
        return n.value

    def synchronize(self):
        self._check(self.lib.zdm_synchronize(self._h))

    def time(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(self.lib.zdm_get_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)))
        return t.value, dt.value, n.value

    def transport_rhs(self, X, A=None, t=0.0):
        X = np.ascontiguousarray(X, dtype=np.float64)
        F = np.empty_like(X)
        Ap = None
        if A is not None:
            A = np.ascontiguousarray(A, dtype=np.float64)
            Ap = _dp(A)
        self._check(self.lib.zdm_transport_rhs(self._h, float(t), _dp(X), Ap, _dp(F)))
        return F

    def transport_rhs_device(self, X_ptr, F_ptr, A_ptr=None, t=0.0, stream=None):
        """The right-hand side on device buffers (integer addresses of float64 AoS arrays, e.g.
        torch tensors' data_ptr()), ordered on `stream` (a hipStream_t as an int, e.g.
        torch.cuda.current_stream().cuda_stream; None: the solver's stream).  No host
        synchronisation."""
        st = self.stream_handle() if stream is None else int(stream)
        self._check(self.lib.zdm_transport_rhs_device(self._h, float(t), ctypes.c_void_p(int(X_ptr)),
                                                      ctypes.c_void_p(int(A_ptr)) if A_ptr else None,
                                                      ctypes.c_void_p(int(F_ptr)), ctypes.c_void_p(st)))

    def source_step(self, dt):
        self._check(self.lib.zdm_source_step(self._h, float(dt)))

    def enforce_constraints(self):
        self._check(self.lib.zdm_enforce_constraints(self._h))

    def cfl_dt(self):
        d = ctypes.c_double()
        self._check(self.lib.zdm_cfl_dt(self._h, ctypes.byref(d)))
        return d.value

    def enable_timing(self, on=True):
        """HIP-event kernel timing: False off, True every step, k > 1 every k-th step."""
        self._check(self.lib.zdm_enable_timing(self._h, int(on)))

    def fuse_post(self, on=True):
        """True (the default): the final SSP stage and the post-step run as one kernel where the
        step allows it (no MOOD, no derivative-aux refresh, no well-balanced term); False:
        separate kernels.  Bitwise the same results."""
        self._check(self.lib.zdm_fuse_post(self._h, int(bool(on))))

    def flux_nontemporal(self, mode=-1):
        """Cache policy of the stage and final flux launches: 1 non-temporal accesses for their
        read-once / write-once streams, 0 ordinary ones, -1 (the default) by mesh size (on where
        one flux pass exceeds 256 MiB).  Returns whether the non-temporal form is now in use.
        Bitwise the same results either way."""
        active = ctypes.c_int()
        self._check(self.lib.zdm_flux_nontemporal(self._h, int(mode), ctypes.byref(active)))
        return bool(active.value)

    def kernel_time(self, kernel):
        """(average ms, launches) of kernel id 0 prep, 1 flux + stage update, 2 post-step,
        3 final stage fused with the post-step."""
        ms, n = ctypes.c_double(), ctypes.c_int()
        self._check(self.lib.zdm_kernel_time(self._h, int(kernel), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # ---- discrete adjoint (calibration gradients)
    def adjoint_setup(self, nsteps, gauges=(), obs_every=0, obs=None, checkpoint_every=0,
                      exact_limiter=False, chorin=False, stretch_tape=None):
        """Loss = sum over observation steps k*obs_every and gauges of (h - obs)^2.
        chorin=True (VAM models): the steps are Chorin pressure-split steps with the Krylov
        settings of chorin_setup; no tape layouts.
        exact_limiter=False holds each face's limiter value fixed (frozen-limiter
        linearisation); True differentiates it through its active branch.
        checkpoint_every: -2 / -1 tape with / without face states, 1 every state, k > 1 a
        checkpoint every k steps, -2-c (c >= 1) the binomial schedule on c state slots, 0 the
        richest layout that fits in device memory.
        stretch_tape (checkpoint_every >= 2 or <= -3): None the automatic choice, 0 none, L a
        tape of L steps (one segment for checkpoint_every >= 2), or (L, buffers); see
        adjoint_stretch_tape."""
        g = np.ascontiguousarray(gauges, dtype=np.int32)
        nobs = nsteps // obs_every if len(g) and obs_every else 0
        o = np.zeros((nobs, len(g))) if obs is None else np.ascontiguousarray(obs, dtype=np.float64)
        if o.size != nobs * len(g):
            raise ZdmError("adjoint_setup: obs must have nsteps // obs_every rows of len(gauges)")
        self._adj = (int(nsteps), g, int(obs_every), o.reshape(nobs, len(g)))
        fn = self.lib.zdm_adjoint_setup_chorin if chorin else self.lib.zdm_adjoint_setup
        self._check(fn(self._h, int(nsteps), int(checkpoint_every), int(obs_every), len(g),
                       g.ctypes.data_as(_c_int_p), _dp(o) if o.size else None, int(bool(exact_limiter))))
        if stretch_tape is not None:
            st = stretch_tape if isinstance(stretch_tape, tuple) else (stretch_tape, 0)
            self.adjoint_stretch_tape(int(st[0]), int(st[1]))

    def adjoint_stretch_tape(self, steps=-1, buffers=0):
        """Stretch tape of a checkpointed layout: each stretch of up to `steps` steps is re-run
        once with its stages recorded, then reversed with transposed kernels only; with two
        buffers the re-runs overlap the reversal on a second stream.  steps >= 0 sets it (0:
        none; buffers 0 automatic, 1 or 2), -1 only queries; returns (steps, level, buffers),
        (0, 0, 0) for none."""
        n, lv, nb = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.zdm_adjoint_stretch_tape(self._h, int(steps), int(buffers), ctypes.byref(n),
                                                      ctypes.byref(lv), ctypes.byref(nb)))
        return n.value, lv.value, nb.value

    def implicit_setup(self, **kw):
        """Newton / Krylov settings of time_integration "implicit" (NewtonOptions fields)."""
        o = NewtonOptions(**kw)
        self._check(self.lib.zdm_implicit_setup(self._h, ctypes.byref(o)))

    def implicit_stats(self):
        """(Newton iterations, GMRES iterations, last step's Newton iterations, last reason)."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        c, d = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.zdm_implicit_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                ctypes.byref(d)))
        return a.value, b.value, c.value, d.value

    def adjoint_krylov_its(self):
        """GMRES iterations of the transposed pressure solves since adjoint_setup."""
        v = ctypes.c_longlong()
        self._check(self.lib.zdm_adjoint_krylov_its(self._h, ctypes.byref(v)))
        return v.value

    AUTOTUNE_CANDIDATES = [(256, 1), (128, 1), (64, 1), (256, 2), (128, 2), (64, 2), (256, 4), (128, 4),
                           (64, 4)]

    def set_launch(self, block_size, cells_per_thread=1):
        """Workgroup size (64/128/256) and cells per thread of the transport kernels."""
        self._check(self.lib.zdm_set_launch(self._h, int(block_size), int(cells_per_thread)))

    def launch(self):
        b, c = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.zdm_get_launch(self._h, ctypes.byref(b), ctypes.byref(c)))
        return b.value, c.value

    def autotune(self, steps=3):
        """Time each candidate launch shape on the resident state (restored afterwards), keep
        the fastest; returns (block_size, cells_per_thread, {shape: ms per step})."""
        b, c = ctypes.c_int(0), ctypes.c_int(0)
        ms = np.zeros(9)
        self._check(self.lib.zdm_autotune(self._h, int(steps), ctypes.byref(b), ctypes.byref(c), _dp(ms)))
        return b.value, c.value, dict(zip(self.AUTOTUNE_CANDIDATES, ms.tolist()))

    def troubled_cells(self):
        """MOOD overrides since the last initialize (summed over ensemble members)."""
        n = ctypes.c_longlong(0)
        self._check(self.lib.zdm_troubled_cells(self._h, ctypes.byref(n)))
        return n.value

    def adjoint_layout(self):
        """(checkpoint_every, tape level) chosen by adjoint_setup."""
        ck, tp = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.zdm_adjoint_layout(self._h, ctypes.byref(ck), ctypes.byref(tp)))
        return ck.value, tp.value

    def adjoint_forward(self, X0, A0=None, dts=None):
        """Forward run of the calibration problem; dts = None records the CFL steps."""
        X0 = np.ascontiguousarray(X0, dtype=np.float64).reshape(self.mesh.ncells, self.info.n_q)
        Ap = None
        if A0 is not None:
            A0 = np.ascontiguousarray(A0, dtype=np.float64).reshape(self.mesh.ncells, self.info.n_aux)
            Ap = _dp(A0)
        dp_ = None
        if dts is not None:
            dts = np.ascontiguousarray(dts, dtype=np.float64)
            if dts.size != self._adj[0]:
                raise ZdmError("adjoint_forward: dts must hold nsteps values")
            dp_ = _dp(dts)
        L = ctypes.c_double()
        self._check(self.lib.zdm_adjoint_forward(self._h, _dp(X0), Ap, dp_, ctypes.byref(L)))
        return L.value

    def adjoint_forward_frozen_alpha(self, X0, A0=None, dts=None, record=False):
        """The map the frozen-limiter gradient differentiates: a forward run whose face limiter
        values are recorded (record=True) or replayed from the last recording (dts required),
        for central differences.  Overwrites the forward a gradient would reverse."""
        X0 = np.ascontiguousarray(X0, dtype=np.float64).reshape(self.mesh.ncells, self.info.n_q)
        Ap = None
        if A0 is not None:
            A0 = np.ascontiguousarray(A0, dtype=np.float64).reshape(self.mesh.ncells, self.info.n_aux)
            Ap = _dp(A0)
        dp_ = None
        if dts is not None:
            dts = np.ascontiguousarray(dts, dtype=np.float64)
            if dts.size != self._adj[0]:
                raise ZdmError("adjoint_forward_frozen_alpha: dts must hold nsteps values")
            dp_ = _dp(dts)
        L = ctypes.c_double()
        self._check(self.lib.zdm_adjoint_forward_frozen_alpha(self._h, _dp(X0), Ap, dp_, int(bool(record)),
                                                              ctypes.byref(L)))
        return L.value

    def adjoint_gradient(self):
        """(dL/dX0, dL/dA0) of the last adjoint_forward."""
        gX = np.empty((self.mesh.ncells, self.info.n_q))
        gA = np.empty((self.mesh.ncells, self.info.n_aux))
        self._check(self.lib.zdm_adjoint_gradient(self._h, _dp(gX), _dp(gA)))
        return gX, gA

    def adjoint_forward_device(self, dX0, dA0=None, dts=None):
        """adjoint_forward on device buffers: dX0 / dA0 are device addresses (ints) of
        (ncells, n_q) / (ncells, n_aux) float64 arrays, read on the solver stream; returns the
        loss once it is known."""
        dp_ = None
        if dts is not None:
            dts = np.ascontiguousarray(dts, dtype=np.float64)
            if dts.size != self._adj[0]:
                raise ZdmError("adjoint_forward_device: dts must hold nsteps values")
            dp_ = _dp(dts)
        L = ctypes.c_double()
        self._check(self.lib.zdm_adjoint_forward_device(self._h, ctypes.c_void_p(dX0),
                                                        ctypes.c_void_p(dA0) if dA0 else None, dp_,
                                                        ctypes.byref(L)))
        return L.value

    def adjoint_gradient_device(self, dgX0, dgA0=None):
        """adjoint_gradient into device buffers (addresses as ints, either may be 0), written
        on the solver stream; returns with the copies enqueued."""
        self._check(self.lib.zdm_adjoint_gradient_device(self._h, ctypes.c_void_p(dgX0) if dgX0 else None,
                                                         ctypes.c_void_p(dgA0) if dgA0 else None))

    def stream_handle(self):
        """The solver's hipStream_t as an int (for torch.cuda.ExternalStream)."""
        p = ctypes.c_void_p()
        self._check(self.lib.zdm_solver_stream(self._h, ctypes.byref(p)))
        return p.value or 0

    def adjoint_dts(self):
        d = np.empty(self._adj[0])
        self._check(self.lib.zdm_adjoint_dts(self._h, _dp(d)))
        return d

    def adjoint_residuals(self):
        r = np.empty(self._adj[3].shape)
        if r.size:
            self._check(self.lib.zdm_adjoint_residuals(self._h, _dp(r)))
        return r

    # ---- Chorin pressure split (VAM models, ChorinVAMSolver, ChorinVAM.hpp:86-129)
    def chorin_setup(self, krylov=None, **kw):
        """Krylov settings of the pressure solve (KrylovOptions or its fields by name)."""
        o = krylov if krylov is not None else KrylovOptions(**kw)
        self._check(self.lib.zdm_chorin_setup(self._h, ctypes.byref(o)))
        self.krylov = o

    def chorin_step(self, n=1):
        """n Chorin steps; returns (steps taken, total GMRES iterations, last KSP reason)."""
        st, its, rs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.zdm_chorin_step(self._h, int(n), ctypes.byref(st), ctypes.byref(its),
                                             ctypes.byref(rs)))
        return st.value, its.value, rs.value

    def chorin_stats(self):
        """The last pressure solve: (iterations, KSPConvergedReason, final and initial
        preconditioned residual norms)."""
        it, rs = ctypes.c_int(), ctypes.c_int()
        rn, rn0 = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.zdm_chorin_stats(self._h, ctypes.byref(it), ctypes.byref(rs), ctypes.byref(rn),
                                              ctypes.byref(rn0)))
        return it.value, rs.value, rn.value, rn0.value

    def chorin_pressure(self):
        P = np.empty((self.mesh.ncells, 2))
        self._check(self.lib.zdm_chorin_pressure(self._h, _dp(P)))
        return P

    def vam_frozen_aux(self):
        CA = np.empty((self.mesh.ncells, 20))
        self._check(self.lib.zdm_vam_frozen_aux(self._h, _dp(CA)))
        return CA

    def vam_residual(self, dt, P):
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(self.mesh.ncells, 2)
        R = np.empty_like(P)
        self._check(self.lib.zdm_vam_residual(self._h, float(dt), _dp(P), _dp(R)))
        return R

    def vam_pmat(self, dt):
        """Assembled pressure matrix as 2x2-block CSR: (row pointers, block columns,
        values (nnzb, 2, 2))."""
        z = np.zeros(1, dtype=np.int32)
        nnzb = self.lib.zdm_vam_pmat(self._h, float(dt), 0, None, None, None)
        if nnzb < 0:
            raise ZdmError(self.lib.zdm_last_error().decode())
        rp = np.empty(self.mesh.ncells + 1, dtype=np.int32)
        col = np.empty(nnzb, dtype=np.int32)
        val = np.empty((nnzb, 2, 2))
        r = self.lib.zdm_vam_pmat(self._h, float(dt), nnzb, rp.ctypes.data_as(_c_int_p),
                                  col.ctypes.data_as(_c_int_p), _dp(val))
        if r != nnzb:
            raise ZdmError(self.lib.zdm_last_error().decode())
        del z
        return rp, col, val


class Ensemble(Solver):
    """Many independent runs of one model in one set of launches (the batching form): the
    members' meshes are joined into a disjoint union, each member keeps its own dt, time and
    step count.  Member states are views into the union state."""

    def __init__(self, model, meshes, **kw):
        union, off = Mesh.concat(meshes)
        self.members = list(meshes)
        super().__init__(model, union, cell_offsets=off, **kw)

    def split(self, X):
        o = self.cell_offsets
        return [X[o[i]:o[i + 1]] for i in range(len(o) - 1)]

    def join(self, Xs):
        return np.concatenate(Xs, axis=0)


def bsr2_gmres(rowptr, col, val, b, x0=None, krylov=None, model="sme2", **kw):
    """GMRES + point-block Jacobi on a 2x2-block CSR system on the device; returns
    (x, iterations, KSPConvergedReason)."""
    lib = model_lib(model)
    o = krylov if krylov is not None else KrylovOptions(**kw)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    nb = rowptr.shape[0] - 1
    if val.size != 4 * col.size or b.size != 2 * nb:
        raise ZdmError("bsr2_gmres: inconsistent sizes")
    x = np.zeros(2 * nb) if x0 is None else np.array(x0, dtype=np.float64)
    its, rs = ctypes.c_int(), ctypes.c_int()
    if lib.zdm_bsr2_gmres(nb, rowptr.ctypes.data_as(_c_int_p), col.ctypes.data_as(_c_int_p),
                          _dp(val), _dp(b), _dp(x), ctypes.byref(o), ctypes.byref(its),
                          ctypes.byref(rs)) != 0:
        raise ZdmError(lib.zdm_last_error().decode())
    return x, its.value, rs.value


def stream_copy_gbs(model="sme2", nbytes=1 << 30, iters=20):
    """Achievable HBM bandwidth of a plain device copy (GB/s, read + write)."""
    lib = model_lib(model)
    g, ms = ctypes.c_double(), ctypes.c_double()
    if lib.zdm_stream_copy(nbytes, iters, ctypes.byref(g), ctypes.byref(ms)) != 0:
        raise ZdmError(lib.zdm_last_error().decode())
    return g.value, ms.value
