"""ctypes binding of libmhfem_amd.so (include/mhfem.h) and a host-side mirror
of the reference's ``TNL::MHFEM::Solver`` (Solver.h:64-117) and
``control.h::solve`` (control.h:64-102).

There is no fallback: if the HIP library is missing or fails to load, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libmhfem_amd.so"))

# array identifiers (include/mhfem.h)
ARRAYS = ["Z_iF", "Z_iK", "N_ijK", "u_ijKe", "m_iK", "D_ijK", "w_iKe", "a_ijKe", "r_ijK", "f_iK",
          "v_iKe", "m_iE_upw", "Z_ijE_upw", "b_ijK_storage", "R_ijKe", "R_iK",
          "bc_values", "bc_dirichlet", "bc_mobility", "rhs", "cell_measure", "face_measure"]
ARRAY_ID = {name: k for k, name in enumerate(ARRAYS)}

SOLVERS = {"bicgstab": 0, "gmres": 1, "cg": 2, "tfqmr": 3, "jacobi": 4, "bicgstabl": 5, "idrs": 6, "sor": 7}
PRECONDITIONERS = {"none": 0, "diagonal": 1, "jacobi": 1, "ilu0": 2, "amg": 3, "ilu-isai": 4, "parilu-isai": 5, "ilut": 6}
PRECISIONS = {"fp64": 0, "fp32-refined": 1, "fp32-reliable": 2}


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("num_points", C.c_int), ("num_cells", C.c_int),
                ("points", C.c_void_p), ("cell_vertices", C.c_void_p)]


class ModelDesc(C.Structure):
    _fields_ = [("num_equations", C.c_int), ("mass_lumping", C.c_int), ("advection", C.c_int),
                ("do_mobility_upwind", C.c_int)]


class SolverConfig(C.Structure):
    _fields_ = [("solver", C.c_int), ("preconditioner", C.c_int),
                ("convergence_residue", C.c_double), ("max_iterations", C.c_int),
                ("gmres_restart", C.c_int), ("precision", C.c_int),
                ("inner_residue", C.c_double), ("check_every", C.c_int), ("rows_per_block", C.c_int),
                ("face_groups", C.c_int), ("bicgstab_ell", C.c_int), ("idr_s", C.c_int),
                ("sor_omega", C.c_double), ("ilut_p", C.c_int), ("ilut_threshold", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_iterations", C.c_int),
                ("residue", C.c_double), ("converged", C.c_int), ("fallback_iterations", C.c_int)]


class Timers(C.Structure):
    _fields_ = [("pre_iterate", C.c_double), ("update_b_R_Q", C.c_double), ("upwind", C.c_double),
                ("assemble", C.c_double), ("preconditioner", C.c_double),
                ("linear_solver", C.c_double), ("post_iterate", C.c_double),
                ("iterations", C.c_longlong), ("time_steps", C.c_longlong),
                ("spmv_ms", C.c_double), ("spmv_timed", C.c_longlong)]


TIME_KINDS = {"constant": 0, "sine": 1, "table": 2}


class TimeFunction(C.Structure):
    _fields_ = [("kind", C.c_int), ("offset", C.c_double), ("amplitude", C.c_double),
                ("angular_frequency", C.c_double), ("phase", C.c_double), ("points", C.c_int),
                ("times", C.c_void_p), ("values", C.c_void_p)]


GUESS_COMBINE = {"combined": 0, "network": 1, "history": 2}


class GuessConfig(C.Structure):
    _fields_ = [("enabled", C.c_int), ("train_steps", C.c_int), ("epochs", C.c_int),
                ("learning_rate", C.c_float), ("seed", C.c_uint), ("combine", C.c_int),
                ("sample_stride", C.c_int), ("reject_after", C.c_int), ("history_depth", C.c_int)]


class GuessStats(C.Structure):
    _fields_ = [("used", C.c_int), ("r0_norm", C.c_double), ("r_guess_norm", C.c_double),
                ("loss", C.c_double), ("trained_steps", C.c_int), ("coef_network", C.c_double),
                ("network_gain", C.c_double), ("network_active", C.c_int)]


# every symbol declared in include/mhfem.h with its signature
_P, _I, _D, _L = C.c_void_p, C.c_int, C.c_double, C.c_long
SIGNATURES = {
    "mhfem_create": (_P, [C.POINTER(MeshDesc), C.POINTER(ModelDesc)]),
    "mhfem_create_cells": (_P, [C.POINTER(MeshDesc), _I, C.POINTER(ModelDesc)]),
    "mhfem_destroy": (None, [_P]),
    "mhfem_last_error": (C.c_char_p, [_P]),
    "mhfem_get_counts": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "mhfem_get_topology": (_I, [_P, _P, _P]),
    "mhfem_set_array": (_I, [_P, _I, _P, _L]),
    "mhfem_get_array": (_I, [_P, _I, _P, _L]),
    "mhfem_array_size": (_L, [_P, _I]),
    "mhfem_device_array": (_P, [_P, _I]),
    "mhfem_stream": (_P, [_P]),
    "mhfem_set_boundary_conditions": (_I, [_P, _P, _P, _P]),
    "mhfem_set_initial_condition": (_I, [_P, _I, _P, _L]),
    "mhfem_set_boundary_time_function": (_I, [_P, _I, _I, C.POINTER(TimeFunction)]),
    "mhfem_init": (_I, [_P]),
    "mhfem_setup_linear_system": (_I, [_P, C.POINTER(SolverConfig)]),
    "mhfem_pre_iterate": (_I, [_P, _D, _D]),
    "mhfem_assemble": (_I, [_P, _D, _D]),
    "mhfem_solve": (_I, [_P, C.POINTER(SolveStats)]),
    "mhfem_post_iterate": (_I, [_P, _D, _D]),
    "mhfem_solve_interval": (_I, [_P, _D, _D, _D, C.POINTER(SolveStats)]),
    "mhfem_system_size": (_I, [_P, C.POINTER(_I), C.POINTER(_L)]),
    "mhfem_export_system": (_I, [_P, _P, _P, _P, _P]),
    "mhfem_solve_rhs": (_I, [_P, _P, _P, C.POINTER(SolveStats)]),
    "mhfem_solve_rhs_noalloc": (_I, [_P, _P, _P, C.POINTER(SolveStats)]),
    "mhfem_benchmark_spmv": (_I, [_P, _I, _I, C.POINTER(_D)]),
    "mhfem_step": (_I, [_P, _D, _D, C.POINTER(SolveStats)]),
    "mhfem_autotune": (_I, [_P, C.POINTER(_I), C.POINTER(_D)]),
    "mhfem_create_batch": (_P, [C.POINTER(MeshDesc), _I, C.POINTER(ModelDesc)]),
    "mhfem_batch_info": (_I, [_P, C.POINTER(_I), _P, _P]),
    "mhfem_batch_stats": (_I, [_P, _P, _P, _I]),
    "mhfem_guess_configure": (_I, [_P, C.POINTER(GuessConfig)]),
    "mhfem_guess_stats_last": (_I, [_P, C.POINTER(GuessStats)]),
    "mhfem_guess_num_params": (_I, []),
    "mhfem_guess_get_weights": (_I, [_P, _P, _I]),
    "mhfem_guess_set_weights": (_I, [_P, _P, _I]),
    "mhfem_guess_forward": (_I, [_P, _P, _L, _P]),
    "mhfem_guess_gradient": (_I, [_P, _P, _P, _L, _P]),
    "mhfem_get_timers": (_I, [_P, C.POINTER(Timers)]),
    "mhfem_reset_timers": (_I, [_P]),
    "mhfem_set_timing": (_I, [_P, _I]),
    "mhfem_synchronize": (_I, [_P]),
    "mhfem_mesh_file_open": (_P, [C.c_char_p, C.c_char_p]),
    "mhfem_mesh_file_close": (None, [_P]),
    "mhfem_mesh_file_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_L), C.POINTER(_L)]),
    "mhfem_mesh_file_read": (_I, [_P, _P, _P]),
    "mhfem_mesh_file_cell_data": (_L, [_P, C.c_char_p, _P, _L]),
    "mhfem_mesh_file_cell_data_count": (_I, [_P]),
    "mhfem_mesh_file_cell_data_name": (C.c_char_p, [_P, _I]),
    "mhfem_write_vtu": (_I, [C.c_char_p, C.POINTER(MeshDesc), _I, _L, _D, _I, C.POINTER(C.c_char_p),
                             C.POINTER(C.c_void_p), _I]),
    "mhfem_bc_save": (_I, [C.c_char_p, C.c_int64, _P, _P, _P]),
    "mhfem_bc_load": (_I, [C.c_char_p, C.POINTER(C.c_int64), _P, _P, _P, C.c_int64]),
    "mhfem_save_vector": (_I, [C.c_char_p, _P, C.c_int64]),
    "mhfem_load_vector": (_L, [C.c_char_p, _P, C.c_int64]),
    "mhfem_save_linear_system": (_I, [_P, C.c_char_p]),
    "mhfem_amg_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "mhfem_amg_level": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_L), _P]),
    "mhfem_amg_apply": (_I, [_P, _P, _P]),
    "mhfem_isai_info": (_I, [_P, C.POINTER(_L), C.POINTER(_L), C.POINTER(_I)]),
    "mhfem_isai_matrix": (_I, [_P, _I, _P, _P, _P]),
}

_lib = None


def lib():
    """Load libmhfem_amd.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmhfem_amd.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class MHFEMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _global_check(rc):
    if rc < 0:
        raise MHFEMError(rc, lib().mhfem_last_error(None).decode())
    return rc


def _path(p):
    return os.fsencode(os.fspath(p))


# ---- files on either side of the path (host code in the library, no GPU work)
VTK_FORMATS = {"ascii": 0, "binary": 1, "zlib": 2}


class MeshFile:
    """A mesh file of lines, triangles, quadrangles or tetrahedra with its cell
    data, read by the library: .vtu (VTK XML) or .vtk (legacy VTK).  Stands for
    TNL::Meshes::loadMesh (control.h:20-48) and the reader's readCellData
    (Solver.hpp:296-306)."""

    def __init__(self, path, format="auto"):
        L = lib()
        h = L.mhfem_mesh_file_open(_path(path), format.encode())
        if not h:
            raise MHFEMError(-7, L.mhfem_last_error(None).decode())
        try:
            d, nvc, npts, nc = C.c_int(), C.c_int(), C.c_long(), C.c_long()
            _global_check(L.mhfem_mesh_file_info(h, C.byref(d), C.byref(nvc), C.byref(npts), C.byref(nc)))
            self.dim = d.value
            self.points = np.empty((npts.value, self.dim), np.float64)
            self.cells = np.empty((nc.value, nvc.value), np.int32)
            _global_check(L.mhfem_mesh_file_read(h, _ptr(self.points), _ptr(self.cells)))
            self._data = {}
            for k in range(L.mhfem_mesh_file_cell_data_count(h)):
                name = L.mhfem_mesh_file_cell_data_name(h, k)
                size = _global_check(L.mhfem_mesh_file_cell_data(h, name, None, 0))
                v = np.empty(size, np.float64)
                _global_check(L.mhfem_mesh_file_cell_data(h, name, _ptr(v), size))
                self._data[name.decode()] = v
        finally:
            L.mhfem_mesh_file_close(h)

    def cell_data_names(self):
        return list(self._data)

    def cell_data(self, name):
        if name not in self._data:
            raise KeyError(f"cell data array '{name}' was not found in the mesh file")
        return self._data[name]


def write_vtu(path, points, cells, cell_data=None, cycle=-1, time=-1.0, format="zlib"):
    """VTU file of a mesh (cells of 2, 3 or 4 vertices) with float64 cell
    arrays (the snapshot writer of HeatEquationModel.h:164-231)."""
    pts = np.ascontiguousarray(points, np.float64)
    cel = np.ascontiguousarray(cells, np.int32)
    mesh = MeshDesc(pts.shape[1], pts.shape[0], cel.shape[0], _ptr(pts), _ptr(cel))
    cell_data = cell_data or {}
    arrays = [np.ascontiguousarray(v, np.float64).ravel() for v in cell_data.values()]
    for name, v in zip(cell_data, arrays):
        if v.size != cel.shape[0]:
            raise ValueError(f"cell data '{name}' has {v.size} values for {cel.shape[0]} cells")
    names = (C.c_char_p * max(1, len(arrays)))(*[k.encode() for k in cell_data])
    vals = (C.c_void_p * max(1, len(arrays)))(*[v.ctypes.data for v in arrays])
    _global_check(lib().mhfem_write_vtu(_path(path), C.byref(mesh), cel.shape[1], int(cycle), float(time),
                                        len(arrays), names, vals, VTK_FORMATS[format]))


class BoundaryConditionsStorage:
    """TNL::MHFEM::BoundaryConditionsStorage (BoundaryConditionsStorage.h:9-72):
    ``tags``, ``values`` and ``dirichletValues`` of length ``dofSize`` in the
    i*F+E layout, saved and loaded in TNL's binary object format."""

    def __init__(self, dofSize=0):
        self.dofSize = int(dofSize)
        self.tags = np.zeros(self.dofSize, np.uint8)
        self.values = np.zeros(self.dofSize, np.float64)
        self.dirichletValues = np.zeros(self.dofSize, np.float64)

    def save(self, path):
        t = np.ascontiguousarray(self.tags, np.uint8)
        v = np.ascontiguousarray(self.values, np.float64)
        d = np.ascontiguousarray(self.dirichletValues, np.float64)
        if not (t.size == v.size == d.size == self.dofSize):
            raise ValueError("tags, values and dirichletValues must have dofSize elements")
        _global_check(lib().mhfem_bc_save(_path(path), self.dofSize, _ptr(t), _ptr(v), _ptr(d)))

    @classmethod
    def load(cls, path):
        L = lib()
        n = C.c_int64()
        _global_check(L.mhfem_bc_load(_path(path), C.byref(n), None, None, None, 0))
        s = cls(n.value)
        _global_check(L.mhfem_bc_load(_path(path), C.byref(n), _ptr(s.tags), _ptr(s.values),
                                      _ptr(s.dirichletValues), s.dofSize))
        return s


def save_vector(path, x):
    """A float64 vector as TNL::Containers::Array< double >."""
    v = np.ascontiguousarray(x, np.float64).ravel()
    _global_check(lib().mhfem_save_vector(_path(path), _ptr(v), v.size))


def load_vector(path):
    n = _global_check(lib().mhfem_load_vector(_path(path), None, 0))
    v = np.empty(n, np.float64)
    _global_check(lib().mhfem_load_vector(_path(path), _ptr(v), n))
    return v


class Solver:
    """Host mirror of TNL::MHFEM::Solver over the C ABI.

    Method names follow the reference: ``setMesh`` is the constructor,
    ``setup``/``setupLinearSystem`` take the solver options, ``preIterate``,
    ``assembleLinearSystem``, ``solveLinearSystem``, ``postIterate`` form one
    time step and :meth:`solve` is control.h::solve.
    """

    def __init__(self, problem, solver=None, preconditioner=None, tol=None, precision="fp64",
                 inner_tol=1e-4, check_every=8, max_iterations=100000, gmres_restart=30,
                 setup=True, rows_per_block=0, output_directory=None, face_groups=0, bicgstab_ell=0,
                 idr_s=0, sor_omega=0.0, ilut_p=0, ilut_threshold=0.0):
        pr = problem
        self.problem = pr
        # where a failed solve leaves its linear system (the reference's output-directory)
        self.output_directory = output_directory
        model = ModelDesc(pr.n, pr.lumping, pr.advection, pr.do_mobility_upwind)
        L = lib()
        instances = pr.extra.get("instances") if pr.extra else None
        if instances:
            # batched ensemble: one segment per instance (mhfem_create_batch)
            self._keep = [(np.ascontiguousarray(p_, np.float64), np.ascontiguousarray(c_, np.int32))
                          for p_, c_ in instances]
            descs = (MeshDesc * len(instances))(*[MeshDesc(pr.dim, p_.shape[0], c_.shape[0], _ptr(p_), _ptr(c_))
                                                  for p_, c_ in self._keep])
            self.ctx = L.mhfem_create_batch(descs, len(instances), C.byref(model))
        else:
            self._points = np.ascontiguousarray(pr.points, np.float64)
            self._cells = np.ascontiguousarray(pr.cells, np.int32)
            mesh = MeshDesc(pr.dim, self._points.shape[0], self._cells.shape[0],
                            _ptr(self._points), _ptr(self._cells))
            self.ctx = L.mhfem_create_cells(C.byref(mesh), self._cells.shape[1], C.byref(model))
        if not self.ctx:
            raise MHFEMError(-1, L.mhfem_last_error(None).decode())
        c, f, b = C.c_int(), C.c_int(), C.c_int()
        L.mhfem_get_counts(self.ctx, C.byref(c), C.byref(f), C.byref(b))
        self.num_cells, self.num_faces, self.num_boundary_faces = c.value, f.value, b.value
        self.n = pr.n
        self.iterations = []
        self.config = SolverConfig(SOLVERS[solver or pr.solver],
                                   PRECONDITIONERS[preconditioner or pr.preconditioner],
                                   pr.tol if tol is None else tol, max_iterations, gmres_restart,
                                   PRECISIONS[precision], inner_tol, check_every, rows_per_block,
                                   face_groups, bicgstab_ell, idr_s, sor_omega, ilut_p, ilut_threshold)
        if setup:
            self.setInitialCondition()
            self.setupLinearSystem()

    def _check(self, rc):
        if rc != 0:
            raise MHFEMError(rc, lib().mhfem_last_error(self.ctx).decode())

    def close(self):
        if getattr(self, "ctx", None):
            lib().mhfem_destroy(self.ctx)
            self.ctx = None

    __del__ = close

    # ---- model data
    def set_array(self, name, values):
        v = np.ascontiguousarray(values, np.float64).ravel()
        self._check(lib().mhfem_set_array(self.ctx, ARRAY_ID[name], _ptr(v), v.size))

    def get_array(self, name):
        size = lib().mhfem_array_size(self.ctx, ARRAY_ID[name])
        out = np.empty(size, np.float64)
        self._check(lib().mhfem_get_array(self.ctx, ARRAY_ID[name], _ptr(out), size))
        return out

    def topology(self):
        fpc = self.problem.fpc
        cf = np.empty((self.num_cells, fpc), np.int32)
        fc = np.empty((self.num_faces, 2), np.int32)
        self._check(lib().mhfem_get_topology(self.ctx, _ptr(cf), _ptr(fc)))
        return cf, fc

    # ---- Solver::setInitialCondition (Solver.hpp:255-344)
    def setInitialCondition(self):
        pr = self.problem
        tags = np.ascontiguousarray(pr.tags, np.uint8)
        vals = np.ascontiguousarray(pr.bc_values, np.float64)
        dirv = np.ascontiguousarray(pr.bc_dirichlet, np.float64)
        self._check(lib().mhfem_set_boundary_conditions(self.ctx, _ptr(tags), _ptr(vals), _ptr(dirv)))
        for name in ("N_ijK", "u_ijKe", "m_iK", "D_ijK", "w_iKe", "a_ijKe", "r_ijK", "f_iK"):
            v = np.asarray(pr.coeffs[name], np.float64)
            self.set_array(name, v.reshape(1) if v.ndim == 0 else v)
        if pr.bc_mobility is not None:
            self.set_array("bc_mobility", pr.bc_mobility)
        for (which, i), spec in (pr.boundary_functions or {}).items():
            self.set_boundary_time_function(which, i, spec)
        z = np.ascontiguousarray(pr.Z_iK0, np.float64).reshape(self.n, -1)
        for i in range(self.n):
            zi = np.ascontiguousarray(z[i])
            self._check(lib().mhfem_set_initial_condition(self.ctx, i, _ptr(zi), zi.size))
        self._check(lib().mhfem_init(self.ctx))

    def set_boundary_time_function(self, which, species, spec):
        """The boundary model's time function of the Dirichlet ("dirichlet") or
        Neumann ("neumann") values of one species (mhfem_set_boundary_time_function);
        spec None restores the stationary values."""
        w = ARRAY_ID["bc_dirichlet"] if which == "dirichlet" else ARRAY_ID["bc_values"]
        if spec is None:
            self._check(lib().mhfem_set_boundary_time_function(self.ctx, w, species, None))
            return
        tt = np.ascontiguousarray(spec.get("times", []), np.float64)
        vv = np.ascontiguousarray(spec.get("values", []), np.float64)
        f = TimeFunction(TIME_KINDS[spec["kind"]], spec.get("offset", 0.0), spec.get("amplitude", 0.0),
                         spec.get("angular_frequency", 0.0), spec.get("phase", 0.0), int(tt.size),
                         tt.ctypes.data if tt.size else None, vv.ctypes.data if vv.size else None)
        self._check(lib().mhfem_set_boundary_time_function(self.ctx, w, species, C.byref(f)))

    def _host_boundary_values(self):
        """a time-dependent boundary model the library does not evaluate itself"""
        return self.problem.dirichlet_at is not None and not self.problem.boundary_functions

    def setupLinearSystem(self):
        self._check(lib().mhfem_setup_linear_system(self.ctx, C.byref(self.config)))

    def amg_hierarchy(self):
        """(rows, nnz, aggregates or None) of every level of the AMG preconditioner."""
        lv, rb = C.c_int(), C.c_int()
        self._check(lib().mhfem_amg_info(self.ctx, C.byref(lv), C.byref(rb)))
        out = []
        for l in range(lv.value):
            rows, nnz = C.c_int(), C.c_long()
            self._check(lib().mhfem_amg_level(self.ctx, l, C.byref(rows), C.byref(nnz), None))
            agg = None
            if l + 1 < lv.value:
                agg = np.empty(rows.value, np.int32)
                self._check(lib().mhfem_amg_level(self.ctx, l, None, None, _ptr(agg)))
            out.append((rows.value, nnz.value, agg))
        return out, rb.value

    def amg_apply(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        self._check(lib().mhfem_amg_apply(self.ctx, _ptr(b), _ptr(x)))
        return x

    def isai_matrices(self):
        """(M_L, M_U, rebuilds): the ILU-ISAI preconditioner's approximate
        inverses as (rowptr, col, val) triples, and its rebuild count."""
        nl, nu, rb = C.c_long(), C.c_long(), C.c_int()
        self._check(lib().mhfem_isai_info(self.ctx, C.byref(nl), C.byref(nu), C.byref(rb)))
        rows, nnzA = C.c_int(), C.c_long()
        self._check(lib().mhfem_system_size(self.ctx, C.byref(rows), C.byref(nnzA)))
        out = []
        for upper, nnz in ((0, nl.value), (1, nu.value)):
            rp = np.empty(rows.value + 1, np.int32)
            col = np.empty(nnz, np.int32)
            val = np.empty(nnz)
            self._check(lib().mhfem_isai_matrix(self.ctx, upper, _ptr(rp), _ptr(col), _ptr(val)))
            out.append((rp, col, val))
        return out[0], out[1], rb.value

    def saveLinearSystem(self, directory):
        """Solver::saveLinearSystem (Solver.hpp:966-977): matrix.mtx, dof.vec.tnl
        and rhs.vec.tnl in ``directory``."""
        _global_check(lib().mhfem_save_linear_system(self.ctx, _path(directory)))

    # ---- one time step (control.h:83-98)
    def preIterate(self, t, tau):
        self._check(lib().mhfem_pre_iterate(self.ctx, t, tau))

    def assembleLinearSystem(self, t, tau):
        if self._host_boundary_values():
            self.set_array("bc_dirichlet", self.problem.dirichlet_at(t + tau))
        self._check(lib().mhfem_assemble(self.ctx, t, tau))

    def solveLinearSystem(self):
        st = SolveStats()
        rc = lib().mhfem_solve(self.ctx, C.byref(st))
        self.iterations.append(st.iterations)
        self.last_stats = st
        if rc == -5 and self.output_directory:
            # the reference saves the system before it throws (Solver.hpp:1259-1264)
            msg = lib().mhfem_last_error(self.ctx).decode()
            self.saveLinearSystem(self.output_directory)
            raise MHFEMError(rc, msg)
        self._check(rc)
        return st

    def postIterate(self, t, tau):
        self._check(lib().mhfem_post_iterate(self.ctx, t, tau))

    def step(self, t, tau):
        self.preIterate(t, tau)
        self.assembleLinearSystem(t, tau)
        st = self.solveLinearSystem()
        self.postIterate(t, tau)
        return st

    def step_graph(self, t, tau):
        """one time step replayed from captured HIP graphs (mhfem_step)"""
        if self._host_boundary_values():
            self.set_array("bc_dirichlet", self.problem.dirichlet_at(t + tau))
        st = SolveStats()
        self._check(lib().mhfem_step(self.ctx, t, tau, C.byref(st)))
        self.iterations.append(st.iterations)
        self.last_stats = st
        return st

    def autotune(self):
        rpb, ms = C.c_int(), C.c_double()
        self._check(lib().mhfem_autotune(self.ctx, C.byref(rpb), C.byref(ms)))
        return rpb.value, ms.value

    def batch_info(self):
        cnt = C.c_int()
        self._check(lib().mhfem_batch_info(self.ctx, C.byref(cnt), None, None))
        co = np.empty(cnt.value + 1, np.int32)
        fo = np.empty(cnt.value + 1, np.int32)
        self._check(lib().mhfem_batch_info(self.ctx, C.byref(cnt), _ptr(co), _ptr(fo)))
        return co, fo

    def batch_stats(self):
        co, _ = self.batch_info()
        k = co.size - 1
        its = np.empty(k, np.int32)
        res = np.empty(k, np.float64)
        self._check(lib().mhfem_batch_stats(self.ctx, _ptr(its), _ptr(res), k))
        return its, res

    def solve(self, t0, t1, dt):
        """control.h::solve(problem, t0, t1, dt) (control.h:64-102)."""
        if not self._host_boundary_values():
            st = SolveStats()
            self._check(lib().mhfem_solve_interval(self.ctx, t0, t1, dt, C.byref(st)))
            return st
        t = t0
        st = None
        while t1 - t > dt * 1e-6:
            tau = min(dt, t1 - t)
            st = self.step(t, tau)
            t += tau
        return st

    # ---- linear system
    def export_system(self):
        rows, nnz = C.c_int(), C.c_long()
        self._check(lib().mhfem_system_size(self.ctx, C.byref(rows), C.byref(nnz)))
        rp = np.empty(rows.value + 1, np.int32)
        col = np.empty(nnz.value, np.int32)
        val = np.empty(nnz.value, np.float64)
        rhs = np.empty(rows.value, np.float64)
        self._check(lib().mhfem_export_system(self.ctx, _ptr(rp), _ptr(col), _ptr(val), _ptr(rhs)))
        return rp, col, val, rhs

    def solve_rhs(self, rhs, x0):
        rhs = np.ascontiguousarray(rhs, np.float64)
        x = np.array(x0, np.float64, copy=True)
        st = SolveStats()
        self._check(lib().mhfem_solve_rhs(self.ctx, _ptr(rhs), _ptr(x), C.byref(st)))
        return x, st

    def benchmark_spmv(self, precision="fp64", reps=50):
        ms = C.c_double()
        self._check(lib().mhfem_benchmark_spmv(self.ctx, PRECISIONS[precision], reps, C.byref(ms)))
        return ms.value

    # ---- learned initial guess
    def configure_guess(self, enabled=True, train_steps=100, epochs=20, learning_rate=1e-3, seed=1,
                        combine="combined", sample_stride=1, reject_after=0, history_depth=0):
        cfg = GuessConfig(1 if enabled else 0, train_steps, epochs, learning_rate, seed,
                          GUESS_COMBINE[combine], sample_stride, reject_after, history_depth)
        self._check(lib().mhfem_guess_configure(self.ctx, C.byref(cfg)))

    def guess_stats(self):
        st = GuessStats()
        self._check(lib().mhfem_guess_stats_last(self.ctx, C.byref(st)))
        return {k: getattr(st, k) for k, _ in GuessStats._fields_}

    def guess_weights(self):
        n = lib().mhfem_guess_num_params()
        w = np.empty(n, np.float32)
        self._check(lib().mhfem_guess_get_weights(self.ctx, _ptr(w), n))
        return w

    def set_guess_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(lib().mhfem_guess_set_weights(self.ctx, _ptr(w), w.size))

    def guess_forward(self, Xbits):
        X = np.ascontiguousarray(Xbits, np.uint16)
        y = np.empty(X.shape[0], np.float32)
        self._check(lib().mhfem_guess_forward(self.ctx, _ptr(X), X.shape[0], _ptr(y)))
        return y

    def guess_gradient(self, Xbits, t):
        X = np.ascontiguousarray(Xbits, np.uint16)
        t = np.ascontiguousarray(t, np.float32)
        g = np.empty(lib().mhfem_guess_num_params() + 1, np.float32)
        self._check(lib().mhfem_guess_gradient(self.ctx, _ptr(X), _ptr(t), X.shape[0], _ptr(g)))
        return g[:-1], float(g[-1])

    def synchronize(self):
        self._check(lib().mhfem_synchronize(self.ctx))

    def device_array(self, name):
        """device address of an array (mhfem_device_array), or None"""
        return lib().mhfem_device_array(self.ctx, ARRAY_ID[name] if isinstance(name, str) else name)

    def set_timing(self, on=True):
        """False/0: off; True/1: phase timers; 2: phase timers + sampled
        SpMV launch timing (timers()['spmv_ms'] / ['spmv_timed'])"""
        self._check(lib().mhfem_set_timing(self.ctx, int(on)))

    def reset_timers(self):
        self._check(lib().mhfem_reset_timers(self.ctx))

    def timers(self):
        t = Timers()
        self._check(lib().mhfem_get_timers(self.ctx, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timers._fields_}
