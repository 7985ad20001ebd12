"""ctypes wrapper around the CPU oracle (oracle/_build/libmhfem_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MHFEM_ORACLE_LIB lets a mutation check point the pins at a deliberately
# broken copy of the oracle
LIB_PATH = os.environ.get("MHFEM_ORACLE_LIB", os.path.join(_HERE, "_build", "libmhfem_oracle.so"))

SOLVERS = {"bicgstab": 0, "gmres": 1, "cg": 2, "tfqmr": 3, "jacobi": 4, "bicgstabl": 5, "idrs": 6, "sor": 7}
PRECONDITIONERS = {"none": 0, "diagonal": 1, "jacobi": 1, "ilu0": 2, "ilut": 3}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P, I, D, L64 = C.c_void_p, C.c_int, C.c_double, C.c_long
        sig = {
            "or_topology": (I, [I, I, P, I, P, P, P]),
            "or_create": (P, [I, I, P, I, P, I, I, I, I]),
            "or_topology_shape": (I, [I, I, I, P, I, P, P, P]),
            "or_create_shape": (P, [I, I, I, P, I, P, I, I, I, I]),
            "or_destroy": (None, [P]),
            "or_set_threads": (None, [I]),
            "or_error": (C.c_char_p, [P]),
            "or_num_cells": (I, [P]),
            "or_num_faces": (I, [P]),
            "or_array": (P, [P, C.c_char_p, C.POINTER(L64)]),
            "or_init": (None, [P]),
            "or_setup_linear_system": (I, [P]),
            "or_pre_iterate": (None, [P, D, D]),
            "or_assemble": (None, [P, D, D]),
            "or_solve": (I, [P, I, I, D, I, I, C.POINTER(I), C.POINTER(D)]),
            "or_post_iterate": (None, [P, D, D]),
            "or_rows": (I, [P]),
            "or_nnz": (L64, [P]),
            "or_rowptr": (P, [P]),
            "or_colidx": (P, [P]),
            "or_values": (P, [P]),
            "or_ilu0_factorize": (I, [I, P, P, P]),
            "or_ilu0_apply": (None, [I, P, P, P, P, P]),
            "or_spmv": (None, [I, P, P, P, P, P]),
            "or_krylov": (I, [I, P, P, P, P, P, I, I, D, I, I, C.POINTER(I), C.POINTER(D)]),
            "or_local_mass_inverse": (None, [I, I, P, D, D, P]),
            "or_ilut_factorize": (L64, [I, P, P, P, I, D, P, P, P]),
            "or_set_ilut": (None, [I, D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _shape(dim, vertices_per_cell):
    """1 (box: edges, quadrangles, hexahedra) or 0 (simplex)."""
    return 1 if dim == 1 or vertices_per_cell == 2 ** dim else 0


def topology(points, cells):
    points = np.ascontiguousarray(points, np.float64)
    cells = np.ascontiguousarray(cells, np.int32)
    nc, nvc = cells.shape
    dim = points.shape[1]
    box = _shape(dim, nvc)
    fpc = 2 * dim if box else dim + 1
    cf = np.empty((nc, fpc), np.int32)
    fc = np.empty((nc * fpc, 2), np.int32)
    nf = lib().or_topology_shape(dim, box, points.shape[0], _ptr(points), nc, _ptr(cells),
                                 _ptr(cf), _ptr(fc))
    if nf < 0:
        raise ValueError("non-conforming mesh")
    return cf, fc[:nf].copy()


def local_mass_inverse(cell_points, lumping, D=1.0):
    P = np.ascontiguousarray(cell_points, np.float64)
    dim = P.shape[1]
    from math import factorial
    M = np.ones((dim + 1, dim + 1))
    M[:, 1:] = P
    meas = abs(np.linalg.det(M)) / factorial(dim)
    out = np.zeros(14)
    lib().or_local_mass_inverse(dim, lumping, _ptr(P), meas, D, _ptr(out))
    return out[: (9 if dim == 2 else 14)]


def ilu0(rowptr, col, val):
    rp = np.ascontiguousarray(rowptr, np.int32)
    c = np.ascontiguousarray(col, np.int32)
    lu = np.array(val, np.float64, copy=True)
    rc = lib().or_ilu0_factorize(rp.size - 1, _ptr(rp), _ptr(c), _ptr(lu))
    if rc:
        raise ValueError("ILU(0) breakdown")
    return lu


def ilut(rowptr, col, val, p=10, tau=1e-4):
    """ILUT(p, tau) factors as (rowptr, col, values) in one CSR."""
    rp = np.ascontiguousarray(rowptr, np.int32)
    c = np.ascontiguousarray(col, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    prp, pcol, pval = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nnz = lib().or_ilut_factorize(rp.size - 1, _ptr(rp), _ptr(c), _ptr(v), p, tau, C.byref(prp), C.byref(pcol),
                                  C.byref(pval))
    if nnz < 0:
        raise MemoryError("ILUT")
    n = rp.size - 1
    out = (np.frombuffer((C.c_int * (n + 1)).from_address(prp.value), np.int32).copy(),
           np.frombuffer((C.c_int * nnz).from_address(pcol.value), np.int32).copy(),
           np.frombuffer((C.c_double * nnz).from_address(pval.value), np.float64).copy())
    libc = C.CDLL(None)
    for ptr in (prp, pcol, pval):
        libc.free(ptr)
    return out


def set_ilut(p, tau):
    lib().or_set_ilut(p, tau)


def ilu0_apply(rowptr, col, lu, x):
    rp = np.ascontiguousarray(rowptr, np.int32)
    c = np.ascontiguousarray(col, np.int32)
    lu = np.ascontiguousarray(lu, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    lib().or_ilu0_apply(rp.size - 1, _ptr(rp), _ptr(c), _ptr(lu), _ptr(x), _ptr(y))
    return y


def krylov(rowptr, col, val, b, x0, solver="bicgstab", preconditioner="diagonal", tol=1e-10,
           max_iter=10000, restart=30):
    rp = np.ascontiguousarray(rowptr, np.int32)
    c = np.ascontiguousarray(col, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, np.float64, copy=True)
    it = C.c_int(0)
    res = C.c_double(0)
    conv = lib().or_krylov(rp.size - 1, _ptr(rp), _ptr(c), _ptr(v), _ptr(b), _ptr(x),
                           SOLVERS[solver], PRECONDITIONERS[preconditioner], tol, max_iter,
                           restart, C.byref(it), C.byref(res))
    return x, conv, it.value, res.value


class Oracle:
    """One problem instance in the oracle (mirrors the reference's Solver)."""

    def __init__(self, problem, lumping=None, advection=None, do_mobility_upwind=None):
        pr = problem
        self.problem = pr
        self.points = np.ascontiguousarray(pr.points, np.float64)
        self.cells = np.ascontiguousarray(pr.cells, np.int32)
        self.n = pr.n
        lum = pr.lumping if lumping is None else lumping
        adv = pr.advection if advection is None else advection
        dmu = pr.do_mobility_upwind if do_mobility_upwind is None else do_mobility_upwind
        dim = self.points.shape[1]
        self.p = lib().or_create_shape(dim, _shape(dim, self.cells.shape[1]), self.points.shape[0],
                                       _ptr(self.points), self.cells.shape[0], _ptr(self.cells),
                                       pr.n, lum, adv, dmu)
        if not self.p:
            raise ValueError("or_create failed")
        self.nc = lib().or_num_cells(self.p)
        self.nf = lib().or_num_faces(self.p)
        for name in ("N_ijK", "u_ijKe", "m_iK", "D_ijK", "w_iKe", "a_ijKe", "r_ijK", "f_iK"):
            self.array(name)[:] = pr.coefficient_array(name)
        self.array("tags", np.uint8)[:] = pr.tags
        self.array("bc_values")[:] = pr.bc_values
        self.array("bc_dirichlet")[:] = pr.bc_dirichlet
        if pr.bc_mobility is not None:
            self.array("bc_mobility")[:] = pr.bc_mobility
        self.array("Z_iK")[:] = pr.Z_iK0
        if not dmu:
            # constant mobility model: m_iE_upw is the uniform value 1 (HeatEquationModel.h:130)
            self.array("m_iE_upw")[:] = 1.0
        lib().or_init(self.p)
        if lib().or_setup_linear_system(self.p):
            raise RuntimeError(lib().or_error(self.p).decode())
        self.iterations = []

    def __del__(self):
        if getattr(self, "p", None):
            lib().or_destroy(self.p)
            self.p = None

    def array(self, name, dtype=np.float64):
        size = C.c_long(0)
        ptr = lib().or_array(self.p, name.encode(), C.byref(size))
        if not ptr:
            raise KeyError(name)
        if name in ("cell_faces", "face_cells"):
            dtype = np.int32
        if name == "tags":
            dtype = np.uint8
        ct = {np.float64: C.c_double, np.int32: C.c_int, np.uint8: C.c_uint8}[dtype]
        buf = (ct * size.value).from_address(ptr)
        return np.frombuffer(buf, dtype=dtype)

    def csr(self):
        L = lib()
        rows = L.or_rows(self.p)
        nnz = L.or_nnz(self.p)
        rp = np.frombuffer((C.c_int * (rows + 1)).from_address(L.or_rowptr(self.p)), np.int32)
        col = np.frombuffer((C.c_int * nnz).from_address(L.or_colidx(self.p)), np.int32)
        val = np.frombuffer((C.c_double * nnz).from_address(L.or_values(self.p)), np.float64)
        rhs = self.array("rhs")
        return rp, col, val, rhs

    def pre_iterate(self, t, tau):
        lib().or_pre_iterate(self.p, t, tau)

    def assemble(self, t, tau):
        """Solver::assembleLinearSystem; a time-dependent boundary model is
        evaluated here on the host at time + tau (Solver.hpp:948)."""
        pr = self.problem
        if pr.boundary_functions:
            from mhfem_amd import problems as _problems

            F = self.nf
            if not hasattr(self, "_bc_base"):
                self._bc_base = (np.array(pr.bc_dirichlet, np.float64), np.array(pr.bc_values, np.float64))
            for which, base, name in (("dirichlet", self._bc_base[0], "bc_dirichlet"),
                                      ("neumann", self._bc_base[1], "bc_values")):
                self.array(name)[:] = _problems.boundary_values_at(base, self.n, F, pr.boundary_functions,
                                                                   which)(t + tau)
        elif pr.dirichlet_at is not None:
            self.array("bc_dirichlet")[:] = pr.dirichlet_at(t + tau)
        lib().or_assemble(self.p, t, tau)

    def solve(self, solver=None, preconditioner=None, tol=None, max_iter=100000, restart=30):
        pr = self.problem
        it = C.c_int(0)
        res = C.c_double(0)
        conv = lib().or_solve(self.p, SOLVERS[solver or pr.solver],
                              PRECONDITIONERS[preconditioner or pr.preconditioner],
                              pr.tol if tol is None else tol, max_iter, restart,
                              C.byref(it), C.byref(res))
        if conv != 1:
            raise RuntimeError(lib().or_error(self.p).decode())
        self.iterations.append(it.value)
        return it.value, res.value

    def post_iterate(self, t, tau):
        lib().or_post_iterate(self.p, t, tau)

    def step(self, t, tau, **kw):
        self.pre_iterate(t, tau)
        self.assemble(t, tau)
        out = self.solve(**kw)
        self.post_iterate(t, tau)
        return out

    def run(self, t0, t1, dt, **kw):
        """control.h::solve (control.h:64-102)."""
        t = t0
        while t1 - t > dt * 1e-6:
            tau = min(dt, t1 - t)
            self.step(t, tau, **kw)
            t += tau
        return t
