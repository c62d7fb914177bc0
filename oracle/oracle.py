"""ctypes front end of the CPU oracle.

TEST INFRASTRUCTURE ONLY: the oracle is the checker the HIP product path is
compared against.  Only tests/, __graft_entry__.smoke() and the CPU-baseline leg
of bench.py may import this module; the product package never does.

The oracle restates the reference assembly path (include/builder.hpp,
include/curfe.hpp, include/assembly_lhs.hpp, include/assembly_rhs.hpp,
include/dof.hpp, src/mesh.cpp, src/qr.cpp, src/reffe.cpp) on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

FE = {"line_p1": 1, "quad_q1": 2, "quad_q2": 3, "hex_q1": 4, "hex_q2": 5, "tet_p1": 6, "tet_p2": 7}
MESH = {"line": 1, "square": 2, "cube_hex": 3, "cube_tet5": 4, "cube_tet6": 5}
SIDE = {"bottom": 1, "right": 2, "top": 3, "left": 4, "back": 5, "front": 6}
TERM = {"stiffness": 1, "stiffness_fe": 2, "mass": 3, "mass_fe": 4}
RHS = {"none": 0, "const": 1, "sinsin": 2, "sinsinsin": 3, "fevar1d": 4, "qvalues": 5, "projection": 6}

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (plain g++); make rebuilds the
    library when a source is newer.  Where the sources cannot be compiled (no
    g++) an existing library is used as it is."""
    if force or not os.path.exists(_LIB) or shutil.which(os.environ.get("CXX", "g++")):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE, "liboracle.so"])
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        vp, i32, i64, dp, u32p, u8p, i64p = (
            C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_double),
            C.POINTER(C.c_uint32), C.POINTER(C.c_uint8), C.POINTER(C.c_int64))
        L.orc_new.restype = vp
        L.orc_new.argtypes = [i32]
        L.orc_free.argtypes = [vp]
        L.orc_error.restype = C.c_char_p
        L.orc_error.argtypes = [vp]
        L.orc_mesh.argtypes = [vp, i32, dp, dp, i32, i32, i32]
        L.orc_mesh_gmsh.argtypes = [vp, C.c_char_p, i32]
        L.orc_set_coords.argtypes = [vp, dp]
        L.orc_set_qr.argtypes = [vp, i32, dp, dp]
        L.orc_set_threads.argtypes = [vp, i32]
        L.orc_info.restype = i64
        L.orc_info.argtypes = [vp, i32]
        L.orc_get_mesh.argtypes = [vp, dp, u32p]
        L.orc_get_bfacets.argtypes = [vp, C.POINTER(C.c_int32)]
        L.orc_get_elem_map.argtypes = [vp, u32p]
        L.orc_get_dof_coords.argtypes = [vp, dp]
        L.orc_bc_dofs.restype = i64
        L.orc_bc_dofs.argtypes = [vp, i32, u32p]
        L.orc_geometry.argtypes = [vp, i64, i64, dp, dp, dp, dp, dp]
        L.orc_assemble.argtypes = [vp, i32, C.POINTER(C.c_int), dp, dp, u8p, dp, i32, dp, dp, C.c_double]
        L.orc_get_csr.argtypes = [vp, i64p, u32p, dp, dp]
        L.orc_pattern.restype = i64
        L.orc_pattern.argtypes = [vp, i64p, u32p, u32p]
        L.orc_spmv.argtypes = [vp, dp, dp]
        L.orc_jacobi.argtypes = [vp, C.c_double, i32, dp, dp]
        L.orc_mg_new.restype = vp
        L.orc_mg_new.argtypes = [i32, C.POINTER(vp), C.c_double, i32, i32, C.c_double]
        L.orc_mg_free.argtypes = [vp]
        L.orc_mg_vcycle.argtypes = [vp, dp, dp]
        L.orc_mg_pcg.argtypes = [vp, dp, dp, C.c_double, i32, C.POINTER(C.c_int), dp]
        L.orc_mg_prolongation.restype = i64
        L.orc_mg_prolongation.argtypes = [vp, i32, i64p, u32p, dp]
        L.orc_transfer.restype = i64
        L.orc_transfer.argtypes = [vp, vp, i32, i64p, u32p, dp]
        L.orc_mf_apply.argtypes = [vp, dp, u8p, i32, dp, dp]
        L.orc_mf_grad_kappa.argtypes = [vp, u8p, i32, dp, dp, dp]
        L.orc_stokes.argtypes = [vp, i32] + [C.c_double] * 5 + [u8p, dp, i32, C.c_double, i64p, i64p]
        _lib = L
    return _lib


def _p(a, t=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Space:
    """Mesh + FE space + Builder state of the oracle for one problem."""

    def __init__(self, fe: str, mesh: str, n, origin=(0.0, 0.0, 0.0), length=(1.0, 1.0, 1.0), qr=None,
                 gmsh=None):
        """mesh: a structured kind with n elements per axis, or "gmsh" with gmsh = path."""
        self.L = lib()
        self.fe = fe
        self.h = self.L.orc_new(FE[fe])
        if not self.h:
            raise ValueError(f"unknown FE {fe}")
        n = list(n) + [0] * (3 - len(n))
        self.n = n
        o = _f64(origin)
        ln = _f64(length)
        if mesh == "gmsh":
            geo = {"line_p1": 1, "quad_q1": 2, "quad_q2": 2, "hex_q1": 3, "hex_q2": 3, "tet_p1": 4, "tet_p2": 4}[fe]
            rc = self.L.orc_mesh_gmsh(self.h, os.fsencode(gmsh), geo)
        else:
            rc = self.L.orc_mesh(self.h, MESH[mesh], _p(o), _p(ln), int(n[0]), int(n[1]), int(n[2]))
        if rc != 0:
            raise RuntimeError(self.L.orc_error(self.h).decode())
        if qr is not None:
            nodes, weights = _f64(qr[0]), _f64(qr[1])
            self.L.orc_set_qr(self.h, len(weights), _p(nodes), _p(weights))
        self.npts = self.L.orc_info(self.h, 0)
        self.nelem = self.L.orc_info(self.h, 1)
        self.npe = self.L.orc_info(self.h, 2)
        self.ndofs = self.L.orc_info(self.h, 3)
        self.nloc = self.L.orc_info(self.h, 4)
        self.nq = self.L.orc_info(self.h, 5)

    def __del__(self):
        try:
            self.L.orc_free(self.h)
        except Exception:
            pass

    # ---- mesh / dofs
    def mesh(self):
        xyz = np.empty((self.npts, 3))
        conn = np.empty((self.nelem, self.npe), dtype=np.uint32)
        self.L.orc_get_mesh(self.h, _p(xyz), _p(conn, C.c_uint32))
        return xyz, conn

    def set_coords(self, xyz):
        xyz = _f64(xyz)
        assert xyz.shape == (self.npts, 3)
        self.L.orc_set_coords(self.h, _p(xyz))

    def set_threads(self, n: int):
        """Threads of the Builder's element loop and close (1: serial, as the
        reference); the results are bitwise the same for any count."""
        self.L.orc_set_threads(self.h, int(n))

    def bfacets(self):
        nb = self.L.orc_info(self.h, 7)
        out = np.empty((nb, 3), dtype=np.int32)
        self.L.orc_get_bfacets(self.h, _p(out, C.c_int32))
        return out

    def elem_map(self):
        out = np.empty((self.nelem, self.nloc), dtype=np.uint32)
        self.L.orc_get_elem_map(self.h, _p(out, C.c_uint32))
        return out

    def dof_coords(self):
        out = np.empty((self.ndofs, 3))
        self.L.orc_get_dof_coords(self.h, _p(out))
        return out

    def bc_dofs(self, side: str):
        m = SIDE[side]
        k = self.L.orc_bc_dofs(self.h, m, None)
        out = np.empty(k, dtype=np.uint32)
        self.L.orc_bc_dofs(self.h, m, _p(out, C.c_uint32))
        return out

    def interpolate(self, f):
        x = self.dof_coords()
        return np.asarray(f(x), dtype=np.float64)

    # ---- geometry
    def geometry(self, e0=0, e1=None):
        e1 = self.nelem if e1 is None else e1
        ne, nq, n = e1 - e0, self.nq, self.nloc
        dim = {"line_p1": 1, "quad_q1": 2, "quad_q2": 2}.get(self.fe, 3)
        jac = np.empty((ne, nq, 3, dim))
        detJ = np.empty((ne, nq))
        JxW = np.empty((ne, nq))
        qpt = np.empty((ne, nq, 3))
        dphi = np.empty((ne, nq, n, 3))
        rc = self.L.orc_geometry(self.h, e0, e1, _p(jac), _p(detJ), _p(JxW), _p(qpt), _p(dphi))
        assert rc == 0
        return dict(jac=jac, detJ=detJ, JxW=JxW, qpt=qpt, dphi=dphi)

    # ---- assembly
    def assemble(self, terms, kappa=None, bc=None, rhs="none", rhs_params=(), qvalues=None, prune=1e-16):
        """terms: list of (kind, coef); bc: (flag u8[ndofs], val f64[ndofs]) or None."""
        kinds = (C.c_int * max(1, len(terms)))(*[TERM[t[0]] for t in terms])
        coefs = _f64([t[1] for t in terms] or [0.0])
        kap = _f64(kappa) if kappa is not None else None
        flag = val = None
        if bc is not None:
            flag = np.ascontiguousarray(bc[0], dtype=np.uint8)
            val = _f64(bc[1])
        prm = _f64(list(rhs_params) + [0.0] * 4)
        qv = _f64(qvalues) if qvalues is not None else None
        rc = self.L.orc_assemble(self.h, len(terms), kinds, _p(coefs), _p(kap), _p(flag, C.c_uint8),
                                 _p(val), RHS[rhs], _p(prm), _p(qv), prune)
        if rc != 0:
            raise RuntimeError(self.L.orc_error(self.h).decode())
        nnz = self.L.orc_info(self.h, 6)
        rowptr = np.empty(self.ndofs + 1, dtype=np.int64)
        col = np.empty(nnz, dtype=np.uint32)
        valv = np.empty(nnz)
        b = np.empty(self.ndofs)
        self.L.orc_get_csr(self.h, _p(rowptr, C.c_int64), _p(col, C.c_uint32), _p(valv), _p(b))
        return rowptr, col, valv, b

    def stokes(self, interleaved, coefs, flag, val, mode=0, prune=1e-16):
        """Stokes block system of the reference's test/assemblyvector.cpp,
        test/stokes2dquad.cpp, stokes3dhex.cpp and stokes3dtet.cpp on this
        QUAD_Q2 / HEX_Q2 / TET_P2 space (velocity: dim components, BLOCK or
        INTERLEAVED dofs; pressure Q1 / P1 after them).  coefs: dict with
        keys stiffness, tensor, mass (velocity block), grad, div (missing: 0).
        flag / val: Dirichlet data per dof of the velocity (2 * ndofs entries)
        or of the whole system; a constrained pressure dof acts as
        buildLhs(AssemblyDummy{p}, bcsP) plus the couplings' bcsP.  mode 0:
        vector spaces, mode 1: the scalar component spaces.  Returns
        (rowptr, col, val, b)."""
        nd = C.c_int64()
        nnz = C.c_int64()
        k = {"stiffness": 0.0, "tensor": 0.0, "mass": 0.0, "grad": 0.0, "div": 0.0}
        k.update(coefs)
        dim = {"quad_q2": 2, "hex_q2": 3, "tet_p2": 3}[self.fe]
        f = np.zeros(max(len(flag), dim * self.ndofs + self.nelem * 8), dtype=np.uint8)
        v = np.zeros(len(f))
        f[: len(flag)] = np.asarray(flag, dtype=np.uint8)
        v[: len(val)] = np.asarray(val, dtype=np.float64)
        rc = self.L.orc_stokes(self.h, int(interleaved), float(k["stiffness"]), float(k["tensor"]),
                               float(k["mass"]), float(k["grad"]), float(k["div"]), _p(f, C.c_uint8), _p(v),
                               int(mode), prune, C.byref(nd), C.byref(nnz))
        if rc != 0:
            raise RuntimeError(self.L.orc_error(self.h).decode())
        rowptr = np.empty(nd.value + 1, dtype=np.int64)
        col = np.empty(nnz.value, dtype=np.uint32)
        vv = np.empty(nnz.value)
        b = np.empty(nd.value)
        self.L.orc_get_csr(self.h, _p(rowptr, C.c_int64), _p(col, C.c_uint32), _p(vv), _p(b))
        return rowptr, col, vv, b

    def pattern(self):
        nnz = self.L.orc_pattern(self.h, None, None, None)
        rowptr = np.empty(self.ndofs + 1, dtype=np.int64)
        col = np.empty(nnz, dtype=np.uint32)
        emap = np.empty((self.nelem, self.nloc, self.nloc), dtype=np.uint32)
        self.L.orc_pattern(self.h, _p(rowptr, C.c_int64), _p(col, C.c_uint32), _p(emap, C.c_uint32))
        return rowptr, col, emap

    # ---- sparse kernels on the last assembled matrix
    def spmv(self, x):
        x = _f64(x)
        y = np.empty(self.ndofs)
        self.L.orc_spmv(self.h, _p(x), _p(y))
        return y

    def jacobi(self, b, x, omega=2.0 / 3.0, sweeps=1):
        x = _f64(x).copy()
        b = _f64(b)
        self.L.orc_jacobi(self.h, omega, sweeps, _p(b), _p(x))
        return x

    def mf_apply(self, kappa, x, cons=None):
        x = _f64(x)
        nb = 1 if x.ndim == 1 else x.shape[1]
        y = np.empty_like(x)
        c = None if cons is None else np.ascontiguousarray(cons, dtype=np.uint8)
        rc = self.L.orc_mf_apply(self.h, _p(_f64(kappa)), _p(c, C.c_uint8), nb, _p(x), _p(y))
        assert rc == 0
        return y

    def mf_grad_kappa(self, x, ybar, cons=None):
        x = _f64(x)
        ybar = _f64(ybar)
        nb = 1 if x.ndim == 1 else x.shape[1]
        g = np.empty(self.ndofs)
        c = None if cons is None else np.ascontiguousarray(cons, dtype=np.uint8)
        rc = self.L.orc_mf_grad_kappa(self.h, _p(c, C.c_uint8), nb, _p(x), _p(ybar), _p(g))
        assert rc == 0
        return g


def transfer(fine: Space, coarse: Space, which=0):
    L = lib()
    nnz = L.orc_transfer(fine.h, coarse.h, which, None, None, None)
    nr = fine.ndofs if which == 0 else coarse.ndofs
    rowptr = np.empty(nr + 1, dtype=np.int64)
    col = np.empty(nnz, dtype=np.uint32)
    val = np.empty(nnz)
    L.orc_transfer(fine.h, coarse.h, which, _p(rowptr, C.c_int64), _p(col, C.c_uint32), _p(val))
    return rowptr, col, val


class Multigrid:
    """Geometric MG (weighted Jacobi, R = P^T, CG coarse solve) over assembled levels."""

    def __init__(self, levels, omega=2.0 / 3.0, nu1=2, nu2=2, coarse_rtol=1e-10):
        self.L = lib()
        self.levels = levels
        arr = (C.c_void_p * len(levels))(*[s.h for s in levels])
        self.h = self.L.orc_mg_new(len(levels), arr, omega, nu1, nu2, coarse_rtol)
        assert self.h

    def __del__(self):
        try:
            self.L.orc_mg_free(self.h)
        except Exception:
            pass

    def vcycle(self, b):
        b = _f64(b)
        x = np.zeros_like(b)
        self.L.orc_mg_vcycle(self.h, _p(b), _p(x))
        return x

    def pcg(self, b, x0=None, rtol=1e-8, maxit=500):
        b = _f64(b)
        x = np.zeros_like(b) if x0 is None else _f64(x0).copy()
        it = C.c_int(0)
        rr = C.c_double(0)
        self.L.orc_mg_pcg(self.h, _p(b), _p(x), rtol, maxit, C.byref(it), C.byref(rr))
        return x, it.value, rr.value

    def prolongation(self, level):
        nnz = self.L.orc_mg_prolongation(self.h, level, None, None, None)
        nr = self.levels[level].ndofs
        rowptr = np.empty(nr + 1, dtype=np.int64)
        col = np.empty(nnz, dtype=np.uint32)
        val = np.empty(nnz)
        self.L.orc_mg_prolongation(self.h, level, _p(rowptr, C.c_int64), _p(col, C.c_uint32), _p(val))
        return rowptr, col, val
