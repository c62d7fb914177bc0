"""ctypes bindings of libproxpde_amd.so (C ABI in include/proxpde_amd.h and
include/proxpde_amd_host.h).

The library is required: importing the bindings without a built library
raises, there is no Python fallback for any kernel.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libproxpde_amd.so")

FE = {"line_p1": 1, "quad_q1": 2, "quad_q2": 3, "hex_q1": 4, "hex_q2": 5, "tet_p1": 6, "tet_p2": 7,
      "quad_q2q1": 8, "hex_q2q1": 9, "tet_p2p1": 10}
FE_DIM = {1: 1, 2: 2, 3: 2, 4: 3, 5: 3, 6: 3, 7: 3, 8: 2, 9: 3, 10: 3}
MESH = {"line": 1, "square": 2, "cube_hex": 3, "cube_tet5": 4, "cube_tet6": 5}
SIDE = {"bottom": 1, "right": 2, "top": 3, "left": 4, "back": 5, "front": 6}
TERM = {"stiffness": 1, "stiffness_fe": 2, "mass": 3, "mass_fe": 4, "grad": 5, "div": 6, "tensor_stiffness": 7}
RHS = {"none": 0, "const": 1, "sinsin": 2, "sinsinsin": 3, "fevar1d": 4, "qvalues": 5, "projection": 6}
FP64, FP32 = 0, 1
FP32_MF = 2  # pxg_mg_solve: FP32 refined, all levels but the coarsest matrix-free

_lib = None


class PxgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"proxpde_amd error {code}: {msg}")
        self.code = code


class Term(C.Structure):
    _fields_ = [("kind", C.c_int), ("coef", C.c_double)]


def lib():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() or python -m proxpde_amd.build")
    # One HIP runtime per process: when PyTorch is present its runtime is
    # loaded first, so the library binds to it (same soname) and device
    # pointers / streams can be shared with torch tensors.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i32, i64 = C.c_void_p, C.c_int, C.c_int64
    dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
    u32p, i32p, i64p = C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    pp = C.POINTER(C.c_void_p)
    sig = {
        "pxh_mesh_structured": [pp, i32, dp, dp, i32, i32, i32],
        "pxh_mesh_gmsh": [pp, C.c_char_p, i32],
        "pxh_mesh_destroy": [vp],
        "pxh_mesh_info": [vp, i64p, i64p, C.POINTER(C.c_int), i64p],
        "pxh_mesh_get": [vp, dp, u32p],
        "pxh_mesh_set_coords": [vp, dp],
        "pxh_mesh_bfacets": [vp, i32p],
        "pxh_dofmap": [vp, i32, i64p, C.POINTER(C.c_int), u32p],
        "pxh_bc_dofs": [vp, i32, u32p, i32, u32p, i64p],
        "pxh_dof_coords": [vp, i32, u32p, i64, dp],
        "pxg_create": [pp, i32],
        "pxg_destroy": [vp],
        "pxg_set_stream": [vp, vp],
        "pxg_synchronize": [vp],
        "pxg_get_sizes": [vp, i64p, i64p, C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "pxg_profile": [vp, i32],
        "pxg_profile_get": [vp, C.c_char_p, dp, i64p],
        "pxg_set_fe": [vp, i32, i32, dp, dp],
        "pxg_set_mesh": [vp, i64, dp, i64, i32, u32p],
        "pxg_set_dofmap": [vp, i64, u32p],
        "pxg_set_dirichlet": [vp, i64, u32p, dp],
        "pxg_set_coefficient": [vp, dp],
        "pxg_build_pattern": [vp, i64p],
        "pxg_assemble": [vp, i32, C.POINTER(Term), i32],
        "pxg_assemble_rhs": [vp, i32, dp, dp],
        "pxg_get_qpoints": [vp, dp],
        "pxg_close": [vp, C.c_double, i64p],
        "pxg_clear": [vp],
        "pxg_get_csr": [vp, i64p, i32p, dp, fp, dp],
        "pxg_get_pattern": [vp, i64p, i32p, u32p],
        "pxg_get_geometry": [vp, dp, dp],
        "pxg_spmv": [vp, i32, vp, vp],
        "pxg_jacobi": [vp, i32, C.c_double, i32, vp, vp],
        "pxg_jacobi_noalloc": [vp, i32, C.c_double, i32, vp, vp],
        "pxg_csr_device": [vp, C.POINTER(i64p), C.POINTER(i32p), C.POINTER(dp), C.POINTER(fp), i64p],
        "pxg_rhs_device": [vp, C.POINTER(dp)],
        "pxg_matrix_stamp": [vp, C.POINTER(C.c_uint64)],
        "pxg_csr_pack": [vp, i32, C.POINTER(C.c_int)],
        "pxg_mf_apply": [vp, vp, i32, vp, vp, i32],
        "pxg_mf_grad_kappa": [vp, i32, vp, vp, vp, i32],
        "pxg_mg_create": [pp, i32, pp, C.POINTER(C.c_int), C.c_double, i32, i32, C.c_double],
        "pxg_mg_destroy": [vp],
        "pxg_mg_update": [vp],
        "pxg_mg_solve": [vp, i32, vp, vp, C.c_double, i32, C.POINTER(C.c_int), dp],
        "pxg_mg_vcycle": [vp, vp, vp],
        "pxg_mg_prolongation": [vp, i32, i64p, i32p, dp, i64p],
        "pxg_mg_restriction": [vp, i32, i64p, i32p, dp, i64p],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.pxg_last_error.restype = C.c_char_p
    L.pxg_last_error.argtypes = [vp]
    L.pxg_mg_last_error.restype = C.c_char_p
    L.pxg_mg_last_error.argtypes = [vp]
    L.pxh_last_error.restype = C.c_char_p
    L.pxh_last_error.argtypes = []
    _lib = L
    return L


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(rc, ctx=None):
    if rc != 0:
        msg = lib().pxg_last_error(ctx).decode() if ctx else "host layer"
        raise PxgError(rc, msg)


# ---------------------------------------------------------------------------
GEO = {"line": 1, "quad": 2, "hex": 3, "tet": 4}


class HostMesh:
    """Mesh + dof numbering of the host layer (pxh_*): structured
    (HostMesh(kind, n, ...)) or read from a GMSH 2.2 file (HostMesh.gmsh)."""

    def __init__(self, kind: str, n, origin=(0.0, 0.0, 0.0), length=(1.0, 1.0, 1.0), _handle=None):
        L = lib()
        n = list(n) + [0] * (3 - len(n))
        self.kind = kind
        self.n = n
        self.origin = _f64(origin)
        self.length = _f64(length)
        if _handle is None:
            h = C.c_void_p()
            _check(L.pxh_mesh_structured(C.byref(h), MESH[kind], _p(self.origin), _p(self.length), n[0], n[1],
                                         n[2]))
        else:
            h = _handle
        self.h = h
        npts, nelem, nbf = C.c_int64(), C.c_int64(), C.c_int64()
        npe = C.c_int()
        L.pxh_mesh_info(h, C.byref(npts), C.byref(nelem), C.byref(npe), C.byref(nbf))
        self.npts, self.nelem, self.npe, self.nbfacets = npts.value, nelem.value, npe.value, nbf.value

    @classmethod
    def gmsh(cls, path: str, geo: str):
        """readGMSH: volume elements of geometry `geo` ("line", "quad", "hex", "tet")."""
        L = lib()
        h = C.c_void_p()
        rc = L.pxh_mesh_gmsh(C.byref(h), os.fsencode(path), GEO[geo])
        if rc != 0:
            raise PxgError(rc, L.pxh_last_error().decode())
        return cls("gmsh:" + geo, (0, 0, 0), _handle=h)

    def __del__(self):
        try:
            lib().pxh_mesh_destroy(self.h)
        except Exception:
            pass

    def arrays(self):
        xyz = np.empty((self.npts, 3))
        conn = np.empty((self.nelem, self.npe), dtype=np.uint32)
        lib().pxh_mesh_get(self.h, _p(xyz), _p(conn, C.c_uint32))
        return xyz, conn

    def set_coords(self, xyz):
        xyz = _f64(xyz)
        assert xyz.shape == (self.npts, 3)
        _check(lib().pxh_mesh_set_coords(self.h, _p(xyz)))

    def bfacets(self):
        out = np.empty((self.nbfacets, 3), dtype=np.int32)
        lib().pxh_mesh_bfacets(self.h, _p(out, C.c_int32))
        return out

    def dofmap(self, fe: str):
        L = lib()
        nd, nl = C.c_int64(), C.c_int()
        _check(L.pxh_dofmap(self.h, FE[fe], C.byref(nd), C.byref(nl), None))
        emap = np.empty((self.nelem, nl.value), dtype=np.uint32)
        _check(L.pxh_dofmap(self.h, FE[fe], C.byref(nd), C.byref(nl), _p(emap, C.c_uint32)))
        return nd.value, emap

    def bc_dofs(self, fe: str, emap, side: str):
        L = lib()
        emap = np.ascontiguousarray(emap, dtype=np.uint32)
        cnt = C.c_int64(0)
        _check(L.pxh_bc_dofs(self.h, FE[fe], _p(emap, C.c_uint32), SIDE[side], None, C.byref(cnt)))
        out = np.empty(cnt.value, dtype=np.uint32)
        _check(L.pxh_bc_dofs(self.h, FE[fe], _p(emap, C.c_uint32), SIDE[side], _p(out, C.c_uint32), C.byref(cnt)))
        return out

    def dof_coords(self, fe: str, emap, ndofs):
        emap = np.ascontiguousarray(emap, dtype=np.uint32)
        out = np.empty((ndofs, 3))
        _check(lib().pxh_dof_coords(self.h, FE[fe], _p(emap, C.c_uint32), ndofs, _p(out)))
        return out


# ---------------------------------------------------------------------------
class Assembler:
    """One pxg context: FE space on a mesh, its CSR pattern and matrices."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        rc = L.pxg_create(C.byref(h), device)
        if rc != 0:
            raise PxgError(rc, "pxg_create failed: " + L.pxg_last_error(None).decode())
        self.h = h
        self.ndofs = 0
        self.nnz = 0

    def __del__(self):
        try:
            lib().pxg_destroy(self.h)
        except Exception:
            pass

    def _ck(self, rc):
        _check(rc, self.h)

    def set_fe(self, fe: str, qr=None):
        self.fe = fe
        if qr is None:
            self._ck(lib().pxg_set_fe(self.h, FE[fe], 0, None, None))
        else:
            nodes, w = _f64(qr[0]), _f64(qr[1])
            self._ck(lib().pxg_set_fe(self.h, FE[fe], len(w), _p(nodes), _p(w)))

    def set_mesh(self, xyz, conn):
        xyz = _f64(xyz)
        conn = np.ascontiguousarray(conn, dtype=np.uint32)
        self.nelem, npe = conn.shape
        self._ck(lib().pxg_set_mesh(self.h, xyz.shape[0], _p(xyz), self.nelem, npe, _p(conn, C.c_uint32)))

    def set_dofmap(self, ndofs, emap):
        emap = np.ascontiguousarray(emap, dtype=np.uint32)
        self.ndofs = int(ndofs)
        self.nloc = emap.shape[1]
        self._ck(lib().pxg_set_dofmap(self.h, self.ndofs, _p(emap, C.c_uint32)))

    def set_dirichlet(self, dofs, values=None):
        dofs = np.ascontiguousarray(dofs, dtype=np.uint32)
        vals = None if values is None else _f64(values)
        self._ck(lib().pxg_set_dirichlet(self.h, len(dofs), _p(dofs, C.c_uint32), _p(vals)))

    def set_coefficient(self, kappa):
        kappa = _f64(kappa)
        self._ck(lib().pxg_set_coefficient(self.h, _p(kappa)))

    def build_pattern(self):
        nnz = C.c_int64()
        self._ck(lib().pxg_build_pattern(self.h, C.byref(nnz)))
        self.nnz_structural = nnz.value
        return nnz.value

    def assemble(self, terms, precision=FP64):
        arr = (Term * len(terms))(*[Term(TERM[k], float(c)) for k, c in terms])
        self._ck(lib().pxg_assemble(self.h, len(terms), arr, precision))

    def assemble_rhs(self, kind, params=(), qvalues=None):
        prm = _f64(list(params) + [0.0] * 4)
        qv = None if qvalues is None else _f64(qvalues)
        self._ck(lib().pxg_assemble_rhs(self.h, RHS[kind], _p(prm), _p(qv)))

    def qpoints(self, nq):
        out = np.empty((self.nelem, nq, 3))
        self._ck(lib().pxg_get_qpoints(self.h, _p(out)))
        return out

    def geometry(self, nq):
        detJ = np.empty((self.nelem, nq))
        JxW = np.empty((self.nelem, nq))
        self._ck(lib().pxg_get_geometry(self.h, _p(detJ), _p(JxW)))
        return detJ, JxW

    def close(self, eps=1e-16):
        nnz = C.c_int64()
        self._ck(lib().pxg_close(self.h, eps, C.byref(nnz)))
        self.nnz = nnz.value
        return nnz.value

    def clear(self):
        self._ck(lib().pxg_clear(self.h))

    def stamp(self):
        """Changes whenever the closed matrix is replaced or invalidated."""
        st = C.c_uint64()
        self._ck(lib().pxg_matrix_stamp(self.h, C.byref(st)))
        return st.value

    def pack(self, enable=True):
        """SpMV / Jacobi on a block-packed copy of the closed matrix (pxg_csr_pack)."""
        self._ck(lib().pxg_csr_pack(self.h, 1 if enable else 0, None))

    def packed(self):
        """True if SpMV / Jacobi of the current closed matrix read the packed copy."""
        act = C.c_int()
        self._ck(lib().pxg_csr_pack(self.h, -1, C.byref(act)))
        return bool(act.value)

    def csr(self, fp32=False):
        rowptr = np.empty(self.ndofs + 1, dtype=np.int64)
        col = np.empty(self.nnz, dtype=np.int32)
        val = np.empty(self.nnz, dtype=np.float32 if fp32 else np.float64)
        b = np.empty(self.ndofs)
        if fp32:
            self._ck(lib().pxg_get_csr(self.h, _p(rowptr, C.c_int64), _p(col, C.c_int32), None,
                                       _p(val, C.c_float), _p(b)))
        else:
            self._ck(lib().pxg_get_csr(self.h, _p(rowptr, C.c_int64), _p(col, C.c_int32), _p(val), None, _p(b)))
        return rowptr, col, val, b

    def rhs_host(self):
        b = np.empty(self.ndofs)
        self._ck(lib().pxg_get_csr(self.h, None, None, None, None, _p(b)))
        return b

    def pattern(self):
        rowptr = np.empty(self.ndofs + 1, dtype=np.int64)
        col = np.empty(self.nnz_structural, dtype=np.int32)
        emap = np.empty((self.nelem, self.nloc, self.nloc), dtype=np.uint32)
        self._ck(lib().pxg_get_pattern(self.h, _p(rowptr, C.c_int64), _p(col, C.c_int32), _p(emap, C.c_uint32)))
        return rowptr, col, emap

    # device-pointer entry points (pointers are ints, e.g. torch data_ptr())
    def spmv_dev(self, x_ptr, y_ptr, precision=FP64):
        self._ck(lib().pxg_spmv(self.h, precision, x_ptr, y_ptr))

    def jacobi_dev(self, b_ptr, x_ptr, omega, sweeps, precision=FP64):
        self._ck(lib().pxg_jacobi(self.h, precision, omega, sweeps, b_ptr, x_ptr))

    def jacobi_noalloc_dev(self, b_ptr, x_ptr, omega, sweeps, precision=FP64):
        """pxg_jacobi_noalloc: asynchronous on the context stream, no allocation."""
        self._ck(lib().pxg_jacobi_noalloc(self.h, precision, omega, sweeps, b_ptr, x_ptr))

    def mf_apply_dev(self, kappa_ptr, nb, x_ptr, y_ptr, use_bc=False):
        self._ck(lib().pxg_mf_apply(self.h, kappa_ptr, nb, x_ptr, y_ptr, int(use_bc)))

    def mf_grad_kappa_dev(self, nb, x_ptr, ybar_ptr, gk_ptr, use_bc=False):
        self._ck(lib().pxg_mf_grad_kappa(self.h, nb, x_ptr, ybar_ptr, gk_ptr, int(use_bc)))

    def rhs_ptr(self):
        p = C.POINTER(C.c_double)()
        self._ck(lib().pxg_rhs_device(self.h, C.byref(p)))
        return C.cast(p, C.c_void_p).value

    def synchronize(self):
        self._ck(lib().pxg_synchronize(self.h))

    def sizes(self):
        """(n_dofs, n_local, n_elem, device) of the context (pxg_get_sizes)."""
        nd, ne = C.c_int64(), C.c_int64()
        nl, dev = C.c_int(), C.c_int()
        self._ck(lib().pxg_get_sizes(self.h, C.byref(nd), C.byref(ne), C.byref(nl), C.byref(dev)))
        return nd.value, nl.value, ne.value, dev.value

    def profile(self, enable=True):
        self._ck(lib().pxg_profile(self.h, int(enable)))

    def profile_get(self, name):
        ms = C.c_double()
        cnt = C.c_int64()
        self._ck(lib().pxg_profile_get(self.h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def set_stream(self, stream_ptr):
        self._ck(lib().pxg_set_stream(self.h, stream_ptr))


class Multigrid:
    """pxg_mg: geometric MG-preconditioned CG over closed level contexts."""

    def __init__(self, levels, dims, omega=2.0 / 3.0, nu1=2, nu2=2, coarse_rtol=1e-10):
        L = lib()
        self.levels = levels
        arr = (C.c_void_p * len(levels))(*[a.h.value for a in levels])
        d = (C.c_int * (3 * len(dims)))(*[int(v) for dd in dims for v in dd])
        h = C.c_void_p()
        rc = L.pxg_mg_create(C.byref(h), len(levels), arr, d, omega, nu1, nu2, coarse_rtol)
        if rc != 0:
            raise PxgError(rc, "pxg_mg_create failed: " + levels[0].__class__.__name__)
        self.h = h

    def __del__(self):
        try:
            lib().pxg_mg_destroy(self.h)
        except Exception:
            pass

    def _ck(self, rc):
        if rc != 0:
            raise PxgError(rc, lib().pxg_mg_last_error(self.h).decode())

    def update(self):
        self._ck(lib().pxg_mg_update(self.h))

    def solve_dev(self, b_ptr, x_ptr, rtol=1e-8, maxit=500, mode=FP64):
        it = C.c_int(0)
        rr = C.c_double(0)
        self._ck(lib().pxg_mg_solve(self.h, mode, b_ptr, x_ptr, rtol, maxit, C.byref(it), C.byref(rr)))
        return it.value, rr.value

    def vcycle_dev(self, b_ptr, x_ptr):
        self._ck(lib().pxg_mg_vcycle(self.h, b_ptr, x_ptr))

    def prolongation(self, level):
        L = lib()
        nnz = C.c_int64()
        self._ck(L.pxg_mg_prolongation(self.h, level, None, None, None, C.byref(nnz)))
        nr = self.levels[level].ndofs
        rowptr = np.empty(nr + 1, dtype=np.int64)
        col = np.empty(nnz.value, dtype=np.int32)
        val = np.empty(nnz.value)
        self._ck(L.pxg_mg_prolongation(self.h, level, _p(rowptr, C.c_int64), _p(col, C.c_int32), _p(val),
                                       C.byref(nnz)))
        return rowptr, col, val

    def restriction(self, level):
        """The reference's Restrictor (rows of P^T scaled to sum one), host CSR."""
        L = lib()
        nnz = C.c_int64()
        self._ck(L.pxg_mg_restriction(self.h, level, None, None, None, C.byref(nnz)))
        nr = self.levels[level + 1].ndofs
        rowptr = np.empty(nr + 1, dtype=np.int64)
        col = np.empty(nnz.value, dtype=np.int32)
        val = np.empty(nnz.value)
        self._ck(L.pxg_mg_restriction(self.h, level, _p(rowptr, C.c_int64), _p(col, C.c_int32), _p(val),
                                      C.byref(nnz)))
        return rowptr, col, val
