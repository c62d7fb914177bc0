"""The native library loads on the host and exports every function declared
in include/*.h (no device calls)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in ("proxpde_amd.h", "proxpde_amd_host.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?(?:int|char)\s*\*?\s*(px[gh]_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_all_declared_symbols_exported(native):
    L = native.lib()
    names = _declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_codes_without_device(native):
    """Host-side argument checks return PXG_E_INVALID without touching a GPU."""
    import ctypes as C
    L = native.lib()
    assert L.pxg_destroy(None) == 0
    assert L.pxh_mesh_structured(None, 3, None, None, 1, 1, 1) == -1
    h = C.c_void_p()
    o = (C.c_double * 3)(0, 0, 0)
    ln = (C.c_double * 3)(1, 1, 1)
    assert L.pxh_mesh_structured(C.byref(h), 99, o, ln, 1, 1, 1) == -1
    assert L.pxh_dofmap(None, 4, None, None, None) == -1
