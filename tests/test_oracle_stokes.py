"""CPU pins of the oracle's vector spaces and coupling terms (reference
include/builder.hpp:156-277 buildCoupling, include/assembly_coupling.hpp:42-132
AssemblyGrad / AssemblyDiv, include/dof.hpp:112-329 vector dof orderings).

The first check is the reference's own test/assemblyvector.cpp: the Stokes
system assembled on a 2-component Q2 velocity space equals the one assembled
component by component on scalar spaces whose dof maps are shifted by hand, for
the BLOCK and the INTERLEAVED orderings.  The others pin the coupling integrals
against closed forms: the grad rows of interior velocity functions sum to zero
over the pressure basis, div applied to the interpolant of u = (x, 0) gives the
integrals of the pressure basis (sum 1 on the unit square), and grad = div^T."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O


def _sides(o):
    return {s: o.bc_dofs(s) for s in ("bottom", "right", "top", "left")}


def _vel_bc(o, interleaved):
    """test/assemblyvector.cpp: u = 0 on bottom, right, left; u = (1, 0) on top."""
    nU = o.ndofs
    flag = np.zeros(2 * nU, np.uint8)
    val = np.zeros(2 * nU)

    def gid(i, d):
        return 2 * i + d if interleaved else i + d * nU
    sides = _sides(o)
    for s in ("bottom", "right", "top", "left"):
        for i in sides[s]:
            for d in range(2):
                flag[gid(i, d)] = 1
    # later conditions win (BCEss list order): top after bottom/right, left last
    for s in ("bottom", "right", "top", "left"):
        for i in sides[s]:
            val[gid(i, 0)] = 1.0 if s == "top" else 0.0
            val[gid(i, 1)] = 0.0
    return flag, val


def _csr(rp, col, val):
    n = len(rp) - 1
    return sp.csr_matrix((val, col.astype(np.int64), rp), shape=(n, n))


@pytest.mark.parametrize("interleaved", [False, True])
@pytest.mark.parametrize("n", [(3, 3), (4, 2)])
def test_vector_spaces_match_component_spaces(interleaved, n):
    O.build()
    o = O.Space("quad_q2", "square", n)
    flag, val = _vel_bc(o, interleaved)
    rv, cv, vv, bv = o.stokes(interleaved, (1.0, -1.0, -1.0), flag, val, mode=0)
    rs, cs, vs, bs = o.stokes(interleaved, (1.0, -1.0, -1.0), flag, val, mode=1)
    assert np.array_equal(rv, rs) and np.array_equal(cv, cs)
    assert np.array_equal(vv, vs)
    assert np.abs(bv - bs).max() <= 1e-15 * max(1.0, np.abs(bv).max())


def test_coupling_closed_forms():
    O.build()
    n = (3, 4)
    o = O.Space("quad_q2", "square", n)
    nU = o.ndofs
    no = np.zeros(2 * nU, np.uint8)
    zero = np.zeros(2 * nU)
    cg, cd = -1.5, 0.75
    rp, col, v, b = o.stokes(False, (0.0, cg, cd), no, zero, mode=0, prune=-1.0)
    A = _csr(rp, col, v).toarray()
    nd = A.shape[0]
    nP = nd - 2 * nU
    G = A[:2 * nU, 2 * nU:]
    D = A[2 * nU:, :2 * nU]
    assert np.abs(A[:2 * nU, :2 * nU]).max() == 0.0 and np.abs(A[2 * nU:, 2 * nU:]).max() == 0.0
    # grad = (cg / cd) div^T
    assert np.abs(G - (cg / cd) * D.T).max() <= 1e-14 * np.abs(G).max()
    # interior velocity functions: sum over the pressure basis of d phi / dx_c = 0
    x = o.dof_coords()
    interior = (x[:, 0] > 1e-12) & (x[:, 0] < 1 - 1e-12) & (x[:, 1] > 1e-12) & (x[:, 1] < 1 - 1e-12)
    rows = np.concatenate([np.where(interior)[0], nU + np.where(interior)[0]])
    assert np.abs(G[rows].sum(axis=1)).max() <= 1e-14
    # div of the interpolant of u = (x, 0): cd * integral of each pressure basis
    u = np.concatenate([x[:, 0], np.zeros(nU)])
    w = D @ u
    assert abs(w.sum() - cd * 1.0) <= 1e-14
    h = (1.0 / n[0]) * (1.0 / n[1])
    assert abs(w.max() - cd * h) <= 1e-14
    assert nP == (n[0] + 1) * (n[1] + 1)
