"""CPU pins of the oracle's vector spaces and coupling terms (reference
include/builder.hpp:156-277 buildCoupling, include/assembly_coupling.hpp:42-132
AssemblyGrad / AssemblyDiv, include/dof.hpp:112-329 vector dof orderings).

The first check is the reference's own test/assemblyvector.cpp: the Stokes
system assembled on a 2-component Q2 velocity space equals the one assembled
component by component on scalar spaces whose dof maps are shifted by hand, for
the BLOCK and the INTERLEAVED orderings.  The others pin the coupling integrals
against closed forms: the grad rows of interior velocity functions sum to zero
over the pressure basis, div applied to the interpolant of u = (x, 0) gives the
integrals of the pressure basis (sum 1 on the unit square), and grad = div^T;
the cross term of the tensor stiffness is the gradient of the divergence, and
a pinned pressure dof acts as AssemblyDummy plus the couplings' bcsP."""
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
@pytest.mark.parametrize("n,mass", [((3, 3), 0.0), ((4, 2), 0.0), ((3, 2), 2.5)])
def test_vector_spaces_match_component_spaces(interleaved, n, mass):
    O.build()
    o = O.Space("quad_q2", "square", n)
    flag, val = _vel_bc(o, interleaved)
    k = {"stiffness": 1.0, "mass": mass, "grad": -1.0, "div": -1.0}
    rv, cv, vv, bv = o.stokes(interleaved, k, flag, val, mode=0)
    rs, cs, vs, bs = o.stokes(interleaved, k, flag, val, mode=1)
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
    rp, col, v, b = o.stokes(False, {"grad": cg, "div": cd}, no, zero, mode=0, prune=-1.0)
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


def test_tensor_stiffness_is_stiffness_plus_grad_div():
    """AssemblyTensorStiffness = AssemblyStiffness + X with X_(a,di),(j,dj) =
    int d phi_a/dx_di d phi_j/dx_dj: X u = int grad phi_a div u, so X applied
    to the interpolant of a divergence-free field vanishes and for u = (x, 0)
    it equals D^T 1 / cd (div coupling, sum over the pressure basis)."""
    O.build()
    n = (3, 2)
    o = O.Space("quad_q2", "square", n)
    nU = o.ndofs
    no = np.zeros(2 * nU, np.uint8)
    zero = np.zeros(2 * nU)
    ct, cd = 0.7, -1.25
    At = _csr(*o.stokes(False, {"tensor": ct}, no, zero, prune=-1.0)[:3]).toarray()[:2 * nU, :2 * nU]
    As = _csr(*o.stokes(False, {"stiffness": ct}, no, zero, prune=-1.0)[:3]).toarray()[:2 * nU, :2 * nU]
    Ad = _csr(*o.stokes(False, {"div": cd}, no, zero, prune=-1.0)[:3]).toarray()
    D = Ad[2 * nU:, :2 * nU]
    X = At - As
    scale = np.abs(At).max()
    assert np.abs(At - At.T).max() <= 1e-14 * scale
    x = o.dof_coords()
    for u in (np.concatenate([x[:, 1], np.zeros(nU)]), np.concatenate([x[:, 0], -x[:, 1]]),
              np.concatenate([np.zeros(nU), x[:, 0]])):
        assert np.abs(X @ u).max() <= 1e-13 * scale
    u = np.concatenate([x[:, 0], np.zeros(nU)])
    ref = (ct / cd) * (D.T @ np.ones(D.shape[0]))
    assert np.abs(ref).max() > 0.1 * ct
    assert np.abs(X @ u - ref).max() <= 1e-13 * scale


def test_pinned_pressure_acts_as_dummy_with_bcs():
    """A constrained pressure dof: buildLhs(AssemblyDummy{p}, bcsP) puts 1 on
    its diagonal per element and h in b per element; the div rows of that dof
    are cleared, its grad column is lifted into the velocity rhs."""
    O.build()
    n = (3, 3)
    o = O.Space("quad_q2", "square", n)
    nU = o.ndofs
    vflag, vval = _vel_bc(o, False)
    k = {"stiffness": 1.0, "grad": -1.0, "div": -1.0}
    r0, c0, v0, b0 = o.stokes(False, k, vflag, vval)
    A0 = _csr(r0, c0, v0).toarray()
    nd = A0.shape[0]
    flag = np.zeros(nd, np.uint8)
    val = np.zeros(nd)
    flag[:2 * nU] = vflag
    val[:2 * nU] = vval
    pin = 2 * nU + 5  # an interior-edge pressure dof of the 4 x 4 vertex grid (2 elements)
    flag[pin] = 1
    val[pin] = 2.5
    r1, c1, v1, b1 = o.stokes(False, k, flag, val)
    A1 = _csr(r1, c1, v1).toarray()
    count = A1[pin, pin]
    assert count in (1.0, 2.0, 4.0)
    row = A1[pin].copy()
    row[pin] = 0.0
    assert np.abs(row).max() == 0.0 and np.abs(A1[:, pin]).sum() == count
    assert b1[pin] == 2.5 * count
    free = np.where(flag[:2 * nU] == 0)[0]
    assert np.abs((b1 - b0)[free] + 2.5 * A0[free, pin]).max() <= 1e-14
    others = [r for r in range(2 * nU, nd) if r != pin]
    assert np.array_equal(b1[others], b0[others])
    A0[:, pin] = 0.0
    A0[pin, :] = 0.0
    A0[pin, pin] = count
    assert np.abs(A1 - A0).max() == 0.0


def lu_solve(A, b):
    """Direct solve standing in for the reference's LUSolver: SuperLU and one
    step of iterative refinement (without it the pressure error of the
    stokes2dquad case is 7e-14, above the reference's 3e-14, from the
    factorisation's rounding alone)."""
    import scipy.sparse.linalg as spl
    A = A.tocsc()
    lu = spl.splu(A)
    x = lu.solve(b)
    return x + lu.solve(b - A @ x)


def stokes_known_answer(space, pspace, solve_system, dim, bcs, nu, ly=1.0):
    """Known answer of the reference's Stokes tests: with the inlet
    u = (0, (1 - x^2) / 2 [, 0]) and the velocity conditions `bcs` (side,
    constrained components; the inlet on bottom, zero elsewhere), the Q2-Q1 /
    P2-P1 solution is the interpolant of that u and of p = nu (ly - y).
    solve_system(flag, val, nu) returns the solution (BLOCK ordering).
    Returns the 2-norms of the nodal errors of every component and of p."""
    x = space.dof_coords()
    nU = len(x)
    xp = pspace.dof_coords()
    nd = dim * nU + len(xp)
    flag = np.zeros(nd, np.uint8)
    val = np.zeros(nd)
    exact = np.zeros((nU, dim))
    exact[:, 1] = 0.5 * (1 - x[:, 0] ** 2)
    for side, comps in bcs:
        ids = np.asarray(space.bc_dofs(side), np.int64)
        for d in comps:
            flag[ids + d * nU] = 1
            val[ids + d * nU] = exact[ids, d] if side == "bottom" else 0.0
    sol = solve_system(flag, val, nu)
    errs = [np.linalg.norm(sol[d * nU:(d + 1) * nU] - exact[:, d]) for d in range(dim)]
    errs.append(np.linalg.norm(sol[dim * nU:] - nu * (ly - xp[:, 1])))
    return errs


S2D_SIDES = (("bottom", (0, 1)), ("right", (0, 1)), ("top", (0,)), ("left", (0,)))
S3D_SIDES = (("bottom", (0, 1, 2)), ("right", (0, 1, 2)), ("top", (0, 2)), ("left", (0, 2)),
             ("back", (2,)), ("front", (2,)))


def stokes2dquad_known_answer(space, pspace, solve_system, n=(10, 10)):
    """test/stokes2dquad.cpp: errors below 1e-14, 2e-14, 3e-14 for u, v, p
    (stokes2dquad.cpp:142-152) on 10^2 quads, nu = 0.1."""
    return tuple(stokes_known_answer(space, pspace, solve_system, 2, S2D_SIDES, 0.1))


def test_stokes2dquad_known_answer():
    O.build()
    n = (10, 10)
    o = O.Space("quad_q2", "square", n)
    op = O.Space("quad_q1", "square", n)

    def solve(flag, val, nu):
        rp, col, v, b = o.stokes(False, {"tensor": nu, "grad": -1.0, "div": -1.0}, flag, val)
        return lu_solve(_csr(rp, col, v), b)
    eu, ev, ep = stokes2dquad_known_answer(o, op, solve, n)
    print(f"errors u {eu:.2e} v {ev:.2e} p {ep:.2e}")
    assert eu <= 1e-14 and ev <= 2e-14 and ep <= 3e-14


@pytest.mark.parametrize("fe,mesh", [("tet_p2", "cube_tet5"), ("hex_q2", "cube_hex")])
@pytest.mark.parametrize("interleaved", [False, True])
def test_vector_spaces_match_component_spaces_3d(fe, mesh, interleaved):
    """assemblyvector.cpp's equivalence on the 3D pairs (P2-P1 tets, Q2-Q1 hexes)."""
    O.build()
    n = (2, 2, 2)
    o = O.Space(fe, mesh, n)
    nU = o.ndofs
    flag = np.zeros(3 * nU, np.uint8)
    val = np.zeros(3 * nU)
    x = o.dof_coords()
    for side in ("bottom", "left", "front"):
        ids = np.asarray(o.bc_dofs(side), np.int64)
        for d in range(3):
            g = 3 * ids + d if interleaved else ids + d * nU
            flag[g] = 1
            val[g] = x[ids, d] + 0.5 * d
    k = {"stiffness": 0.7, "mass": 1.5, "grad": -1.0, "div": -2.0}
    rv, cv, vv, bv = o.stokes(interleaved, k, flag, val, mode=0)
    rs, cs, vs, bs = o.stokes(interleaved, k, flag, val, mode=1)
    assert np.array_equal(rv, rs) and np.array_equal(cv, cs)
    assert np.array_equal(vv, vs)
    assert np.abs(bv - bs).max() <= 1e-15 * max(1.0, np.abs(bv).max())


# test/stokes3dtet.cpp:122-127 and test/stokes3dhex.cpp:126-131: checkError
# (include/def.hpp:357-367) asks |error - ref| < 1e-12 for every norm
STOKES3D_REF = {
    "tet_p2": (1.191610761242353e-15, 4.347071660478249e-15, 1.168951507916659e-15, 5.36156027085171e-14),
    "hex_q2": (1.309442419567852e-15, 1.021847252919293e-14, 1.800953966048908e-15, 1.097009563518918e-14),
}
STOKES3D_MESH = {"tet_p2": ("cube_tet5", "tet_p1"), "hex_q2": ("cube_hex", "hex_q1")}


@pytest.mark.parametrize("fe", ["tet_p2", "hex_q2"])
def test_stokes3d_known_answer(fe):
    O.build()
    n = (4, 4, 4)
    mesh, pfe = STOKES3D_MESH[fe]
    ln = (1.0, 2.0, 1.0)
    o = O.Space(fe, mesh, n, length=ln)
    op = O.Space(pfe, mesh, n, length=ln)

    def solve(flag, val, nu):
        rp, col, v, b = o.stokes(False, {"tensor": nu, "grad": -1.0, "div": -1.0}, flag, val)
        return lu_solve(_csr(rp, col, v), b)
    errs = stokes_known_answer(o, op, solve, 3, S3D_SIDES, 0.1, ly=2.0)
    print(fe, ["%.3e" % e for e in errs])
    for e, r in zip(errs, STOKES3D_REF[fe]):
        assert abs(e - r) < 1e-12
