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


def stokes2dquad_known_answer(space, pspace, solve_system, n=(10, 10)):
    """The reference's stokes2dquad known answer: the Q2-Q1 solution equals the
    interpolant of u = (0, (1 - x^2) / 2), p = nu (1 - y) (errors below 1e-14,
    2e-14, 3e-14 in the 2-norm of the nodal vectors, stokes2dquad.cpp:142-152).
    solve_system(flag, val) returns the solution vector."""
    nu = 0.1
    x = space.dof_coords()
    nU = len(x)
    xp = pspace.dof_coords()
    nd = 2 * nU + len(xp)
    flag = np.zeros(nd, np.uint8)
    val = np.zeros(nd)
    for side, f, comps in (("bottom", lambda p: np.stack([0 * p[:, 0], 0.5 * (1 - p[:, 0] ** 2)], 1), (0, 1)),
                           ("right", lambda p: np.zeros((len(p), 2)), (0, 1)),
                           ("top", lambda p: np.zeros((len(p), 2)), (0,)),
                           ("left", lambda p: np.zeros((len(p), 2)), (0,))):
        ids = np.asarray(space.bc_dofs(side), np.int64)
        v = f(x[ids])
        for d in comps:
            flag[ids + d * nU] = 1
            val[ids + d * nU] = v[:, d]
    sol = solve_system(flag, val, nu)
    u_ex = np.zeros(nU)
    v_ex = 0.5 * (1 - x[:, 0] ** 2)
    p_ex = nu * (1 - xp[:, 1])
    eu = np.linalg.norm(sol[:nU] - u_ex)
    ev = np.linalg.norm(sol[nU:2 * nU] - v_ex)
    ep = np.linalg.norm(sol[2 * nU:] - p_ex)
    return eu, ev, ep


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
