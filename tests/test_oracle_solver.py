"""CPU pins of the oracle's operator-apply half: SpMV, weighted Jacobi,
prolongation / restriction, the multigrid cycle and PCG, and the matrix-free
kappa operator with its gradient.  Each check is against something the
oracle's formula does not restate: the reference's own known answers
(test/prolongation.cpp:209, test/restriction.cpp:102, carried to an extruded
hex mesh by an exact identity), polynomial reproduction, scipy routines,
symmetry, and exact linearity in kappa."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import oracle as O


def _csr(rowptr, col, val, n, m=None):
    return sp.csr_matrix((val, col.astype(np.int64), rowptr), shape=(n, n if m is None else m))


def _space(fe, n, kappa=None, sides=True, terms=(("stiffness_fe", 1.0),)):
    o = O.Space(fe, "cube_hex", n)
    flag = np.zeros(o.ndofs, np.uint8)
    if sides:
        for s in ("left", "right", "bottom", "top", "back", "front"):
            flag[o.bc_dofs(s)] = 1
    kap = np.ones(o.ndofs) if kappa is None else kappa(o.dof_coords())
    rp, col, val, b = o.assemble(list(terms), kappa=kap, bc=(flag, np.zeros(o.ndofs)), rhs="const",
                                 rhs_params=(1.0,))
    return o, _csr(rp, col, val, o.ndofs), b, flag, kap


def _kappa(x):
    return np.exp(0.5 * np.sin(3 * x[:, 0]) * np.cos(2 * x[:, 1] + x[:, 2]))


def test_transfer_known_answers_extruded():
    """Q1 on 3x2 -> 6x4 with u = (1-x)^2 (reference test/prolongation.cpp:209 and
    test/restriction.cpp:102, Q1 quad).  Extruded to hex 3x2x1 -> 6x4x2: u does
    not depend on z and the Q1 transfer is a tensor product, so every z-layer of
    nodes repeats the 2D values: |P u_c| = sqrt(3) * 3.009757793463 (3 fine
    layers) and |R u_f| = sqrt(2) * 1.759929209669 (2 coarse layers, R = rows of
    P^T scaled to sum one)."""
    O.build()
    fine = O.Space("hex_q1", "cube_hex", (6, 4, 2))
    coarse = O.Space("hex_q1", "cube_hex", (3, 2, 1))
    f = lambda x: (1.0 - x[:, 0]) ** 2
    P = _csr(*O.transfer(fine, coarse, 0), fine.ndofs, coarse.ndofs)
    R = _csr(*O.transfer(fine, coarse, 1), coarse.ndofs, fine.ndofs)
    uf = P @ coarse.interpolate(f)
    uc = R @ fine.interpolate(f)
    assert abs(np.linalg.norm(uf) - np.sqrt(3.0) * 3.009757793463) <= 1e-11
    assert abs(np.linalg.norm(uc) - np.sqrt(2.0) * 1.759929209669) <= 1e-11


@pytest.mark.parametrize("fe,f", [
    ("hex_q1", lambda x: 1 + 2 * x[:, 0] - x[:, 1] + 3 * x[:, 0] * x[:, 1] * x[:, 2]),
    ("hex_q2", lambda x: (x[:, 0] ** 2) * (1 - x[:, 1]) * x[:, 2] ** 2 + x[:, 1] ** 2 - 0.5 * x[:, 0] * x[:, 2])])
def test_prolongation_reproduces_coarse_space(fe, f):
    """Every coarse-space function is a fine-space function: P interpolates it exactly."""
    O.build()
    fine = O.Space(fe, "cube_hex", (4, 6, 2))
    coarse = O.Space(fe, "cube_hex", (2, 3, 1))
    P = _csr(*O.transfer(fine, coarse, 0), fine.ndofs, coarse.ndofs)
    assert np.abs(P @ coarse.interpolate(f) - fine.interpolate(f)).max() <= 1e-13
    # partition of unity of the restriction rows
    R = _csr(*O.transfer(fine, coarse, 1), coarse.ndofs, fine.ndofs)
    assert np.abs(R @ np.ones(fine.ndofs) - 1.0).max() <= 1e-14


def test_spmv_and_jacobi_against_scipy():
    O.build()
    o, A, b, flag, _ = _space("hex_q2", (3, 2, 2), _kappa)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, o.ndofs)
    y = A @ x
    assert np.abs(o.spmv(x) - y).max() <= 1e-14 * np.abs(y).max()
    d = A.diagonal()
    x1 = x + (2.0 / 3.0) * (b - A @ x) / d
    x2 = x1 + (2.0 / 3.0) * (b - A @ x1) / d
    assert np.abs(o.jacobi(b, x, 2.0 / 3.0, 2) - x2).max() <= 1e-13 * np.abs(x2).max()


def _mg(fe, ns, kappa=_kappa):
    levels = [_space(fe, (n, n, n), kappa) for n in ns]
    return levels, O.Multigrid([lv[0] for lv in levels])


@pytest.mark.parametrize("fe,ns", [("hex_q1", (8, 4, 2)), ("hex_q2", (4, 2))])
def test_multigrid_cycle_and_pcg(fe, ns):
    O.build()
    levels, mg = _mg(fe, ns)
    o0, A, b, flag, _ = levels[0]
    u = spl.spsolve(A.tocsc(), b)
    # one level: the cycle is the direct coarse solve
    _, one = _mg(fe, ns[-1:])
    oc, Ac, bc, _, _ = _space(fe, (ns[-1],) * 3, _kappa)
    assert np.abs(one.vcycle(bc) - spl.spsolve(Ac.tocsc(), bc)).max() <= 1e-12 * np.abs(bc).max()
    # the cycle with nu1 = nu2 Jacobi sweeps and R = P^T is a symmetric operator
    rng = np.random.default_rng(5)
    b1, b2 = rng.uniform(-1, 1, (2, o0.ndofs))
    b1[flag == 1] = 0.0
    b2[flag == 1] = 0.0
    v1, v2 = mg.vcycle(b1), mg.vcycle(b2)
    assert abs(v1 @ b2 - b1 @ v2) <= 1e-12 * abs(v1 @ b2)
    # and a good preconditioner: PCG reaches 1e-10 in a few iterations
    x, it, rr = mg.pcg(b, rtol=1e-10)
    assert rr <= 1e-10 and it <= 20
    assert np.linalg.norm(x - u) <= 1e-8 * np.linalg.norm(u)


@pytest.mark.parametrize("use_bc", [False, True])
def test_matrix_free_operator_and_gradient(use_bc):
    """mf_apply equals the assembled kappa-stiffness (with the Dirichlet form when
    constrained dofs are given); the gradient is exact because x.K(kappa)y is
    linear in kappa: g_n = x.K(e_n)y, assembled with kappa = e_n."""
    O.build()
    n = (2, 1, 2)
    o, A, _, flag, kap = _space("hex_q2", n, _kappa, sides=use_bc)
    cons = flag if use_bc else None
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, o.ndofs)
    y = rng.uniform(-1, 1, o.ndofs)
    Ax = A @ x
    assert np.abs(o.mf_apply(kap, x, cons) - Ax).max() <= 1e-13 * np.abs(Ax).max()
    g = o.mf_grad_kappa(x, y, cons)
    ge = np.empty(o.ndofs)
    for i in range(o.ndofs):
        e = np.zeros(o.ndofs)
        e[i] = 1.0
        rp, col, val, _ = o.assemble([("stiffness_fe", 1.0)], kappa=e, prune=-1.0)
        Ki = _csr(rp, col, val, o.ndofs)
        if use_bc:
            # derivative of C K C + D_c: the constrained part does not depend on kappa
            m = (flag == 0).astype(float)
            ge[i] = (m * y) @ (Ki @ (m * x))
        else:
            ge[i] = y @ (Ki @ x)
    assert np.abs(g - ge).max() <= 1e-12 * np.abs(ge).max()
