"""LibTorch autograd operators: y = K(kappa) x and u = K(kappa)^{-1} b.
Gradients are checked against central finite differences (step 1e-6, FP64)
and against the oracle's adjoint gradient."""
import numpy as np
import pytest
import torch

from proxpde_amd import native as N
from proxpde_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ag(gpu):
    from proxpde_amd import autograd as A
    A.load()
    return A


def test_kappa_apply_gradcheck(ag, gpu):
    prob = W.Problem("hex_q2", (2, 2, 1))
    a = prob.assembler()
    op = ag.KappaOperator(a)
    g = torch.Generator(device=gpu).manual_seed(3)
    kappa = (0.2 * torch.randn(prob.ndofs, dtype=torch.float64, device=gpu, generator=g)).exp().requires_grad_()
    x = torch.randn(prob.ndofs, 2, dtype=torch.float64, device=gpu, generator=g).requires_grad_()
    assert torch.autograd.gradcheck(lambda k, v: op(k, v), (kappa, x), eps=1e-6, atol=1e-7, rtol=1e-6)


def test_kappa_apply_matches_oracle_gradient(ag, oracle, gpu):
    """C4 shape in small: Q2 hex, batch of 8 right-hand sides."""
    n = (3, 3, 3)
    prob = W.Problem("hex_q2", n)
    o = oracle.Space("hex_q2", "cube_hex", n)
    a = prob.assembler()
    op = ag.KappaOperator(a)
    kap = np.exp(W.grf(prob.dof_coords(), seed=2, ell=0.3, features=64))
    rng = np.random.default_rng(3)
    x = rng.normal(size=(prob.ndofs, 8))
    w = rng.normal(size=(prob.ndofs, 8))
    kt = torch.from_numpy(kap).to(gpu).requires_grad_()
    xt = torch.from_numpy(x).to(gpu).requires_grad_()
    y = op(kt, xt)
    loss = (y * torch.from_numpy(w).to(gpu)).sum()
    loss.backward()
    g_ref = o.mf_grad_kappa(x, w)
    assert np.abs(kt.grad.cpu().numpy() - g_ref).max() <= 1e-12 * np.abs(g_ref).max()
    gx_ref = o.mf_apply(kap, w)
    assert np.abs(xt.grad.cpu().numpy() - gx_ref).max() <= 1e-12 * np.abs(gx_ref).max()


@pytest.mark.parametrize("fe,ns", [("hex_q1", (8, 4, 2)), ("hex_q2", (4, 2))])
def test_kappa_solve_adjoint_gradient(ag, oracle, gpu, fe, ns):
    solver = ag.KappaSolver(fe, ns)
    x = solver.probs[0].dof_coords()
    kap0 = np.exp(0.5 * np.sin(3 * x[:, 0] + x[:, 1]) * np.cos(2 * x[:, 2]))
    b = torch.from_numpy(solver.rhs(1.0)).to(gpu)
    ut = torch.from_numpy(0.01 * np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1]) * np.sin(np.pi * x[:, 2])).to(gpu)

    def J(k, rtol=1e-12):
        u = solver(k, b, rtol=rtol)
        return 0.5 * ((u - ut) ** 2).sum()

    kt = torch.from_numpy(kap0).to(gpu).requires_grad_()
    Jv = J(kt)
    Jv.backward()
    g = kt.grad.cpu().numpy()
    # central differences along a random direction
    rng = np.random.default_rng(7)
    d = rng.normal(size=len(kap0))
    h = 1e-6
    with torch.no_grad():
        jp = float(J(torch.from_numpy(kap0 + h * d).to(gpu)))
        jm = float(J(torch.from_numpy(kap0 - h * d).to(gpu)))
    fd = (jp - jm) / (2 * h)
    an = float(g @ d)
    print(f"  {fe}: directional derivative adjoint={an:.10e} FD={fd:.10e}")
    assert abs(fd - an) <= 1e-5 * abs(an)
    # oracle adjoint gradient: lambda = K^{-1}(u - ut), g = -dK/dkappa(lambda, u)
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    o = oracle.Space(fe, "cube_hex", (ns[0],) * 3)
    cons = np.zeros(o.ndofs, np.uint8)
    cons[solver.constrained] = 1
    rp, col, v, bo = o.assemble([("stiffness_fe", 1.0)], kappa=kap0, bc=(cons, np.zeros(o.ndofs)), rhs="const",
                                rhs_params=(1.0,))
    A = sp.csr_matrix((v, col.astype(np.int64), rp)).tocsc()
    u = spl.spsolve(A, bo)
    lam = spl.spsolve(A, u - ut.cpu().numpy())
    g_ref = -o.mf_grad_kappa(u, lam, cons)
    assert np.abs(g - g_ref).max() <= 1e-7 * np.abs(g_ref).max()


@pytest.mark.parametrize("fe,ns", [("hex_q1", (16, 8, 4, 2)), ("hex_q2", (8, 4, 2))])
@pytest.mark.parametrize("variant", ["fp32", "fp32_fine64", "fp32_mf"])
def test_kappa_solve_fp32_refined_matches_fp64(ag, gpu, fe, ns, variant):
    """The refined forms (FP32 levels with the matrix-free FP64 residual; the
    finest level in FP64; all levels but the coarsest matrix-free in FP32)
    against the FP64 solve, solution and adjoint gradient."""
    solver = ag.KappaSolver(fe, ns)
    x = solver.probs[0].dof_coords()
    kap = torch.from_numpy(np.exp(np.sin(4 * x[:, 0]) * np.cos(3 * x[:, 1]))).to(gpu)
    b = torch.from_numpy(solver.rhs(1.0)).to(gpu)
    mode = N.FP32_MF if variant == "fp32_mf" else N.FP32
    out = {}
    # the refined form first: straight after solver.rhs() (which clears the
    # finest level), before any FP64 solve has closed it again
    for name, md, f64 in ((variant, mode, variant == "fp32_fine64"), ("fp64", N.FP64, False)):
        k = kap.clone().requires_grad_()
        u = solver(k, b, rtol=1e-8, mode=md, fine_fp64=f64)
        (0.5 * (u ** 2).sum()).backward()
        out[name] = (u.detach(), k.grad.detach())
    err = float((out[variant][0] - out["fp64"][0]).norm() / out["fp64"][0].norm())
    gerr = float((out[variant][1] - out["fp64"][1]).norm() / out["fp64"][1].norm())
    print(f"  {fe} {variant} vs FP64: solution {err:.2e}, gradient {gerr:.2e}")
    assert err < 1e-6 and gerr < 1e-6


def test_kappa_solve_backward_reassembles_only_when_needed(ag, gpu):
    """The backward pass reuses the level matrices of the forward pass unless a
    level was cleared or closed in between; both routes give one gradient."""
    solver = ag.KappaSolver("hex_q2", (4, 2))
    x = solver.probs[0].dof_coords()
    kap = torch.from_numpy(np.exp(np.sin(3 * x[:, 0]) * np.cos(2 * x[:, 1]))).to(gpu)
    b = torch.from_numpy(solver.rhs(1.0)).to(gpu)
    grads = []
    for intervene in (False, True):
        k = kap.clone().requires_grad_()
        u = solver(k, b, rtol=1e-12)
        stamps = [c.stamp() for c in solver.ctx]
        if intervene:
            solver.rhs(1.0)  # clears the finest level: the matrices must be rebuilt
            assert solver.ctx[0].stamp() != stamps[0]
        (0.5 * (u ** 2).sum()).backward()
        if not intervene:
            assert [c.stamp() for c in solver.ctx] == stamps
        grads.append(k.grad.clone())
    assert float((grads[0] - grads[1]).abs().max()) <= 1e-12 * float(grads[0].abs().max())


def test_operators_check_lengths_and_batch(ag, gpu):
    """Tensors whose length is not the context's dof count, or with more than
    16 right-hand sides, are rejected before any kernel runs."""
    prob = W.Problem("hex_q1", (2, 2, 2))
    a = prob.assembler()
    op = ag.KappaOperator(a)
    n = prob.ndofs
    k = torch.ones(n, dtype=torch.float64, device=gpu)
    with pytest.raises(RuntimeError, match="rows"):
        op(torch.ones(n - 1, dtype=torch.float64, device=gpu), torch.ones(n, 2, dtype=torch.float64, device=gpu))
    with pytest.raises(RuntimeError, match="rows"):
        op(k, torch.ones(n + 3, 2, dtype=torch.float64, device=gpu))
    with pytest.raises(RuntimeError, match="right-hand sides"):
        op(k, torch.ones(n, 17, dtype=torch.float64, device=gpu))
    y = op(k, torch.ones(n, 16, dtype=torch.float64, device=gpu))
    assert y.shape == (n, 16)
    solver = ag.KappaSolver("hex_q1", [4, 2])
    ks = torch.ones(solver.ndofs, dtype=torch.float64, device=gpu)
    with pytest.raises(RuntimeError, match="rows"):
        solver(ks, torch.ones(solver.ndofs + 1, dtype=torch.float64, device=gpu))
