"""Operator-apply parity through the C ABI: CSR SpMV and weighted Jacobi on the
assembled matrix, prolongation / V-cycle / MG-PCG (FP64 and FP32 + FP64
refinement), and the matrix-free kappa operator with its gradients, all
against the oracle on the same inputs."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from proxpde_amd import workloads as W

pytestmark = pytest.mark.gpu


def _torch_stream(a):
    """Run the context on torch's current stream: the device-pointer entry
    points read torch tensors made on that stream (the context's own stream
    is non-blocking and would not wait for them)."""
    import torch
    a.set_stream(torch.cuda.current_stream().cuda_stream)
    return a


def _hex_level(fe, n, kappa_fn, sides=True, own_stream=False):
    prob = W.Problem(fe, (n, n, n))
    a = prob.assembler() if own_stream else _torch_stream(prob.assembler())
    dofs = prob.all_sides() if sides else np.zeros(0, np.uint32)
    a.set_dirichlet(dofs, np.zeros(len(dofs)))
    kap = kappa_fn(prob.dof_coords())
    a.set_coefficient(kap)
    a.build_pattern()
    a.assemble([("stiffness_fe", 1.0)])
    a.assemble_rhs("const", (1.0,))
    a.close()
    return prob, a, kap, dofs


def _oracle_level(oracle, fe, n, kap, dofs):
    o = oracle.Space(fe, "cube_hex", (n, n, n))
    flag = np.zeros(o.ndofs, np.uint8)
    flag[dofs] = 1
    o.assemble([("stiffness_fe", 1.0)], kappa=kap, bc=(flag, np.zeros(o.ndofs)), rhs="const", rhs_params=(1.0,))
    return o


def _kappa(x):
    return np.exp(0.5 * np.sin(3 * x[:, 0]) * np.cos(2 * x[:, 1] + x[:, 2]))


def test_spmv_and_jacobi(oracle, gpu):
    import torch
    prob, a, kap, dofs = _hex_level("hex_q2", 4, _kappa)
    o = _oracle_level(oracle, "hex_q2", 4, kap, dofs)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, prob.ndofs)
    xt = torch.from_numpy(x).to(gpu)
    yt = torch.empty_like(xt)
    a.spmv_dev(xt.data_ptr(), yt.data_ptr())
    a.synchronize()
    y_ref = o.spmv(x)
    assert np.abs(yt.cpu().numpy() - y_ref).max() <= 1e-13 * np.abs(y_ref).max()
    # FP32 SpMV
    x32 = xt.float()
    y32 = torch.empty_like(x32)
    a.spmv_dev(x32.data_ptr(), y32.data_ptr(), precision=1)
    a.synchronize()
    assert np.abs(y32.cpu().numpy() - y_ref).max() <= 1e-5 * np.abs(y_ref).max()
    # weighted Jacobi
    b = rng.uniform(-1, 1, prob.ndofs)
    bt = torch.from_numpy(b).to(gpu)
    x0 = torch.zeros_like(bt)
    a.jacobi_dev(bt.data_ptr(), x0.data_ptr(), 2.0 / 3.0, 3)
    a.synchronize()
    x_ref = o.jacobi(b, np.zeros(prob.ndofs), 2.0 / 3.0, 3)
    assert np.abs(x0.cpu().numpy() - x_ref).max() <= 1e-13 * np.abs(x_ref).max()


@pytest.mark.parametrize("fe,n,packed", [("hex_q2", 6, True), ("hex_q1", 60, True), ("hex_q2", 32, False)])
def test_packed_spmv_bitwise(gpu, fe, n, packed):
    """SpMV and Jacobi on the block-packed copy equal the plain CSR kernels bit
    for bit.  The dofs are renumbered at random, so row blocks span up to all
    columns: Q1 60^3 (226,981 dofs) uses every high-bit pattern of the 18-bit
    offsets; Q2 32^3 (274,625 dofs) has blocks spanning 2^18 columns or more,
    where the pack is skipped and the plain kernel must run."""
    import torch
    prob = W.Problem(fe, (n, n, n))
    perm = np.random.default_rng(1).permutation(prob.ndofs).astype(np.uint32)
    a = _torch_stream(prob.assembler())
    a.set_dofmap(prob.ndofs, perm[prob.emap])
    dofs = perm[prob.all_sides()]
    a.set_dirichlet(dofs, np.zeros(len(dofs)))
    a.set_coefficient(_kappa(prob.dof_coords())[np.argsort(perm)])
    a.build_pattern()
    a.assemble([("stiffness_fe", 1.0)])
    a.close()
    rng = np.random.default_rng(2)
    x = torch.from_numpy(rng.uniform(-1, 1, prob.ndofs)).to(gpu)
    bt = torch.from_numpy(rng.uniform(-1, 1, prob.ndofs)).to(gpu)
    out = {}
    for pack in (False, True):
        a.pack(pack)
        for prec, dt in ((0, torch.float64), (1, torch.float32)):
            xv = x.to(dt)
            y = torch.empty_like(xv)
            a.spmv_dev(xv.data_ptr(), y.data_ptr(), precision=prec)
            xj = torch.zeros_like(xv)
            a.jacobi_dev(bt.to(dt).data_ptr(), xj.data_ptr(), 2.0 / 3.0, 3, precision=prec)
            a.synchronize()
            out[pack, prec] = (y.cpu().numpy(), xj.cpu().numpy())
    for prec in (0, 1):
        for k in range(2):
            assert np.array_equal(out[False, prec][k], out[True, prec][k])
    # the packed copy exists only where every block fits the 18-bit offsets
    assert a.packed() == packed


def _levels(oracle, fe, ns):
    gl, ol = [], []
    for n in ns:
        prob, a, kap, dofs = _hex_level(fe, n, _kappa)
        gl.append((prob, a))
        ol.append(_oracle_level(oracle, fe, n, kap, dofs))
    return gl, ol


@pytest.mark.parametrize("fe,ns", [("hex_q1", (8, 4, 2)), ("hex_q2", (4, 2))])
def test_prolongation_and_vcycle(oracle, native, gpu, fe, ns):
    import torch
    gl, ol = _levels(oracle, fe, ns)
    mg = native.Multigrid([a for _, a in gl], [(n, n, n) for n in ns])
    omg = oracle.Multigrid(ol)
    for l in range(len(ns) - 1):
        rp, col, val = mg.prolongation(l)
        rpo, colo, valo = omg.prolongation(l)
        assert np.array_equal(rp, rpo) and np.array_equal(col, colo.astype(np.int32))
        assert np.array_equal(val, valo)
    rng = np.random.default_rng(1)
    b = rng.uniform(-1, 1, gl[0][0].ndofs)
    b[gl[0][0].all_sides()] = 0.0
    bt = torch.from_numpy(b).to(gpu)
    xt = torch.zeros_like(bt)
    mg.vcycle_dev(bt.data_ptr(), xt.data_ptr())
    x_ref = omg.vcycle(b)
    # both coarsest solves are direct (dense inverse / dense Cholesky)
    assert np.abs(xt.cpu().numpy() - x_ref).max() <= 1e-11 * np.abs(x_ref).max()


def test_coarse_direct_solve_and_stamp(native, gpu):
    """A one-level hierarchy is the coarsest solve itself: the exact solution;
    re-assembly changes the matrix stamp and the next cycle refactors."""
    import torch
    prob, a, kap, dofs = _hex_level("hex_q2", 3, _kappa)
    mg = native.Multigrid([a], [(3, 3, 3)])
    for scale in (1.0, 3.0):
        rp, col, val, b = a.csr()
        A = sp.csr_matrix((val, col, rp), shape=(prob.ndofs, prob.ndofs))
        u = spl.spsolve(A.tocsc(), b)
        bt = torch.from_numpy(b).to(gpu)
        xt = torch.zeros_like(bt)
        mg.vcycle_dev(bt.data_ptr(), xt.data_ptr())
        assert np.abs(xt.cpu().numpy() - u).max() <= 1e-12 * np.abs(u).max()
        s0 = a.stamp()
        a.set_coefficient(scale * 3.0 * kap)
        a.clear()
        assert a.stamp() != s0
        a.assemble([("stiffness_fe", 1.0)])
        a.assemble_rhs("const", (1.0,))
        a.close()


def test_prolongation_known_answer(native, gpu):
    """Reference test/prolongation.cpp:209 (Q1, 3x2 -> 6x4, u = (1-x)^2, |P u_c| =
    3.009757793463), extruded to hex 3x2x1 -> 6x4x2 where each of the 3 fine
    z-layers repeats the 2D values: |P u_c| = sqrt(3) * 3.009757793463.  No
    Dirichlet dofs, so no rows or columns of P are removed."""
    ctxs = []
    for n in ((6, 4, 2), (3, 2, 1)):
        prob = W.Problem("hex_q1", n)
        a = _torch_stream(prob.assembler())
        a.set_dirichlet(np.zeros(0, np.uint32), np.zeros(0))
        a.set_coefficient(np.ones(prob.ndofs))
        a.build_pattern()
        a.assemble([("stiffness_fe", 1.0), ("mass", 1.0)])
        a.close()
        ctxs.append((prob, a))
    mg = native.Multigrid([a for _, a in ctxs], [(6, 4, 2), (3, 2, 1)])
    rp, col, val = mg.prolongation(0)
    P = sp.csr_matrix((val, col, rp), shape=(ctxs[0][0].ndofs, ctxs[1][0].ndofs))
    uc = (1.0 - ctxs[1][0].dof_coords()[:, 0]) ** 2
    assert abs(np.linalg.norm(P @ uc) - np.sqrt(3.0) * 3.009757793463) <= 1e-11


def test_restriction_known_answer_and_oracle(oracle, native, gpu):
    """Reference test/restriction.cpp:102 (Q1, 6x4 -> 3x2, u = (1-x)^2,
    |R u_f| = 1.759929209669, R = rows of P^T scaled to sum one), extruded to
    hex 6x4x2 -> 3x2x1 where each of the 2 coarse z-layers repeats the 2D
    values: |R u_f| = sqrt(2) * 1.759929209669.  The exported restriction is
    also the oracle's (pattern bit-exact, values within 1e-15)."""
    ctxs = []
    for n in ((6, 4, 2), (3, 2, 1)):
        prob = W.Problem("hex_q1", n)
        a = _torch_stream(prob.assembler())
        a.set_dirichlet(np.zeros(0, np.uint32), np.zeros(0))
        a.set_coefficient(np.ones(prob.ndofs))
        a.build_pattern()
        a.assemble([("stiffness_fe", 1.0), ("mass", 1.0)])
        a.close()
        ctxs.append((prob, a))
    mg = native.Multigrid([a for _, a in ctxs], [(6, 4, 2), (3, 2, 1)])
    rp, col, val = mg.restriction(0)
    nf, nc = ctxs[0][0].ndofs, ctxs[1][0].ndofs
    R = sp.csr_matrix((val, col, rp), shape=(nc, nf))
    uf = (1.0 - ctxs[0][0].dof_coords()[:, 0]) ** 2
    assert abs(np.linalg.norm(R @ uf) - np.sqrt(2.0) * 1.759929209669) <= 1e-11
    assert np.abs(R @ np.ones(nf) - 1.0).max() <= 1e-14
    fine = oracle.Space("hex_q1", "cube_hex", (6, 4, 2))
    coarse = oracle.Space("hex_q1", "cube_hex", (3, 2, 1))
    orp, ocol, oval = oracle.transfer(fine, coarse, 1)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol.astype(np.int32))
    assert np.abs(val - oval).max() <= 1e-15


@pytest.mark.parametrize("fe,ns", [("hex_q1", (16, 8, 4, 2)), ("hex_q2", (8, 4, 2)), ("hex_q2", (6, 3))])
def test_mg_matrix_free_levels_through_element_cubes(oracle, native, gpu, fe, ns, monkeypatch):
    """The FP32 matrix-free levels applied through the element cubes (one
    lane per element, slot-major metrics, cube nodes summed across lanes,
    the Jacobi / residual epilogue fused) precondition like the per-element
    kernel with its gather: same iteration count, solutions within the
    solve tolerance, FP64 residual at the tolerance.  (Levels of any size go
    through the cubes here; by default only large ones do.)"""
    import torch
    out = {}
    for cubes in (False, True):
        monkeypatch.setenv("PXG_MF_CUBE_MIN_ELEMS", "0" if cubes else str(1 << 60))
        gl, _ = _levels(oracle, fe, ns)
        prob, a0 = gl[0]
        mg = native.Multigrid([a for _, a in gl], [(n, n, n) for n in ns])
        rp, col, val, b = a0.csr()
        xt = torch.zeros(prob.ndofs, dtype=torch.float64, device=gpu)
        it, rr = mg.solve_dev(a0.rhs_ptr(), xt.data_ptr(), rtol=1e-8, maxit=500, mode=2)
        A = sp.csr_matrix((val, col, rp), shape=(prob.ndofs, prob.ndofs))
        u = xt.cpu().numpy()
        assert np.linalg.norm(b - A @ u) / np.linalg.norm(b) <= 1.05e-8
        out[cubes] = (it, u)
        del mg
    assert abs(out[True][0] - out[False][0]) <= 1
    assert np.linalg.norm(out[True][1] - out[False][1]) <= 1e-7 * np.linalg.norm(out[False][1])


@pytest.mark.parametrize("fe,ns", [("hex_q1", (16, 8, 4, 2)), ("hex_q2", (8, 4, 2))])
def test_mg_pcg_fp64_and_fp32_refined(oracle, native, gpu, fe, ns):
    import torch
    gl, ol = _levels(oracle, fe, ns)
    prob, a0 = gl[0]
    mg = native.Multigrid([a for _, a in gl], [(n, n, n) for n in ns])
    rp, col, val, b = a0.csr()
    A = sp.csr_matrix((val, col, rp), shape=(prob.ndofs, prob.ndofs))
    u_exact = spl.spsolve(A.tocsc(), b)
    omg = oracle.Multigrid(ol)
    u_o, it_o, rr_o = omg.pcg(b, rtol=1e-8)
    bt = torch.from_numpy(b).to(gpu)
    # mode 2: FP32-refined with every level but the coarsest applied
    # matrix-free from kappa (their operators are the assembled matrices up to
    # rounding)
    its = {}
    for mode in (0, 1, 2):
        xt = torch.zeros_like(bt)
        it, rr = mg.solve_dev(a0.rhs_ptr(), xt.data_ptr(), rtol=1e-8, maxit=500, mode=mode)
        its[mode] = it
        u = xt.cpu().numpy()
        res = np.linalg.norm(b - A @ u) / np.linalg.norm(b)
        err = np.linalg.norm(u - u_exact) / np.linalg.norm(u_exact)
        print(f"  {fe} mode={('fp64', 'fp32-ir', 'fp32-ir-mf')[mode]} iters={it} (oracle {it_o}) "
              f"relres={res:.2e} err vs exact={err:.2e}")
        assert res <= 1.05e-8
        assert err <= 1e-6
        if mode == 0:
            assert abs(it - it_o) <= 1
    # the matrix-free levels (element cubes, FP32) precondition like the
    # assembled FP32 levels they stand for
    assert abs(its[2] - its[1]) <= 1


@pytest.mark.parametrize("fe,mesh,n,pert", [("hex_q2", "cube_hex", (3, 3, 3), 0.0),
                                            ("tet_p2", "cube_tet6", (3, 3, 3), 0.2),
                                            ("hex_q1", "cube_hex", (4, 3, 5), 0.0)])
@pytest.mark.parametrize("use_bc", [False, True])
def test_matrix_free_apply_and_gradient(oracle, gpu, fe, mesh, n, pert, use_bc):
    import torch
    prob = W.Problem(fe, n, mesh=mesh)
    o = oracle.Space(fe, mesh, n)
    if pert:
        xyz = W.perturb_interior(prob.xyz, 1.0 / max(n), seed=1, amount=pert)
        prob.set_coords(xyz)
        o.set_coords(xyz)
    a = _torch_stream(prob.assembler())
    dofs = prob.all_sides()
    a.set_dirichlet(dofs, np.zeros(len(dofs)))
    cons = np.zeros(prob.ndofs, np.uint8)
    cons[dofs] = 1
    kap = np.exp(W.grf(prob.dof_coords(), seed=2, ell=0.3, features=64))
    rng = np.random.default_rng(3)
    nb = 8
    x = rng.normal(size=(prob.ndofs, nb))
    ybar = rng.normal(size=(prob.ndofs, nb))
    kt = torch.from_numpy(kap).to(gpu)
    xt = torch.from_numpy(x).to(gpu)
    yt = torch.empty_like(xt)
    a.mf_apply_dev(kt.data_ptr(), nb, xt.data_ptr(), yt.data_ptr(), use_bc)
    y_ref = o.mf_apply(kap, x, cons if use_bc else None)
    assert np.abs(yt.cpu().numpy() - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
    gt = torch.empty_like(kt)
    yb = torch.from_numpy(ybar).to(gpu)
    a.mf_grad_kappa_dev(nb, xt.data_ptr(), yb.data_ptr(), gt.data_ptr(), use_bc)
    g_ref = o.mf_grad_kappa(x, ybar, cons if use_bc else None)
    assert np.abs(gt.cpu().numpy() - g_ref).max() <= 1e-12 * np.abs(g_ref).max()
    # matrix-free operator == assembled matrix (same Dirichlet form)
    if use_bc:
        a.set_coefficient(kap)
        a.build_pattern()
        a.assemble([("stiffness_fe", 1.0)])
        a.close(-1.0)
        rp, col, val, _ = a.csr()
        A = sp.csr_matrix((val, col, rp), shape=(prob.ndofs, prob.ndofs))
        assert np.abs(A @ x - yt.cpu().numpy()).max() <= 1e-12 * np.abs(y_ref).max()


@pytest.mark.parametrize("fe,n", [("hex_q2", (3, 2, 2)), ("hex_q1", (5, 3, 2))])
@pytest.mark.parametrize("nb", [1, 2, 3, 4, 5, 11])
def test_matrix_free_batch_sizes(oracle, gpu, fe, n, nb):
    """The apply at every batch size: one vector, batches filling 2, 4 or 8
    lanes per element, lanes without a vector (3, 5, 11) and a second chunk
    of vectors (11); element counts that leave a wavefront's last element
    slots empty.  Against the oracle, with the Dirichlet form."""
    import torch
    prob = W.Problem(fe, n)
    o = oracle.Space(fe, "cube_hex", n)
    xyz = W.perturb_interior(prob.xyz, 1.0 / max(n), seed=5, amount=0.2)
    prob.set_coords(xyz)
    o.set_coords(xyz)
    a = _torch_stream(prob.assembler())
    dofs = prob.all_sides()
    a.set_dirichlet(dofs, np.zeros(len(dofs)))
    cons = np.zeros(prob.ndofs, np.uint8)
    cons[dofs] = 1
    kap = np.exp(W.grf(prob.dof_coords(), seed=6, ell=0.3, features=64))
    x = np.random.default_rng(7).normal(size=(prob.ndofs, nb))
    kt = torch.from_numpy(kap).to(gpu)
    xt = torch.from_numpy(x).to(gpu)
    yt = torch.full_like(xt, np.nan)
    a.mf_apply_dev(kt.data_ptr(), nb, xt.data_ptr(), yt.data_ptr(), True)
    y_ref = o.mf_apply(kap, x, cons)
    assert np.abs(yt.cpu().numpy() - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
    # every entry written (the output starts as NaN), the same bits again
    y2 = torch.full_like(xt, np.nan)
    a.mf_apply_dev(kt.data_ptr(), nb, xt.data_ptr(), y2.data_ptr(), True)
    assert torch.equal(yt, y2)


@pytest.mark.parametrize("fe,n", [("hex_q2", (4, 4, 4)), ("hex_q2", (6, 4, 2)), ("hex_q2", (5, 4, 3)),
                                  ("hex_q1", (8, 6, 4)), ("hex_q1", (7, 3, 2))])
@pytest.mark.parametrize("nb", [8, 4, 2, 11])
def test_matrix_free_element_cubes(oracle, gpu, fe, n, nb):
    """Batches go through 2 x 2 x 2 element cubes (shared nodes summed across
    lanes, cube-interior dofs written directly, the others from partial
    values): whole cubes (even sizes), cubes cut at odd sizes, perturbed
    interior nodes, with and without the Dirichlet form; against the oracle
    and bitwise repeatable."""
    import torch
    prob = W.Problem(fe, n)
    o = oracle.Space(fe, "cube_hex", n)
    xyz = W.perturb_interior(prob.xyz, 1.0 / max(n), seed=11, amount=0.25)
    prob.set_coords(xyz)
    o.set_coords(xyz)
    a = _torch_stream(prob.assembler())
    dofs = prob.all_sides()
    a.set_dirichlet(dofs, np.zeros(len(dofs)))
    cons = np.zeros(prob.ndofs, np.uint8)
    cons[dofs] = 1
    kap = np.exp(W.grf(prob.dof_coords(), seed=12, ell=0.3, features=64))
    x = np.random.default_rng(13).normal(size=(prob.ndofs, nb))
    kt = torch.from_numpy(kap).to(gpu)
    xt = torch.from_numpy(x).to(gpu)
    for use_bc in (True, False):
        yt = torch.full_like(xt, np.nan)
        a.mf_apply_dev(kt.data_ptr(), nb, xt.data_ptr(), yt.data_ptr(), use_bc)
        y_ref = o.mf_apply(kap, x, cons if use_bc else None)
        assert np.abs(yt.cpu().numpy() - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
        y2 = torch.full_like(xt, np.nan)
        a.mf_apply_dev(kt.data_ptr(), nb, xt.data_ptr(), y2.data_ptr(), use_bc)
        assert torch.equal(yt, y2)


def test_gradient_central_differences(gpu):
    """d/dkappa (ybar . K(kappa) x) against central differences, step 1e-6."""
    import torch
    prob = W.Problem("hex_q2", (2, 2, 2))
    a = _torch_stream(prob.assembler())
    rng = np.random.default_rng(4)
    kap = np.exp(0.3 * rng.normal(size=prob.ndofs))
    x = rng.normal(size=(prob.ndofs, 2))
    yb = rng.normal(size=(prob.ndofs, 2))
    d = rng.normal(size=prob.ndofs)
    kt = torch.from_numpy(kap).to(gpu)
    xt = torch.from_numpy(x).to(gpu)
    ybt = torch.from_numpy(yb).to(gpu)
    gt = torch.empty_like(kt)
    a.mf_grad_kappa_dev(2, xt.data_ptr(), ybt.data_ptr(), gt.data_ptr())

    def L(k):
        kk = torch.from_numpy(k).to(gpu)
        y = torch.empty_like(xt)
        a.mf_apply_dev(kk.data_ptr(), 2, xt.data_ptr(), y.data_ptr())
        a.synchronize()
        return float((y.cpu().numpy() * yb).sum())

    h = 1e-6
    fd = (L(kap + h * d) - L(kap - h * d)) / (2 * h)
    an = float(gt.cpu().numpy() @ d)
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))


def test_mg_levels_on_their_own_streams(native, gpu):
    """Every level context keeps the non-blocking stream it was created with:
    the values each level casts and packs on its own stream on first use must
    be complete before the cycle (on the finest level's stream) reads them.
    The solves equal, bit for bit, those of levels sharing one stream."""
    import torch
    out = []
    for own in (False, True):
        gl = [_hex_level("hex_q1", n, _kappa, own_stream=own) for n in (16, 8, 4, 2)]
        mg = native.Multigrid([a for _, a, _, _ in gl], [(n, n, n) for n in (16, 8, 4, 2)])
        a0 = gl[0][1]
        res = []
        for mode in (0, 1, 0):
            xt = torch.zeros(gl[0][0].ndofs, dtype=torch.float64, device=gpu)
            torch.cuda.synchronize()
            it, rr = mg.solve_dev(a0.rhs_ptr(), xt.data_ptr(), rtol=1e-8, maxit=200, mode=mode)
            assert rr <= 1e-8
            res.append((it, xt.cpu().numpy()))
        out.append(res)
    for (it0, x0), (it1, x1) in zip(out[0], out[1]):
        assert it0 == it1 and np.array_equal(x0, x1)
