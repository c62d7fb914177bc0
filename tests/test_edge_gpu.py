"""Edge cases of the HIP assembly path through the C ABI, against the oracle.

* one-element meshes of every element type and Taylor-Hood pair, with and
  without Dirichlet data, FP64 (bar of the other parity tests: pattern
  bit-exact, values within 1e-12 relative) and FP32 (2e-6 in max norm); the
  persistent element kernels then run with fewer blocks than work counters;
* a rebuild after clear on those meshes (the staged rebuild into the closed
  layout on a single stage block) equal bit for bit to the first build;
* every dof constrained: the closed matrix is the diagonal of the element
  counts and b the values times the counts (Builder::buildLhs puts a unit
  row into every element of a constrained dof, include/builder.hpp:83-124);
* SpMV and Jacobi on the closed one-element matrices against the oracle.
"""
import numpy as np
from _tol import entrywise
import pytest
import scipy.sparse as sp

from proxpde_amd import native as N
from proxpde_amd import workloads as W

pytestmark = pytest.mark.gpu

SINGLE = [
    ("line_p1", "line", (1,)),
    ("quad_q1", "square", (1, 1)),
    ("quad_q2", "square", (1, 1)),
    ("hex_q1", "cube_hex", (1, 1, 1)),
    ("hex_q2", "cube_hex", (1, 1, 1)),
    ("tet_p1", "cube_tet6", (1, 1, 1)),
    ("tet_p2", "cube_tet5", (1, 1, 1)),
]
SIDES = {1: ("left",), 2: ("bottom", "left"), 3: ("bottom", "left", "front")}
ALL = {1: ("left", "right"), 2: ("bottom", "right", "top", "left"),
       3: ("bottom", "right", "top", "left", "back", "front")}


def _rel_ok(a, b, rowptr=None, tol=1e-12):
    """Entry-wise bound of tests/_tol.py (per matrix row when rowptr is given)."""
    return entrywise(a, b, rowptr, tol)[0]


def _setup(oracle, fe, mesh, n, sides):
    prob = W.Problem(fe, n, mesh=mesh)
    o = oracle.Space(fe, mesh, n, length=prob.length)
    x = prob.dof_coords()
    kappa = 1.0 + 0.5 * x[:, 0] + 0.25 * x[:, 1] ** 2
    flag = np.zeros(prob.ndofs, np.uint8)
    val = np.zeros(prob.ndofs)
    dofs = prob.bc_dofs(sides) if sides else np.zeros(0, np.uint32)
    flag[dofs] = 1
    val[dofs] = 1.0 + x[dofs, 0] - x[dofs, 1]
    a = prob.assembler()
    a.set_dirichlet(dofs, val[dofs])
    a.set_coefficient(kappa)
    a.build_pattern()
    return prob, o, a, kappa, (flag, val)


@pytest.mark.parametrize("bc", [False, True], ids=["free", "dirichlet"])
@pytest.mark.parametrize("fe,mesh,n", SINGLE, ids=[c[0] for c in SINGLE])
def test_single_cell_meshes(oracle, gpu, fe, mesh, n, bc):
    dim = len(n)
    print(f"single cell {fe} {mesh} bc={bc}", flush=True)
    prob, o, a, kappa, bcd = _setup(oracle, fe, mesh, n, SIDES[dim] if bc else ())
    terms = [("stiffness_fe", 1.0), ("mass", 0.5)]
    rp_o, col_o, v_o, b_o = o.assemble(terms, kappa=kappa, bc=bcd, rhs="const", rhs_params=(1.0,))
    a.assemble(terms)
    a.assemble_rhs("const", (1.0,))
    a.close(1e-16)
    first = a.csr()
    rp, col, v, b = first
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o.astype(np.int32))
    assert _rel_ok(v, v_o, rp_o) and _rel_ok(b, b_o)
    # rebuild after clear: the staged path into the closed layout
    for _ in range(2):
        a.clear()
        a.assemble(terms)
        a.assemble_rhs("const", (1.0,))
        a.close(1e-16)
        for x, y in zip(first, a.csr()):
            assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    # FP32 element kernels
    a.clear()
    a.assemble(terms, precision=N.FP32)
    a.close(1e-16)
    rp32, col32, v32, _ = a.csr(fp32=True)
    nd = prob.ndofs
    A32 = sp.csr_matrix((v32.astype(np.float64), col32, rp32), shape=(nd, nd)).toarray()
    A64 = sp.csr_matrix((v_o, col_o.astype(np.int64), rp_o), shape=(nd, nd)).toarray()
    assert np.abs(A32 - A64).max() <= 2e-6 * np.abs(A64).max()


# meshes whose dofs all lie on the boundary (the 5-tet split has no interior
# edge, so its P2 dofs qualify; Q2 elements have interior nodes)
@pytest.mark.parametrize("fe,mesh,n", [("quad_q1", "square", (2, 1)), ("hex_q1", "cube_hex", (2, 1, 1)),
                                      ("tet_p1", "cube_tet6", (1, 1, 1)), ("tet_p2", "cube_tet5", (1, 1, 1))],
                         ids=["quad_q1", "hex_q1", "tet_p1", "tet_p2"])
def test_every_dof_constrained(oracle, gpu, fe, mesh, n):
    """All dofs on the boundary and constrained: A = diag(elements per dof),
    b = values x elements per dof, the element terms leave no trace."""
    dim = len(n)
    prob, o, a, kappa, (flag, val) = _setup(oracle, fe, mesh, n, ALL[dim])
    assert flag.all(), "every dof of these meshes lies on the boundary"
    a.assemble([("stiffness_fe", 1.0), ("mass", 2.0)])
    a.close(1e-16)
    rp, col, v, b = a.csr()
    count = np.bincount(prob.emap.ravel().astype(np.int64), minlength=prob.ndofs).astype(np.float64)
    assert np.array_equal(rp, np.arange(prob.ndofs + 1))
    assert np.array_equal(col, np.arange(prob.ndofs, dtype=np.int32))
    assert np.array_equal(v, count)
    assert np.allclose(b, val * count, rtol=0, atol=0)


@pytest.mark.parametrize("pair,n", [("quad_q2q1", (1, 1)), ("tet_p2p1", (1, 1, 1)), ("hex_q2q1", (1, 1, 1))],
                         ids=["quad_q2q1", "tet_p2p1", "hex_q2q1"])
def test_single_cell_taylor_hood(oracle, gpu, pair, n):
    vfe, mesh = {"quad_q2q1": ("quad_q2", "square"), "tet_p2p1": ("tet_p2", "cube_tet5"),
                 "hex_q2q1": ("hex_q2", "cube_hex")}[pair]
    print(f"single cell {pair}", flush=True)
    prob = W.StokesProblem(n, pair=pair)
    o = oracle.Space(vfe, mesh, n, length=prob.vel.length)
    dim = prob.dim
    bcs = [("bottom", tuple(0.5 + d for d in range(dim)))]
    dofs, vals, flag, full = prob.dirichlet(bcs, pressure=((0, 0.25),))
    coefs = {"stiffness": 1.0, "tensor": 0.5, "grad": -1.0, "div": -1.0}
    rp_o, col_o, v_o, b_o = o.stokes(False, coefs, flag, full)
    a = prob.assembler()
    a.set_dirichlet(dofs, vals)
    a.build_pattern()
    terms = [("stiffness", 1.0), ("tensor_stiffness", 0.5), ("grad", -1.0), ("div", -1.0)]
    a.assemble(terms)
    a.close(1e-16)
    first = a.csr()
    nd = prob.ndofs
    G = sp.csr_matrix((first[2], first[1], first[0]), shape=(nd, nd)).toarray()
    R = sp.csr_matrix((v_o, col_o.astype(np.int64), rp_o), shape=(nd, nd)).toarray()
    scale = np.abs(R).max()
    only = (G != 0) != (R != 0)
    assert np.abs(G[only]).max(initial=0.0) <= 1e-14 * scale and np.abs(R[only]).max(initial=0.0) <= 1e-14 * scale
    assert np.abs(G - R).max() <= 1e-12 * scale
    assert _rel_ok(first[3], b_o)
    a.clear()
    a.assemble(terms)
    a.close(1e-16)
    for x, y in zip(first, a.csr()):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


@pytest.mark.parametrize("fe,mesh,n", [("quad_q1", "square", (1, 1)), ("hex_q2", "cube_hex", (1, 1, 1))],
                         ids=["quad_q1", "hex_q2"])
def test_spmv_and_jacobi_single_cell(oracle, gpu, fe, mesh, n):
    import torch
    dim = len(n)
    prob, o, a, kappa, bcd = _setup(oracle, fe, mesh, n, SIDES[dim])
    terms = [("stiffness_fe", 1.0), ("mass", 0.5)]
    o.assemble(terms, kappa=kappa, bc=bcd, rhs="const", rhs_params=(1.0,))
    a.assemble(terms)
    a.assemble_rhs("const", (1.0,))
    a.close(1e-16)
    # the context works on torch's stream, where the inputs are written
    a.set_stream(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, prob.ndofs)
    b = rng.uniform(-1, 1, prob.ndofs)
    xt = torch.from_numpy(x).to(gpu)
    bt = torch.from_numpy(b).to(gpu)
    yt = torch.empty_like(xt)
    a.spmv_dev(xt.data_ptr(), yt.data_ptr())
    a.synchronize()
    assert np.abs(yt.cpu().numpy() - o.spmv(x)).max() <= 1e-14 * max(1.0, np.abs(x).max())
    yj = o.jacobi(b, x, omega=2.0 / 3.0, sweeps=2)
    zt = xt.clone()
    a.jacobi_dev(bt.data_ptr(), zt.data_ptr(), 2.0 / 3.0, 2)
    a.synchronize()
    assert np.abs(zt.cpu().numpy() - yj).max() <= 1e-13


def test_new_mesh_fe_or_dofmap_drops_derived_state(gpu):
    """A mesh with another element count drops the dof map and everything
    built from it; a new dof map drops the pattern and the dof adjacency of
    the matrix-free operator; another element drops the mesh.  The next calls
    report PXG_E_STATE (-5) or rebuild, instead of launching over buffers
    sized for the old mesh."""
    import torch
    small, big = W.Problem("hex_q1", (3, 3, 3)), W.Problem("hex_q1", (4, 3, 3))
    a = small.assembler()
    a.build_pattern()
    a.assemble([("stiffness", 1.0)])
    a.close()
    a.set_mesh(big.xyz, big.conn)
    with pytest.raises(N.PxgError) as err:
        a.assemble([("stiffness", 1.0)])
    assert err.value.code == -5
    a.set_dofmap(big.ndofs, big.emap)
    a.build_pattern()
    a.assemble([("stiffness", 1.0)])
    a.close()
    fresh = big.assembler()
    fresh.build_pattern()
    fresh.assemble([("stiffness", 1.0)])
    fresh.close()
    for x, y in zip(a.csr(), fresh.csr()):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    # matrix-free: first on the small mesh, then on the big one through the
    # same context (its dof adjacency must be rebuilt for the new map)
    m = small.assembler()
    m.set_stream(torch.cuda.current_stream().cuda_stream)
    for prob, first in ((small, True), (big, False)):
        if not first:
            m.set_mesh(prob.xyz, prob.conn)
            m.set_dofmap(prob.ndofs, prob.emap)
        k = torch.ones(prob.ndofs, dtype=torch.float64, device=gpu)
        x = torch.linspace(-1.0, 1.0, 2 * prob.ndofs, dtype=torch.float64, device=gpu).reshape(prob.ndofs, 2)
        y = torch.empty_like(x)
        m.mf_apply_dev(k.data_ptr(), 2, x.data_ptr(), y.data_ptr())
        m.synchronize()
        ref = prob.assembler()
        ref.set_stream(torch.cuda.current_stream().cuda_stream)
        yr = torch.empty_like(x)
        ref.mf_apply_dev(k.data_ptr(), 2, x.data_ptr(), yr.data_ptr())
        ref.synchronize()
        assert torch.equal(y, yr)
    # another element shape: the mesh goes too
    a.set_fe("tet_p1")
    with pytest.raises(N.PxgError) as err:
        a.assemble([("stiffness", 1.0)])
    assert err.value.code == -5
    nd, nl, ne, dev = a.sizes()
    assert nd == 0 and ne == 0 and nl == 4
