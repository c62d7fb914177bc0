"""pxg_jacobi_noalloc (the entry point that allocates nothing) against
pxg_jacobi through the C ABI: bitwise the same sweeps, on two calls in a
row, the second on a smaller matrix; sizes beyond its scratch vector are
rejected before any launch."""
import numpy as np
import pytest
import torch

from proxpde_amd import native as N
from proxpde_amd import workloads as W

pytestmark = pytest.mark.gpu


def _level(n, fe="hex_q1"):
    prob = W.Problem(fe, n)
    a = prob.assembler()
    a.set_stream(torch.cuda.current_stream().cuda_stream)
    dofs = prob.all_sides()
    a.set_dirichlet(dofs, np.zeros(len(dofs)))
    x = prob.dof_coords()
    a.set_coefficient(np.exp(np.sin(3.0 * x[:, 0]) * np.cos(2.0 * x[:, 1])))
    a.build_pattern()
    a.assemble([("stiffness_fe", 1.0), ("mass", 0.5)])
    a.close()
    return prob, a


@pytest.mark.parametrize("precision,sweeps", [(N.FP64, 3), (N.FP32, 2), (N.FP64, 0)])
def test_noalloc_jacobi_matches_jacobi(gpu, precision, sweeps):
    dt = torch.float64 if precision == N.FP64 else torch.float32
    for n in ((12, 12, 12), (6, 5, 4)):  # the second call on a smaller matrix
        print(f"no-alloc jacobi hex_q1 {n} precision={precision} sweeps={sweeps}", flush=True)
        prob, a = _level(n)
        g = torch.Generator(device=gpu).manual_seed(7)
        b = torch.randn(prob.ndofs, dtype=dt, device=gpu, generator=g)
        x0 = torch.randn(prob.ndofs, dtype=dt, device=gpu, generator=g)
        ref = x0.clone()
        out = x0.clone()
        torch.cuda.synchronize()
        a.jacobi_dev(b.data_ptr(), ref.data_ptr(), 2.0 / 3.0, sweeps, precision)
        a.jacobi_noalloc_dev(b.data_ptr(), out.data_ptr(), 2.0 / 3.0, sweeps, precision)
        a.synchronize()
        assert torch.equal(out, ref)
        if sweeps == 0:
            assert torch.equal(out, x0)


def test_noalloc_jacobi_rejects_large_matrices(gpu):
    """4,198,401 rows (Q1 quads 2048 x 2048) exceed the 2^22-entry scratch
    vector: PXG_E_INVALID, nothing launched, x untouched."""
    prob = W.Problem("quad_q1", (2048, 2048))
    a = prob.assembler()
    a.set_stream(torch.cuda.current_stream().cuda_stream)
    a.build_pattern()
    a.assemble([("mass", 1.0)])
    a.close()
    assert prob.ndofs > 1 << 22
    b = torch.ones(prob.ndofs, dtype=torch.float64, device=gpu)
    x = torch.zeros(prob.ndofs, dtype=torch.float64, device=gpu)
    with pytest.raises(N.PxgError) as err:
        a.jacobi_noalloc_dev(b.data_ptr(), x.data_ptr(), 2.0 / 3.0, 2)
    assert err.value.code == -1
    a.synchronize()
    assert not x.any()
