"""PyTorch front end of the LibTorch autograd operators (torch_ext/ops.cpp).

    op = KappaOperator(assembler, use_bc=False)
    y = op(kappa, x)                # y = K(kappa) x, differentiable in kappa and x

    solver = KappaSolver("hex_q1", [64, 32, 16, 8])
    u = solver(kappa, b)            # u = K(kappa)^{-1} b (MG-PCG), differentiable in kappa and b

Both run the HIP kernels of libproxpde_amd.so through the C ABI; the
backward passes are the hand-written adjoint kernels (no torch autograd of
the element loop).
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import native as N
from . import workloads as W

_loaded = False


def load():
    global _loaded
    if not _loaded:
        N.lib()
        path = os.path.join(N.HERE, "libproxpde_torch.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build()")
        torch.ops.load_library(path)
        _loaded = True


def _stream():
    return torch.cuda.current_stream().cuda_stream


class KappaOperator:
    """y = K(kappa) x on one assembler context (matrix-free)."""

    def __init__(self, assembler: N.Assembler, use_bc: bool = False):
        load()
        self.a = assembler
        self.use_bc = use_bc

    def __call__(self, kappa, x):
        return torch.ops.proxpde.kappa_apply(kappa, x, self.a.h.value, _stream(), self.use_bc)


def injection(coarse_pts, fine_pts, scale):
    """Index of the fine dof at the position of every coarse dof."""
    def key(p):
        k = np.rint(p * scale).astype(np.int64)
        return (k[:, 0] * (4 * scale + 8) + k[:, 1]) * (4 * scale + 8) + k[:, 2]
    kf = key(fine_pts)
    order = np.argsort(kf)
    pos = np.searchsorted(kf[order], key(coarse_pts))
    idx = order[np.clip(pos, 0, len(order) - 1)]
    if not np.array_equal(kf[idx], key(coarse_pts)):
        raise ValueError("coarse dofs are not nested in the fine mesh")
    return idx


class KappaSolver:
    """u = K(kappa)^{-1} b with homogeneous Dirichlet on all sides of the unit
    cube, geometric MG-PCG over hex levels n, n/2, ... (Jacobi omega=2/3, 2+2)."""

    def __init__(self, fe: str, ns, omega=2.0 / 3.0, nu1=2, nu2=2, coarse_rtol=1e-10, device=0):
        load()
        self.fe = fe
        self.ns = list(ns)
        self.probs, self.ctx = [], []
        for n in self.ns:
            p = W.Problem(fe, (n, n, n))
            a = p.assembler(device=device)
            d = p.all_sides()
            a.set_dirichlet(d, np.zeros(len(d)))
            a.set_coefficient(np.ones(p.ndofs))
            a.build_pattern()
            a.assemble([("stiffness_fe", 1.0)])
            a.close()
            self.probs.append(p)
            self.ctx.append(a)
        fine = self.probs[0].dof_coords()
        scale = 2 * self.ns[0]
        dev = torch.device(f"cuda:{device}")
        self.inject = [torch.from_numpy(injection(p.dof_coords(), fine, scale)).to(dev) for p in self.probs[1:]]
        self.mg = N.Multigrid(self.ctx, [(n, n, n) for n in self.ns], omega, nu1, nu2, coarse_rtol)
        self.ndofs = self.probs[0].ndofs
        self.constrained = self.probs[0].all_sides()

    def rhs(self, f=1.0):
        """Assembled right-hand side of f (constant) on the finest level."""
        a = self.ctx[0]
        a.clear()
        a.assemble_rhs("const", (float(f),))
        return a.rhs_host()

    def __call__(self, kappa, b, rtol=1e-8, mode=N.FP64, maxit=1000, fp32_levels=None, fine_fp64=False):
        """mode FP32: FP32 element matrices on all but the coarsest level (fp32_levels)
        and the FP64 refinement residual through the matrix-free operator; with
        fine_fp64 the finest level is assembled in FP64, its FP32 copy rounded from
        it, and the residual uses that matrix (cheaper for low-order elements).
        mode FP32_MF: as FP32, with only the coarsest level assembled and the
        others applied matrix-free in FP32 from kappa (per-point metrics
        computed once per kappa)."""
        if fp32_levels is None:
            fp32_levels = mode in (N.FP32, N.FP32_MF)
        return torch.ops.proxpde.kappa_solve(kappa, b, [c.h.value for c in self.ctx], self.inject,
                                             self.mg.h.value, _stream(), rtol, int(mode), maxit, bool(fp32_levels),
                                             bool(fine_fp64))
