"""Benchmark of the assembly / operator-apply path on one MI355X.

Workload (the headline configuration): Q2 triquadratic hexahedra on the unit
cube, 64^3 elements (2,146,689 dofs), 3x3x3 Gauss, lognormal kappa = exp(g),
g a Gaussian random field (std 1, correlation length 0.1, seed 2) at the dofs,
homogeneous Dirichlet on all sides.  One step = one Builder pass:
clear + buildLhs(AssemblyStiffnessFE(kappa)) + closeMatrix, i.e. geometry,
element matrices, Dirichlet elimination, deterministic CSR scatter and prune.
Reported beside it: CSR SpMV bandwidth on the assembled matrix, per-kernel
times and the roofline of the dominant kernel, and the oracle's serial CPU
assembly rate on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def q2hex_flops_per_elem(nq=27, n=27, dim=3, g=27, kappa=True):
    """Reference-formulation flops of one Q2-hex kappa-stiffness element:
    geometry 2*g*3*dim + 6 dim^2 + 60 + 6 dim^2, gradients 6 n dim,
    the n x n outer products 6 n^2 + 3 n, kappa interpolation 2 n (per q)."""
    per_q = 2 * g * 3 * dim + 6 * dim * dim + 60 + 6 * dim * dim + 6 * n * dim + 6 * n * n + 3 * n
    if kappa:
        per_q += 2 * n
    return nq * per_q


def measured_traffic(n):
    """HBM bytes per launch of the kernels of the n = 64 workload, from the
    FETCH_SIZE / WRITE_SIZE counter runs recorded in profiles/ (calibrated
    against a plain device copy); {} for other sizes."""
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_q2hex64.json")
    if n != 64 or not os.path.exists(tpath):
        return {}
    with open(tpath) as fh:
        return json.load(fh).get("kernels", {})


def kappa_apply_batch(args, dev, a, kappa_host, nb=8):
    """y = K(kappa) x for a batch of nb vectors (x ~ N(0,1), seed 3, layout
    [ndofs][nb]) through the autograd operator, forward and backward (dx = K dy,
    dkappa by the element contraction), on the assembler's mesh."""
    import torch

    from proxpde_amd import autograd as A

    op = A.KappaOperator(a)
    a.set_stream(torch.cuda.current_stream().cuda_stream)
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    kap = torch.from_numpy(kappa_host).to(dev)
    x = torch.randn(a.ndofs, nb, dtype=torch.float64, device=dev, generator=g)
    w = torch.randn(a.ndofs, nb, dtype=torch.float64, device=dev, generator=g)
    k = kap.clone().requires_grad_()
    xr = x.clone().requires_grad_()
    for _ in range(2):
        y = op(k, xr)
        (y * w).sum().backward()
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        y = op(kap, x)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        k.grad = None
        xr.grad = None
        y = op(k, xr)
        (y * w).sum().backward()
    torch.cuda.synchronize()
    t_fb = (time.perf_counter() - t0) / reps
    # the operator's own launches (element kernel + per-dof gather), HIP events
    # on the context stream, without the autograd wrapper and its host wait
    kap_c = kap.contiguous()
    y = torch.empty_like(x)
    gk = torch.empty_like(kap)
    a.profile(True)
    for _ in range(reps):
        a.mf_apply_dev(kap_c.data_ptr(), nb, x.data_ptr(), y.data_ptr())
        a.mf_grad_kappa_dev(nb, x.data_ptr(), w.data_ptr(), gk.data_ptr())
    ap_ms, ap_n = a.profile_get("mf_apply")
    gr_ms, gr_n = a.profile_get("mf_gradk")
    a.profile(False)
    E = a.nelem
    N_ = a.ndofs
    # algorithmic work of the forward apply (reference formulation): Q [1140 geometry + 2n kappa + nb (12 n + 5)]
    fl = E * 27 * (1140 + 2 * 27 + nb * (12 * 27 + 5))
    # compulsory bytes of the apply: x in and y out (nb vectors), kappa, the
    # vertices (65^3) and the vertex and dof ids of every element
    nv = (round(E ** (1 / 3)) + 1) ** 3
    by = 2 * nb * N_ * 8 + 8 * N_ + 24 * nv + 4 * E * (8 + 27)
    t_ap = ap_ms / max(ap_n, 1) * 1e-3
    tr = measured_traffic(args.n)
    traffic = (tr["mf_apply_batch8"]["bytes"] + tr["mf_gather_batch8"]["bytes"]
               if nb == 8 and "mf_apply_batch8" in tr and "mf_gather_batch8" in tr else None)
    # (the sum-factorised kernels do far fewer flops than the reference
    # formulation, so the flop rate below is an equivalent, not a hardware rate)
    return {"batch": nb, "forward_ms": 1e3 * t_fwd, "forward_backward_ms": 1e3 * t_fb,
            "forward_reference_equiv_gflops": fl / t_fwd / 1e9, "apply_kernels_ms": 1e3 * t_ap,
            "gradk_kernels_ms": gr_ms / max(gr_n, 1), "apply_bytes": by,
            "roofline": {"kernel": "k_sf_batch<3,8,true> + k_sf_part", "bound": "hbm",
                         "achieved": by / t_ap / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": by / t_ap / 1e9 / HBM_PEAK_GBS, "traffic": traffic}}


def solve_and_adjoint(args, dev, N, W):
    """Q2 hex n^3: u = K(kappa)^{-1} f (MG-PCG to rel. residual 1e-8, levels
    n, n/2, ...), J = 1/2 |u - u_t|^2 and dJ/dkappa by the adjoint solve.
    kappa = exp(GRF seed 2); u_t = K(kappa_t)^{-1} f with kappa_t = exp(GRF seed 4).
    Timed: assembly of all levels + solve + adjoint (forward and backward)."""
    import torch

    from proxpde_amd import autograd as A

    n = args.n
    ns = [n >> l for l in range(args.solve_levels)]
    t0 = time.time()
    solver = A.KappaSolver("hex_q2", ns)
    t_setup = time.time() - t0
    x = solver.probs[0].dof_coords()
    kap = torch.from_numpy(np.exp(W.grf(x, seed=2, ell=0.1, features=1024, device=dev))).to(dev)
    kap_t = torch.from_numpy(np.exp(W.grf(x, seed=4, ell=0.1, features=1024, device=dev))).to(dev)
    b = torch.from_numpy(solver.rhs(1.0)).to(dev)
    with torch.no_grad():
        ut = solver(kap_t, b, rtol=1e-10)
    res = {}
    sol = {}
    # fp32_ir: FP32 element matrices on all but the coarsest level, FP64
    # refinement residual by the matrix-free operator; fp32_ir_fine64: the
    # finest level assembled in FP64 (FP32 copy rounded), residual by that matrix
    # fp32_ir_mf: the finest level not assembled, applied matrix-free in FP32
    # (per-point metrics once per kappa), FP64 residual by the FP64 operator
    for name, mode, fine64 in (("fp64", N.FP64, False), ("fp32_ir", N.FP32, False),
                               ("fp32_ir_fine64", N.FP32, True), ("fp32_ir_mf", N.FP32_MF, False)):
        progress(f"solve+adjoint {name}")
        times = []
        for rep in range(args.solve_reps + 1):
            k = kap.clone().requires_grad_()
            torch.cuda.synchronize()
            ts = time.perf_counter()
            u = solver(k, b, rtol=1e-8, mode=mode, fine_fp64=fine64)
            J = 0.5 * ((u - ut) ** 2).sum()
            J.backward()
            torch.cuda.synchronize()
            if rep > 0:
                times.append(time.perf_counter() - ts)
        sol[name] = (u.detach(), k.grad.detach(), float(J))
        res[name] = {"ms": 1e3 * min(times)}
        # the bare solve on the matrices of this mode (still assembled): time and
        # CG iterations (FP32 mode: inner iterations summed over refinements)
        xs = torch.zeros_like(b)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        its, _ = solver.mg.solve_dev(b.data_ptr(), xs.data_ptr(), 1e-8, 1000, mode)
        torch.cuda.synchronize()
        res[name]["solve_only_ms"] = 1e3 * (time.perf_counter() - ts)
        res[name]["cg_iterations"] = its
    # FP64 residual of each solution with the matrix-free FP64 operator
    op = A.KappaOperator(solver.ctx[0], use_bc=True)
    for name in sol:
        u = sol[name][0]
        with torch.no_grad():
            r = b - op(kap, u.unsqueeze(1)).squeeze(1)
        res[name]["relres_fp64"] = float(r.norm() / b.norm())
    u64, g64, J64 = sol["fp64"]
    for name in ("fp32_ir", "fp32_ir_fine64", "fp32_ir_mf"):
        u32, g32, J32 = sol[name]
        res[name]["u_rel_err_vs_fp64"] = float((u32 - u64).norm() / u64.norm())
        res[name]["grad_rel_err_vs_fp64"] = float((g32 - g64).norm() / g64.norm())
        res[f"speedup_{name}_over_fp64"] = res["fp64"]["ms"] / res[name]["ms"]
    best = min(("fp32_ir", "fp32_ir_fine64", "fp32_ir_mf"), key=lambda k: res[k]["ms"])
    res["fastest_fp32_refined"] = best
    res["speedup_fp32_refined_over_fp64"] = res["fp64"]["ms"] / res[best]["ms"]
    res["levels"] = ns
    res["setup_s"] = t_setup
    res["J"] = J64
    return res


def _sync_time(fn, reps=1):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, out


def config_c2(args, dev, N, W):
    """Q1 hex 128^3, 2x2x2 Gauss, stiffness (kappa = 1), no Dirichlet; FP64 and
    FP32 element kernels; then 100 SpMV on x ~ U(-1, 1) seed 0, FP64 and FP32."""
    import torch
    n = 128
    prob = W.Problem("hex_q1", (n, n, n))
    a = prob.assembler()
    a.set_dirichlet(np.zeros(0, np.uint32), np.zeros(0))
    a.build_pattern()
    res = {}
    for name, prec in (("fp64", N.FP64), ("fp32", N.FP32)):
        def build():
            a.clear()
            a.assemble([("stiffness", 1.0)], precision=prec)
            a.close(1e-16)
        build()
        t, _ = _sync_time(build, 3)
        res[f"build_{name}_ms"] = 1e3 * t
        res[f"elements_per_s_{name}"] = prob.mesh.nelem / t
    # FP32 element kernels against the FP64 matrix (the prune of cancellation
    # zeros may differ between the precisions: compared on the union pattern)
    import scipy.sparse as sp
    rp32, col32, v32, _ = a.csr(fp32=True)
    a.clear()
    a.assemble([("stiffness", 1.0)], precision=N.FP64)
    a.close(1e-16)
    rp, col, v, _ = a.csr()
    A64 = sp.csr_matrix((v, col, rp), shape=(prob.ndofs, prob.ndofs))
    A32 = sp.csr_matrix((v32.astype(np.float64), col32, rp32), shape=(prob.ndofs, prob.ndofs))
    res["nnz"] = int(len(v))
    res["nnz_fp32"] = int(len(v32))
    res["fp32_values_max_rel_err"] = float(abs(A32 - A64).max() / np.abs(v).max())
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    x = torch.rand(prob.ndofs, dtype=torch.float64, device=dev, generator=g) * 2 - 1
    y = torch.empty_like(x)
    a.set_stream(torch.cuda.current_stream().cuda_stream)
    for name, prec, xv, yv, s in (("fp64", N.FP64, x, y, 8), ("fp32", N.FP32, x.float(), y.float(), 4)):
        a.spmv_dev(xv.data_ptr(), yv.data_ptr(), precision=prec)
        t, _ = _sync_time(lambda: a.spmv_dev(xv.data_ptr(), yv.data_ptr(), precision=prec), 100)
        by = len(v) * (s + 4) + 8 * (prob.ndofs + 1) + 2 * s * prob.ndofs
        res[f"spmv_{name}_ms"] = 1e3 * t
        res[f"spmv_{name}_GB/s"] = by / t / 1e9
    return {"metric": "C2: Q1 hex 128^3 stiffness build + 100 SpMV", "config": {"workload": "Q1 hex 128^3, kappa=1, "
            "no Dirichlet", "elements": prob.mesh.nelem, "dofs": prob.ndofs}, "results": res}


def config_c3(args, dev, N, W):
    """P2 tets, 48^3 cubes x 6 (Kuhn), interior vertices perturbed U(-0.2h, 0.2h)
    seed 1, Gauss-11, kappa = exp(GRF seed 2) at the P2 nodes; stiffness(kappa) +
    mass(1); FP64 and FP32 element kernels, FP32 error against FP64."""
    n = 48
    prob = W.Problem("tet_p2", (n, n, n), mesh="cube_tet6")
    prob.set_coords(W.perturb_interior(prob.xyz, 1.0 / n, seed=1, amount=0.2))
    kap = np.exp(W.grf(prob.dof_coords(), seed=2, std=1.0, ell=0.1, features=4096))
    a = prob.assembler()
    d = prob.all_sides()
    a.set_dirichlet(d, np.zeros(len(d)))
    a.set_coefficient(kap)
    a.build_pattern()
    terms = [("stiffness_fe", 1.0), ("mass", 1.0)]
    res = {}
    out = {}
    for name, prec in (("fp64", N.FP64), ("fp32", N.FP32)):
        def build():
            a.clear()
            a.assemble(terms, precision=prec)
            a.close(1e-16)
        build()
        t, _ = _sync_time(build, 3)
        res[f"build_{name}_ms"] = 1e3 * t
        res[f"elements_per_s_{name}"] = prob.mesh.nelem / t
        out[name] = a.csr(fp32=prec == N.FP32)
    import scipy.sparse as sp
    nd = prob.ndofs
    A64 = sp.csr_matrix((out["fp64"][2], out["fp64"][1], out["fp64"][0]), shape=(nd, nd))
    A32 = sp.csr_matrix((out["fp32"][2].astype(np.float64), out["fp32"][1], out["fp32"][0]), shape=(nd, nd))
    res["nnz"] = int(A64.nnz)
    res["fp32_values_max_rel_err"] = float(abs(A32 - A64).max() / abs(A64).max())
    return {"metric": "C3: P2 tet (6-tet Kuhn) 48^3 stiffness(kappa)+mass build", "config": {
            "workload": "P2 tet 48^3 x 6, perturbed 0.2h, Gauss-11, lognormal kappa", "elements": prob.mesh.nelem,
            "dofs": prob.ndofs}, "results": res}


def config_c5(args, dev, N, W):
    """Q1 hex 256^3, f = 1, Dirichlet 0 on all sides, Jacobi 2/3 with 2+2 sweeps,
    PCG to 1e-8; J = 1/2 |u - u_t|^2 with u_t = K(kappa_t)^{-1} f, kappa_t
    lognormal seed 4; J and dJ/dkappa at kappa = 1; FP64 vs FP32-refined.
    Two hierarchies: the configuration's four levels 256/128/64/32 (the 33^3
    coarsest level solved by FP64 Jacobi-CG to 1e-10 inside every cycle), and
    six levels 256/.../8 (the 9^3 coarsest level inverted densely once per
    coefficient)."""
    import torch

    from proxpde_amd import autograd as A
    n = 256
    out = {}
    for tag, nlev in (("levels4", 4), ("levels6", 6)):
        ns = [n >> l for l in range(nlev)]
        progress(f"C5 hierarchy {ns}")
        t0 = time.time()
        solver = A.KappaSolver("hex_q1", ns)
        t_setup = time.time() - t0
        x = solver.probs[0].dof_coords()
        kap_t = torch.from_numpy(np.exp(W.grf(x, seed=4, ell=0.1, features=1024, device=dev))).to(dev)
        b = torch.from_numpy(solver.rhs(1.0)).to(dev)
        with torch.no_grad():
            ut = solver(kap_t, b, rtol=1e-10)
        kap = torch.ones(solver.ndofs, dtype=torch.float64, device=dev)
        res = {"levels": ns, "setup_s": t_setup}
        sol = {}
        # fp32_ir: FP32 element matrices on the levels above the coarsest, the
        # finest level assembled in FP64 (its FP32 copy rounded) for the
        # refinement residual; fp32_ir_mf: the residual by the matrix-free operator
        for name, mode, fine64 in (("fp64", N.FP64, False), ("fp32_ir", N.FP32, True), ("fp32_ir_mf", N.FP32, False)):
            progress(f"C5 {tag} solve+adjoint {name}")
            times = []
            for rep in range(3):
                k = kap.clone().requires_grad_()
                torch.cuda.synchronize()
                ts = time.perf_counter()
                u = solver(k, b, rtol=1e-8, mode=mode, fine_fp64=fine64)
                J = 0.5 * ((u - ut) ** 2).sum()
                J.backward()
                torch.cuda.synchronize()
                if rep > 0:
                    times.append(time.perf_counter() - ts)
            sol[name] = (u.detach(), k.grad.detach())
            xs = torch.zeros_like(b)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            its, _ = solver.mg.solve_dev(b.data_ptr(), xs.data_ptr(), 1e-8, 1000, mode)
            torch.cuda.synchronize()
            res[name] = {"ms": 1e3 * min(times), "solve_only_ms": 1e3 * (time.perf_counter() - ts),
                         "cg_iterations": its}
        op = A.KappaOperator(solver.ctx[0], use_bc=True)
        for name in sol:
            with torch.no_grad():
                r = b - op(kap, sol[name][0].unsqueeze(1)).squeeze(1)
            res[name]["relres_fp64"] = float(r.norm() / b.norm())
        for name in ("fp32_ir", "fp32_ir_mf"):
            res[name]["u_rel_err_vs_fp64"] = float((sol[name][0] - sol["fp64"][0]).norm() / sol["fp64"][0].norm())
            res[name]["grad_rel_err_vs_fp64"] = float((sol[name][1] - sol["fp64"][1]).norm() /
                                                      sol["fp64"][1].norm())
            res[f"speedup_{name}_over_fp64"] = res["fp64"]["ms"] / res[name]["ms"]
        out[tag] = res
        del solver, op, sol
        torch.cuda.empty_cache()
    return {"metric": "C5: Q1 hex 256^3 solve + adjoint gradient, FP64 vs FP32-refined", "config": {
            "workload": "Q1 hex 256^3, f=1, Dirichlet 0, MG 256/128/64/32 (and 6 levels), PCG 1e-8, kappa=1, "
                        "u_t from kappa_t seed 4", "dofs": 257 ** 3}, "results": out}


STOKES_SIZES = {"quad_q2q1": (1024, 1024), "tet_p2p1": (40, 40, 40), "hex_q2q1": (48, 48, 48)}


def config_stokes(args, dev, N, W):
    """Taylor-Hood Stokes system (--pair quad_q2q1: Q2-Q1 on 1024^2 quads;
    tet_p2p1: P2-P1 on 40^3 cubes x 5 tets; hex_q2q1: Q2-Q1 on 48^3 hexes),
    BLOCK velocity dofs, interior vertices perturbed U(-0.2h, 0.2h) seed 3:
    tensor stiffness 0.1 + grad -1 + div -1 (the operator of the reference's
    stokes2dquad / stokes3dhex / stokes3dtet tests), Dirichlet velocity on
    every side; full builds (clear + assemble + close) in FP64 and FP32, the
    element kernel and the staged scatter timed with HIP events."""
    pair = args.pair
    n = STOKES_SIZES[pair]
    prob = W.StokesProblem(n, pair=pair)
    ln = prob.vel.length
    prob.set_coords(W.perturb_interior(prob.xyz, 1.0 / max(n), seed=3, amount=0.2, length=ln))
    dim = prob.dim
    sides = ("bottom", "right", "top", "left") + (("back", "front") if dim == 3 else ())
    inflow = tuple(1.0 if d == 1 else 0.0 for d in range(dim))
    bcs = [("bottom", inflow)] + [(sd, (0.0,) * dim) for sd in sides[1:]]
    dofs, vals, _, _ = prob.dirichlet(bcs)
    a = prob.assembler()
    a.set_dirichlet(dofs, vals)
    t0 = time.time()
    nnz_s = a.build_pattern()
    t_pattern = time.time() - t0
    E = prob.mesh.nelem
    nl = prob.emap.shape[1]
    terms = [("tensor_stiffness", 0.1), ("grad", -1.0), ("div", -1.0)]
    res = {"pattern_s": t_pattern, "nnz_structural": int(nnz_s), "local_dofs": int(nl)}
    for name, prec in (("fp64", N.FP64), ("fp32", N.FP32)):
        def build():
            a.clear()
            a.assemble(terms, precision=prec)
            a.close(1e-16)
        build()
        build()
        a.profile(True)
        t, _ = _sync_time(build, 5)
        el = "element" if prec == N.FP64 else "element32"
        sc = "scatter" if prec == N.FP64 else "scatter32"
        pe, ps = a.profile_get(el), a.profile_get(sc)
        a.profile(False)
        res[f"build_{name}_ms"] = 1e3 * t
        res[f"elements_per_s_{name}"] = E / t
        s = 8 if prec == N.FP64 else 4
        ke_ms = pe[0] / max(pe[1], 1)
        sc_ms = ps[0] / max(ps[1], 1)
        npe = prob.conn.shape[1]
        # element kernel, algorithmic bytes per element: the element matrix and
        # its lifting values written, the vertex ids and coordinates, the dof
        # ids, flags and values read
        ke_bytes = E * (nl * nl * s + nl * 8 + npe * 4 + npe * 24 + nl * (4 + 1 + 8))
        # staged scatter: per structural entry its info word and value written,
        # per contribution its slot and element-matrix value, per element row
        # its pair id and lifting value, per row the segment offset, b read and
        # written and both inverse diagonals
        sc_bytes = nnz_s * (2 + s) + E * nl * nl * (2 + s) + E * nl * (4 + 8) + prob.ndofs * (4 + 16 + 12)
        res[f"element_{name}_ms"] = ke_ms
        res[f"element_{name}_GB/s"] = ke_bytes / (ke_ms * 1e-3) / 1e9 if ke_ms else None
        res[f"scatter_{name}_ms"] = sc_ms
        res[f"scatter_{name}_GB/s"] = sc_bytes / (sc_ms * 1e-3) / 1e9 if sc_ms else None
    res["nnz"] = int(a.nnz)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        m = {"quad_q2q1": (128, 128), "tet_p2p1": (8, 8, 8), "hex_q2q1": (8, 8, 8)}[pair]
        vfe, mesh = {"quad_q2q1": ("quad_q2", "square"), "tet_p2p1": ("tet_p2", "cube_tet5"),
                     "hex_q2q1": ("hex_q2", "cube_hex")}[pair]
        o = O.Space(vfe, mesh, m, length=ln)
        sp_ = W.StokesProblem(m, pair=pair)
        _, _, flag, full = sp_.dirichlet(bcs)
        coefs = {"tensor": 0.1, "grad": -1.0, "div": -1.0}
        o.stokes(False, coefs, flag, full)
        tc = time.perf_counter()
        reps = 3
        for _ in range(reps):
            o.stokes(False, coefs, flag, full)
        tcpu = (time.perf_counter() - tc) / reps
        cpu = {"value": o.nelem / tcpu, "unit": "elements/s", "cores": 1, "kind": "port",
               "sample": f"oracle buildLhs + 2 x buildCoupling + close, {'x'.join(map(str, m))} {mesh}, {reps} reps"}
    return {"metric": f"Stokes {pair} {'x'.join(map(str, n))} build (tensor stiffness + grad + div), elements/s",
            "value": res["elements_per_s_fp64"], "unit": "elements/s",
            "config": {"workload": f"Taylor-Hood {pair} on {'x'.join(map(str, n))} perturbed {prob.vel.mesh_kind}, "
                                   "BLOCK dofs, Dirichlet velocity", "elements": E, "dofs": prob.ndofs},
            "results": res, "cpu_baseline": cpu}


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64, help="elements per axis")
    ap.add_argument("--spmv", type=int, default=100, help="SpMV applications timed")
    ap.add_argument("--cpu-sample", type=int, default=64,
                    help="elements per axis of the CPU baseline's workload (64: the benchmarked one)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--solve-levels", type=int, default=5)
    ap.add_argument("--solve-reps", type=int, default=2)
    ap.add_argument("--config", choices=["c2", "c3", "c5", "stokes"], default=None,
                    help="run one of the other workloads instead of the headline one (one JSON line)")
    ap.add_argument("--pair", choices=["quad_q2q1", "tet_p2p1", "hex_q2q1"], default="quad_q2q1",
                    help="Taylor-Hood pair of --config stokes")
    ap.add_argument("--copy-gib", type=float, default=1.0,
                    help="size of a plain device copy timed for reference (0: none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import proxpde_amd.native as N
    from proxpde_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.init()
    if args.config:
        fn = {"c2": config_c2, "c3": config_c3, "c5": config_c5, "stokes": config_stokes}[args.config]
        progress(f"config {args.config}")
        out = fn(args, dev, N, W)
        out.update({"n_gpus": 1, "data": "synthetic", "dtype": "f64 / f32"})
        print(json.dumps(out))
        return
    n = args.n
    t0 = time.time()
    prob = W.Problem("hex_q2", (n, n, n))
    xd = prob.dof_coords()
    kappa = np.exp(W.grf(xd, seed=2, std=1.0, ell=0.1, features=1024, device=dev))
    dofs = prob.all_sides()
    a = prob.assembler()
    a.set_dirichlet(dofs, np.zeros(len(dofs)))
    a.set_coefficient(kappa)
    t_setup_host = time.time() - t0
    torch.cuda.synchronize()
    t1 = time.time()
    nnz_s = a.build_pattern()
    t_pattern = time.time() - t1
    E = prob.mesh.nelem
    terms = [("stiffness_fe", 1.0)]

    def step():
        a.clear()
        a.assemble(terms)
        a.close(1e-16)

    progress(f"setup done (host {t_setup_host:.1f} s, pattern {t_pattern:.2f} s); assembling")
    # the first build after the pattern takes the full close (keep flags, scan,
    # compaction); later rebuilds write into that pruned layout when the prune
    # set is unchanged, which close verifies
    t_first = None
    for w in range(args.warmup):
        torch.cuda.synchronize()
        tw = time.perf_counter()
        step()
        torch.cuda.synchronize()
        if w == 0:
            t_first = 1e3 * (time.perf_counter() - tw)
    a.profile(True)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - ts
    ms_step = 1e3 * elapsed / args.steps
    prof = {k: a.profile_get(k) for k in ("geometry", "element", "scatter", "rhs_scatter", "close")}
    a.profile(False)
    nnz = a.nnz

    progress(f"assembly {ms_step:.2f} ms/step; SpMV")
    # --- SpMV on the assembled matrix, on a torch stream so torch events time it
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    a.set_stream(stream.cuda_stream)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    x = torch.rand(prob.ndofs, dtype=torch.float64, device=dev, generator=gen) * 2 - 1
    y = torch.empty_like(x)
    x32 = x.float()
    y32 = torch.empty_like(x32)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def time_spmv(xv, yv, prec):
        for _ in range(5):
            a.spmv_dev(xv.data_ptr(), yv.data_ptr(), precision=prec)
        ev0.record(stream)
        for _ in range(args.spmv):
            a.spmv_dev(xv.data_ptr(), yv.data_ptr(), precision=prec)
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / args.spmv

    # algorithmic bytes: values + column indices, row pointers, x and y once
    spmv_bytes = nnz * 12 + 8 * (prob.ndofs + 1) + 16 * prob.ndofs
    spmv32_bytes = nnz * 8 + 8 * (prob.ndofs + 1) + 8 * prob.ndofs
    spmv_ms = time_spmv(x, y, N.FP64)
    spmv32_ms = time_spmv(x32, y32, N.FP32)
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    spmv32_gbs = spmv32_bytes / (spmv32_ms * 1e-3) / 1e9
    # the block-packed copy (18-bit column offsets, what the multigrid levels use)
    a.pack(True)
    pk_ms = time_spmv(x, y, N.FP64)
    pk32_ms = time_spmv(x32, y32, N.FP32)
    pk_active = a.packed()

    # one weighted-Jacobi sweep (the multigrid smoother, omega 2/3) on the packed
    # copy: matrix + x, b, the inverse diagonal read and x written.  Two sweeps
    # per call of pxg_jacobi_noalloc (x -> its scratch vector -> x, no copy and
    # no host wait between calls), the time per sweep reported
    def time_jacobi(xv, bv, prec):
        for _ in range(3):
            a.jacobi_noalloc_dev(bv.data_ptr(), xv.data_ptr(), 2.0 / 3.0, 2, precision=prec)
        ev0.record(stream)
        for _ in range(args.spmv):
            a.jacobi_noalloc_dev(bv.data_ptr(), xv.data_ptr(), 2.0 / 3.0, 2, precision=prec)
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / (2 * args.spmv)

    bj = torch.rand(prob.ndofs, dtype=torch.float64, device=dev, generator=gen) * 2 - 1
    jx, jx32 = y.clone(), y32.clone()
    jac_ms = time_jacobi(jx, bj, N.FP64)
    jac32_ms = time_jacobi(jx32, bj.float(), N.FP32)
    jac_bytes = nnz * 10.25 + 4 * 8 * prob.ndofs
    jac32_bytes = nnz * 6.25 + 4 * 4 * prob.ndofs
    del jx, jx32, bj
    a.pack(False)
    pk_bytes = nnz * 10.25 + 8 * (prob.ndofs + 1) + 16 * prob.ndofs
    pk32_bytes = nnz * 6.25 + 8 * (prob.ndofs + 1) + 8 * prob.ndofs
    # the row-group kernel on the same matrix, for comparison
    os.environ["PXG_SPMV_KERNEL"] = "rowgroup"
    rg_ms = time_spmv(x, y, N.FP64)
    rg32_ms = time_spmv(x32, y32, N.FP32)
    del os.environ["PXG_SPMV_KERNEL"]

    # --- a plain device-to-device copy of known size: the achievable HBM rate,
    # and the calibration of the FETCH_SIZE / WRITE_SIZE counters
    copy_cal = None
    if args.copy_gib > 0:
        nel = int(args.copy_gib * 2 ** 30) // 8
        src = torch.ones(nel, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        ev0.record(stream)
        for _ in range(3):
            dst.copy_(src)
        ev1.record(stream)
        torch.cuda.synchronize()
        cms = ev0.elapsed_time(ev1) / 3
        # the same copy as a compute kernel (an elementwise product by one),
        # which the FETCH_SIZE / WRITE_SIZE counter runs always see
        torch.mul(src, 1.0, out=dst)
        ev0.record(stream)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        ev1.record(stream)
        torch.cuda.synchronize()
        kms = ev0.elapsed_time(ev1) / 3
        copy_cal = {"bytes_each_way": nel * 8, "ms": cms, "GB/s_read_plus_write": 2 * nel * 8 / (cms * 1e-3) / 1e9,
                    "kernel_ms": kms, "kernel_GB/s_read_plus_write": 2 * nel * 8 / (kms * 1e-3) / 1e9}
        del src, dst

    # --- rooflines: the dominant kernel of the build step (the JSON "roofline"),
    # and the scatter and SpMV beside it.  achieved = algorithmic work per launch
    # / the launch's average duration (HIP events on the launch stream);
    # traffic = HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter
    # runs recorded in profiles/ (calibrated against a plain device copy).
    per = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in prof.items()}
    traffic = measured_traffic(n)

    def tr(key):
        return traffic[key]["bytes"] if key in traffic else None

    # element kernel: reference-formulation flops (q2hex_flops_per_elem) per element
    el_fl = q2hex_flops_per_elem() * E
    el_ach = el_fl / (per["element"] * 1e-3) / 1e12 if per["element"] else None
    rf_element = {"kernel": "k_element_mfma<double,27,3,8,27,false>", "bound": "mfma", "achieved": el_ach,
                  "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                  "frac": el_ach / FP64_PEAK_TFLOPS if el_ach else None, "traffic": tr("element")}
    # rebuild into the pruned layout (staged by row blocks, element rows in
    # pair order): per structural entry its 16-bit info word and the value
    # written (8 B, kept or pruned); per element-matrix entry its 16-bit LDS
    # slot and its value (8 B); per element-matrix row its lifting value
    # (8 B); per row its pair offset, b read and written, both inverse
    # diagonals (4 + 16 + 12 B)
    sc_by = nnz_s * (2 + 8) + E * 27 * 27 * (2 + 8) + E * 27 * 8 + prob.ndofs * (4 + 16 + 12)
    sc_ach = sc_by / (per["scatter"] * 1e-3) / 1e9 if per["scatter"] else None
    rf_scatter = {"kernel": "k_scatter_staged<double,27,32,64,true>", "bound": "hbm", "achieved": sc_ach,
                  "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": sc_ach / HBM_PEAK_GBS if sc_ach else None, "traffic": tr("scatter")}
    rf_spmv = {"kernel": "k_csr_stream<double,0>", "bound": "hbm", "achieved": spmv_gbs, "peak": HBM_PEAK_GBS,
               "unit": "GB/s", "frac": spmv_gbs / HBM_PEAK_GBS, "traffic": tr("spmv_fp64")}
    roof = rf_element if per["element"] >= per["scatter"] else rf_scatter
    jac_gbs = jac_bytes / (jac_ms * 1e-3) / 1e9
    rf_jacobi = {"kernel": "k_csr_packed<double,2,8>", "bound": "hbm", "achieved": jac_gbs, "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": jac_gbs / HBM_PEAK_GBS, "traffic": tr("jacobi_packed_fp64")}
    rooflines = {"element": rf_element, "scatter": rf_scatter, "spmv_fp64": rf_spmv, "jacobi_fp64": rf_jacobi}

    # --- solve + adjoint gradient, plain FP64 vs FP32 + FP64 refinement
    progress("matrix-free K(kappa) x, batch 8")
    mf = kappa_apply_batch(args, dev, a, np.asarray(kappa))

    solve = None
    progress("solve + adjoint")
    if not args.no_solve:
        solve = solve_and_adjoint(args, dev, N, W)

    # --- CPU baseline: the oracle's Builder (buildLhs + closeMatrix) on the
    # same workload (same mesh, dof numbering, kappa and Dirichlet data), on
    # all host cores (element loop and per-row close threaded with OpenMP,
    # bitwise the serial result) and on one core as the reference runs it
    cpu = None
    progress("CPU baseline")
    if not args.no_cpu:
        from oracle import oracle as O
        m = args.cpu_sample
        s = O.Space("hex_q2", "cube_hex", (m, m, m))
        if m == n:
            kap, dflag = np.asarray(kappa), dofs
        else:
            ps = W.Problem("hex_q2", (m, m, m))
            kap = np.exp(W.grf(ps.dof_coords(), seed=2, std=1.0, ell=0.1, features=1024))
            dflag = ps.all_sides()
        flag = np.zeros(s.ndofs, np.uint8)
        flag[dflag] = 1
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        model = None
        try:
            with open("/proc/cpuinfo") as fh:
                model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), None)
        except OSError:
            pass
        times = {}
        for th in (cores, 1):
            s.set_threads(th)
            tc = time.perf_counter()
            s.assemble(terms, kappa=kap, bc=(flag, np.zeros(s.ndofs)))
            times[th] = time.perf_counter() - tc
        cpu = {"value": (m ** 3) / times[cores], "unit": "elements/s", "cores": cores, "kind": "port",
               "sample": f"oracle Builder (buildLhs + closeMatrix), Q2 hex {m}^3 = {m ** 3} elements, same kappa "
                         f"and Dirichlet data, element loop and per-row close on {cores} OpenMP threads, one pass "
                         f"{times[cores]:.2f} s",
               "serial": {"value": (m ** 3) / times[1], "cores": 1, "seconds": times[1]},
               "host": {"nproc": os.cpu_count(), "affinity": cores, "cpu_model": model}}

    value = E / (ms_step * 1e-3)
    out = {
        "metric": "elements assembled/s (Q2 hex 64^3, kappa stiffness, buildLhs+closeMatrix)",
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Q2 hex {n}^3 unit cube, 3x3x3 Gauss, lognormal kappa (GRF std 1, ell 0.1, seed 2), "
                               "all-sides Dirichlet", "elements": E, "dofs": prob.ndofs,
                   "nnz_structural": nnz_s, "nnz": nnz, "parallelism": "single GPU"},
        "kernels_ms": {k: per[k] for k in per},
        "first_build_ms": t_first,
        "pattern_build_s": t_pattern,
        "host_setup_s": t_setup_host,
        "spmv": {"ms": spmv_ms, "GB/s": spmv_gbs, "frac_hbm": spmv_gbs / HBM_PEAK_GBS,
                 "fp32_ms": spmv32_ms, "fp32_GB/s": spmv32_gbs, "bytes": spmv_bytes,
                 "rowgroup_kernel_ms": rg_ms, "rowgroup_kernel_fp32_ms": rg32_ms,
                 "jacobi_packed": {"ms": jac_ms, "GB/s": jac_bytes / (jac_ms * 1e-3) / 1e9, "fp32_ms": jac32_ms,
                                   "fp32_GB/s": jac32_bytes / (jac32_ms * 1e-3) / 1e9, "bytes": jac_bytes,
                                   "fp32_bytes": jac32_bytes},
                 "packed": {"active": pk_active, "ms": pk_ms, "GB/s": pk_bytes / (pk_ms * 1e-3) / 1e9,
                            "fp32_ms": pk32_ms, "fp32_GB/s": pk32_bytes / (pk32_ms * 1e-3) / 1e9,
                            "bytes": pk_bytes, "fp32_bytes": pk32_bytes}},
        "device_copy": copy_cal,
        "reference_flops_per_elem": q2hex_flops_per_elem(),
        "reference_gflops_per_s": q2hex_flops_per_elem() * value / 1e9,
        "kappa_apply": mf,
        "solve_adjoint": solve,
        "roofline": roof,
        "rooflines": rooflines,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
