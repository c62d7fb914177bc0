"""Synthetic inputs of the assembly / solve workloads.

Everything here is deterministic from explicit seeds:
  * uniform draws come from numpy's PCG64 (numpy.random.default_rng(seed));
  * the lognormal coefficient kappa = exp(g) uses a Gaussian random field g
    with squared-exponential covariance (std `std`, correlation length `ell`)
    sampled by random Fourier features: g(x) = std*sqrt(2/M) sum_m cos(w_m.x + b_m),
    w_m ~ N(0, I/ell^2), b_m ~ U(0, 2 pi), drawn from default_rng(seed).
The field is evaluated once on the host (or on the GPU through torch for large
node sets) and the same array is handed to every implementation.
"""
from __future__ import annotations

import numpy as np

from . import native as N

ALL_SIDES = {1: ("left", "right"), 2: ("bottom", "right", "top", "left"),
             3: ("bottom", "right", "top", "left", "back", "front")}

MESH_OF_FE = {"line_p1": "line", "quad_q1": "square", "quad_q2": "square", "hex_q1": "cube_hex",
              "hex_q2": "cube_hex", "tet_p1": "cube_tet6", "tet_p2": "cube_tet6"}


def perturb_interior(xyz, h, seed, amount=0.2, origin=(0.0, 0.0, 0.0), length=(1.0, 1.0, 1.0)):
    """Move vertices off the boundary by U(-amount*h, amount*h) per coordinate."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(-amount * h, amount * h, size=xyz.shape)
    o = np.asarray(origin)
    ln = np.asarray(length)
    tol = 1e-9 * max(ln.max(), 1.0)
    on_bd = np.zeros(len(xyz), dtype=bool)
    for k in range(3):
        if ln[k] > 0:
            on_bd |= np.abs(xyz[:, k] - o[k]) < tol
            on_bd |= np.abs(xyz[:, k] - o[k] - ln[k]) < tol
    out = xyz.copy()
    out[~on_bd] += d[~on_bd]
    return out


def grf(points, seed, std=1.0, ell=0.1, features=1024, device=None):
    """Gaussian random field at `points` (n x 3) by random Fourier features."""
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, 1.0 / ell, size=(features, 3))
    b = rng.uniform(0.0, 2.0 * np.pi, size=features)
    scale = std * np.sqrt(2.0 / features)
    pts = np.ascontiguousarray(points, dtype=np.float64)
    if device is not None:
        import torch
        P = torch.from_numpy(pts).to(device)
        W = torch.from_numpy(w).to(device)
        B = torch.from_numpy(b).to(device)
        out = torch.empty(len(pts), dtype=torch.float64, device=device)
        step = 1 << 18
        for s in range(0, len(pts), step):
            out[s:s + step] = torch.cos(P[s:s + step] @ W.T + B).sum(dim=1) * scale
        return out.cpu().numpy()
    out = np.empty(len(pts))
    step = 1 << 14
    for s in range(0, len(pts), step):
        out[s:s + step] = np.cos(pts[s:s + step] @ w.T + b).sum(axis=1) * scale
    return out


class Problem:
    """Mesh, dof map and boundary data of one structured workload (host layer)."""

    def __init__(self, fe: str, n, mesh: str | None = None, origin=(0.0, 0.0, 0.0), length=None, gmsh=None):
        """Structured mesh `mesh` with n elements per axis, or (gmsh = path) the
        mesh of a GMSH 2.2 file with the geometry of `fe`."""
        self.fe = fe
        self.mesh_kind = mesh or MESH_OF_FE[fe]
        dim = N.FE_DIM[N.FE[fe]]
        if length is None:
            length = {1: (1.0, 0.0, 0.0), 2: (1.0, 1.0, 0.0), 3: (1.0, 1.0, 1.0)}[dim]
        self.dim = dim
        self.origin = origin
        self.length = length
        self.n = list(n)
        if gmsh is not None:
            geo = {"line_p1": "line", "quad_q1": "quad", "quad_q2": "quad", "hex_q1": "hex", "hex_q2": "hex",
                   "tet_p1": "tet", "tet_p2": "tet"}[fe]
            self.mesh_kind = "gmsh"
            self.mesh = N.HostMesh.gmsh(gmsh, geo)
        else:
            self.mesh = N.HostMesh(self.mesh_kind, n, origin, length)
        self.xyz, self.conn = self.mesh.arrays()
        self.ndofs, self.emap = self.mesh.dofmap(fe)

    def set_coords(self, xyz):
        self.mesh.set_coords(xyz)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)

    def bc_dofs(self, sides):
        out = [self.mesh.bc_dofs(self.fe, self.emap, s) for s in sides]
        return np.unique(np.concatenate(out)) if out else np.zeros(0, np.uint32)

    def all_sides(self):
        return self.bc_dofs(ALL_SIDES[self.dim])

    def dof_coords(self):
        return self.mesh.dof_coords(self.fe, self.emap, self.ndofs)

    def assembler(self, qr=None, device=0):
        a = N.Assembler(device)
        a.set_fe(self.fe, qr)
        a.set_mesh(self.xyz, self.conn)
        a.set_dofmap(self.ndofs, self.emap)
        return a
