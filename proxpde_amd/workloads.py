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
    """Move vertices off the boundary by U(-amount*h, amount*h) per coordinate
    (coordinates whose length is 0, the normal of a planar mesh, stay put)."""
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
    d[:, ln == 0] = 0.0  # a planar (or line) mesh stays in its plane
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


PAIRS = {"quad_q2q1": ("quad_q2", "quad_q1", 2, "square"), "hex_q2q1": ("hex_q2", "hex_q1", 3, "cube_hex"),
         "tet_p2p1": ("tet_p2", "tet_p1", 3, "cube_tet5")}


class StokesProblem:
    """Taylor-Hood system on a structured mesh (host layer).

    Velocity: the Q2 / P2 space with dim components, BLOCK (id + d nU) or
    INTERLEAVED (dim id + d) dof ordering (reference include/dof.hpp:139-166);
    pressure: the Q1 / P1 space after the velocity (offset dim nU), as in the
    reference's test/assemblyvector.cpp, stokes2dquad.cpp, stokes3dhex.cpp and
    stokes3dtet.cpp.  `emap` is the dof map of the pair's element
    (PXG_FE_QUAD_Q2Q1 / HEX_Q2Q1 / TET_P2P1): [u_0 (NV) | ... | u_dim-1 | p]."""

    def __init__(self, n, interleaved=False, pair="quad_q2q1", mesh=None, length=None):
        vfe, pfe, dim, kind = PAIRS[pair]
        self.pair = pair
        self.dim = dim
        self.interleaved = bool(interleaved)
        if length is None:
            length = (1.0, 1.0, 0.0) if dim == 2 else (1.0, 1.0, 1.0)
        self.vel = Problem(vfe, n, mesh=mesh or kind, length=length)
        self.mesh = self.vel.mesh
        self.xyz, self.conn = self.vel.xyz, self.vel.conn
        self.nU = self.vel.ndofs
        self.pfe = pfe
        self.nP, emapP = self.mesh.dofmap(pfe)
        self.emapP = emapP
        ev = self.vel.emap.astype(np.int64)
        E, nv = ev.shape
        emap = np.empty((E, dim * nv + emapP.shape[1]), dtype=np.uint32)
        for d in range(dim):
            emap[:, nv * d:nv * (d + 1)] = self.vel_id(ev, d)
        emap[:, dim * nv:] = emapP.astype(np.int64) + dim * self.nU
        self.emap = emap
        self.ndofs = dim * self.nU + self.nP

    def vel_id(self, i, d):
        """Global id of component d of velocity dof i."""
        i = np.asarray(i, dtype=np.int64)
        return self.dim * i + d if self.interleaved else i + d * self.nU

    def set_coords(self, xyz):
        self.vel.set_coords(xyz)
        self.xyz = self.vel.xyz

    def dof_coords(self):
        """Coordinates of the velocity dofs."""
        return self.vel.dof_coords()

    def pressure_coords(self):
        return self.mesh.dof_coords(self.pfe, self.emapP, self.nP)

    def dirichlet(self, bcs, pressure=()):
        """Dirichlet data of the whole system from BCEss-like entries
        (side, value, comps): value is a dim-tuple or a function of the dof
        coordinates returning an (n, dim) array, comps the constrained
        components (default all).  Later entries win where sides share dofs,
        as the reference's BCEss list does.  pressure: (pressure dof, value)
        pairs.  Returns (dofs, values) for set_dirichlet and the full flag /
        value arrays."""
        flag = np.zeros(self.ndofs, np.uint8)
        val = np.zeros(self.ndofs)
        x = self.dof_coords()
        for entry in bcs:
            side, value = entry[0], entry[1]
            comps = entry[2] if len(entry) > 2 else tuple(range(self.dim))
            ids = self.vel.bc_dofs((side,)).astype(np.int64)
            vv = value(x[ids]) if callable(value) else np.tile(np.asarray(value, np.float64), (len(ids), 1))
            for d in comps:
                g = self.vel_id(ids, d)
                flag[g] = 1
                val[g] = vv[:, d]
        for k, v in pressure:
            flag[self.dim * self.nU + k] = 1
            val[self.dim * self.nU + k] = v
        dofs = np.nonzero(flag)[0].astype(np.uint32)
        return dofs, val[dofs], flag, val

    def assembler(self, qr=None, device=0):
        a = N.Assembler(device)
        a.set_fe(self.pair, qr)
        a.set_mesh(self.xyz, self.conn)
        a.set_dofmap(self.ndofs, self.emap)
        return a
