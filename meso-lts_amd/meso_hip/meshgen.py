"""Seeded generators for meso inputs: gambit ``.neu`` meshes, DVS files and
case files.

The written files follow the subsets meso reads:
  * gambit neutral file -- src/mesh/mesh_reader.cpp:8-139 (header counts line
    after ``NUMNP``, ``NODAL COORDINATES``, ``ELEMENTS/CELLS`` with at most 7
    node ids per line, ``ELEMENT GROUP``, ``BOUNDARY CONDITIONS`` entries
    ``cell type local-face``, all 1-based); local face k of a quad/triangle is
    the edge (k, k+1), brick faces follow src/mesh/object.cpp:32 with
    lexicographic brick nodes.
  * DVS files -- src/solver/dvs/dvs_generator.cpp (``Gauss-Hermite`` x/w lines,
    ``Newton-Cotes`` ``n mount scale``, ``Sparse`` xi_x xi_y xi_z w lines).
  * case files -- src/solver/config.cpp:30-108 ([settings]/[group]/[mark]).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

QUAD, TRIA, BRICK = 2, 3, 4


@dataclass
class GambitMesh:
    nodes: np.ndarray                 # (NN, dim) float64
    cells: list                       # list of node-id lists (0-based)
    kinds: list                       # gambit element kind per cell
    dim: int
    group: str = "fluid"
    bsets: dict = field(default_factory=dict)  # name -> list of (cell, local_face) 0-based

    @property
    def n_cells(self):
        return len(self.cells)


def write_neu(mesh: GambitMesh, path: str, title: str = "meso-hip mesh") -> None:
    nn = len(mesh.nodes)
    out = []
    out.append("        CONTROL INFO 2.4.6")
    out.append("** GAMBIT NEUTRAL FILE")
    out.append(title)
    out.append("PROGRAM:  meso-hip meshgen     VERSION:  1")
    out.append("generated")
    out.append("     NUMNP     NELEM     NGRPS    NBSETS     NDFCD     NDFVL")
    out.append(f"{nn:10d}{mesh.n_cells:10d}{1:10d}{len(mesh.bsets):10d}{mesh.dim:10d}{mesh.dim:10d}")
    out.append("ENDOFSECTION")
    out.append("   NODAL COORDINATES 2.4.6")
    for i, p in enumerate(mesh.nodes):
        out.append(f"{i + 1:10d}" + "".join(f"{v:24.17e}" for v in p))
    out.append("ENDOFSECTION")
    out.append("      ELEMENTS/CELLS 2.4.6")
    for i, (kind, nl) in enumerate(zip(mesh.kinds, mesh.cells)):
        ids = [n + 1 for n in nl]
        head = f"{i + 1:8d} {kind:2d} {len(ids):2d} " + " ".join(f"{n:8d}" for n in ids[:7])
        out.append(head)
        if len(ids) > 7:
            out.append(" " * 15 + " ".join(f"{n:8d}" for n in ids[7:]))
    out.append("ENDOFSECTION")
    out.append("       ELEMENT GROUP 2.4.6")
    out.append(f"GROUP: {1:10d} ELEMENTS: {mesh.n_cells:10d} MATERIAL: {2:10d} NFLAGS: {1:10d}")
    out.append(f"{mesh.group:>32s}")
    out.append("       0")
    ids = np.arange(1, mesh.n_cells + 1)
    for s in range(0, len(ids), 10):
        out.append("".join(f"{v:8d}" for v in ids[s:s + 10]))
    out.append("ENDOFSECTION")
    for name, entries in mesh.bsets.items():
        out.append(" BOUNDARY CONDITIONS 2.4.6")
        out.append(f"{name:>32s}{1:8d}{len(entries):8d}{0:8d}{6:8d}")
        for c, lf in entries:
            out.append(f"{c + 1:10d}{mesh.kinds[c]:5d}{lf + 1:5d}")
        out.append("ENDOFSECTION")
    with open(path, "w") as fp:
        fp.write("\n".join(out) + "\n")


def _jitter_2d(xy: np.ndarray, h: float, amp: float, rng: np.random.Generator, eps=1e-12) -> np.ndarray:
    """Move interior nodes by U(-amp*h, amp*h) per coordinate; boundary nodes
    slide along their edge; corners stay."""
    out = xy.copy()
    d = rng.uniform(-amp * h, amp * h, size=xy.shape)
    on_x = (np.abs(xy[:, 0]) < eps) | (np.abs(xy[:, 0] - 1.0) < eps)
    on_y = (np.abs(xy[:, 1]) < eps) | (np.abs(xy[:, 1] - 1.0) < eps)
    d[on_x, 0] = 0.0
    d[on_y, 1] = 0.0
    out += d
    return out


def quad_cavity(nx: int, ny: int | None = None, jitter: float = 0.0, seed: int = 0) -> GambitMesh:
    """Unit-square quad mesh; marks ``lid`` (y=1) and ``wall`` (other sides)."""
    ny = nx if ny is None else ny
    xs, ys = np.linspace(0.0, 1.0, nx + 1), np.linspace(0.0, 1.0, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    if jitter > 0:
        nodes = _jitter_2d(nodes, min(1.0 / nx, 1.0 / ny), jitter, np.random.default_rng(seed))
    nid = lambda i, j: j * (nx + 1) + i
    cells, kinds = [], []
    lid, wall = [], []
    for j in range(ny):
        for i in range(nx):
            c = len(cells)
            cells.append([nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)])
            kinds.append(QUAD)
            if j == 0:
                wall.append((c, 0))
            if i == nx - 1:
                wall.append((c, 1))
            if j == ny - 1:
                lid.append((c, 2))
            if i == 0:
                wall.append((c, 3))
    return GambitMesh(nodes, cells, kinds, 2, bsets={"lid": lid, "wall": wall})


def _boundary_edge_count(tri, nodes, eps=1e-12):
    cnt = 0
    for k in range(3):
        a, b = nodes[tri[k]], nodes[tri[(k + 1) % 3]]
        if (abs(a[0] - b[0]) < eps and (abs(a[0]) < eps or abs(a[0] - 1) < eps)) or \
           (abs(a[1] - b[1]) < eps and (abs(a[1]) < eps or abs(a[1] - 1) < eps)):
            cnt += 1
    return cnt


def tri_cavity(n: int, jitter: float = 0.2, seed: int = 0) -> GambitMesh:
    """Delaunay triangulation of an (n+1)^2 lattice on the unit square with
    seed-jittered nodes. Triangles with two boundary edges (whose single
    neighbour makes the 2D least-squares matrix singular) are removed by
    flipping their interior edge."""
    from scipy.spatial import Delaunay

    xs = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    nodes = _jitter_2d(nodes, 1.0 / n, jitter, np.random.default_rng(seed))
    tris = [list(t) for t in Delaunay(nodes).simplices]
    # orient counter-clockwise
    for t in tris:
        a, b, c = nodes[t[0]], nodes[t[1]], nodes[t[2]]
        if (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]) < 0:
            t[1], t[2] = t[2], t[1]
    for _ in range(8):
        changed = False
        edge_map = {}
        for ti, t in enumerate(tris):
            for k in range(3):
                e = tuple(sorted((t[k], t[(k + 1) % 3])))
                edge_map.setdefault(e, []).append(ti)
        for ti, t in enumerate(tris):
            if _boundary_edge_count(t, nodes) < 2:
                continue
            for k in range(3):
                e = tuple(sorted((t[k], t[(k + 1) % 3])))
                others = [o for o in edge_map[e] if o != ti]
                if not others:
                    continue
                tj = others[0]
                a, b = e
                c = [v for v in t if v not in e][0]
                d = [v for v in tris[tj] if v not in e][0]
                tris[ti] = [c, a, d]
                tris[tj] = [c, d, b]
                for tt in (ti, tj):
                    p0, p1, p2 = (nodes[v] for v in tris[tt])
                    if (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p1[1] - p0[1]) * (p2[0] - p0[0]) < 0:
                        tris[tt][1], tris[tt][2] = tris[tt][2], tris[tt][1]
                changed = True
                break
            if changed:
                break
        if not changed:
            break
    lid, wall = [], []
    eps = 1e-12
    for ci, t in enumerate(tris):
        for k in range(3):
            a, b = nodes[t[k]], nodes[t[(k + 1) % 3]]
            if abs(a[1] - 1) < eps and abs(b[1] - 1) < eps:
                lid.append((ci, k))
            elif (abs(a[1]) < eps and abs(b[1]) < eps) or (abs(a[0]) < eps and abs(b[0]) < eps) or \
                 (abs(a[0] - 1) < eps and abs(b[0] - 1) < eps):
                wall.append((ci, k))
    return GambitMesh(nodes, tris, [TRIA] * len(tris), 2, bsets={"lid": lid, "wall": wall})


def hex_cavity(n: int, jitter: float = 0.0, seed: int = 0) -> GambitMesh:
    """Unit-cube brick mesh; ``lid`` is the y=1 face, ``wall`` the other five.
    Brick nodes are lexicographic: 0 (0,0,0), 1 (1,0,0), 2 (0,1,0), 3 (1,1,0),
    then the same at z=1."""
    xs = np.linspace(0.0, 1.0, n + 1)
    Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        h = 1.0 / n
        d = rng.uniform(-jitter * h, jitter * h, size=nodes.shape)
        on = (np.abs(nodes) < 1e-12) | (np.abs(nodes - 1.0) < 1e-12)
        d[on] = 0.0
        nodes = nodes + d
    nid = lambda i, j, k: (k * (n + 1) + j) * (n + 1) + i
    cells, lid, wall = [], [], []
    # brick local faces (object.cpp:32): 0 y-, 1 x+, 2 y+, 3 x-, 4 z-, 5 z+
    for k in range(n):
        for j in range(n):
            for i in range(n):
                c = len(cells)
                cells.append([nid(i, j, k), nid(i + 1, j, k), nid(i, j + 1, k), nid(i + 1, j + 1, k),
                              nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i, j + 1, k + 1), nid(i + 1, j + 1, k + 1)])
                if j == 0:
                    wall.append((c, 0))
                if i == n - 1:
                    wall.append((c, 1))
                if j == n - 1:
                    lid.append((c, 2))
                if i == 0:
                    wall.append((c, 3))
                if k == 0:
                    wall.append((c, 4))
                if k == n - 1:
                    wall.append((c, 5))
    return GambitMesh(nodes, cells, [BRICK] * len(cells), 3, bsets={"lid": lid, "wall": wall})


# ---------------------------------------------------------------- DVS files
def write_gauss_hermite(path: str, n: int) -> None:
    x, w = np.polynomial.hermite.hermgauss(n)
    with open(path, "w") as fp:
        fp.write("Gauss-Hermite\n")
        for a, b in zip(x, w):
            fp.write(f"{a:.18f} {b:.18f}\n")


def write_newton_cotes(path: str, degree: int, mount: int, scale: float) -> None:
    with open(path, "w") as fp:
        fp.write(f"Newton-Cotes\n{degree} {mount} {scale!r}\n")


def write_sparse_gh3d(path: str, n: int, RT: float) -> None:
    """Full-range 3D Gauss-Hermite tensor set written as a ``Sparse`` list:
    xi = sqrt(2RT) x_i, weight = prod over axes of sqrt(2RT) w_i exp(x_i^2),
    so that sum_v w_v f(xi_v) approximates the integral of f."""
    x, w = np.polynomial.hermite.hermgauss(n)
    s = math.sqrt(2.0 * RT)
    wt = s * w * np.exp(x * x)
    with open(path, "w") as fp:
        fp.write("Sparse\n")
        for k in range(n):
            for j in range(n):
                for i in range(n):
                    fp.write(f"{s * x[i]:.17e} {s * x[j]:.17e} {s * x[k]:.17e} {wt[i] * wt[j] * wt[k]:.17e}\n")


# ---------------------------------------------------------------- case files
def viscosity_from_kn(kn: float, rho: float = 1.0, L: float = 1.0, omega: float = 0.5, R: float = 0.5,
                      T: float = 1.0) -> float:
    """mu = 15 rho lambda / (2 (5-2w)(7-2w)) sqrt(2 pi R T), lambda = L Kn
    (files/scripts/para_nodim.py:20-25)."""
    return 15.0 * rho * L * kn / (2.0 * (5.0 - 2.0 * omega) * (7.0 - 2.0 * omega)) * math.sqrt(2.0 * math.pi * R * T)


def write_case(path: str, settings: dict, groups: list, marks: list) -> None:
    """settings: ordered key->value; groups/marks: lists of dicts of
    key -> token string (``name``/``type`` plus patch lines)."""
    lines = ["[settings]"]
    for k, v in settings.items():
        lines.append(f"{k:20s} {v}")
    for g in groups:
        lines.append("")
        lines.append("[group]")
        for k, v in g.items():
            lines.append(f"{k:16s} {v}")
    for m in marks:
        lines.append("")
        lines.append("[mark]")
        for k, v in m.items():
            lines.append(f"{k:16s} {v}")
    with open(path, "w") as fp:
        fp.write("\n".join(lines) + "\n")


def fmt(x: float) -> str:
    return repr(float(x))


def shakhov_cavity_case(case_dir: str, mesh: GambitMesh, dvs: tuple, mu0: float, lid_u: float, dt: float | None = None,
                        cfl: float = 0.5, Pr: float = 2.0 / 3.0, max_step: int = 1000, residual_limit: float = 1e-7,
                        name: str = "cavity", omega: float = 0.5, T_wall: float = 1.0, limiter: bool = False,
                        mask: bool = False) -> str:
    """Write mesh, DVS and a cdugks@shakhov case into case_dir; returns the case path.
    dvs = ("nc", degree, mount, scale) or ("sparse-gh", n, RT). dt defaults to
    cfl * min_cell_size / max|xi|. mask: the fluid group sets `mask switch True`
    (bad cells are repaired instead of stopping the run)."""
    os.makedirs(case_dir, exist_ok=True)
    write_neu(mesh, os.path.join(case_dir, "mesh.neu"))
    if dvs[0] == "nc":
        write_newton_cotes(os.path.join(case_dir, "dvs.dvs"), dvs[1], dvs[2], dvs[3])
        dvs_type = "Newton-Cotes"
        vmax = dvs[3] * math.sqrt(mesh.dim)
    elif dvs[0] == "sparse-gh":
        write_sparse_gh3d(os.path.join(case_dir, "dvs.dvs"), dvs[1], dvs[2])
        dvs_type = "Sparse"
        x, _ = np.polynomial.hermite.hermgauss(dvs[1])
        vmax = math.sqrt(2 * dvs[2]) * abs(x).max() * math.sqrt(3)
    else:
        raise ValueError(dvs)
    if dt is None:
        dt = cfl * min_cell_size(mesh) / vmax
    settings = {
        "solver": "cdugks@shakhov", "case-name": name, "mesh-file": "mesh.neu", "dvs-file": "dvs.dvs",
        "dvs-type": dvs_type, "write-interval": 100000, "max-step": max_step, "residual-limit": fmt(residual_limit),
        "Pr": fmt(Pr), "gradient-switch": "True", "limiter-switch": "True" if limiter else "False", "timeStep": fmt(dt),
        "gas-constant": "0.5", "gas-k": "0", "ref-density": "1.0", "ref-length": "1.0", "ref-temperature": "1.0",
        "ref-viscosity": fmt(mu0), "vhs-index": fmt(omega),
    }
    groups = [{"name": "fluid", "density": "fixedValue scalar 1.0", "temperature": "fixedValue scalar 1.0"}]
    if mask:
        groups[0]["mask"] = "switch True"
    zero = "0.0 0.0 0.0"
    marks = [
        {"name": "lid", "type": "wall", "velocity": f"fixedValue vector {fmt(lid_u)} 0.0 0.0",
         "temperature": f"fixedValue scalar {fmt(T_wall)}"},
        {"name": "wall", "type": "wall", "velocity": f"fixedValue vector {zero}",
         "temperature": f"fixedValue scalar {fmt(T_wall)}"},
    ]
    path = os.path.join(case_dir, "case.txt")
    write_case(path, settings, groups, marks)
    return path


def bgk_cavity_case(case_dir: str, mesh: GambitMesh, solver: str = "cdugks@incompressible", gh_points: int = 3,
                    Re: float = 400.0, Ma: float = 0.1414213562373095, dt: float = 2.5e-3, lid_u: float = 0.1,
                    max_step: int = 2000, residual_limit: float = 1e-30, name: str = "re400") -> str:
    """Write mesh, a Gauss-Hermite DVS and an incompressible BGK cavity case
    with the settings of files/case-re400.txt."""
    os.makedirs(case_dir, exist_ok=True)
    write_neu(mesh, os.path.join(case_dir, "mesh.neu"))
    write_gauss_hermite(os.path.join(case_dir, "dvs.dvs"), gh_points)
    settings = {
        "solver": solver, "case-name": name, "mesh-file": "mesh.neu", "dvs-file": "dvs.dvs",
        "dvs-type": "Gauss-Hermite", "write-interval": 100000, "max-step": max_step,
        "residual-limit": fmt(residual_limit), "Re": fmt(Re), "Ma": fmt(Ma), "gradient-switch": "True",
        "timeStep": fmt(dt), "gas-constant": "0.5", "ref-density": "1.0", "ref-length": "1.0",
        "ref-temperature": "1.0", "CFL": "0.8",
    }
    groups = [{"name": "fluid", "density": "fixedValue scalar 1.0"}]
    marks = [
        {"name": "lid", "type": "wall", "velocity": f"fixedValue vector {fmt(lid_u)} 0.0 0.0"},
        {"name": "wall", "type": "wall"},
    ]
    path = os.path.join(case_dir, "case.txt")
    write_case(path, settings, groups, marks)
    return path


def min_cell_size(mesh: GambitMesh) -> float:
    """Equivalent diameter of the smallest cell (src/mesh/object.cpp:118-119)."""
    best = None
    for kind, nl in zip(mesh.kinds, mesh.cells):
        p = mesh.nodes[nl]
        if mesh.dim == 2:
            def tri(a, b, c):
                return 0.5 * abs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]))
            v = tri(p[0], p[1], p[2]) + (tri(p[0], p[3], p[2]) if kind == QUAD else 0.0)
            d = 2.0 * math.sqrt(v / math.pi)
        else:
            lo, hi = p.min(axis=0), p.max(axis=0)
            v = float(np.prod(hi - lo))
            d = 2.0 * (3.0 * v / (4.0 * math.pi)) ** (1.0 / 3.0)
        best = d if best is None else min(best, d)
    return best
