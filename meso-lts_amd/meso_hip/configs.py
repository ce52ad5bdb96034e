"""The benchmark workloads as seeded meso cases.

  cavity64_bgk      64x64 uniform quads, D2Q9 Gauss-Hermite (RT=0.5),
                    cdugks@incompressible with the case-re400 settings
                    (Re 400, Ma 0.1414213562373095, dt 2.5e-3, lid u = 0.1).
  cavity_kn0075_tri Delaunay of a 101^2 node lattice jittered by 0.2h, 48^2
                    Newton-Cotes on [-4.5, 4.5]^2, Shakhov Pr = 2/3, Kn 0.075,
                    lid u = 0.2, dt = 0.5 * min cell size / max |xi|.
  cavity_kn1_quad   200x200 quads jittered by 10%, 101^2 Newton-Cotes on
                    [-5, 5]^2, Shakhov Kn = 1, lid u = 0.2, dt by CFL 0.5.
  cube_kn01_hex     32^3 bricks jittered by 5%, 16^3 Gauss-Hermite (Sparse),
                    Shakhov Kn = 0.1, lid u = 0.1 on y = 1.
  batch instances   (batch_instances) seeded independent Shakhov cavities:
                    N x N quads with N ~ U{32..128} (10% jitter), Kn
                    log-uniform on [0.01, 10], lid u ~ U[0.05, 0.3], M ~
                    U{16..64} Newton-Cotes points per axis on [-4.5, 4.5].
Viscosity from Kn by mu = 15 rho Kn / (2 (5-2w)(7-2w)) sqrt(2 pi R T) with
w = 0.5, R = 0.5, T = 1 (files/scripts/para_nodim.py:20-25).
"""
from __future__ import annotations

import os

from . import meshgen as mg

WORKLOADS = {
    "cavity64_bgk": dict(kind="bgk", n=64, dvs=3, steps=2000),
    "cavity_kn0075_tri": dict(kind="shak", mesh="tri", n=100, jitter=0.2, seed=2, dvs=("nc", 1, 47, 4.5), kn=0.075,
                              lid=0.2),
    "cavity_kn1_quad": dict(kind="shak", mesh="quad", n=200, jitter=0.1, seed=3, dvs=("nc", 1, 100, 5.0), kn=1.0,
                            lid=0.2),
    "cube_kn01_hex": dict(kind="shak", mesh="hex", n=32, jitter=0.05, seed=4, dvs=("sparse-gh", 16, 0.5), kn=0.1,
                          lid=0.1),
}


def build(name: str, root: str, scale: float = 1.0) -> str:
    """Write workload `name` under root/name and return its case path. `scale`
    (<1) shrinks the mesh resolution for quick runs."""
    w = WORKLOADS[name]
    d = os.path.join(root, name if scale == 1.0 else f"{name}_s{scale:g}")
    case = os.path.join(d, "case.txt")
    if os.path.exists(case):
        return case
    if w["kind"] == "bgk":
        n = max(4, int(w["n"] * scale))
        return mg.bgk_cavity_case(d, mg.quad_cavity(n), gh_points=w["dvs"], max_step=w["steps"])
    n = max(4, int(w["n"] * scale))
    if w["mesh"] == "quad":
        mesh = mg.quad_cavity(n, jitter=w["jitter"], seed=w["seed"])
    elif w["mesh"] == "tri":
        mesh = mg.tri_cavity(n, jitter=w["jitter"], seed=w["seed"])
    else:
        mesh = mg.hex_cavity(n, jitter=w["jitter"], seed=w["seed"])
    return mg.shakhov_cavity_case(d, mesh, w["dvs"], mg.viscosity_from_kn(w["kn"]), w["lid"], name=name)


def batch_instances(count: int, base_seed: int = 1000) -> list:
    """Parameters of `count` independent cavity instances, instance i drawn
    from seed base_seed + i."""
    import numpy as np

    out = []
    for i in range(count):
        rng = np.random.default_rng(base_seed + i)
        out.append(dict(seed=base_seed + i, n=int(rng.integers(32, 129)), kn=float(10 ** rng.uniform(-2.0, 1.0)),
                        lid=float(rng.uniform(0.05, 0.3)), m=int(rng.integers(16, 65))))
    return out


def build_instance(inst: dict, root: str, scale: float = 1.0) -> str:
    """Write one instance of batch_instances under root and return its case path."""
    n = max(4, int(inst["n"] * scale))
    d = os.path.join(root, f"batch_{inst['seed']}_{n}")
    case = os.path.join(d, "case.txt")
    if os.path.exists(case):
        return case
    mesh = mg.quad_cavity(n, jitter=0.1, seed=inst["seed"])
    return mg.shakhov_cavity_case(d, mesh, ("nc", 1, inst["m"] - 1, 4.5), mg.viscosity_from_kn(inst["kn"]), inst["lid"],
                                  name=f"batch_{inst['seed']}")
