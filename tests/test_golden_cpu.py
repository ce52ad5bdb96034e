"""The host front-end and the FP64 oracle against golden vectors written by
meso itself (tests/golden/make_golden.py runs the reference compiled from its
sources). CPU only.

The oracle restates meso's loops in the same order and is compiled without
FMA contraction, so on these inputs it reproduces meso bit for bit; the
tolerance below (1e-13 relative to the field's max) leaves room only for a
different libm.
"""
import os

import numpy as np
import pytest

from conftest import max_rel
from meso_hip import capi
from oracle.oracle import Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(d for d in os.listdir(GOLDEN) if os.path.exists(os.path.join(GOLDEN, d, "golden.npz")))
TOL = 1e-13


def load(name):
    g = np.load(os.path.join(GOLDEN, name, "golden.npz"), allow_pickle=False)
    return {k: g[k] for k in g.files}


@pytest.mark.parametrize("name", CASES)
def test_frontend_geometry_matches_meso(name):
    g = load(name)
    cs = capi.Case(os.path.join(GOLDEN, name, "case.txt"))
    A = cs.arrays()
    assert (cs.n_cells, cs.n_faces, cs.n_vel, cs.mesh.dim) == (int(g["nc"]), int(g["nf"]), int(g["nv"]), int(g["dim"]))
    assert np.array_equal(A["face_cells"], g["face_cells"])
    assert np.array_equal(np.diff(A["nbr_ptr"]), g["nbr_count"])
    for k, gk in (("cell_pos", "cell_pos"), ("cell_vol", "cell_vol"), ("face_pos", "face_pos"),
                  ("face_area", "face_area"), ("face_normal", "face_normal"), ("lsq_C", "lsq_C"),
                  ("xi", "dvs_xi"), ("w", "dvs_w")):
        assert max_rel(A[k], g[gk]) <= TOL, k
    assert abs(cs.params.dt - float(g["dt"])) <= TOL * abs(float(g["dt"]))
    assert abs(A["total_volume"] - g["mesh_scalars"][0]) <= 1e-14
    assert abs(A["min_cell_size"] - g["mesh_scalars"][1]) <= 1e-14


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_meso(name):
    g = load(name)
    cs = capi.Case(os.path.join(GOLDEN, name, "case.txt"))
    orc = Oracle(cs)
    orc.init_equilibrium()
    keys = ["rho", "u", "T", "q"] if cs.model == capi.MODEL_CDUGKS_SHAKHOV else ["rho", "u"]
    m = orc.macros()
    for k in keys:
        assert max_rel(m[k], g[f"step0_{k}"]) <= TOL, f"step 0 {k}"
    done = 0
    for s in g["steps"]:
        s = int(s)
        orc.step(s - done)
        done = s
        m = orc.macros()
        for k in keys:
            assert max_rel(m[k], g[f"step{s}_{k}"]) <= TOL, f"step {s} {k}"
        if cs.model == capi.MODEL_CDUGKS_SHAKHOV:
            fm = orc.flux_moments()
            for i, k in enumerate(("fluxM0", "fluxM1", "fluxM2")):
                ref = g[f"step{s}_{k}"]
                scale = max(np.abs(ref).max(), 1e-300)
                assert np.abs(fm[i] - ref).max() <= 1e-12 * scale + 1e-18, f"step {s} {k}"


@pytest.mark.parametrize("name", [c for c in CASES if c in ("bgk_quad8", "shak_quad6")])
def test_oracle_state_matches_meso(name):
    """Full distribution state after a few steps, velocity-major like meso."""
    g = load(name)
    ss = int(g["state_step"])
    cs = capi.Case(os.path.join(GOLDEN, name, "case.txt"))
    orc = Oracle(cs)
    orc.init_equilibrium()
    orc.step(ss)
    f = orc.get_state()
    if cs.model == capi.MODEL_CDUGKS_SHAKHOV:
        assert max_rel(f[0], g[f"state{ss}_g"]) <= TOL
        assert max_rel(f[1], g[f"state{ss}_h"]) <= TOL
    else:
        assert max_rel(f[0], g[f"state{ss}_f"]) <= TOL
