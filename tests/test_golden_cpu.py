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
    orc.set_mask()  # the case's `mask switch True` groups (shak_tri6_mask)
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
                assert max_rel(fm[i], g[f"step{s}_{k}"]) <= 1e-12, f"step {s} {k}"


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


@pytest.mark.parametrize("name", [c for c in CASES if "step100_residual" in np.load(os.path.join(GOLDEN, c, "golden.npz")).files])
def test_oracle_residual_matches_meso(name):
    """meso's residuals at its check steps (Solver::residual, field.cpp:222-249,
    on the fields do_step compares; every 100 steps). The division by the old
    value gives inf / NaN where that is 0 (at rest, and w in 2D): those match
    as they are."""
    g = load(name)
    cs = capi.Case(os.path.join(GOLDEN, name, "case.txt"))
    orc = Oracle(cs)
    orc.init_equilibrium()
    done = 0
    for s in range(100, int(max(g["steps"])) + 1, 100):
        orc.step(s - done)
        done = s
        r, ref = orc.residual(), g[f"step{s}_residual"]
        fin = np.isfinite(ref)
        assert np.array_equal(np.isnan(r), np.isnan(ref)) and np.array_equal(np.isinf(r), np.isinf(ref)), s
        assert np.all(np.abs(r[fin] - ref[fin]) <= 1e-12 * np.abs(ref[fin])), (s, r, ref)


def test_oracle_mask_repair_matches_meso():
    """shak_tri6 with `mask switch True`: cell 10 goes NaN at step 26 and meso
    repairs it (cell_value_interp, cdugks_shakhov.cpp:766-789) instead of
    stopping; the full state after that step and the macros after the next
    ones match meso's (NaN heat fluxes included)."""
    g = load("shak_tri6_mask")
    cs = capi.Case(os.path.join(GOLDEN, "shak_tri6_mask", "case.txt"))
    assert cs.cell_mask().all()
    orc = Oracle(cs)
    orc.set_mask()
    orc.init_equilibrium()
    rep = orc.step(25)
    assert rep.bad_cells == 0 and rep.masked_cells == 0
    rep = orc.step(1)
    assert rep.bad_cells == 0 and rep.masked_cells >= 1
    f = orc.get_state()
    assert max_rel(f[0], g["state26_g"]) <= TOL
    assert max_rel(f[1], g["state26_h"]) <= TOL
    m = orc.macros()
    assert np.isnan(m["q"][9]).all() and np.isfinite(m["rho"][9])  # q of the repaired cell is its NaN sum
    # without the mask the same run stops at step 26 with that cell
    orc2 = Oracle(cs)
    orc2.init_equilibrium()
    rep = orc2.step(30)
    assert rep.steps_done == 26 and rep.first_bad_cell == 9
