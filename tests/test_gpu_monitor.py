"""The NaN / T < 0 monitor and the mask repair on the GPU (cdugks_shakhov.cpp:
698-719, 766-789), and meso's residuals, against golden vectors of meso itself
and the oracle.

The coarse triangle cavity shak_tri6 (unlimited second-order reconstruction)
drives cell 10 to a NaN density at step 26 in meso, the oracle and here. That
is a numerical event of the scheme: every kernel runs to its end and the
device flags the cell.
"""
import os

import numpy as np
import pytest

from conftest import max_rel
from meso_hip import capi
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-10


def _say(*a):
    print(*a, flush=True)


def load(name):
    g = np.load(os.path.join(GOLDEN, name, "golden.npz"), allow_pickle=False)
    return {k: g[k] for k in g.files}


@pytest.mark.parametrize("graphs", [False, True])
def test_monitor_stops_at_the_bad_step(case_factory, graphs):
    """Without a mask the step that makes a cell bad completes (as in meso),
    later steps of the call are skipped, and the report names the cell."""
    cs = capi.Case(case_factory("shak_tri6"))
    _say(f"[gpu] monitor shak_tri6 graphs={graphs}")
    dev = capi.Device(cs)
    dev.set_graphs(graphs)
    orc = Oracle(cs)
    dev.init_equilibrium()
    orc.init_equilibrium()
    rep = dev.step(20)
    assert rep.bad_cells == 0 and rep.steps_done == 20
    orc.step(20)
    rep = dev.step(10)
    ro = orc.step(10)
    assert ro.steps_done == 6 and ro.first_bad_cell == 9
    assert rep.steps_done == ro.steps_done and rep.bad_cells == ro.bad_cells
    assert rep.first_bad_cell == ro.first_bad_cell
    # the state is the one after the bad step, buffers in the right parity
    assert max_rel(dev.get_state(), orc.get_state()) < TOL
    md, mo = dev.macros(), orc.macros()
    for k in ("rho", "u", "T", "q"):
        a, b = (md[k], mo[k]) if k in ("rho", "T") else (md[k][:, :2], mo[k][:, :2])  # 2D: D components
        assert max_rel(a, b) < TOL, k


def test_mask_repair_matches_meso(case_factory):
    """shak_tri6 with `mask switch True`: the GPU repairs the cell as meso's
    cell_value_interp does and runs on; state and macros against meso's own
    golden vectors (NaN heat fluxes of the repaired cells included)."""
    name = "shak_tri6_mask"
    g = load(name)
    cs = capi.Case(os.path.join(GOLDEN, name, "case.txt"))
    _say(f"[gpu] mask repair {name} NC={cs.n_cells} NV={cs.n_vel}")
    dev = capi.Device(cs)
    dev.set_mask()
    dev.init_equilibrium()
    done = 0
    masked = 0
    for s in [int(x) for x in g["steps"]]:
        rep = dev.step(s - done)
        done = s
        assert rep.bad_cells == 0
        masked += rep.masked_cells
        m = dev.macros()
        for k in ("rho", "u", "T", "q"):
            # the device keeps D components of u and q (meso's q_z is the NaN
            # sum 0 * NaN in the repaired cells of this 2D case)
            a, b = (m[k], g[f"step{s}_{k}"]) if k in ("rho", "T") else (m[k][:, :2], g[f"step{s}_{k}"][:, :2])
            err = max_rel(a, b)
            assert err < TOL, f"step {s} {k}: {err:.2e}"
        if s == 26:
            assert masked >= 1
            f = dev.get_state()
            assert max_rel(f[0], g["state26_g"]) < TOL
            assert max_rel(f[1], g["state26_h"]) < TOL


def test_mask_repair_fp32x_close_to_oracle(case_factory):
    """The repair writes the FP32X form (word and extension byte) of the
    equilibrium; one step after it the state is close to the oracle's."""
    name = "shak_tri6_mask"
    cs = capi.Case(os.path.join(GOLDEN, name, "case.txt"))
    dev = capi.Device(cs, capi.STORAGE_FP32X)
    orc = Oracle(cs)
    for x in (dev, orc):
        x.set_mask()
        x.init_equilibrium()
    orc.step(25)
    m = orc.macros()
    dev.set_state(orc.get_state(), m["rho"], m["u"], m["T"], m["q"])
    rep = dev.step(1)
    ro = orc.step(1)
    assert rep.masked_cells == ro.masked_cells >= 1
    md, mo = dev.macros(), orc.macros()
    for k in ("rho", "u", "T"):
        a, b = (md[k], mo[k]) if k != "u" else (md[k][:, :2], mo[k][:, :2])
        assert max_rel(a, b) < 1e-6, k
    f, fo = dev.get_state(), orc.get_state()
    assert max_rel(f[:, :, 9], fo[:, :, 9]) < 1e-6  # the repaired cell's rows


@pytest.mark.parametrize("name", ["shak_quad6_res", "bgk_quad8", "re400_cavity"])
def test_residual_matches_meso(name):
    """meso_hip_residual at meso's check steps against the residuals meso's
    do_step formed there (inf / NaN where it divides by 0)."""
    g = load(name)
    cs = capi.Case(os.path.join(GOLDEN, name, "case.txt"))
    _say(f"[gpu] residual {name}")
    dev = capi.Device(cs)
    dev.init_equilibrium()
    done = 0
    for s in range(100, int(max(g["steps"])) + 1, 100):
        dev.step(s - done)
        done = s
        r, ref = dev.residual(), g[f"step{s}_residual"]
        # at the first check the old velocity is the rest state: meso divides
        # by its round-off (~1e-17, giving ~1e15) where the device's exact
        # zero gives inf; both are "huge" (no limit can pass them)
        huge = lambda x: np.isinf(x) | (np.abs(np.nan_to_num(x)) > 1e10)
        assert np.array_equal(np.isnan(r), np.isnan(ref)), (s, r, ref)
        assert np.array_equal(huge(r), huge(ref)), (s, r, ref)
        fin = ~np.isnan(ref) & ~huge(ref)
        assert np.allclose(r[fin], ref[fin], rtol=1e-8, atol=0), (s, r, ref)
