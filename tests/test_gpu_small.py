"""The small-velocity-set step (k_small_faces + k_small_cells: velocity sets of
at most 64, FP64 storage, CDUGKS models without the limiter) against the FP64
oracle, against the per-stage kernels of the same case, and through the NaN /
T < 0 monitor and the mask repair.

Tolerances: 1e-10 relative (max norm) against the oracle, as for every FP64
case; 1e-12 between the two GPU step paths (they differ only in the order of
the velocity and face sums); captured step pairs bit-identical to eager steps.
"""
import os

import numpy as np
import pytest

from conftest import max_rel
from meso_hip import capi
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10

# (case, lanes per face/cell, steps compared): 9 velocities (BGK, D2Q9), 16
# (Shakhov triangles: the unlimited scheme drives a cell bad at step 13, so
# this case stops at 10), 27 (3D Shakhov), 36 (Shakhov quads and the farfield
# and pressure channels)
CASES = [("bgk_quad8", 16, [1, 2, 10, 50]), ("bgk_tri6", 16, [1, 2, 10, 50]),
         ("bgk_channel_farfield", 16, [1, 2, 10, 50]), ("re400_cavity", 16, [1, 2, 10]),
         ("shak_tri6_nc4", 16, [1, 2, 10]), ("shak_hex3_gh3", 32, [1, 2, 10, 50]),
         ("shak_quad6_nc6", 64, [1, 2, 10, 50]), ("shak_channel_pressure_nc6", 64, [1, 2, 10, 50]),
         ("shak_channel_farfield_nc6", 64, [1, 2, 10, 50])]


def _say(*a):
    print(*a, flush=True)


def keys_of(cs):
    return ["rho", "u", "T", "q"] if cs.model == capi.MODEL_CDUGKS_SHAKHOV else ["rho", "u"]


def kernel_names(dev):
    dev.profile(True)
    dev.step(1)
    names = set(dev.kernel_times())
    dev.profile(False)
    return names


def device(cs, small=True, storage=capi.STORAGE_FP64):
    """A context with (small=True, the default) or without the small-set step
    (MESO_HIP_SMALL is read when the context is created)."""
    old = os.environ.get("MESO_HIP_SMALL")
    os.environ["MESO_HIP_SMALL"] = "1" if small else "0"
    try:
        return capi.Device(cs, storage)
    finally:
        if old is None:
            del os.environ["MESO_HIP_SMALL"]
        else:
            os.environ["MESO_HIP_SMALL"] = old


@pytest.mark.parametrize("name,lp,steps", CASES, ids=[c[0] for c in CASES])
def test_small_step_matches_oracle(case_factory, name, lp, steps):
    cs = capi.Case(case_factory(name))
    _say(f"[gpu] small step {name} NC={cs.n_cells} NV={cs.n_vel} lanes={lp}")
    probe = device(cs)
    probe.init_equilibrium()
    names = kernel_names(probe)
    assert {"small_faces", "small_cells"} <= names and "face_moments" not in names, names
    dev = device(cs)
    orc = Oracle(cs)
    dev.init_equilibrium()
    orc.init_equilibrium()
    done = 0
    for s in steps:
        rep = dev.step(s - done)
        orc.step(s - done)
        done = s
        assert rep.bad_cells == 0
        md, mo = dev.macros(), orc.macros()
        for k in keys_of(cs):
            err = max_rel(md[k], mo[k])
            assert err < TOL, f"{name} step {s} {k}: {err:.2e}"
        for a, b in zip(dev.flux_moments(), orc.flux_moments()):
            assert np.allclose(a, b, rtol=1e-9, atol=1e-14), f"{name} step {s} flux moments"
    assert max_rel(dev.get_state(), orc.get_state()) < TOL


@pytest.mark.parametrize("name", ["bgk_channel_farfield", "shak_hex3_gh3", "shak_channel_pressure_nc6"])
def test_small_step_matches_stage_kernels(case_factory, name):
    cs = capi.Case(case_factory(name))
    _say(f"[gpu] small vs stage kernels {name}")
    a, b = device(cs, True), device(cs, False)
    assert "small_faces" not in kernel_names(b)
    for d in (a, b):
        d.init_equilibrium()
        rep = d.step(40)
        assert rep.bad_cells == 0 and rep.steps_done == 40
    assert max_rel(a.get_state(), b.get_state()) < 1e-12
    ma, mb = a.macros(), b.macros()
    for k in keys_of(cs):
        assert max_rel(ma[k], mb[k]) < 1e-12, k


def test_small_graphs_bitwise(case_factory):
    """Captured step pairs replay the two small-set launches bit for bit, and
    the device step counter (advanced by the face kernel) stays in step."""
    cs = capi.Case(case_factory("shak_quad6_nc6"))
    a, b = device(cs), device(cs)
    for d in (a, b):
        d.init_equilibrium()
    b.set_graphs(True)
    for n in (1, 4, 3, 6):
        a.step(n)
        b.step(n)
        assert np.array_equal(a.get_state(), b.get_state()), n
        assert np.array_equal(a.residual(), b.residual(), equal_nan=True)


@pytest.mark.parametrize("graphs", [False, True])
def test_small_monitor_stops_at_the_bad_step(case_factory, graphs):
    """shak_tri6_nc4 (16 velocities) drives cell 9 to a NaN density at its 13th
    step in the oracle: the small-set step reports the same step and cell,
    skips the rest of the call, and returns the state after the bad step."""
    cs = capi.Case(case_factory("shak_tri6_nc4"))
    _say(f"[gpu] small monitor graphs={graphs}")
    dev = device(cs)
    dev.set_graphs(graphs)
    orc = Oracle(cs)
    dev.init_equilibrium()
    orc.init_equilibrium()
    rep = dev.step(8)
    orc.step(8)
    assert rep.bad_cells == 0 and rep.steps_done == 8
    rep = dev.step(10)
    ro = orc.step(10)
    assert ro.steps_done == 5 and ro.first_bad_cell == 9
    assert rep.steps_done == ro.steps_done and rep.bad_cells == ro.bad_cells
    assert rep.first_bad_cell == ro.first_bad_cell
    assert max_rel(dev.get_state(), orc.get_state()) < TOL


def test_small_mask_repair_matches_oracle(case_factory):
    """The same case with every cell masked: the bad cell is repaired (inverse
    distance average of its neighbours, Shakhov equilibrium) and the run goes
    on, as in the oracle."""
    cs = capi.Case(case_factory("shak_tri6_nc4"))
    dev = device(cs)
    orc = Oracle(cs)
    mask = np.ones(cs.n_cells, dtype=np.int32)
    dev.set_mask(mask)
    orc.set_mask(mask)
    dev.init_equilibrium()
    orc.init_equilibrium()
    rep = dev.step(16)
    ro = orc.step(16)
    assert ro.masked_cells > 0 and ro.bad_cells == 0
    assert rep.steps_done == 16 and rep.bad_cells == 0 and rep.masked_cells == ro.masked_cells
    md, mo = dev.macros(), orc.macros()
    for k in ("rho", "u", "T"):
        assert max_rel(md[k], mo[k]) < TOL, k
