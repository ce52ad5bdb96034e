"""The fused batch (meso_hip_batch_*): several contexts advanced with one
launch per stage (GPU, through the C ABI).

The fused launches run, for every context, the blocks its own launches run,
on the same data with the same arithmetic, so a fused step must give bit for
bit what meso_hip_step gives on each context alone: macros and the full
distribution state are compared with array_equal, for FP32, FP32X and FP64
storage, with and without the captured step pairs, across plan rebuilds
(a context stepped on its own between batch calls, a parameter change), with
contexts of several kernel instances and contexts that keep their own
launches (velocity sets under 64, mirror boundaries) in the same batch, and
through the NaN / T < 0 monitor. Building the plan sets the fused devices'
velocity chunks for the batch grid; the devices stepped alone are given the
same tiles (meso_hip_set_vchunk) before the comparison.
"""
import numpy as np
import pytest

from meso_hip import capi, configs

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def cases(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("fused"))
    return [capi.Case(configs.build_instance(i, root, 0.4)) for i in configs.batch_instances(5, base_seed=3000)]


def devices(cases, storage):
    out = []
    for cs in cases:
        st = storage(cs) if callable(storage) else storage
        d = capi.Device(cs, st)
        d.init_equilibrium()
        out.append(d)
    return out


def same_tiles(alone, fused):
    """The batch plan set the fused devices' velocity chunks for its grid:
    give the devices stepped alone the same tiles."""
    for a, f in zip(alone, fused):
        t = f.tiles()
        if t["patches"]:
            a.set_vchunk(t["vchunk"])


def assert_same(alone, fused, what):
    for k, (a, b) in enumerate(zip(alone, fused)):
        ma, mb = a.macros(), b.macros()
        for key in ("rho", "u", "T", "q"):
            assert np.array_equal(ma[key], mb[key], equal_nan=True), f"{what}: instance {k} {key}"
        assert np.array_equal(a.get_state(), b.get_state(), equal_nan=True), f"{what}: instance {k} state"


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("storage", [capi.STORAGE_FP32, capi.STORAGE_FP32E, capi.STORAGE_FP32X, capi.STORAGE_FP64])
def test_fused_batch_equals_one_by_one(cases, storage, graphs):
    _say(f"[gpu] fused batch storage={storage} graphs={graphs}:",
         ", ".join(f"{c.n_cells}x{c.n_vel}" for c in cases))
    alone, fused = devices(cases, storage), devices(cases, storage)
    b = capi.Batch(fused)
    b.set_graphs(graphs)
    nf, ng, nl = b.info()
    _say(f"    fused {nf} contexts in {ng} group(s), {nl} launches per step; velocity chunks "
         f"{[a.tiles()['vchunk'] for a in alone]} -> {[f.tiles()['vchunk'] for f in fused]}")
    assert nf == len(cases) and ng == min(4, len(cases))  # one kernel instance in each of 4 sub-batches
    # the batch grid takes fewer, larger velocity chunks than each case alone
    assert all(f.tiles()["vchunk"] >= a.tiles()["vchunk"] for a, f in zip(alone, fused))
    same_tiles(alone, fused)
    for d in alone:
        d.step(14)
    for n in (5, 7, 2):  # odd counts: graph pairs and a single eager step
        reps = b.step(n)
        assert all(r.steps_done == n and r.bad_cells == 0 for r in reps)
    assert_same(alone, fused, "fused")
    b.close()
    for d in alone + fused:
        d.close()


@pytest.mark.parametrize("streams", [1, 4])
def test_fused_batch_plan_follows_outside_changes(cases, streams):
    """A context stepped on its own, and a parameter change, between batch
    calls: the plan is rebuilt and the results stay those of single steps
    (one stream for the whole batch, or four sub-batches)."""
    alone, fused = devices(cases, capi.STORAGE_FP32), devices(cases, capi.STORAGE_FP32)
    b = capi.Batch(fused)
    b.set_streams(streams)
    b.set_graphs(True)
    b.info()
    same_tiles(alone, fused)
    b.step(4)
    fused[1].step(1)  # parity of one context moves outside the batch
    b.step(3)
    p = capi.StepParams.from_buffer_copy(cases[2].params)
    p.dt *= 0.9
    fused[2].set_params(p)
    b.step(4)
    for k, d in enumerate(alone):
        if k == 2:
            d.step(7)
            d.set_params(p)
            d.step(4)
        else:
            d.step(12 if k == 1 else 11)
    assert_same(alone, fused, "outside changes")
    b.close()
    for d in alone + fused:
        d.close()


def test_fused_batch_mixed_instances(cases, case_factory):
    """Two storages (two kernel instances) and contexts that keep their own
    launches (36 velocities: per-cell kernels; a symmetry/slip channel: mirror
    cells) in one batch."""
    extra = [capi.Case(case_factory("shak_quad6_nc6_lim")), capi.Case(case_factory("shak_sym_slip"))]
    allc = list(cases[:3]) + extra
    stor = [capi.STORAGE_FP32, capi.STORAGE_FP64, capi.STORAGE_FP32, capi.STORAGE_FP64, capi.STORAGE_FP32]
    it = iter(stor)
    alone = devices(allc, lambda cs: next(it))
    it = iter(stor)
    fused = devices(allc, lambda cs: next(it))
    b = capi.Batch(fused)
    nf, ng, nl = b.info()
    same_tiles(alone, fused)
    for d in alone:
        d.step(9)
    _say(f"[gpu] mixed batch: fused {nf} in {ng} groups, {nl} launches per step")
    assert nf == 3 and ng == 3  # three sub-batches of one context each (two kernel instances)
    b.step(4)
    b.set_graphs(True)
    b.step(5)
    assert_same(alone, fused, "mixed")
    b.close()
    for d in alone + fused:
        d.close()


def test_fused_batch_monitor(cases, case_factory):
    """shak_tri6 turns a cell bad at step 26: in a fused batch its report and
    state are those of meso_hip_step, and the healthy contexts run on."""
    allc = [capi.Case(case_factory("shak_tri6")), cases[0]]
    alone, fused = devices(allc, capi.STORAGE_FP64), devices(allc, capi.STORAGE_FP64)
    b = capi.Batch(fused)
    assert b.info()[0] == 2
    same_tiles(alone, fused)
    r_b = b.step(20)
    r_a = [d.step(20) for d in alone]
    assert all(r.bad_cells == 0 for r in r_a + r_b)
    r_b = b.step(10)
    r_a = [d.step(10) for d in alone]
    _say(f"[gpu] monitor in a batch: {[(r.steps_done, r.bad_cells, r.first_bad_cell) for r in r_b]}")
    for ra, rb in zip(r_a, r_b):
        assert (ra.steps_done, ra.bad_cells, ra.first_bad_cell) == (rb.steps_done, rb.bad_cells, rb.first_bad_cell)
    assert r_b[0].bad_cells > 0 and r_b[0].steps_done == 6 and r_b[1].steps_done == 10
    assert_same(alone, fused, "monitor")
    b.close()
    for d in alone + fused:
        d.close()


def test_fused_batch_keeps_own_tiles(cases):
    """With re-chunking off the plan leaves every device's tiles as they are:
    the same results as meso_hip_step with the default tiles."""
    alone, fused = devices(cases, capi.STORAGE_FP32), devices(cases, capi.STORAGE_FP32)
    b = capi.Batch(fused)
    b.set_rechunk(False)
    b.info()
    assert [a.tiles() for a in alone] == [f.tiles() for f in fused]
    for d in alone:
        d.step(6)
    b.step(6)
    assert_same(alone, fused, "own tiles")
    b.close()
    for d in alone + fused:
        d.close()
