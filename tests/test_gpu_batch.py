"""The batched sweep (meso_hip_step_batch) and the tile autotuner
(meso_hip_autotune) on seeded cavity instances (GPU, through the C ABI).

- A batch step gives, bit for bit, what stepping each instance alone gives
  (the same kernels on the same data, only issued together).
- Autotuning runs trial steps that are not committed: the state is unchanged,
  and the FP64 step afterwards matches the oracle to 1e-10 (the velocity chunk
  only changes the order of the partial sums).
"""
import numpy as np
import pytest

from meso_hip import capi, configs
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def cases(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("batch"))
    return [capi.Case(configs.build_instance(i, root, 0.4)) for i in configs.batch_instances(5, base_seed=2000)]


@pytest.mark.parametrize("storage", [capi.STORAGE_FP32, capi.STORAGE_FP64])
def test_batch_equals_one_by_one(cases, storage):
    _say(f"[gpu] batch of {len(cases)} storage={storage}:",
         ", ".join(f"{c.n_cells}x{c.n_vel}" for c in cases))
    alone, batch = [], []
    for cs in cases:
        for lst in (alone, batch):
            d = capi.Device(cs, storage)
            d.init_equilibrium()
            lst.append(d)
    for d in alone:
        d.step(12)
    reps = capi.step_batch(batch, 5)
    reps = capi.step_batch(batch, 7)
    assert all(r.steps_done == 7 and r.bad_cells == 0 for r in reps)
    for k, (a, b) in enumerate(zip(alone, batch)):
        ma, mb = a.macros(), b.macros()
        for key in ("rho", "u", "T", "q"):
            assert np.array_equal(ma[key], mb[key]), f"instance {k} {key}"
        ga, gb = a.get_state(), b.get_state()
        assert np.array_equal(ga, gb), f"instance {k} state"
    for d in alone + batch:
        d.close()


def test_autotune_keeps_state_and_parity(cases):
    cs = max(cases, key=lambda c: c.n_cells * c.n_vel)
    _say(f"[gpu] autotune NC={cs.n_cells} NV={cs.n_vel}")
    dev = capi.Device(cs, capi.STORAGE_FP64)
    orc = Oracle(cs)
    dev.init_equilibrium()
    orc.init_equilibrium()
    dev.step(2)
    orc.step(2)
    before = dev.get_state()
    m_before = dev.macros()
    t0 = dev.tiles()
    vc = dev.autotune(2)
    t1 = dev.tiles()
    _say(f"    tiles before {t0}\n    tiles after  {t1}")
    assert vc == 0 or (vc % 64 == 0 and vc > 0)
    assert t1["vchunk"] == vc and t1["patches"] == t0["patches"]
    if t1["patches"]:
        assert 1 <= t1["update_cells"] <= 32 and t1["moment_size"] >= 1
    assert np.array_equal(dev.get_state(), before)
    for key in ("rho", "u", "T", "q"):
        assert np.array_equal(dev.macros()[key], m_before[key])
    dev.step(3)
    orc.step(3)
    md, mo = dev.macros(), orc.macros()
    for key in ("rho", "u", "T", "q"):
        err = np.abs(md[key] - mo[key]).max() / max(np.abs(mo[key]).max(), 1e-300)
        assert err < 1e-10, f"{key}: {err:.2e}"
    dev.close()


@pytest.mark.parametrize("name,storage", [("shak_kn1_quad10", capi.STORAGE_FP32), ("shak_hex3", capi.STORAGE_FP64),
                                          ("shak_quad6_lim", capi.STORAGE_FP64)])
def test_autotuned_tiles_keep_parity(case_factory, name, storage):
    """After the autotuner rebuilt the patch sets (the tiles it chose), the
    step still matches the oracle: FP64 to 1e-10, FP32 to 2e-6 rel. L2."""
    cs = capi.Case(case_factory(name))
    dev = capi.Device(cs, storage)
    orc = Oracle(cs)
    dev.init_equilibrium()
    orc.init_equilibrium()
    dev.autotune(1)
    _say(f"[gpu] autotuned {name} storage={storage}: {dev.tiles()}")
    dev.step(6)
    orc.step(6)
    md, mo = dev.macros(), orc.macros()
    for key in ("rho", "u", "T", "q"):
        if storage == capi.STORAGE_FP64:
            err = np.abs(md[key] - mo[key]).max() / max(np.abs(mo[key]).max(), 1e-300)
            assert err < 1e-10, f"{key}: {err:.2e}"
        else:
            err = np.linalg.norm(md[key] - mo[key]) / max(np.linalg.norm(mo[key]), 1e-300)
            assert err < 2e-6, f"{key}: {err:.2e}"
    dev.close()


def test_tiles_without_patches(case_factory):
    """Velocity sets under 64 run the per-cell kernels: no tiles to report or
    tune (meso_hip_tiles patches = 0, meso_hip_autotune returns 0)."""
    cs = capi.Case(case_factory("bgk_quad8"))  # D2Q9
    dev = capi.Device(cs, capi.STORAGE_FP32)
    dev.init_equilibrium()
    t = dev.tiles()
    assert t["patches"] == 0 and t["vchunk"] == 0
    assert dev.autotune(1) == 0
    dev.close()
