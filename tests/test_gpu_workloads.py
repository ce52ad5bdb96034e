"""The benchmark workloads at their full sizes (GPU, through the C ABI).

Size-independent checks on every workload:
  - total mass sum(rho V) of the closed cavities is conserved to 1e-12
    (conservative update, telescoping face fluxes), for FP64 and FP32 storage;
  - FP32 storage against the FP64 step of the same GPU: relative L2 <= 1e-6
    on every macro component after N steps;
  - a cavity at rest (lid speed 0, uniform T) stays at rest.
Against the FP64 oracle at full size (1e-10 max-norm, as the small cases):
  the 64^2 D2Q9 cavity (20 steps), the 101^2 Delaunay cavity, the 200^2 quad
  cavity and the 32^3 hex cube (1-2 steps each).
Batched instances (seeded cavities of random size, Kn, lid speed and velocity
set): FP64 against the oracle after 2 steps, FP32 against FP64 after 50
(relative L2 <= 2e-6).
"""
import os

import numpy as np
import pytest

from conftest import rel_l2
from meso_hip import capi, configs
from meso_hip import meshgen as mg
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_TOL = 1e-6
MASS_TOL = 1e-12

# steps of the mass / FP32 checks, steps of the oracle check
WORKLOADS = {
    "cavity64_bgk": (200, [1, 20]),
    "cavity_kn0075_tri": (20, [1, 2]),
    "cavity_kn1_quad": (10, [1]),
    "cube_kn01_hex": (10, [1, 2]),
}


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def wroot(tmp_path_factory):
    return str(tmp_path_factory.mktemp("workloads"))


def field_errors(cs, a, b):
    """Max-norm errors of a against b, each field relative to its own maximum
    (vectors: the maximum over all components, so that a component that is
    zero to round-off, such as v after one step of a lid-driven cavity, is
    measured against the flow's scale)."""
    out = {}
    for k in ("rho", "u", "T", "q"):
        if k in ("T", "q") and cs.model != capi.MODEL_CDUGKS_SHAKHOV:
            continue
        x, y = np.asarray(a[k], dtype=np.float64), np.asarray(b[k], dtype=np.float64)
        if x.ndim == 2:
            x, y = x[:, :cs.mesh.dim], y[:, :cs.mesh.dim]
        out[k] = np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)
    return out


def components(cs, m):
    """Macro components compared: rho, u_d, (T, q_d) for Shakhov."""
    dim = cs.mesh.dim
    out = {"rho": m["rho"]}
    for d in range(dim):
        out[f"u{d}"] = m["u"][:, d]
    if cs.model == capi.MODEL_CDUGKS_SHAKHOV:
        out["T"] = m["T"]
        for d in range(dim):
            out[f"q{d}"] = m["q"][:, d]
    return out


def run_storage(cs, storage, steps):
    vol = cs.arrays()["cell_vol"]
    dev = capi.Device(cs, storage)
    dev.init_equilibrium()
    m0 = float((dev.macros()["rho"] * vol).sum())
    rep = dev.step(steps)
    m = dev.macros()
    dev.close()
    m1 = float((m["rho"] * vol).sum())
    return m, abs(m1 - m0) / m0, rep


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_workload_mass_and_fp32(wroot, name):
    steps = WORKLOADS[name][0]
    cs = capi.Case(configs.build(name, wroot))
    _say(f"[gpu] {name} NC={cs.n_cells} NV={cs.n_vel} steps={steps}")
    m64, dm64, _ = run_storage(cs, capi.STORAGE_FP64, steps)
    m32, dm32, _ = run_storage(cs, capi.STORAGE_FP32, steps)
    assert dm64 < MASS_TOL, f"{name} FP64 mass drift {dm64:.2e}"
    assert dm32 < MASS_TOL, f"{name} FP32 mass drift {dm32:.2e}"
    c64, c32 = components(cs, m64), components(cs, m32)
    for k in c64:
        assert np.all(np.isfinite(c32[k])), f"{name} {k} not finite"
        err = rel_l2(c32[k], c64[k])
        _say(f"    {name} FP32 vs FP64 after {steps} steps: {k} {err:.2e}")
        assert err < FP32_TOL, f"{name} {k}: {err:.2e}"


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_workload_fp64_matches_oracle(wroot, name):
    steps = WORKLOADS[name][1]
    cs = capi.Case(configs.build(name, wroot))
    _say(f"[gpu] {name} NC={cs.n_cells} NV={cs.n_vel} oracle steps={steps}")
    dev = capi.Device(cs, capi.STORAGE_FP64)
    orc = Oracle(cs)
    dev.init_equilibrium()
    orc.init_equilibrium()
    done = 0
    for s in steps:
        dev.step(s - done)
        orc.step(s - done)
        done = s
        for k, err in field_errors(cs, dev.macros(), orc.macros()).items():
            _say(f"    {name} FP64 vs oracle after {s} steps: {k} {err:.2e}")
            assert err < FP64_TOL, f"{name} step {s} {k}: {err:.2e}"
    dev.close()


def test_cavity_at_rest_stays_at_rest(wroot):
    """The 200^2 jittered quad cavity with the lid at rest stays at rest. Not
    to round-off: the discrete moments of the 101^2 Newton-Cotes Maxwellian
    miss rho = T = 1 by ~1e-11, so the walls (T = 1) drive a flow of that size,
    as in meso (the oracle on 50^2 gives |u| 9e-13, |rho - 1| 3e-12,
    |T - 1| 6e-11 after 5 steps)."""
    d = os.path.join(wroot, "rest_quad200")
    case = mg.shakhov_cavity_case(d, mg.quad_cavity(200, jitter=0.1, seed=3), ("nc", 1, 100, 5.0),
                                  mg.viscosity_from_kn(1.0), 0.0, name="rest")
    cs = capi.Case(case)
    for storage in (capi.STORAGE_FP64, capi.STORAGE_FP32):
        dev = capi.Device(cs, storage)
        dev.init_equilibrium()
        dev.step(5)
        m = dev.macros()
        dev.close()
        assert np.abs(m["u"]).max() < 1e-10, f"storage {storage}: |u| {np.abs(m['u']).max():.2e}"
        assert np.abs(m["rho"] - 1.0).max() < 1e-9
        assert np.abs(m["T"] - 1.0).max() < 1e-8


BATCH = configs.batch_instances(6)
# 50 steps on coarse velocity sets at low Kn: the small heat flux reaches 1.1e-6
# (seed 1004: 16^2 velocities, Kn 0.025); the bound of the small FP32 cases
FP32_TOL_BATCH = 2e-6


@pytest.mark.parametrize("inst", BATCH, ids=[f"seed{i['seed']}" for i in BATCH])
def test_batch_instance(wroot, inst):
    cs = capi.Case(configs.build_instance(inst, wroot))
    _say(f"[gpu] batch {inst} NC={cs.n_cells} NV={cs.n_vel}")
    dev = capi.Device(cs, capi.STORAGE_FP64)
    orc = Oracle(cs)
    dev.init_equilibrium()
    orc.init_equilibrium()
    dev.step(2)
    orc.step(2)
    for k, err in field_errors(cs, dev.macros(), orc.macros()).items():
        assert err < FP64_TOL, f"{inst['seed']} {k}: {err:.2e}"
    dev.step(48)
    m64 = components(cs, dev.macros())
    dev.close()
    m32, dm32, _ = run_storage(cs, capi.STORAGE_FP32, 50)
    assert dm32 < MASS_TOL
    m32 = components(cs, m32)
    for k in m64:
        err = rel_l2(m32[k], m64[k])
        _say(f"    batch {inst['seed']} FP32 vs FP64 after 50 steps: {k} {err:.2e}")
        assert err < FP32_TOL_BATCH, f"{inst['seed']} FP32 {k}: {err:.2e}"
