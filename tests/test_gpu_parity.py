"""GPU parity: the gfx950 step (through the C ABI) against the FP64 oracle on
the same seeded inputs.

Tolerances (relative, against the oracle):
  FP64 storage: max-norm 1e-10 on rho, u, T, q and flux moments after 10-20
  steps (summation order of the velocity reductions differs from meso's).
  FP32 storage: relative L2 <= 1e-6 on every macro component after 25-200
  steps (channels with a small cross-flow: 5x the FP32 storage-rounding floor
  of the same case, storage_floor). The run to meso's convergence (FP32 and
  FP32X against FP64 at 1e-6) is in tests/test_gpu_precision.py.
"""
import numpy as np
import pytest

from conftest import max_rel, rel_l2
from meso_hip import capi
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_TOL = 1e-6


def _say(*a):
    print(*a, flush=True)


def run_pair(case_path, steps, storage=capi.STORAGE_FP64):
    cs = capi.Case(case_path)
    _say(f"[gpu] {case_path} NC={cs.n_cells} NV={cs.n_vel} model={cs.model} storage={storage} steps={steps}")
    dev = capi.Device(cs, storage)
    orc = Oracle(cs)
    dev.init_equilibrium()
    orc.init_equilibrium()
    out = []
    done = 0
    for s in steps:
        dev.step(s - done)
        orc.step(s - done)
        done = s
        out.append((s, dev.macros(), orc.macros(), dev.flux_moments(), orc.flux_moments()))
    return cs, dev, orc, out


def keys_of(cs):
    return ["rho", "u", "T", "q"] if cs.model == capi.MODEL_CDUGKS_SHAKHOV else ["rho", "u"]


@pytest.mark.parametrize("name", ["bgk_quad8", "bgk_tri6", "dugks_quad8", "shak_quad6", "shak_tri6",
                                  "shak_hex3", "shak_channel_pressure", "shak_channel_farfield",
                                  "bgk_channel_farfield", "shak_sym_slip", "shak_quad6_lim", "shak_tri6_lim",
                                  "shak_hex3_lim", "shak_quad6_nc6_lim", "shak_sym_slip_lim", "neumann",
                                  "kn0075_cavity", "re400_cavity"])
def test_fp64_matches_oracle(case_factory, name):
    cs, dev, orc, out = run_pair(case_factory(name), [1, 2, 10])
    for s, md, mo, fd, fo in out:
        for k in keys_of(cs):
            err = max_rel(md[k], mo[k])
            assert err < FP64_TOL, f"{name} step {s} {k}: {err:.2e}"
        if cs.model != capi.MODEL_DUGKS_BGK:
            for a, b in zip(fd, fo):
                # flux moments vanish at rest: absolute floor of 1e-14
                assert np.allclose(a, b, rtol=1e-9, atol=1e-14), f"{name} step {s} flux moments"


# Steps per case: 200 (a developed flow) except the coarse triangle cavity,
# where meso's unlimited second-order reconstruction drives one corner cell to
# T < 0 after ~30 steps (the oracle and the reference do the same), and meso's
# four-state Riemann problem (2990 unstructured velocities, no axis tables), 50.
FP32_CASES = {"bgk_quad8": 200, "shak_quad6": 200, "shak_tri6": 25, "shak_hex3": 200, "shak_quad12_nc24": 200,
              "shak_quad6_lim": 200, "shak_hex3_lim": 200, "neumann": 50, "kn0075_cavity": 200,
              # the other boundary kinds in FP32 storage: farfield and pressure
              # inlets/outlets, symmetry and slip walls (mirror-boundary cells)
              "shak_channel_pressure": 200, "shak_channel_farfield": 200, "bgk_channel_farfield": 200,
              "shak_sym_slip": 200, "shak_sym_slip_lim": 200}
# The channels carry a small cross-flow V (and heat flux) against a large
# stream, so their small components see the FP32 rounding of the stored state
# at a larger relative size. Their tolerance is taken from the same case run
# on the CPU with that rounding alone (storage_floor: the FP64 oracle, the
# stored deviation from f_ref rounded to FP32 after every step): the GPU must
# stay within 5x that floor per component (at least FP32_TOL). On the symmetry
# and slip channel the floor is qy 3.0e-5 after 200 steps (the GPU: 3.5e-5);
# a wrong sign or term in a boundary gives 1e-3 or more. FP32X storage holds
# these cases to 1e-7 (tests/test_gpu_precision.py).
FLOOR_CASES = ("shak_channel_pressure", "shak_channel_farfield", "shak_sym_slip", "shak_sym_slip_lim")


def f_ref_of(cs, orc):
    """The reference equilibrium of the FP32 deviation storage: the rest
    Maxwellian pair (Shakhov) or the constant rho_ref (BGK) at the mean
    initial density and temperature, per velocity ([2, nv, 1])."""
    m = orc.macros()
    p = cs.params
    xi = cs.arrays()["xi"]
    rho = m["rho"].mean()
    if cs.model != capi.MODEL_CDUGKS_SHAKHOV:
        g = np.full(xi.shape[0], rho)
        return np.stack([g, np.zeros_like(g)])[:, :, None]
    RT = p.R * m["T"].mean()
    D = cs.mesh.dim
    A = rho * (0.5 / RT) / np.pi if D == 2 else rho / (2.0 * np.pi * RT) ** 1.5
    g = A * np.exp(-(0.5 / RT) * (xi ** 2).sum(1))
    return np.stack([g, (p.K + 3.0 - D) * RT * g])[:, :, None]


def storage_floor(cs, steps):
    """Relative L2 of every macro component when only the stored state is
    rounded to FP32 (deviation from f_ref, after every step), against the
    unrounded oracle run."""
    a, b = Oracle(cs), Oracle(cs)
    a.init_equilibrium()
    b.init_equilibrium()
    ref = f_ref_of(cs, b)
    for _ in range(steps):
        a.step(1)
        b.step(1)
        f = ((b.get_state() - ref).astype(np.float32).astype(np.float64)) + ref
        m = b.macros()
        b.set_state(f, m["rho"], m["u"], m["T"], m["q"])
    ma, mb = a.macros(), b.macros()
    return {k: [rel_l2(mb[k][:, i], ma[k][:, i]) for i in range(cs.mesh.dim)] if ma[k].ndim == 2
            else rel_l2(mb[k], ma[k]) for k in keys_of(cs)}


@pytest.mark.parametrize("name", list(FP32_CASES))
def test_fp32_storage_close_to_oracle(case_factory, name):
    """FP32 storage of the distributions against the FP64 oracle after many steps."""
    cs, dev, orc, out = run_pair(case_factory(name), [FP32_CASES[name]], storage=capi.STORAGE_FP32)
    s, md, mo, _, _ = out[-1]
    _say(f"[gpu] fp32 {name}: " + " ".join(
        f"{k}={rel_l2(md[k].ravel(), mo[k].ravel()):.2e}" for k in keys_of(cs)))
    floor = storage_floor(cs, FP32_CASES[name]) if name in FLOOR_CASES else None
    if floor is not None:
        _say(f"[gpu] fp32 {name} storage-rounding floor: {floor}")
    for k in keys_of(cs):
        if k == "q" or k == "u":
            for i in range(cs.mesh.dim):
                err = rel_l2(md[k][:, i], mo[k][:, i])
                tol = FP32_TOL if floor is None else max(FP32_TOL, 5.0 * floor[k][i])
                assert err < tol, f"{name} {k}[{i}]: {err:.2e} (tolerance {tol:.2e})"
        else:
            err = rel_l2(md[k], mo[k])
            tol = FP32_TOL if floor is None else max(FP32_TOL, 5.0 * floor[k])
            assert err < tol, f"{name} {k}: {err:.2e} (tolerance {tol:.2e})"


def test_state_roundtrip_and_restart(case_factory):
    cs = capi.Case(case_factory("shak_quad6"))
    dev = capi.Device(cs)
    orc = Oracle(cs)
    orc.init_equilibrium()
    orc.step(3)
    f = orc.get_state()
    m = orc.macros()
    dev.set_state(f, m["rho"], m["u"], m["T"], m["q"])
    back = dev.get_state()
    assert np.array_equal(back, f)
    fc = dev.get_state(capi.LAYOUT_CELL_MAJOR)
    assert np.array_equal(fc, np.transpose(f, (0, 2, 1)))
    dev.step(4)
    orc.step(4)
    md, mo = dev.macros(), orc.macros()
    for k in ["rho", "u", "T", "q"]:
        assert max_rel(md[k], mo[k]) < FP64_TOL


def test_init_from_macros_matches_oracle(case_factory):
    cs = capi.Case(case_factory("shak_tri6"))
    rng = np.random.default_rng(0)
    nc = cs.n_cells
    rho = 1.0 + 0.01 * rng.standard_normal(nc)
    u = 0.05 * rng.standard_normal((nc, 3))
    u[:, 2] = 0.0
    T = 1.0 + 0.01 * rng.standard_normal(nc)
    q = 0.01 * rng.standard_normal((nc, 3))
    q[:, 2] = 0.0
    dev, orc = capi.Device(cs), Oracle(cs)
    dev.init_from_macros(rho, u, T, q)
    orc.init_from_macros(rho, u, T, q)
    assert max_rel(dev.get_state(), orc.get_state()) < 1e-14
    dev.step(5)
    orc.step(5)
    md, mo = dev.macros(), orc.macros()
    for k in ["rho", "u", "T", "q"]:
        assert max_rel(md[k], mo[k]) < FP64_TOL


def test_residual_matches_oracle(case_factory):
    """Both sides start from the same developed state, so the stored previous
    copies agree bit for bit (a rest state has u == 0 and inf residuals)."""
    cs = capi.Case(case_factory("shak_quad6"))
    dev, orc = capi.Device(cs), Oracle(cs)
    orc.init_equilibrium()
    orc.step(20)
    m = orc.macros()
    dev.set_state(orc.get_state(), m["rho"], m["u"], m["T"], m["q"])
    orc.set_state(orc.get_state(), m["rho"], m["u"], m["T"], m["q"])
    dev.step(5)
    orc.step(5)
    rd, ro = dev.residual(), orc.residual()
    finite = np.isfinite(ro)
    assert np.array_equal(np.isfinite(rd), finite)
    assert np.allclose(rd[finite], ro[finite], rtol=1e-9, atol=0)


def test_live_param_change(case_factory):
    """dt / mu0 edits between steps (update_config() re-reads them every step)."""
    cs = capi.Case(case_factory("shak_quad6"))
    dev, orc = capi.Device(cs), Oracle(cs)
    dev.init_equilibrium()
    orc.init_equilibrium()
    dev.step(2)
    orc.step(2)
    p = capi.StepParams.from_buffer_copy(cs.params)
    p.dt *= 0.5
    p.mu0 *= 1.5
    dev.set_params(p)
    orc.set_params(p)
    dev.step(3)
    orc.step(3)
    md, mo = dev.macros(), orc.macros()
    for k in ["rho", "u", "T", "q"]:
        assert max_rel(md[k], mo[k]) < FP64_TOL


def test_mass_conserved_in_closed_cavity(case_factory):
    cs = capi.Case(case_factory("shak_quad12_nc24"))
    dev = capi.Device(cs)
    dev.init_equilibrium()
    vol = cs.arrays()["cell_vol"]
    m0 = float((dev.macros()["rho"] * vol).sum())
    dev.step(50)
    m1 = float((dev.macros()["rho"] * vol).sum())
    assert abs(m1 - m0) / m0 < 1e-12


@pytest.mark.parametrize("storage", [capi.STORAGE_FP32, capi.STORAGE_FP64])
@pytest.mark.parametrize("name", ["shak_kn1_quad10", "shak_hex3", "shak_quad6_lim"])
def test_moment_patch_kinds_agree(case_factory, monkeypatch, name, storage):
    """Face-based and cell-based moment patches (MESO_HIP_MOMENT_CELLS) give
    the same step: the same f-bar+ values, summed in another order (FP64
    moments), so the macros agree to 1e-12 (FP64 storage). FP32 storage rounds
    f-bar+ stochastically with random bits drawn per lane and tile, which the
    two patch layouts assign differently: there the two runs differ by that
    rounding noise, and they are held to the FP32 bound against the oracle
    (relative L2 <= 1e-6; measured 1e-8 to 1e-7 after 10 steps; a dropped
    moment term shows at 1e-3)."""
    cs = capi.Case(case_factory(name))
    _say(f"[gpu] moment patch kinds {name} NC={cs.n_cells} NV={cs.n_vel} storage={storage}")
    out = []
    for kind in ("0", "1"):
        monkeypatch.setenv("MESO_HIP_MOMENT_CELLS", kind)
        dev = capi.Device(cs, storage)
        dev.init_equilibrium()
        dev.step(10)
        out.append(dev.macros())
        dev.close()
    for k in keys_of(cs):
        if storage == capi.STORAGE_FP64:
            err, tol = max_rel(out[0][k], out[1][k]), 1e-12
        else:
            err, tol = rel_l2(out[0][k], out[1][k]), FP32_TOL
        assert err < tol, f"{name} {k}: {err:.2e}"
