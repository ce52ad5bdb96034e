"""meso_hip_init_equilibrium_noalloc against meso_hip_init_equilibrium, through
the C ABI (GPU).

The allocation-free entry point stages the per-cell macros in the library's
static 16 MiB device array instead of allocating them. Two calls in a row on
one GPU, the second on a smaller mesh (so that it reads a scratch array still
holding the first call's larger contents), must each give, bit for bit, the
state and macros of the allocating entry point, and the same steps after it.
A mesh beyond the array's 419,430-cell limit is rejected before any device
work.
"""
import numpy as np
import pytest

from meso_hip import capi
from meso_hip import meshgen as mg

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def perturbed_patch(cs, seed):
    rho, u, T = cs.initial_patch()
    rng = np.random.default_rng(seed)
    rho = rho * (1.0 + 0.02 * rng.standard_normal(rho.shape))
    u = u + 0.03 * rng.standard_normal(u.shape)
    u[:, 2] = 0.0 if cs.mesh.dim == 2 else u[:, 2]
    T = T * (1.0 + 0.02 * rng.standard_normal(T.shape))
    return rho, u, T


def test_noalloc_init_matches_the_allocating_entry_point(case_factory):
    for k, name in enumerate(("shak_kn1_quad10", "shak_quad6")):  # the second mesh is smaller
        cs = capi.Case(case_factory(name))
        _say(f"[gpu] noalloc call {k + 1}: {name} NC={cs.n_cells} NV={cs.n_vel}")
        rho, u, T = perturbed_patch(cs, k)
        ref = capi.Device(cs)
        ref.init_equilibrium(rho, u, T)
        dev = capi.Device(cs)
        dev.init_equilibrium_noalloc(rho, u, T)
        assert np.array_equal(dev.get_state(), ref.get_state())
        for d in (ref, dev):
            d.step(3)
        md, mr = dev.macros(), ref.macros()
        for key in ("rho", "u", "T", "q"):
            assert np.array_equal(md[key], mr[key]), f"{name} {key}"
        ref.close()
        dev.close()


def test_noalloc_init_fp32_storage(case_factory):
    cs = capi.Case(case_factory("shak_quad12_nc24"))
    rho, u, T = perturbed_patch(cs, 7)
    ref, dev = capi.Device(cs, capi.STORAGE_FP32), capi.Device(cs, capi.STORAGE_FP32)
    ref.init_equilibrium(rho, u, T)
    dev.init_equilibrium_noalloc(rho, u, T)
    assert np.array_equal(dev.get_state(), ref.get_state())


def test_noalloc_init_rejects_meshes_beyond_its_limit(tmp_path):
    # 650^2 = 422,500 cells > 419,430; D2Q9 keeps the context small
    case = mg.bgk_cavity_case(str(tmp_path / "big"), mg.quad_cavity(650))
    cs = capi.Case(case)
    assert cs.n_cells > 419430
    dev = capi.Device(cs)
    with pytest.raises(capi.MesoError) as e:
        dev.init_equilibrium_noalloc()
    assert e.value.code == capi.MESO_E_INVALID and "limit" in str(e.value)
    dev.init_equilibrium()  # the allocating entry point still takes it
