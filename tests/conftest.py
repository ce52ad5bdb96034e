import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "meso-lts_amd"))
sys.path.insert(0, ROOT)

from meso_hip import meshgen as mg  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def max_rel(a, b):
    """Largest |a - b| relative to max |b|. NaN entries must sit in the same
    places in both (meso's mask repair leaves NaN heat fluxes); inf otherwise."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return np.inf
    a, b = a[~nb], b[~nb]
    if a.size == 0:
        return 0.0
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="session")
def case_factory(tmp_path_factory):
    """Writes small seeded meso cases; returns their case-file paths by name."""
    root = tmp_path_factory.mktemp("cases")
    cache = {}

    def make(name):
        if name in cache:
            return cache[name]
        d = str(root / name)
        mu075 = mg.viscosity_from_kn(0.075)
        if name == "bgk_quad8":
            p = mg.bgk_cavity_case(d, mg.quad_cavity(8))
        elif name == "bgk_tri6":
            p = mg.bgk_cavity_case(d, mg.tri_cavity(6, seed=4))
        elif name == "dugks_quad8":
            p = mg.bgk_cavity_case(d, mg.quad_cavity(8), solver="dugks@incompressible")
        elif name == "bgk_hex3":
            p = mg.bgk_cavity_case(d, mg.hex_cavity(3, jitter=0.05, seed=1))
        elif name == "shak_quad6":
            p = mg.shakhov_cavity_case(d, mg.quad_cavity(6, jitter=0.1, seed=3), ("nc", 1, 7, 4.5), mu075, 0.2)
        elif name == "shak_quad6_lim":  # the golden cases with the Venkatakrishnan limiter
            p = mg.shakhov_cavity_case(d, mg.quad_cavity(6, jitter=0.1, seed=3), ("nc", 1, 7, 4.5), mu075, 0.2,
                                       limiter=True)
        elif name == "shak_tri6_lim":
            p = mg.shakhov_cavity_case(d, mg.tri_cavity(6, jitter=0.2, seed=1), ("nc", 2, 4, 4.5), mu075, 0.2,
                                       limiter=True)
        elif name == "shak_hex3_lim":
            p = mg.shakhov_cavity_case(d, mg.hex_cavity(3, jitter=0.05, seed=2), ("sparse-gh", 4, 0.5),
                                       mg.viscosity_from_kn(0.1), 0.1, limiter=True)
        elif name == "shak_quad6_nc6_lim":  # 36 velocities: the per-cell kernels
            p = mg.shakhov_cavity_case(d, mg.quad_cavity(6, jitter=0.1, seed=3), ("nc", 1, 5, 4.5), mu075, 0.2,
                                       limiter=True)
        elif name == "shak_quad12_nc24":
            p = mg.shakhov_cavity_case(d, mg.quad_cavity(12, jitter=0.1, seed=5), ("nc", 1, 23, 4.5), mu075, 0.2)
        elif name == "shak_tri6":
            p = mg.shakhov_cavity_case(d, mg.tri_cavity(6, jitter=0.2, seed=1), ("nc", 2, 4, 4.5), mu075, 0.2)
        elif name == "shak_hex3":
            p = mg.shakhov_cavity_case(d, mg.hex_cavity(3, jitter=0.05, seed=2), ("sparse-gh", 4, 0.5),
                                       mg.viscosity_from_kn(0.1), 0.1)
        elif name == "shak_quad6_nc6":  # 36 velocities: 64-lane segments of the small-set step
            p = mg.shakhov_cavity_case(d, mg.quad_cavity(6, jitter=0.1, seed=3), ("nc", 1, 5, 4.5), mu075, 0.2)
        elif name == "shak_tri6_nc4":  # 16 velocities: 16-lane segments
            p = mg.shakhov_cavity_case(d, mg.tri_cavity(6, jitter=0.2, seed=1), ("nc", 1, 3, 4.5), mu075, 0.2)
        elif name == "shak_hex3_gh3":  # 27 velocities in 3D: 32-lane segments
            p = mg.shakhov_cavity_case(d, mg.hex_cavity(3, jitter=0.05, seed=2), ("sparse-gh", 3, 0.5),
                                       mg.viscosity_from_kn(0.1), 0.1)
        elif name == "shak_kn1_quad10":
            p = mg.shakhov_cavity_case(d, mg.quad_cavity(10, jitter=0.1, seed=7), ("nc", 1, 16, 5.0),
                                       mg.viscosity_from_kn(1.0), 0.2)
        elif name in ("shak_channel_pressure", "shak_channel_farfield", "shak_sym_slip", "bgk_channel_farfield",
                      "shak_sym_slip_lim", "shak_channel_pressure_nc6", "shak_channel_farfield_nc6",
                      "shak_sym_slip_nc6"):
            p = channel_case(d, name)
        elif name in ("neumann", "kn0075_cavity", "re400_cavity"):
            # meso's own example cases with its shipped meshes and velocity sets
            p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name, "case.txt")
        else:
            raise KeyError(name)
        cache[name] = p
        return p

    return make


def split_sides(m, nx, ny):
    """Re-split the boundary sets of a quad_cavity(nx, ny) mesh into left/right/bottom/top."""
    sides = {"left": [], "right": [], "bottom": [], "top": []}
    for j in range(ny):
        for i in range(nx):
            c = j * nx + i
            if j == 0:
                sides["bottom"].append((c, 0))
            if i == nx - 1:
                sides["right"].append((c, 1))
            if j == ny - 1:
                sides["top"].append((c, 2))
            if i == 0:
                sides["left"].append((c, 3))
    m.bsets = sides
    return m


def channel_case(d, name):
    """8x4 channel: left/right inlet+outlet, bottom wall, top wall / symmetry / slip
    (name suffix _lim: with the Venkatakrishnan limiter; _nc6: 6x6 Newton-Cotes
    velocities instead of 12x12)."""
    limiter = name.endswith("_lim")
    name = name[:-4] if limiter else name
    small = name.endswith("_nc6")
    name = name[:-4] if small else name
    os.makedirs(d, exist_ok=True)
    m = split_sides(mg.quad_cavity(8, 4, jitter=0.05, seed=11), 8, 4)
    mg.write_neu(m, os.path.join(d, "mesh.neu"))
    if name.startswith("bgk"):
        mg.write_gauss_hermite(os.path.join(d, "dvs.dvs"), 3)
        settings = {"solver": "cdugks@incompressible", "mesh-file": "mesh.neu", "dvs-file": "dvs.dvs",
                    "dvs-type": "Gauss-Hermite", "Re": "100.0", "Ma": "0.1", "timeStep": "2.0e-3",
                    "gas-constant": "0.5", "ref-temperature": "1.0", "gradient-switch": "True"}
        groups = [{"name": "fluid", "density": "fixedValue scalar 1.0"}]
        marks = [{"name": "left", "type": "farfield-inlet", "density": "fixedValue scalar 1.01",
                  "velocity": "fixedValue vector 0.05 0.0 0.0"},
                 {"name": "right", "type": "farfield-outlet"},
                 {"name": "bottom", "type": "wall"}, {"name": "top", "type": "wall"}]
        mg.write_case(os.path.join(d, "case.txt"), settings, groups, marks)
        return os.path.join(d, "case.txt")
    mg.write_newton_cotes(os.path.join(d, "dvs.dvs"), 1, 5 if small else 11, 4.5)
    settings = {"solver": "cdugks@shakhov", "mesh-file": "mesh.neu", "dvs-file": "dvs.dvs", "dvs-type": "Newton-Cotes",
                "Pr": repr(2.0 / 3.0), "timeStep": "1.0e-3", "gas-constant": "0.5", "gas-k": "0",
                "ref-temperature": "1.0", "ref-viscosity": repr(mg.viscosity_from_kn(0.1)), "vhs-index": "0.5",
                "gradient-switch": "True", "limiter-switch": "True" if limiter else "False"}
    groups = [{"name": "fluid", "density": "fixedValue scalar 1.0", "temperature": "fixedValue scalar 1.0",
               "velocity": "fixedValue vector 0.05 0.0 0.0"}]
    wall = {"type": "wall", "velocity": "fixedValue vector 0.0 0.0 0.0", "temperature": "fixedValue scalar 1.0"}
    if name == "shak_channel_pressure":
        marks = [{"name": "left", "type": "pressure-inlet", "pressure": "fixedValue scalar 0.52",
                  "temperature": "fixedValue scalar 1.0"},
                 {"name": "right", "type": "pressure-outlet", "pressure": "fixedValue scalar 0.5",
                  "temperature": "fixedValue scalar 1.0"},
                 dict(name="bottom", **wall),
                 {"name": "top", "type": "wall", "velocity": "fixedValue vector 0.0 0.0 0.0",
                  "temperature": "fixedValue scalar 0.98"}]
    elif name == "shak_channel_farfield":
        marks = [{"name": "left", "type": "farfield-inlet", "density": "fixedValue scalar 1.0",
                  "velocity": "fixedValue vector 0.1 0.0 0.0", "temperature": "fixedValue scalar 1.05"},
                 {"name": "right", "type": "farfield-outlet"},
                 dict(name="bottom", **wall), dict(name="top", **wall)]
    else:  # symmetry + slip-wall
        marks = [{"name": "left", "type": "farfield-inlet", "density": "fixedValue scalar 1.0",
                  "velocity": "fixedValue vector 0.1 0.0 0.0", "temperature": "fixedValue scalar 1.0"},
                 {"name": "right", "type": "farfield-outlet"},
                 {"name": "bottom", "type": "slip-wall", "velocity": "fixedValue vector 0.0 0.0 0.0",
                  "temperature": "fixedValue scalar 1.0", "direction": "fixedValue objectType 1",
                  "slip-wall-alpha-u": "fixedValue scalar 0.3", "slip-wall-alpha-T": "fixedValue scalar 0.6"},
                 {"name": "top", "type": "symmetry", "direction": "fixedValue objectType 1"}]
    mg.write_case(os.path.join(d, "case.txt"), settings, groups, marks)
    return os.path.join(d, "case.txt")


def gpu_available():
    try:
        import torch  # noqa: F401
        return torch.cuda.device_count() > 0
    except Exception:
        return False
