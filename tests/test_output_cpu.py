"""The output writers (meso_case_write_output) against the files meso's own
output() wrote (tests/golden/out_*: meso compiled from its sources, run by
tests/golden/make_golden.py). CPU only.

The fixture values are fed back to the writer (text with 18 significant
digits parses back to the same doubles), so the comparison is byte for byte:
file names and layout, headers, number formatting and line breaks, the
Tecplot connectivity (quads, triangles-as-quads, bricks in Tecplot order) and
meso's quirks (the 3D grid header, the line counter of data.plt running on
across variables).
"""
import filecmp
import os

import numpy as np
import pytest

from meso_hip import capi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(d for d in os.listdir(GOLDEN) if d.startswith("out_") and os.path.isdir(os.path.join(GOLDEN, d, "out")))
STEPS = [0, 2]


def load_text(path, width):
    v = np.loadtxt(path, dtype=np.float64, ndmin=2 if width > 1 else 1)
    return v.reshape(-1, width) if width > 1 else v


def result_fields(path, nc, names):
    """Cell variables of a DUGKS result file (15 significant digits)."""
    vals, cur = {}, None
    for line in open(path):
        if line.startswith("## CellVar-"):
            cur = line.strip()[len("## CellVar-"):]
            vals[cur] = []
        elif line.startswith("## "):
            cur = None
        elif cur is not None:
            vals[cur] += [float(x) for x in line.split()]
    return [np.array(vals[n]) for n in names]


def tree(d):
    out = []
    for r, _, fs in os.walk(d):
        out += [os.path.relpath(os.path.join(r, f), d) for f in fs]
    return sorted(out)


@pytest.mark.parametrize("name", CASES)
def test_output_files_match_meso(name, tmp_path):
    fx = os.path.join(GOLDEN, name)
    cs = capi.Case(os.path.join(fx, "case.txt"))
    case_name = cs.setting_str("case-name", "unnamed")
    ref = os.path.join(fx, "out", case_name)
    nc, dim = cs.n_cells, cs.mesh.dim
    t, done = 0.0, 0
    for s in STEPS:
        for _ in range(s - done):
            t += cs.params.dt  # solution_time += dt each step (cdugks_shakhov.cpp:727)
        done = s
        if cs.model == capi.MODEL_DUGKS_BGK:
            names = ["Rho", "U", "V", "W"][:dim + 1]
            if s == STEPS[-1]:  # np-data holds the last output at full precision
                rho = load_text(os.path.join(ref, "np-data", "Rho"), 1)
                u = load_text(os.path.join(ref, "np-data", "vel"), 3)
            else:
                f = result_fields(os.path.join(ref, f"result-{s}.dat.plt"), nc, names)
                rho = f[0]
                u = np.zeros((nc, 3))
                for k in range(dim):
                    u[:, k] = f[1 + k]
            cs.write_output(tmp_path, s, t, rho, u)
        else:
            d = os.path.join(ref, f"step-{s}")
            rho, u = load_text(os.path.join(d, "Rho"), 1), load_text(os.path.join(d, "U"), 3)
            if cs.model == capi.MODEL_CDUGKS_SHAKHOV:
                T, q = load_text(os.path.join(d, "T"), 1), load_text(os.path.join(d, "q"), 3)
                # meso's flux fields are empty (files of 0 bytes) until the first step
                flux = None if s == 0 else (load_text(os.path.join(d, "fluxM0"), 1),
                                            load_text(os.path.join(d, "fluxM1"), 3),
                                            load_text(os.path.join(d, "fluxM2"), 1))
                cs.write_output(tmp_path, s, t, rho, u, T, q, flux)
            else:
                cs.write_output(tmp_path, s, t, rho, u)
    mine = os.path.join(tmp_path, case_name)
    assert tree(mine) == tree(ref)
    for f in tree(ref):
        assert filecmp.cmp(os.path.join(mine, f), os.path.join(ref, f), shallow=False), f


def test_output_needs_shakhov_fields(tmp_path):
    cs = capi.Case(os.path.join(GOLDEN, "out_shak_quad6", "case.txt"))
    nc = cs.n_cells
    with pytest.raises(capi.MesoError):
        cs.write_output(tmp_path, 0, 0.0, np.ones(nc), np.zeros((nc, 3)))
    with pytest.raises(capi.MesoError):  # the flux moments come all three or none
        cs.write_output(tmp_path, 1, 0.0, np.ones(nc), np.zeros((nc, 3)), np.ones(nc), np.zeros((nc, 3)),
                        (np.zeros(nc), None, np.zeros(nc)))


def test_runner_arguments(tmp_path):
    """bin/meso_hip_run rejects bad arguments (2) and an unreadable case (1)
    before it touches a GPU."""
    import subprocess
    run = os.path.join(os.path.dirname(GOLDEN), "..", "meso-lts_amd", "bin", "meso_hip_run")
    assert os.path.exists(run), "meso_hip_run is not built (make -C meso-lts_amd)"
    r = subprocess.run([run, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([run, "--case", str(tmp_path / "none.txt")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "meso_hip_run:" in r.stderr
