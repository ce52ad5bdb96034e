"""bin/meso_hip_run (meso's run of a case: initial, output, the time loop,
output every write-interval steps and at the end) on the cases whose output
meso itself wrote (tests/golden/out_*, steps 0 and 2).

The run is FP64; its files must have meso's names, layout and text, with every
number within 1e-9 of meso's (relative to the largest magnitude in the file,
at least 1e-6: the GPU sums in another order). grid.plt holds geometry only and must match
byte for byte. A resume (read-np-data True, --not-run) writes step 0 from the
latest/ files it read, byte for byte.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from test_output_cpu import GOLDEN, CASES, tree

pytestmark = pytest.mark.gpu

RUN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "meso-lts_amd", "bin", "meso_hip_run")
TOL = 1e-9


def set_keys(case, **kv):
    s = open(case).read()
    for k, v in kv.items():
        key = k.replace("_", "-")
        pat = re.compile(rf"^{re.escape(key)}\s+.*$", re.M)
        line = f"{key:<20} {v}"
        s = pat.sub(line, s, count=1) if pat.search(s) else s.replace("[settings]\n", f"[settings]\n{line}\n", 1)
    open(case, "w").write(s)


def split(path):
    """Words that are not numbers (kept as text), and the numbers."""
    words, nums = [], []
    for tok in open(path).read().split():
        try:
            nums.append(float(tok))
        except ValueError:
            words.append(tok)
    return words, np.array(nums)


def run(args, cwd):
    print("[gpu] meso_hip_run", " ".join(args), flush=True)
    r = subprocess.run([RUN] + args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(RUN):
        pytest.fail(f"{RUN} is not built (make -C meso-lts_amd)")


@pytest.mark.parametrize("name", CASES)
def test_runner_matches_meso_output(name, tmp_path):
    fx = os.path.join(GOLDEN, name)
    for f in os.listdir(fx):
        if os.path.isfile(os.path.join(fx, f)):
            shutil.copy(os.path.join(fx, f), tmp_path / f)
    case = str(tmp_path / "case.txt")
    set_keys(case, max_step=2, write_interval=2)
    run(["--case", case, "--out", str(tmp_path)], tmp_path)
    case_name = [l.split()[1] for l in open(case) if l.startswith("case-name")][0]
    ref, mine = os.path.join(fx, "out", case_name), str(tmp_path / case_name)
    assert tree(mine) == tree(ref)
    for f in tree(ref):
        a, b = os.path.join(mine, f), os.path.join(ref, f)
        if f.endswith("grid.plt"):
            assert open(a).read() == open(b).read(), f
            continue
        wa, na = split(a)
        wb, nb = split(b)
        assert wa == wb, f
        assert na.shape == nb.shape, f
        if nb.size:
            # a field that is zero to round-off (u at step 0: |u| ~ 1e-17 from the
            # discrete moments of the equilibrium) is measured against 1e-6
            err = np.abs(na - nb).max() / max(np.abs(nb).max(), 1e-6)
            assert err < TOL, f"{f}: {err:.2e}"

    if not os.path.isdir(os.path.join(mine, "latest")):
        return  # DUGKS resumes from np-data/*.np.dat, which its output does not write
    saved = str(tmp_path / "saved_latest")
    shutil.copytree(os.path.join(mine, "latest"), saved)
    set_keys(case, read_np_data="True", output_np="True")
    run(["--case", case, "--out", str(tmp_path), "--not-run"], tmp_path)
    for f in os.listdir(saved):
        text = open(os.path.join(mine, "step-0", f)).read()
        if f.startswith("flux"):
            assert text == "", f  # meso's flux fields are empty before the first step
        else:
            assert text == open(os.path.join(saved, f)).read(), f


@pytest.mark.parametrize("storage,tol", [("fp32", 1e-6), ("fp32e", 1e-6), ("fp32x", 1e-7)])
def test_runner_reduced_storages(storage, tol, tmp_path):
    """--storage fp32 / fp32e / fp32x on meso's Shakhov quad case (64 velocities): the
    same files as meso's, every number within the storage's bound of them."""
    fx = os.path.join(GOLDEN, "out_shak_quad6")
    for f in os.listdir(fx):
        if os.path.isfile(os.path.join(fx, f)):
            shutil.copy(os.path.join(fx, f), tmp_path / f)
    case = str(tmp_path / "case.txt")
    set_keys(case, max_step=2, write_interval=2)
    run(["--case", case, "--out", str(tmp_path), "--storage", storage], tmp_path)
    case_name = [l.split()[1] for l in open(case) if l.startswith("case-name")][0]
    ref, mine = os.path.join(fx, "out", case_name), str(tmp_path / case_name)
    assert tree(mine) == tree(ref)
    for f in tree(ref):
        if f.endswith("grid.plt"):
            continue
        wa, na = split(os.path.join(mine, f))
        wb, nb = split(os.path.join(ref, f))
        assert wa == wb and na.shape == nb.shape, f
        if nb.size:
            err = np.abs(na - nb).max() / max(np.abs(nb).max(), 1e-6)
            assert err < tol, f"{storage} {f}: {err:.2e}"
