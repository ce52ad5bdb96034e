"""Cell renumbering of gambit files (meso_mesh_renumber) and the speed of the
case reader, on the CPU.

A renumbered file is the same mesh: every cell keeps its volume and centre
under the returned permutation, every boundary set keeps its faces, and the
oracle (pinned to meso) steps it to the same macros up to the order of its sums
(1e-12). The locality order (recursive coordinate bisection of the centres)
puts neighbouring cells at close ids, a random order does not.
"""
import os
import shutil
import subprocess
import time

import numpy as np
import pytest

from meso_hip import capi
from meso_hip import meshgen as mg
from oracle.oracle import Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESO_TIMER = os.path.join(ROOT, "oracle", "_ref", "meso_mesh_timer")

METHODS = [capi.RENUMBER_RCB, capi.RENUMBER_RANDOM, capi.RENUMBER_KEEP]


def cases(root):
    mu = mg.viscosity_from_kn(0.075)
    return {
        "shak_tri": mg.shakhov_cavity_case(os.path.join(root, "tri"), mg.tri_cavity(6, jitter=0.2, seed=1),
                                           ("nc", 1, 5, 4.5), mu, 0.2),
        "shak_hex": mg.shakhov_cavity_case(os.path.join(root, "hex"), mg.hex_cavity(3, jitter=0.05, seed=2),
                                           ("sparse-gh", 3, 0.5), mg.viscosity_from_kn(0.1), 0.1),
        "bgk_quad": mg.bgk_cavity_case(os.path.join(root, "quad"), mg.quad_cavity(8, jitter=0.1, seed=3)),
    }


def mesh_file(case_path):
    cs = capi.Case(case_path)
    name = cs.setting_str("mesh-file")
    return os.path.join(os.path.dirname(case_path), name)


def renumbered(case_path, root, method, seed=7):
    d = os.path.join(root, f"{os.path.basename(os.path.dirname(case_path))}_m{method}")
    shutil.rmtree(d, ignore_errors=True)
    shutil.copytree(os.path.dirname(case_path), d)
    src = mesh_file(case_path)
    perm = capi.renumber_mesh(src, os.path.join(d, os.path.basename(src)), method, seed)
    return os.path.join(d, os.path.basename(case_path)), perm


@pytest.mark.parametrize("method", METHODS)
def test_renumbered_mesh_is_the_same_mesh(tmp_path, method):
    for name, p in cases(str(tmp_path)).items():
        q, perm = renumbered(p, str(tmp_path), method)
        a, b = capi.Case(p).arrays(), capi.Case(q).arrays()
        nc = len(a["cell_vol"])
        assert np.array_equal(np.sort(perm), np.arange(nc)), name
        if method == capi.RENUMBER_KEEP:
            assert np.array_equal(perm, np.arange(nc))
            for k in ("cell_pos", "face_cells", "face_area", "nbr_ids", "lsq_C"):
                assert np.array_equal(a[k], b[k]), (name, k)
        assert np.array_equal(b["cell_vol"][perm], a["cell_vol"]), name
        assert np.array_equal(b["cell_pos"][perm], a["cell_pos"]), name
        assert len(b["face_area"]) == len(a["face_area"]), name
        assert np.isclose(b["total_volume"], a["total_volume"], rtol=1e-14, atol=0.0)
        # every boundary set keeps its faces (as owner cell and face centre)
        for mark in range(len(a["marks"])):
            fa = np.flatnonzero(a["face_mark"] == mark)
            fb = np.flatnonzero(b["face_mark"] == mark)
            assert len(fa) == len(fb), (name, mark)
            if mark > 0:
                ka = sorted((int(perm[a["face_cells"][f, 0]]), *np.round(a["face_pos"][f], 12)) for f in fa)
                kb = sorted((int(b["face_cells"][f, 0]), *np.round(b["face_pos"][f], 12)) for f in fb)
                assert ka == kb, (name, mark)


@pytest.mark.parametrize("method", [capi.RENUMBER_RCB, capi.RENUMBER_RANDOM])
def test_renumbered_mesh_steps_the_same(tmp_path, method):
    """2D meshes and planar hexes. (A warped hex face is not planar, and meso
    takes its normal from three of its nodes in the order of the cell that owns
    it, the first cell that meets it (src/mesh/object.cpp:38-84): renumbering
    changes that owner and so the face's geometry, for meso as here.)"""
    cs = cases(str(tmp_path))
    cs["shak_hex"] = mg.shakhov_cavity_case(str(tmp_path / "hex0"), mg.hex_cavity(3), ("sparse-gh", 3, 0.5),
                                            mg.viscosity_from_kn(0.1), 0.1)
    for name, p in cs.items():
        q, perm = renumbered(p, str(tmp_path), method)
        o1, o2 = Oracle(capi.Case(p)), Oracle(capi.Case(q))
        o1.init_equilibrium()
        o2.init_equilibrium()
        o1.step(10)
        o2.step(10)
        m1, m2 = o1.macros(), o2.macros()
        for k in m1:
            scale = max(np.abs(m1[k]).max(), 1e-300)
            assert np.abs(m2[k][perm] - m1[k]).max() / scale < 1e-12, (name, k)


def test_meso_cavity_mesh_renumbers(tmp_path):
    """meso's own files/cavity.neu (1928 cells from Pointwise)."""
    src = os.path.join(GOLDEN, "cavity_mesh", "cavity.neu")
    out = str(tmp_path / "cavity_rcb.neu")
    perm = capi.renumber_mesh(src, out, capi.RENUMBER_RCB)
    assert np.array_equal(np.sort(perm), np.arange(len(perm)))
    with open(src) as fa, open(out) as fb:
        la, lb = fa.read().split("\n"), fb.read().split("\n")
    # the node section is copied as it is
    i0 = next(i for i, x in enumerate(la) if "NODAL COORDINATES" in x)
    i1 = next(i for i in range(i0, len(la)) if la[i].strip() == "ENDOFSECTION")
    j0 = next(i for i, x in enumerate(lb) if "NODAL COORDINATES" in x)
    assert la[i0:i1 + 1] == lb[j0:j0 + i1 - i0 + 1]


def test_rcb_order_is_local(tmp_path):
    """Mean |id gap| over LSQ neighbour pairs: the locality order keeps it near
    the neighbours of a structured numbering, a random order spreads it over
    the mesh."""
    p = mg.bgk_cavity_case(str(tmp_path / "q"), mg.quad_cavity(40, jitter=0.1, seed=5))
    gaps = {}
    for method in (capi.RENUMBER_RANDOM, capi.RENUMBER_RCB):
        q, perm = renumbered(p, str(tmp_path), method)
        a = capi.Case(q).arrays()
        ptr, ids = a["nbr_ptr"], a["nbr_ids"]
        rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
        gaps[method] = np.abs(ids - rows).mean()
    assert gaps[capi.RENUMBER_RCB] < gaps[capi.RENUMBER_RANDOM] / 20, gaps


def test_renumber_rejects_bad_input(tmp_path):
    bad = tmp_path / "bad.neu"
    bad.write_text("NUMNP NELEM\n 1 1 1 0 2 2\n")
    with pytest.raises(capi.MesoError):
        capi.renumber_mesh(str(bad), str(tmp_path / "out.neu"))
    with pytest.raises(capi.MesoError):
        capi.renumber_mesh(str(tmp_path / "missing.neu"), str(tmp_path / "out.neu"))


@pytest.mark.skipif(not os.path.exists(MESO_TIMER), reason="meso's reader timer is not built (oracle/Makefile.ref)")
def test_case_reader_is_faster_than_meso(tmp_path):
    """A 300^2 jittered quad mesh (90,000 cells, 10.6 MB): meso_case_open
    (parse, faces, geometry, least squares, the descriptor arrays) against
    meso's own load_gambit (parse, faces, geometry, least squares)."""
    p = mg.bgk_cavity_case(str(tmp_path / "big"), mg.quad_cavity(300, jitter=0.1, seed=3))
    t = time.perf_counter()
    cs = capi.Case(p)
    ours = time.perf_counter() - t
    assert cs.n_cells == 90000
    out = subprocess.run([MESO_TIMER, mesh_file(p)], capture_output=True, text=True, check=True,
                         cwd=os.path.dirname(p)).stdout.split()
    meso = float(out[out.index("seconds") + 1])
    print(f"case reader: {ours:.3f} s, meso load_gambit: {meso:.3f} s")
    assert ours < meso
