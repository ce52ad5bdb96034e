"""Host front-end (case file, gambit mesh, DVS) against closed forms. CPU only.

Each check is chosen so that a plausible slip in the geometry or the DVS
(a transposed index, a lost sign, a wrong weight) breaks it: exact partitions
of the unit square/cube, closure of every cell, outward normals, exactness of
the least-squares gradient on linear fields, numpy's weighted lstsq on random
fields, quadrature identities of the velocity sets and their mirror maps.
"""
import os

import numpy as np
import pytest

from conftest import split_sides
from meso_hip import capi
from meso_hip import meshgen as mg
from oracle.oracle import Oracle


def open_case(tmp_path, mesh, dvs=("nc", 1, 6, 4.0)):
    d = str(tmp_path)
    p = mg.shakhov_cavity_case(d, mesh, dvs, mg.viscosity_from_kn(0.1), 0.1)
    return capi.Case(p)


MESHES = {
    "quad_regular": lambda: mg.quad_cavity(5),
    "quad_jitter": lambda: mg.quad_cavity(6, jitter=0.2, seed=2),
    "tri": lambda: mg.tri_cavity(5, jitter=0.2, seed=3),
    "hex_regular": lambda: mg.hex_cavity(3),
    "hex_jitter": lambda: mg.hex_cavity(3, jitter=0.1, seed=4),
}


def sparse_dvs(name):
    return ("sparse-gh", 3, 0.5) if name.startswith("hex") else ("nc", 1, 6, 4.0)


@pytest.mark.parametrize("name", list(MESHES))
def test_volumes_partition_the_domain(tmp_path, name):
    cs = open_case(tmp_path, MESHES[name](), sparse_dvs(name))
    A = cs.arrays()
    assert np.all(A["cell_vol"] > 0)
    # meso's hexahedron volume is exact for planar faces only (warped faces of
    # a jittered hex mesh leave a small gap between neighbours)
    tol = 1e-3 if name == "hex_jitter" else 1e-13
    assert abs(A["cell_vol"].sum() - 1.0) < tol
    assert abs(A["total_volume"] - A["cell_vol"].sum()) < 1e-13


@pytest.mark.parametrize("name", list(MESHES))
def test_normals_unit_and_outward(tmp_path, name):
    cs = open_case(tmp_path, MESHES[name](), sparse_dvs(name))
    A = cs.arrays()
    n = A["face_normal"]
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-14)
    own = A["face_cells"][:, 0]
    assert np.all(np.einsum("ij,ij->i", A["face_pos"] - A["cell_pos"][own], n) > 0)
    nei = A["face_cells"][:, 1]
    inner = nei != own
    # the neighbour lies on the other side of every interior face
    assert np.all(np.einsum("ij,ij->i", A["cell_pos"][nei[inner]] - A["face_pos"][inner], n[inner]) > 0)


@pytest.mark.parametrize("name", ["quad_regular", "quad_jitter", "tri", "hex_regular"])
def test_cells_are_closed(tmp_path, name):
    """sum over a cell's faces of +-A n vanishes for planar faces (not for warped hexes)."""
    cs = open_case(tmp_path, MESHES[name](), sparse_dvs(name))
    A = cs.arrays()
    S = np.zeros((cs.n_cells, 3))
    for c in range(cs.n_cells):
        for f in A["cell_face_ids"][A["cell_face_ptr"][c]:A["cell_face_ptr"][c + 1]]:
            sgn = 1.0 if A["face_cells"][f, 0] == c else -1.0
            S[c] += sgn * A["face_area"][f] * A["face_normal"][f]
    assert np.abs(S).max() < 1e-14


def test_regular_quad_grid_closed_form(tmp_path):
    n = 5
    cs = open_case(tmp_path, mg.quad_cavity(n))
    A = cs.arrays()
    h = 1.0 / n
    assert np.allclose(A["cell_vol"], h * h, rtol=1e-13)
    assert np.allclose(A["face_area"], h, rtol=1e-13)
    assert cs.n_faces == 2 * n * (n + 1)
    # equivalent diameter of the smallest cell: a disc of the same area (src/mesh/object.cpp:118-119)
    assert abs(A["min_cell_size"] - np.sqrt(4.0 * h * h / np.pi)) < 1e-13
    ij = np.stack(np.meshgrid(np.arange(n), np.arange(n), indexing="xy"), -1).reshape(-1, 2)
    assert np.allclose(A["cell_pos"][:, :2], (ij + 0.5) * h, atol=1e-15)
    # every normal is along an axis (meso's angle form leaves cos(pi/2) ~ 6e-17 off-axis)
    n0 = np.abs(A["face_normal"][:, :2])
    assert np.all(np.minimum(n0[:, 0], n0[:, 1]) < 1e-15)


@pytest.mark.parametrize("name", list(MESHES))
def test_lsq_gradient_exact_on_linear_fields(tmp_path, name):
    cs = open_case(tmp_path, MESHES[name](), sparse_dvs(name))
    orc = Oracle(cs)
    x = cs.arrays()["cell_pos"]
    b = np.array([2.0, -5.0, 0.0 if cs.mesh.dim == 2 else 7.0])
    g = orc.grad(3.0 + x @ b)
    assert np.abs(g - b).max() < 1e-11


@pytest.mark.parametrize("name", list(MESHES))
def test_lsq_gradient_matches_weighted_lstsq(tmp_path, name):
    """grad = argmin sum_j w_j (f_j - f_c - g.dr_j)^2 with w_j = 1/|dr_j|^2
    (include/mesh/mesh.h LeastSquare; mesh_numerical.cpp:24-49)."""
    cs = open_case(tmp_path, MESHES[name](), sparse_dvs(name))
    orc = Oracle(cs)
    A = cs.arrays()
    D = cs.mesh.dim
    f = np.random.default_rng(1).standard_normal(cs.n_cells)
    g = orc.grad(f)
    for c in range(cs.n_cells):
        nb = A["nbr_ids"][A["nbr_ptr"][c]:A["nbr_ptr"][c + 1]]
        dr = A["cell_pos"][nb, :D] - A["cell_pos"][c, :D]
        sw = 1.0 / np.linalg.norm(dr, axis=1)
        ref = np.linalg.lstsq(dr * sw[:, None], (f[nb] - f[c]) * sw, rcond=None)[0]
        assert np.allclose(g[c, :D], ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_lsq_neighbours_are_face_neighbours(tmp_path):
    cs = open_case(tmp_path, mg.quad_cavity(4))
    A = cs.arrays()
    counts = np.diff(A["nbr_ptr"])
    # 4x4 grid: 4 corner cells with 2, 8 edge cells with 3, 4 inner cells with 4 neighbours
    assert sorted(np.bincount(counts).tolist()) == sorted([0, 0, 4, 8, 4])


def test_newton_cotes_dvs(tmp_path):
    L, mount = 4.0, 6
    for deg in (1, 2, 4):
        d = tmp_path / f"nc{deg}"
        d.mkdir()
        cs = open_case(d, mg.quad_cavity(2), ("nc", deg, mount, L))
        A = cs.arrays()
        N = deg * mount + 1
        assert cs.n_vel == N * N
        xi, w = A["xi"], A["w"]
        assert abs(w.sum() - (2 * L) ** 2) < 1e-12
        # a closed Newton-Cotes rule of degree deg integrates x^deg exactly per panel
        assert abs((w * xi[:, 0] ** 2).sum() - (2 * L) * 2 * L ** 3 / 3) < 1e-10 * L ** 4 or deg == 1
        # Gaussian moments to quadrature accuracy
        g = np.exp(-(xi[:, 0] ** 2 + xi[:, 1] ** 2))
        assert abs((w * g).sum() - np.pi) < 0.1
        m = A["mirror"]
        for ax in range(2):
            flip = xi.copy()
            flip[:, ax] *= -1
            assert np.allclose(xi[m[:, ax]], flip, atol=1e-14)
            assert np.array_equal(w[m[:, ax]], w)
        assert np.array_equal(m[:, 2], np.arange(cs.n_vel))


def test_gauss_hermite_dvs_from_meso_file():
    """files/gh-3.dvs (3-point Gauss-Hermite) as meso scales it: sqrt(2RT) x, w/sqrt(pi)."""
    here = os.path.dirname(os.path.abspath(__file__))
    cs = capi.Case(os.path.join(here, "golden", "re400_cavity", "case.txt"))
    A = cs.arrays()
    RT = cs.params.RT
    xi, w = A["xi"], A["w"]
    assert cs.n_vel == 9
    assert abs(w.sum() - 1.0) < 1e-14
    for ax in range(2):
        assert abs((w * xi[:, ax]).sum()) < 1e-15
        assert abs((w * xi[:, ax] ** 2).sum() - RT) < 1e-14
        assert abs((w * xi[:, ax] ** 4).sum() - 3 * RT * RT) < 1e-13  # exact to degree 5
    assert abs((w * xi[:, 0] ** 2 * xi[:, 1] ** 2).sum() - RT * RT) < 1e-14


def test_case_errors(tmp_path):
    m = mg.quad_cavity(3)
    d = str(tmp_path / "ok")
    p = mg.shakhov_cavity_case(d, m, ("nc", 1, 4, 4.0), 0.1, 0.1)
    text = open(p).read()
    with pytest.raises(capi.MesoError) as e:
        capi.Case(str(tmp_path / "missing.txt"))
    assert e.value.code == capi.MESO_E_IO

    def variant(tag, old, new):
        vd = tmp_path / tag
        vd.mkdir()
        for f in ("mesh.neu", "dvs.dvs"):
            os.symlink(os.path.join(d, f), vd / f)
        assert old in text
        (vd / "case.txt").write_text(text.replace(old, new))
        return str(vd / "case.txt")

    # the limiter switch; venkata_k stays 0 with the gradient on (meso reads
    # venkata-limiter-k only with gradient-switch False, cdugks_shakhov.cpp:21-28)
    lim = capi.Case(variant("lim", "limiter-switch       False", "limiter-switch       True"))
    assert lim.params.limiter_switch == 1 and lim.params.venkata_k == 0.0
    lim.reload_settings()
    assert lim.params.limiter_switch == 1 and lim.params.venkata_k == 0.0
    vk = capi.Case(variant("limk", "gradient-switch      True\nlimiter-switch       False",
                           "gradient-switch      False\nlimiter-switch       True\nvenkata-limiter-k    2.5"))
    assert vk.params.venkata_k == 2.5 and vk.params.gradient_switch == 0
    vk.reload_settings()
    assert vk.params.venkata_k == 2.5
    # unknown solver
    with pytest.raises(capi.MesoError):
        capi.Case(variant("solver", "cdugks@shakhov", "cdugks@nosuch"))
    # a boundary set of the mesh without a [mark]
    with pytest.raises(capi.MesoError):
        capi.Case(variant("mark", "name             lid", "name             notlid"))
    # freestream-inlet: the reference reads its values by loop index (not supported)
    with pytest.raises(capi.MesoError):
        capi.Case(variant("free", "type             wall\nvelocity         fixedValue vector 0.1",
                          "type             freestream-inlet\nvelocity         fixedValue vector 0.1"))


def test_settings_and_reload(tmp_path):
    d = str(tmp_path)
    p = mg.shakhov_cavity_case(d, mg.quad_cavity(3), ("nc", 1, 4, 4.0), 0.1, 0.1, dt=1e-3, max_step=77)
    cs = capi.Case(p)
    assert cs.setting_int("max-step", 0) == 77
    assert cs.setting_int("no-such-key", 5) == 5
    assert cs.setting_str("solver") == "cdugks@shakhov"
    assert cs.setting_double("timeStep", 0.0) == 1e-3
    assert cs.params.dt == 1e-3
    text = open(p).read().replace("timeStep             0.001", "timeStep             0.0005")
    open(p, "w").write(text)
    cs.reload_settings()
    assert cs.params.dt == 5e-4


def test_initial_patch(tmp_path):
    d = str(tmp_path)
    p = mg.shakhov_cavity_case(d, mg.quad_cavity(3), ("nc", 1, 4, 4.0), 0.1, 0.1)
    text = open(p).read().replace("density          fixedValue scalar 1.0",
                                  "density          fixedValue scalar 1.25\nvelocity         fixedValue vector 0.1 -0.2 0.0")
    open(p, "w").write(text)
    cs = capi.Case(p)
    rho, u, T = cs.initial_patch()
    assert np.all(rho == 1.25) and np.all(T == 1.0)
    assert np.all(u[:, 0] == 0.1) and np.all(u[:, 1] == -0.2)


def test_side_marks_parse(tmp_path):
    m = split_sides(mg.quad_cavity(4, 3), 4, 3)
    d = str(tmp_path)
    mg.write_neu(m, os.path.join(d, "mesh.neu"))
    mg.write_newton_cotes(os.path.join(d, "dvs.dvs"), 1, 4, 4.0)
    settings = {"solver": "cdugks@shakhov", "mesh-file": "mesh.neu", "dvs-file": "dvs.dvs",
                "dvs-type": "Newton-Cotes", "Pr": "0.6666666666666666", "timeStep": "1e-3", "gas-constant": "0.5",
                "gas-k": "0", "ref-temperature": "1.0", "ref-viscosity": "0.1", "vhs-index": "0.5",
                "gradient-switch": "True", "limiter-switch": "False"}
    groups = [{"name": "fluid", "density": "fixedValue scalar 1.0", "temperature": "fixedValue scalar 1.0"}]
    sym = lambda n, ax: {"name": n, "type": "symmetry", "direction": f"fixedValue objectType {ax}"}
    marks = [sym("left", 0), sym("right", 0), sym("bottom", 1), sym("top", 1)]
    mg.write_case(os.path.join(d, "case.txt"), settings, groups, marks)
    cs = capi.Case(os.path.join(d, "case.txt"))
    A = cs.arrays()
    fm = A["face_mark"]
    bnd = A["face_cells"][:, 0] == A["face_cells"][:, 1]
    assert np.all(fm[~bnd] == 0) and np.all(fm[bnd] > 0)
    assert (bnd.sum(), len(A["marks"])) == (2 * (4 + 3), 5)
    types = {A["marks"][i].type for i in range(1, 5)}
    assert types == {capi.BC["symmetry"]}
