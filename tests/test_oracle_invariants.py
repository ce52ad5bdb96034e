"""The FP64 oracle against invariants the mathematics fixes. CPU only.

These pin the parts of the oracle that the golden vectors leave open (the
symmetry and slip-wall boundaries, whose reference implementation reads
out of bounds) and guard the rest against plausible slips:
  - exact mass conservation in closed boxes (diffuse walls, symmetry planes,
    slip walls): a wrong mirror index or a lost sign in the wall density
    breaks it at O(1e-3), far above rounding;
  - the half-step reconstruction f-bar+ of a linear field is exact:
    f(x_f - xi dt/2) at every interior face, upwind side by sign(xi.n);
  - the residual is the volume-weighted mean relative change (field.cpp:222-249);
  - a uniform state at rest between walls of the same temperature only moves
    by the quadrature error of the velocity set.
"""
import os

import numpy as np
import pytest

from conftest import split_sides
from meso_hip import capi
from meso_hip import meshgen as mg
from oracle.oracle import Oracle


def box_case(d, kinds, solver="cdugks@shakhov", n=(6, 5)):
    """Jittered quad box with one mark per side; kinds: side -> mark dict (type and values)."""
    os.makedirs(d, exist_ok=True)
    nx, ny = n
    m = split_sides(mg.quad_cavity(nx, ny, jitter=0.15, seed=8), nx, ny)
    mg.write_neu(m, os.path.join(d, "mesh.neu"))
    if solver == "cdugks@shakhov":
        mg.write_newton_cotes(os.path.join(d, "dvs.dvs"), 1, 12, 4.5)
        settings = {"solver": solver, "mesh-file": "mesh.neu", "dvs-file": "dvs.dvs", "dvs-type": "Newton-Cotes",
                    "Pr": repr(2.0 / 3.0), "timeStep": "2.0e-3", "gas-constant": "0.5", "gas-k": "0",
                    "ref-temperature": "1.0", "ref-viscosity": repr(mg.viscosity_from_kn(0.1)), "vhs-index": "0.5",
                    "gradient-switch": "True", "limiter-switch": "False"}
        groups = [{"name": "fluid", "density": "fixedValue scalar 1.0", "temperature": "fixedValue scalar 1.0"}]
    else:
        mg.write_gauss_hermite(os.path.join(d, "dvs.dvs"), 3)
        settings = {"solver": solver, "mesh-file": "mesh.neu", "dvs-file": "dvs.dvs", "dvs-type": "Gauss-Hermite",
                    "Re": "100.0", "Ma": "0.1", "timeStep": "2.0e-3", "gas-constant": "0.5",
                    "ref-temperature": "1.0", "gradient-switch": "True"}
        groups = [{"name": "fluid", "density": "fixedValue scalar 1.0"}]
    marks = [dict(name=side, **kinds[side]) for side in ("left", "right", "bottom", "top")]
    mg.write_case(os.path.join(d, "case.txt"), settings, groups, marks)
    return os.path.join(d, "case.txt")


def wavy_init(cs, amp=0.05):
    x = cs.arrays()["cell_pos"]
    rho = 1.0 + amp * np.sin(2 * np.pi * x[:, 0]) * np.cos(np.pi * x[:, 1])
    u = np.zeros((cs.n_cells, 3))
    u[:, 0] = amp * np.sin(np.pi * x[:, 1])
    u[:, 1] = -amp * np.cos(np.pi * x[:, 0]) * 0.5
    T = 1.0 + amp * np.cos(2 * np.pi * x[:, 1])
    return rho, u, T


def mass(orc, cs):
    return float(np.sum(orc.macros()["rho"] * cs.arrays()["cell_vol"]))


WALL = {"type": "wall", "velocity": "fixedValue vector 0.0 0.0 0.0", "temperature": "fixedValue scalar 1.0"}
LID = {"type": "wall", "velocity": "fixedValue vector 0.1 0.0 0.0", "temperature": "fixedValue scalar 1.02"}
SYM = lambda ax: {"type": "symmetry", "direction": f"fixedValue objectType {ax}"}
SLIP = lambda ax: {"type": "slip-wall", "velocity": "fixedValue vector 0.0 0.0 0.0",
                   "temperature": "fixedValue scalar 1.0", "direction": f"fixedValue objectType {ax}",
                   "slip-wall-alpha-u": "fixedValue scalar 0.3", "slip-wall-alpha-T": "fixedValue scalar 0.6"}

BOXES = {
    "walls": {"left": WALL, "right": WALL, "bottom": WALL, "top": LID},
    "symmetry": {"left": SYM(0), "right": SYM(0), "bottom": SYM(1), "top": SYM(1)},
    "slip": {"left": SLIP(0), "right": SLIP(0), "bottom": SLIP(1), "top": SLIP(1)},
    "mixed": {"left": SYM(0), "right": SLIP(0), "bottom": WALL, "top": LID},
}


@pytest.mark.parametrize("box", list(BOXES))
def test_mass_conserved_in_closed_box(tmp_path, box):
    cs = capi.Case(box_case(str(tmp_path), BOXES[box]))
    orc = Oracle(cs)
    orc.init_equilibrium(*wavy_init(cs))
    m0 = mass(orc, cs)
    rep = orc.step(20)
    assert rep.bad_cells == 0
    m1 = mass(orc, cs)
    assert abs(m1 - m0) < 1e-13 * m0, (m0, m1)
    # the flow did change: the check is not vacuous
    assert np.abs(orc.macros()["rho"] - wavy_init(cs)[0]).max() > 1e-3


@pytest.mark.parametrize("solver", ["cdugks@incompressible", "dugks@incompressible"])
def test_mass_conserved_bgk(tmp_path, solver):
    kinds = {"left": {"type": "wall"}, "right": {"type": "wall"}, "bottom": {"type": "wall"},
             "top": {"type": "wall", "velocity": "fixedValue vector 0.1 0.0 0.0"}}
    cs = capi.Case(box_case(str(tmp_path), kinds, solver=solver))
    orc = Oracle(cs)
    rho, u, _ = wavy_init(cs, 0.02)
    orc.init_equilibrium(rho, u, np.ones(cs.n_cells))
    m0 = mass(orc, cs)
    orc.step(20)
    assert abs(mass(orc, cs) - m0) < 1e-13 * m0


def test_mirror_map(tmp_path):
    """The mirror map behind the symmetry and slip-wall values: same weight,
    opposite normal component, zero net mass flux through the plane."""
    cs = capi.Case(box_case(str(tmp_path), BOXES["symmetry"]))
    A = cs.arrays()
    xi, w, mir = A["xi"], A["w"], A["mirror"]
    for ax in range(2):
        # mirrored velocity: same weight, opposite normal component
        assert np.array_equal(w[mir[:, ax]], w)
        assert np.allclose(xi[mir[:, ax], ax], -xi[:, ax], atol=0)
    # a drifting f carries a net normal flux ...
    f = np.exp(-((xi[:, 0] - 0.3) ** 2 + xi[:, 1] ** 2))
    assert abs(np.sum(w * xi[:, 0] * f)) > 1e-3
    # ... which the symmetry value (incoming = mirrored outgoing) cancels exactly
    fin = np.where(xi[:, 0] < 0, f[mir[:, 0]], f)
    assert abs(np.sum(w * xi[:, 0] * fin)) < 1e-14


@pytest.mark.parametrize("mesh", ["quad", "tri", "hex"])
def test_fbar_of_linear_field_is_exact(tmp_path, mesh):
    """f-bar+ = f(x_f - xi dt/2) for a linear field, taken from the upwind
    cell (cdugks_shakhov.cpp:223-298); the LSQ gradient is exact on it."""
    if mesh == "quad":
        m, dvs = mg.quad_cavity(5, jitter=0.2, seed=1), ("nc", 1, 4, 3.0)
    elif mesh == "tri":
        m, dvs = mg.tri_cavity(5, jitter=0.2, seed=2), ("nc", 1, 4, 3.0)
    else:
        m, dvs = mg.hex_cavity(3, jitter=0.05, seed=3), ("sparse-gh", 3, 0.5)
    p = mg.shakhov_cavity_case(str(tmp_path), m, dvs, 0.1, 0.1, dt=0.01)
    cs = capi.Case(p)
    orc = Oracle(cs)
    A = cs.arrays()
    nv, nc = cs.n_vel, cs.n_cells
    rng = np.random.default_rng(5)
    a = rng.standard_normal(nv)
    b = rng.standard_normal((nv, 3))
    if cs.mesh.dim == 2:
        b[:, 2] = 0.0
    cell_f = a[:, None] + b @ A["cell_pos"].T
    ff = orc.fbar(cell_f)
    hdt = 0.5 * cs.params.dt
    own, nei = A["face_cells"][:, 0], A["face_cells"][:, 1]
    inner = own != nei
    xf = A["face_pos"][None, :, :] - hdt * A["xi"][:, None, :]  # [nv, nf, 3]
    exact = a[:, None] + np.einsum("vk,vfk->vf", b, xf)
    assert np.abs(ff[:, inner] - exact[:, inner]).max() < 1e-11 * np.abs(exact).max()
    # boundary faces extrapolate from the owner alone (the value before boundary conditions)
    sn = A["xi"] @ A["face_normal"].T
    outgoing = (sn >= 1e-30) & ~inner[None, :]
    assert np.abs(ff[outgoing] - exact[outgoing]).max() < 1e-11 * np.abs(exact).max()


def test_fbar_first_order_without_gradient(tmp_path):
    m = mg.quad_cavity(4, jitter=0.1, seed=2)
    p = mg.shakhov_cavity_case(str(tmp_path), m, ("nc", 1, 4, 3.0), 0.1, 0.1, dt=0.01)
    text = open(p).read().replace("gradient-switch      True", "gradient-switch      False")
    open(p, "w").write(text)
    cs = capi.Case(p)
    orc = Oracle(cs)
    A = cs.arrays()
    f = np.random.default_rng(2).standard_normal((cs.n_vel, cs.n_cells))
    ff = orc.fbar(f)
    sn = A["xi"] @ A["face_normal"].T
    own, nei = A["face_cells"][:, 0], A["face_cells"][:, 1]
    up = np.where(sn >= 1e-30, own[None, :], nei[None, :])
    want = np.take_along_axis(f, up, axis=1)
    dead = np.abs(sn) < 1e-30
    want[dead] = 0.5 * (f[:, own][dead] + f[:, nei][dead])
    assert np.array_equal(ff, want)


def test_residual_hand_example(tmp_path):
    cs = capi.Case(box_case(str(tmp_path), BOXES["walls"]))
    orc = Oracle(cs)
    orc.init_equilibrium(*wavy_init(cs))
    orc.residual()  # resets the stored copies
    before = orc.macros()
    orc.step(3)
    after = orc.macros()
    r = orc.residual()
    wv = cs.arrays()["cell_vol"] / cs.arrays()["total_volume"]
    rel = lambda a, b: np.sum(np.abs((a - b) / b) * wv)
    assert np.isclose(r[0], rel(after["rho"], before["rho"]), rtol=1e-12)
    for i in range(2):
        assert np.isclose(r[1 + i], rel(after["u"][:, i], before["u"][:, i]), rtol=1e-12)
        assert np.isclose(r[5 + i], rel(after["q"][:, i], before["q"][:, i]), rtol=1e-12)
    assert np.isclose(r[4], rel(after["T"], before["T"]), rtol=1e-12)
    # the copies were reset: an immediate second call compares a state with itself
    assert np.all(orc.residual()[[0, 1, 2, 4, 5, 6]] == 0.0)


def test_rest_state_moves_only_by_quadrature_error(tmp_path):
    kinds = {s: WALL for s in ("left", "right", "bottom", "top")}
    cs = capi.Case(box_case(str(tmp_path), kinds))
    orc = Oracle(cs)
    orc.init_equilibrium()
    m0 = orc.macros()
    orc.step(10)
    m1 = orc.macros()
    # Newton-Cotes 25x25 on [-4.5, 4.5]: Maxwellian moments are exact to ~1e-9
    assert np.abs(m1["rho"] - m0["rho"]).max() < 1e-8
    assert np.abs(m1["u"]).max() < 1e-8
    assert np.abs(m1["T"] - m0["T"]).max() < 1e-8
