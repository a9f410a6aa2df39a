"""Closed-form pins of the three step operations that live in
Simulation::step itself (src/engine/simulation.cpp), which is not compiled
here: the force cap (:588-609), the node clamp in y (:665-675) and the
capsule recirculation in x (:687-700).  Both restatements are checked: the
oracle's (oracle/sfo.cpp) and the reference harness's (oracle/ref_harness.cpp,
over the reference's own CapsuleSystem and Capsule::centroid).

Each property is stated from what the operation must do, not from its
formula, with inputs placed so that a plausible mistake fails:
  cap    |F| exactly at the limit stays (a >= test would rescale it); just
         above it lands on the limit along F (a dropped sqrt, a squared limit
         or a wrong component would not); every cell type is capped (the
         in-step check finds capped SOLID wall cells); the cap follows the body
         force (capped fluid cells sit exactly on the limit); NaN stays and is
         not counted.
  clamp  nodes past either wall line land exactly on 1.5 / ny - 1.5, nodes on
         or inside the band keep their bits; off with periodic y.
  recirc a capsule whose node mean is at or past nx (or below 0) moves by
         -nx (+nx) rigidly; one whose mean is just inside stays, even when some
         of its nodes are outside.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.cases import oracle_sim

WHICH = ["oracle", pytest.param("ref", marks=pytest.mark.skipif(not po.ref_available(),
                                                                 reason="reference build absent"))]
FM = 0.01
SKALAK = dict(model=2, G_s=0.06, C_skalak=10.0)


def exact_above(fx, fy, fm):
    """|F|^2 > fm^2 in exact rational arithmetic."""
    return Fraction(fx) ** 2 + Fraction(fy) ** 2 > Fraction(fm) ** 2


def cap_inputs():
    rng = np.random.default_rng(3)
    th = rng.uniform(0, 2 * np.pi, 40)
    fx, fy = [], []
    # exactly on the limit along the axes: not capped (strict >)
    for a, b in ((FM, 0.0), (0.0, -FM), (-FM, 0.0), (0.0, FM)):
        fx.append(a); fy.append(b)
    # 1e-9 above and below the limit in random directions
    for t in th[:20]:
        fx.append(FM * (1 + 1e-9) * np.cos(t)); fy.append(FM * (1 + 1e-9) * np.sin(t))
    for t in th[20:]:
        fx.append(FM * (1 - 1e-9) * np.cos(t)); fy.append(FM * (1 - 1e-9) * np.sin(t))
    # far above, zero, tiny, and one component large
    fx += [3.0, 0.0, 1e-300, -0.5, 7e-3]
    fy += [-4.0, 0.0, 0.0, 1e-6, 7.2e-3]
    return np.array(fx), np.array(fy)


@pytest.mark.parametrize("which", WHICH)
def test_force_cap_closed_form(which):
    fx, fy = cap_inputs()
    gx, gy, n = po.cap_forces(fx, fy, FM, which)
    above = np.array([exact_above(a, b, FM) for a, b in zip(fx, fy)])
    assert n == int(above.sum())
    assert above.sum() >= 20 and (~above).sum() >= 25
    # untouched cells keep their bits
    assert np.array_equal(gx[~above], fx[~above]) and np.array_equal(gy[~above], fy[~above])
    # rescaled cells lie on the limit, along the original direction
    mag = np.hypot(gx[above], gy[above])
    assert np.max(np.abs(mag / FM - 1.0)) < 4e-16
    cross = gx[above] * fy[above] - gy[above] * fx[above]
    assert np.max(np.abs(cross) / (np.hypot(fx[above], fy[above]) * mag)) < 4e-16
    assert np.all(gx[above] * fx[above] + gy[above] * fy[above] > 0)
    # the limit is the argument: at 0.01 (1 + 2e-9) the cells 1e-9 above 0.01 stay
    _, _, n2 = po.cap_forces(fx, fy, FM * (1 + 2e-9), which)
    assert n2 == int(sum(exact_above(a, b, FM * (1 + 2e-9)) for a, b in zip(fx, fy))) == n - 20


@pytest.mark.parametrize("which", WHICH)
def test_force_cap_leaves_nan_uncounted(which):
    fx = np.array([np.nan, 1.0, 0.0])
    fy = np.array([0.0, 0.0, np.nan])
    gx, gy, n = po.cap_forces(fx, fy, FM, which)
    assert n == 1
    assert np.isnan(gx[0]) and np.isnan(gy[2])
    assert gx[1] == FM and gy[1] == 0.0


@pytest.mark.parametrize("which", WHICH)
def test_node_clamp_closed_form(which):
    ny = 24
    lo, hi = 1.5, ny - 1.5
    y = np.array([-3.0, 0.0, 1.4999999999999998, 1.5, 1.5000000000000002, 12.0,
                  22.499999999999996, 22.5, 22.500000000000004, 23.9, 40.0])
    got = po.clamp_nodes_y(y, ny, which)
    inside = (y >= lo) & (y <= hi)
    assert np.array_equal(got[inside], y[inside])
    assert np.all(got[y < lo] == lo) and np.all(got[y > hi] == hi)
    assert (y < lo).sum() == 3 and (y > hi).sum() == 3


@pytest.mark.parametrize("which", WHICH)
def test_recirculation_closed_form(which):
    nx = 50
    # capsule 0: mean exactly nx -> moves by -nx; capsule 1: mean just below nx
    # with two nodes past it -> stays; capsule 2: mean below 0 -> +nx;
    # capsule 3: mean exactly 0 -> stays
    caps = [np.array([49.0, 50.0, 51.0]), np.array([48.0, 50.5, 51.4999]),
            np.array([-0.2, -0.1, 0.1]), np.array([-1.0, 0.0, 1.0])]
    x = np.concatenate(caps)
    y = np.arange(x.size, dtype=float)
    off = np.cumsum([0] + [c.size for c in caps]).astype(np.int32)
    got, moved = po.recirculate_x(x, y, off, nx, which)
    assert moved == 2
    assert np.array_equal(got[0:3], caps[0] - nx)
    assert np.array_equal(got[3:6], caps[1])
    assert np.array_equal(got[6:9], caps[2] + nx)
    assert np.array_equal(got[9:12], caps[3])


# capsules overlapping both walls: their outer nodes start beyond the wall
# lines, and the contact forces near the walls exceed the force limit
WALL_CASE = dict(nx=48, ny=24, tau=0.8, bfx=3e-5, coll=0,
                 capsules=[dict(centers=[(12.0, 2.0), (30.0, 23.0)], rmin=3.5, nodes=20, mem=SKALAK)])


@pytest.mark.parametrize("which", WHICH)
def test_cap_in_step(which):
    """One step from the same start with and without the limit: the limited
    step's force field is the unlimited one with exactly the cells above the
    limit (wall cells included) rescaled onto it."""
    free = oracle_sim(dict(WALL_CASE, fmax=1e9), which)
    free.step(1)
    capd = oracle_sim(WALL_CASE, which)
    capd.step(1)
    F0, F1 = free.field(), capd.field()
    fx, fy = F0["fx"].ravel(), F0["fy"].ravel()
    gx, gy = F1["fx"].ravel(), F1["fy"].ravel()
    above = np.array([exact_above(a, b, FM) for a, b in zip(fx, fy)])
    assert capd.last_capped() == int(above.sum()) > 0
    assert free.last_capped() == 0
    assert np.array_equal(gx[~above], fx[~above]) and np.array_equal(gy[~above], fy[~above])
    mag = np.hypot(gx[above], gy[above])
    assert np.max(np.abs(mag / FM - 1.0)) < 4e-16
    # capped cells in the wall rows: the cap is not limited to fluid cells
    rows = np.nonzero(above)[0] // WALL_CASE["nx"]
    assert np.any((rows == 0) | (rows == WALL_CASE["ny"] - 1))
    # and in fluid rows, where the body force was added before the cap
    assert np.any((rows > 0) & (rows < WALL_CASE["ny"] - 1))


@pytest.mark.parametrize("which", WHICH)
@pytest.mark.parametrize("periodic_y", [False, True])
def test_clamp_in_step(which, periodic_y):
    """After one step every node sits at its start plus its interpolated
    velocity (dt = 1), clamped into [1.5, ny - 1.5] without periodic y."""
    case = dict(WALL_CASE, periodic_y=periodic_y)
    s = oracle_sim(case, which)
    y0 = s.nodes()["y"].copy()
    s.step(1)
    n = s.nodes()
    moved = y0 + n["vy"]
    lo, hi = 1.5, case["ny"] - 1.5
    if periodic_y:
        assert np.array_equal(n["y"], moved)
        assert np.any(n["y"] < lo) and np.any(n["y"] > hi)
    else:
        want = np.where(moved < lo, lo, np.where(moved > hi, hi, moved))
        assert np.array_equal(n["y"], want)
        assert np.sum(moved < lo) >= 3 and np.sum(moved > hi) >= 3
