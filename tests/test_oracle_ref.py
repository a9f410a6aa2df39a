"""The oracle against the reference's own components, run side by side.

oracle/_ref/libsfref.so is compiled from the reference's sources where they
lie (``make -C oracle ref``: the lbm, coupling and membrane translation units at
-O2 -ffp-contract=off, OpenMP at one thread) and driven in Simulation::step
order.  It exists only where the reference tree is present; elsewhere these
tests skip and tests/test_oracle_golden.py carries the same check through the
committed vectors.  Every comparison is bitwise.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")


def run(which, coll=0, mdf=1, model=2, helf=0, ncap=8, steps=150, scalar=0, nx=100, ny=40,
        seed=42, gvisc=0.01, damp=0.0, fric=0.0, shape=0, boundary=1, periodic_y=0, nodes=0,
        eta=0.0, k_area=None, k_per=None, walls=(0.0, 0.0)):
    p = po.default_params(nx=nx, ny=ny, tau=0.8, boundary_type=boundary, body_force_x=3e-5,
                          collision_model=coll, ibm_iterations=mdf, scalar_enabled=scalar,
                          diffusivity=[0.05], rep_damping_normal=damp, rep_friction_coeff=fric,
                          periodic_y=periodic_y, top_wall_velocity=walls[0],
                          bottom_wall_velocity=walls[1])
    s = po.Sim(p, which)
    kw = dict(model=model, G_s=0.06, C_skalak=10.0, use_helfrich=helf,
              k_bend=1e-3 if helf else 0.005, gamma_visc=gvisc, shape=shape, aspect_ratio=0.6,
              eta_membrane=eta)
    if k_area is not None:
        kw["k_area"] = k_area
    if k_per is not None:
        kw["k_perimeter"] = k_per
    m = po.default_membrane(**kw)
    placed = s.add_capsule_random(ncap, 10, 8, nx - 10, ny - 8, 3.0, 4.5, nodes, m, 0, seed, 1.0, 0)
    s.initialize()
    if scalar:
        s.scalar_initialize(0.0)
        for y in range(ny):
            for x in range(nx):
                s.set_scalar_point(x, y, float(np.exp(-((x - nx / 2) ** 2 + (y - ny / 2) ** 2) / 128.0)), 0)
    s.step(steps)
    out = dict(s.field())
    out.update(s.nodes())
    out["f"] = s.f()
    if scalar:
        out["C"] = s.scalar()
    return out, placed, s.last_capped()


CASES = dict(
    bgk=dict(coll=0, ncap=0),
    mrt=dict(coll=1, ncap=0),
    closed_box=dict(coll=2, ncap=0, boundary=2),
    periodic_y=dict(coll=0, ncap=0, periodic_y=1),
    reg_caps_mdf3=dict(coll=2, mdf=3),
    bgk_mdf5=dict(coll=0, mdf=5, ncap=4),
    bgk_caps=dict(coll=0, ncap=12),
    mrt_helfrich_adr=dict(coll=1, helf=1, scalar=1, ncap=10),
    wlc=dict(model=3, ncap=6),
    neo_hookean=dict(model=1, ncap=6),
    hookean_ellipse=dict(model=0, ncap=6, shape=1),
    biconcave=dict(model=2, ncap=4, shape=2),
    kelvin_voigt=dict(model=2, ncap=6, eta=0.05),
    area_perimeter_penalty=dict(model=2, ncap=6, k_area=0.2, k_per=0.1),
    allpairs_few_nodes=dict(ncap=2, nodes=12),
    dem_contact=dict(ncap=12, damp=0.01, fric=0.1),
    moving_walls=dict(coll=0, ncap=0, walls=(0.02, -0.01)),
    moving_walls_closed_mrt=dict(coll=1, ncap=0, boundary=2, walls=(0.03, 0.0)),
    moving_walls_caps_mdf3=dict(coll=2, mdf=3, ncap=6, walls=(0.01, -0.01)),
    inlet_outlet_bgk=dict(coll=0, ncap=0, boundary=0),
    inlet_outlet_mrt_caps=dict(coll=1, ncap=8, boundary=0),
    inlet_outlet_reg_mdf3_adr=dict(coll=2, mdf=3, ncap=6, boundary=0, scalar=1),
)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_equals_reference_components(name):
    a, pa, ca = run("oracle", **CASES[name])
    b, pb, cb = run("ref", **CASES[name])
    assert pa == pb and ca == cb
    for k in a:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, k
        assert np.array_equal(x, y), (k, float(np.max(np.abs(x - y))) if x.size else 0.0)


@pytest.mark.parametrize("model", [0, 1, 2, 3])
@pytest.mark.parametrize("helf", [0, 1])
def test_membrane_forces_equal_reference(model, helf):
    # random deformations of a placed capsule; every force law in isolation
    outs = []
    for which in ("oracle", "ref"):
        p = po.default_params(nx=64, ny=48, boundary_type=1)
        s = po.Sim(p, which)
        m = po.default_membrane(model=model, G_s=0.05, k_stretch=0.05, use_helfrich=helf,
                                k_bend=2e-3, k_area=0.1, k_perimeter=0.05, gamma_visc=0.02,
                                shape=1, aspect_ratio=0.7)
        s.add_capsule(30.2, 21.7, 5.0, 26, m, 0)
        s.initialize()
        rng = np.random.default_rng(model * 10 + helf)
        nd = s.nodes()
        s.set_nodes(x=nd["x"] + 0.3 * rng.standard_normal(26), y=nd["y"] + 0.3 * rng.standard_normal(26),
                    vx=0.01 * rng.standard_normal(26), vy=0.01 * rng.standard_normal(26))
        s.membrane_forces_only()
        outs.append(s.nodes())
    for k in ("fx", "fy"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("coll", [0, 1, 2])
def test_solid_obstacles_equal_reference(coll):
    # extra SOLID cells set after initialize (LatticeField::setCellType +
    # rebuildBoundaryNodeLists): a rectangular block and a disc in a channel
    from tests.cases import oracle_sim, oracle_state
    case = dict(nx=60, ny=30, tau=0.8, bfx=3e-5, coll=coll, scalar=dict(D=0.05, blob=(20, 15, 4.0)),
                solids=[("rect", 20, 10, 24, 16), ("circle", 42.3, 18.6, 4.2)])
    a = oracle_sim(case)
    b = oracle_sim(case, which="ref")
    a.step(300)
    b.step(300)
    sa, sb = oracle_state(a), oracle_state(b)
    assert int(np.sum(a.flags() == 1)) > 60 * 2 + 20
    for k in ("rho", "ux", "uy", "fx", "fy", "C"):
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(a.f(), b.f())


# The channel builder's pillars (geometry/channel_builder.cpp, circle / rect /
# polygon obstacles): cell flags, capsule placement around them, the obstacle
# contact forces (repulsion.cpp:180-236) and, with use_interpolated_bb, the
# Bouzidi links (interpolated_bounce_back.cpp).  Each case runs the oracle and
# the reference's own classes and compares every field bitwise.
from tests.cases import OBSTACLE_CASES, PILLARS  # noqa: E402


@pytest.mark.parametrize("name", sorted(OBSTACLE_CASES))
def test_obstacles_equal_reference(name):
    from tests.cases import oracle_sim, oracle_state
    case = dict(dict(tau=0.8, bfx=3e-5), **OBSTACLE_CASES[name])
    if case.get("boundary") == "inlet_outlet":
        case["bfx"] = 0.0
    steps = case.pop("steps", 200)
    a = oracle_sim(case)
    b = oracle_sim(case, which="ref")
    assert a.num_capsules == b.num_capsules
    assert np.array_equal(a.flags(), b.flags())
    a.step(steps)
    b.step(steps)
    sa, sb = oracle_state(a), oracle_state(b)
    for k in sa:
        assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), k
    assert np.array_equal(a.f(), b.f())


def test_obstacle_geometry_equals_reference():
    # signedDistance and normalAt of each shape on a grid that crosses every
    # face, edge, vertex and the interiors
    from tests.cases import oracle_sim
    case = dict(nx=130, ny=40, obstacles=PILLARS + [("rect", 20.0, 30.0, 14.0, 25.0)])
    a, b = oracle_sim(case), oracle_sim(case, which="ref")
    for i in range(len(case["obstacles"])):
        for x in np.linspace(0.0, 130.0, 131 * 2 + 1):
            for y in np.linspace(0.0, 40.0, 41 * 2 + 1):
                da, na = a.obstacle_distance(i, x, y)
                db, nb = b.obstacle_distance(i, x, y)
                assert da == db and np.array_equal(na, nb), (i, x, y)


# The step operations of Simulation::step itself (force cap, node clamp,
# capsule recirculation), restated in both libraries, on runs that exercise
# them; see tests/test_oracle_step_pins.py for their closed-form pins.
from tests.cases import STEP_PIN_CASES  # noqa: E402


@pytest.mark.parametrize("name", sorted(STEP_PIN_CASES))
def test_step_pin_cases_equal_reference(name):
    from tests.cases import oracle_sim, oracle_state
    case = dict(STEP_PIN_CASES[name])
    steps = case.pop("steps")
    a = oracle_sim(case)
    b = oracle_sim(case, which="ref")
    a.step(steps)
    b.step(steps)
    sa, sb = oracle_state(a), oracle_state(b)
    for k in sa:
        assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), k
    assert np.array_equal(a.f(), b.f())
    if "walls" in name:
        assert sa["capped"] > 0
        assert np.any(sa["y"] == 1.5) or np.any(sa["y"] == case["ny"] - 1.5)


def test_concentration_write_equals_reference():
    # a write into the concentration array between steps (what the
    # reference's concentration() view allows): the next collision relaxes
    # towards the written C while g keeps its values
    from tests.cases import oracle_sim, oracle_state
    case = dict(nx=40, ny=20, tau=0.8, bfx=3e-5, coll=1, scalar=dict(D=0.05, blob=(20, 10, 4.0)))
    sims = [oracle_sim(case), oracle_sim(case, which="ref")]
    for s in sims:
        s.step(20)
        C = s.scalar().reshape(case["ny"], case["nx"]).copy()
        C[5:9, 10:30] += 0.25
        s.set_concentration(C)
        s.step(30)
    a, b = oracle_state(sims[0]), oracle_state(sims[1])
    for k in ("C", "rho", "ux", "uy"):
        assert np.array_equal(a[k], b[k]), k
