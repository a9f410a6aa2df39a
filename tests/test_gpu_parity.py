"""Parity of the HIP path with the oracle on the same seeded inputs.

FP64 mode evaluates the reference's expressions in the same order, so the
bar is relative L2 <= 1e-12 on rho, u, node positions and velocities (in
practice the fields agree bit for bit).  FP32 mode (f - w_q storage, FP32
collide) is held to the tolerances written in each test.
"""
import math

import numpy as np
import pytest

from tests.cases import oracle_sim, oracle_state, product_sim, product_state, rel_l2

pytestmark = pytest.mark.gpu

SKALAK = dict(model=2, G_s=0.06, C_skalak=10.0)


# Parity metrics: relative L2 of the density field, of the velocity field
# (u_x, u_y together), of node positions and node velocities (x, y together)
# and of the scalar concentration.
METRICS = {"rho": ("rho",), "u": ("ux", "uy"), "pos": ("x", "y"), "vel": ("vx", "vy"), "C": ("C",)}


def metrics(got, ref, keys=("rho", "u", "pos", "vel", "C")):
    out = {}
    for m in keys:
        parts = METRICS[m]
        if any(p not in ref or np.asarray(ref[p]).size == 0 for p in parts):
            continue
        out[m] = rel_l2(np.concatenate([np.ravel(got[p]) for p in parts]),
                        np.concatenate([np.ravel(ref[p]) for p in parts]))
    return out


def _compare(case, steps, tol, keys=("rho", "u", "pos", "vel", "C"), graph=True):
    print(f"case {case.get('name')} steps={steps}", flush=True)
    o = oracle_sim(case)
    o.step(steps)
    ref = oracle_state(o)
    sim = product_sim(case, graph=graph)
    sim.run(steps)
    got = product_state(sim)
    worst = metrics(got, ref, keys)
    print(case.get("name"), worst, "capped", got["capped"], ref["capped"], flush=True)
    for k, v in worst.items():
        assert v <= tol, f"{case.get('name')}: {k} rel L2 {v:.3e} > {tol:g}"
    assert got["capped"] == ref["capped"]
    return got, ref


@pytest.mark.parametrize("coll", [0, 1, 2])
def test_lbm_channel_fp64(coll):
    case = dict(name=f"channel_coll{coll}", nx=96, ny=34, tau=0.8, bfx=3e-5, coll=coll)
    got, ref = _compare(case, 2000, 1e-12)
    assert np.array_equal(got["ux"], ref["ux"]), "FP64 channel should match bit for bit"


def test_lbm_closed_box_fp64():
    case = dict(name="closed", nx=40, ny=30, tau=0.7, bfx=1e-5, coll=0, boundary="closed")
    _compare(case, 500, 1e-12)


def test_lbm_doubly_periodic_fp64():
    case = dict(name="periodic_y", nx=32, ny=24, tau=0.8, bfx=2e-5, coll=1, periodic_y=True)
    _compare(case, 300, 1e-12)


def test_capsules_regularized_mdf3_fp64():
    # the 400 x 80 suspension's physics (regularized, 3 MDF iterations, Skalak C=10) on a shorter channel
    case = dict(name="c2_small", nx=200, ny=80, tau=0.8, bfx=3e-5, coll=2, mdf=3,
                capsules=[dict(count=12, region=(20, 10, 180, 70), rmin=3.0, rmax=3.0, nodes=20,
                               seed=42, mem=SKALAK)])
    _compare(case, 1000, 1e-12)


def test_capsules_bgk_celllist_dem_fp64():
    case = dict(name="celllist_dem", nx=100, ny=40, tau=0.8, bfx=3e-5, coll=0,
                rep=dict(damping_normal=0.01, friction_coeff=0.1),
                capsules=[dict(count=12, region=(8, 8, 92, 32), rmin=3.0, rmax=4.5, nodes=0,
                               seed=7, min_gap=0.3, mem=SKALAK)])
    _compare(case, 800, 1e-12)


def test_capsules_allpairs_fp64():
    case = dict(name="allpairs", nx=60, ny=30, tau=0.8, bfx=3e-5, coll=0,
                capsules=[dict(count=3, region=(8, 8, 52, 22), rmin=3.0, rmax=3.0, nodes=14,
                               seed=3, min_gap=0.2, mem=SKALAK)])
    _compare(case, 800, 1e-12)


@pytest.mark.parametrize("model,shape", [(0, 1), (1, 0), (3, 2)])
def test_membrane_laws_fp64(model, shape):
    mem = dict(model=model, G_s=0.05, k_stretch=0.05, shape=shape, aspect_ratio=0.6)
    case = dict(name=f"law{model}_shape{shape}", nx=80, ny=36, tau=0.8, bfx=3e-5, coll=0,
                capsules=[dict(count=5, region=(8, 8, 72, 28), rmin=3.5, rmax=4.0, nodes=0,
                               seed=11, mem=mem)])
    _compare(case, 600, 1e-12)


def test_helfrich_mrt_scalar_fp64():
    # atan2 on the device and in glibc may differ in the last ulp; the bar
    # stays 1e-12 over this horizon
    case = dict(name="helfrich_adr", nx=96, ny=40, tau=0.7, bfx=1e-5, coll=1,
                scalar=dict(D=0.05, blob=(48, 20, 6.0)),
                capsules=[dict(count=6, region=(8, 8, 88, 32), rmin=3.0, rmax=4.0, nodes=0, seed=5,
                               mem=dict(model=2, G_s=0.06, C_skalak=10.0, use_helfrich=1,
                                        k_bend=1e-3))])
    _compare(case, 600, 1e-12)


@pytest.mark.parametrize("variant", ["channel_bgk", "closed_mrt", "caps_reg_mdf3"])
def test_moving_walls_fp64(variant):
    # moving wall rows (lbm/moving_wall.cpp) after the bounce-back
    case = {"channel_bgk": dict(nx=64, ny=30, tau=0.8, bfx=1e-5, coll=0, walls=(0.02, -0.01)),
            "closed_mrt": dict(nx=48, ny=30, tau=0.7, coll=1, boundary="closed", walls=(0.03, 0.0)),
            "caps_reg_mdf3": dict(nx=96, ny=34, tau=0.8, bfx=3e-5, coll=2, mdf=3, walls=(0.01, -0.01),
                                  capsules=[dict(count=5, region=(8, 6, 88, 28), rmin=3.0, rmax=3.5,
                                                 nodes=20, seed=4, mem=SKALAK)])}[variant]
    case["name"] = "walls_" + variant
    got, ref = _compare(case, 1500, 1e-12)
    assert np.array_equal(got["ux"], ref["ux"]) or "caps" in variant


@pytest.mark.parametrize("variant", ["bgk", "mrt_caps", "reg_mdf3_scalar"])
def test_inlet_outlet_fp64(variant):
    # Zou-He inlet/outlet (lbm/zou_he_boundary.cpp): INLET and OUTLET columns
    # stream without colliding and have their entering populations rebuilt.
    # The reference's scheme is unstable over long runs, so the runs are short.
    caps = [dict(count=4, region=(20, 5, 60, 19), rmin=3.0, rmax=3.5, nodes=20, seed=8, mem=SKALAK)]
    case = {"bgk": dict(nx=80, ny=24, tau=0.8, coll=0, boundary="inlet_outlet"),
            "mrt_caps": dict(nx=80, ny=24, tau=0.8, coll=1, boundary="inlet_outlet", capsules=caps),
            "reg_mdf3_scalar": dict(nx=80, ny=24, tau=0.8, coll=2, mdf=3, boundary="inlet_outlet",
                                    capsules=caps, scalar=dict(D=0.05, blob=(40, 12, 5.0)))}[variant]
    case["name"] = "inlet_outlet_" + variant
    got, ref = _compare(case, 300, 1e-12)
    if variant == "bgk":                     # the plain channel matches bit for bit
        assert np.array_equal(got["ux"], ref["ux"])


def test_inlet_outlet_fp32():
    case = dict(name="inlet_outlet_fp32", nx=80, ny=24, tau=0.8, coll=2, boundary="inlet_outlet")
    o = oracle_sim(case)
    o.step(200)
    ref = oracle_state(o)
    sim = product_sim(dict(case, precision="fp32"))
    sim.run(200)
    e = metrics(product_state(sim), ref, ("rho", "u"))
    print("inlet/outlet fp32", e, flush=True)
    assert e["u"] < 1e-3 and e["rho"] < 1e-5


def test_graph_replay_matches_eager():
    case = dict(name="graph", nx=64, ny=30, tau=0.8, bfx=3e-5, coll=2, mdf=2,
                capsules=[dict(count=4, region=(8, 8, 56, 22), rmin=3.0, rmax=3.0, nodes=16,
                               seed=9, mem=SKALAK)])
    a = product_sim(case, graph=True)
    a.run(37)
    b = product_sim(case, graph=False)
    b.run(37)
    sa, sb = product_state(a), product_state(b)
    for k in ("rho", "ux", "uy", "x", "y"):
        assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("coll", [0, 1, 2])
def test_fp32_deviation_storage(coll):
    case = dict(name=f"fp32_coll{coll}", nx=96, ny=34, tau=0.8, bfx=3e-5, coll=coll)
    o = oracle_sim(case)
    o.step(3000)
    ref = oracle_state(o)
    c32 = dict(case, precision="fp32")
    sim = product_sim(c32)
    sim.run(3000)
    got = product_state(sim)
    e = rel_l2(got["ux"], ref["ux"])
    print("fp32 ux rel L2", e)
    assert e < 2e-4


def test_fp32_capsules():
    case = dict(name="fp32_caps", nx=120, ny=40, tau=0.8, bfx=3e-5, coll=2,
                capsules=[dict(count=10, region=(10, 8, 110, 32), rmin=3.0, rmax=3.0, nodes=20,
                               seed=42, mem=SKALAK)])
    o = oracle_sim(case)
    o.step(1000)
    ref = oracle_state(o)
    sim = product_sim(dict(case, precision="fp32"))
    sim.run(1000)
    got = product_state(sim)
    ex, eu = rel_l2(got["x"], ref["x"]), rel_l2(got["ux"], ref["ux"])
    print("fp32 capsules: x", ex, "ux", eu)
    assert ex < 1e-5 and eu < 1e-3


# The configurations the parity bar is stated on, over 10^4 steps.
def test_poiseuille_400x80_10k_steps():
    case = dict(name="poiseuille_400x80", nx=400, ny=80, tau=0.8, bfx=3e-5, coll=0)
    got, ref = _compare(case, 10000, 1e-12)
    assert np.array_equal(got["ux"], ref["ux"])


def test_capsule_suspension_400x80_10k_steps():
    case = dict(name="suspension_400x80", nx=400, ny=80, tau=0.8, bfx=3e-5, coll=2, mdf=3,
                capsules=[dict(count=30, region=(20, 10, 380, 70), rmin=3.0, rmax=3.0, nodes=20,
                               seed=42, mem=SKALAK)])
    _compare(case, 10000, 1e-12)


@pytest.mark.parametrize("coll", [0, 1, 2])
def test_solid_obstacles_fp64(coll):
    # SOLID cells beyond the channel walls (LatticeField::setCellType): the
    # generic-flags variant of the collide-stream kernel and the scalar's
    # bounce-back at solids
    case = dict(name=f"solids_coll{coll}", nx=60, ny=30, tau=0.8, bfx=3e-5, coll=coll,
                scalar=dict(D=0.05, blob=(20, 15, 4.0)),
                solids=[("rect", 20, 10, 24, 16), ("circle", 42.3, 18.6, 4.2)])
    got, ref = _compare(case, 1500, 1e-12)
    assert np.array_equal(got["ux"], ref["ux"])


def test_fp32_error_report():
    """FP32 deviation storage against FP64 on a capsule suspension: velocity
    field, capsule centroid trajectories and Taylor deformation."""
    case = dict(name="fp32_report", nx=400, ny=80, tau=0.8, bfx=3e-5, coll=2, mdf=3,
                capsules=[dict(count=30, region=(20, 10, 380, 70), rmin=3.0, rmax=3.0, nodes=20,
                               seed=42, mem=SKALAK)])
    a = product_sim(case)                            # FP64 (equal to the oracle, see above)
    b = product_sim(dict(case, precision="fp32"))
    ecen, eD = 0.0, 0.0
    for _ in range(10):
        a.run(200)
        b.run(200)
        ca, cb = a.capsules(), b.capsules()
        for i in range(len(ca)):
            pa, pb = ca[i].centroid(), cb[i].centroid()
            ecen = max(ecen, math.hypot(pa.x - pb.x, pa.y - pb.y))
            eD = max(eD, abs(ca[i].deformationIndex() - cb[i].deformationIndex()))
    sa, sb = product_state(a), product_state(b)
    eu = metrics(sb, sa, ("u",))["u"]
    print(f"fp32 vs fp64 after 2000 steps: velocity rel L2 {eu:.3e}, centroid max |dx| {ecen:.3e}, "
          f"Taylor D max |dD| {eD:.3e}", flush=True)
    assert eu < 1e-3 and ecen < 1e-2 and eD < 1e-3


# The operations of Simulation::step itself on the device: the force cap
# (firing in every step, wall cells included), the node clamp at the wall
# lines and capsule recirculation in x over a non-periodic fluid.  Compared
# with the oracle after the first step, the third and the last, FP64.
from tests.cases import STEP_PIN_CASES  # noqa: E402


@pytest.mark.parametrize("name", sorted(STEP_PIN_CASES))
def test_step_operations_match_oracle(name):
    case = dict(STEP_PIN_CASES[name], name=name)
    last = case.pop("steps")
    o = oracle_sim(case)
    sim = product_sim(case)
    done = 0
    for upto in (1, 3, last):
        print(f"{name} to step {upto}", flush=True)
        o.step(upto - done)
        sim.run(upto - done)
        done = upto
        ref, got = oracle_state(o), product_state(sim)
        worst = metrics(got, ref)
        for k, v in worst.items():
            assert v <= 1e-12, f"{name} step {upto}: {k} rel L2 {v:.3e}"
        assert got["capped"] == ref["capped"], (upto, got["capped"], ref["capped"])
        lo, hi = 1.5, case["ny"] - 1.5
        assert np.sum(got["y"] == lo) == np.sum(ref["y"] == lo)
        assert np.sum(got["y"] == hi) == np.sum(ref["y"] == hi)
        if "walls" in name:
            assert got["capped"] > 0
            if upto == 1:
                assert np.sum(got["y"] == lo) > 0 and np.sum(got["y"] == hi) > 0
        if "recirculate" in name:
            # the device moved the same capsules across the seam as the oracle
            assert np.array_equal(np.floor(got["x"] / case["nx"]), np.floor(ref["x"] / case["nx"]))
