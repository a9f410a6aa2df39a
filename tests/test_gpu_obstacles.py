"""The channel builder's pillars on the device path.

- cell flags from the pillars (ChannelBuilder::applyToField), capsule placement
  around them, the obstacle contact forces after the wall images
  (RepulsionForce::computeObstacleRepulsion, repulsion.cpp:180-236) and, with
  use_interpolated_bb, the Bouzidi links (interpolated_bounce_back.cpp) applied
  after each collide-stream: FP64 against the oracle at the parity bar
  (relative L2 <= 1e-12), bit for bit on the lattice-only cases;
- FP32 storage against the FP64 oracle within the tolerances written below;
- graph replay equals eager launches with the serial link kernel inside;
- a link nearer than half a cell whose upstream cell is SOLID (a fluid gap one
  cell wide) is refused, and batches refuse members with obstacles.
"""
import numpy as np
import pytest

from tests.cases import OBSTACLE_CASES, oracle_sim, oracle_state, product_sim, product_state, rel_l2

pytestmark = pytest.mark.gpu

LATTICE_ONLY = ("ibb_walls_only", "ibb_periodic_seam", "ibb_two_cell_gap", "ibb_inlet_outlet")


def _case(name):
    case = dict(dict(tau=0.8, bfx=3e-5), **OBSTACLE_CASES[name])
    if case.get("boundary") == "inlet_outlet":
        case["bfx"] = 0.0
    steps = case.pop("steps", 200)
    case["name"] = name
    return case, steps


@pytest.mark.parametrize("name", sorted(OBSTACLE_CASES))
def test_obstacles_fp64(name):
    # the step counts stay short of where the reference's own scheme loses
    # stability (the regularized MDF case diverges after ~450 steps)
    case, steps = _case(name)
    print(f"obstacles {name} fp64 steps={steps}", flush=True)
    o = oracle_sim(case)
    o.step(steps)
    ref = oracle_state(o)
    sim = product_sim(case)
    sim.run(steps)
    got = product_state(sim)
    assert len(got["x"]) == len(ref["x"])
    assert np.array_equal(np.asarray(sim.field().cell_types()).ravel(), o.flags().ravel())
    assert np.isfinite(ref["rho"]).all()
    for k in ("rho", "ux", "uy", "x", "y", "vx", "vy", "C"):
        if k in ref and np.asarray(ref[k]).size:
            e = rel_l2(got[k], ref[k])
            print(f"  {k}: rel L2 {e:.3e} bitwise {np.array_equal(got[k], ref[k])}", flush=True)
            assert e <= 1e-12, (name, k, e)
    assert got["capped"] == ref["capped"]
    if name in LATTICE_ONLY:
        for k in ("rho", "ux", "uy"):
            assert np.array_equal(got[k], ref[k]), (name, k)


@pytest.mark.parametrize("name", ["ibb_pillars_mrt_caps", "ibb_two_cell_gap"])
def test_obstacles_fp32(name):
    case, steps = _case(name)
    print(f"obstacles {name} fp32 steps={steps}", flush=True)
    o = oracle_sim(case)
    o.step(steps)
    ref = oracle_state(o)
    sim = product_sim(dict(case, precision="fp32"))
    sim.run(steps)
    got = product_state(sim)
    eu = rel_l2(np.concatenate([got["ux"].ravel(), got["uy"].ravel()]),
                np.concatenate([ref["ux"].ravel(), ref["uy"].ravel()]))
    er = rel_l2(got["rho"], ref["rho"])
    print(f"  fp32: u rel L2 {eu:.3e}, rho rel L2 {er:.3e}", flush=True)
    assert eu < 2e-3 and er < 1e-5
    if len(ref["x"]):
        ex = rel_l2(np.concatenate([got["x"], got["y"]]), np.concatenate([ref["x"], ref["y"]]))
        print(f"  fp32: node positions rel L2 {ex:.3e}", flush=True)
        assert ex < 1e-5


def test_obstacles_graph_matches_eager():
    case, _ = _case("ibb_two_cell_gap")
    case = dict(case, obstacles=case["obstacles"] + [("circle", 36.0, 4.0, 1.5)],
                capsules=[dict(centers=[(6.0, 4.0)], rmin=1.8, nodes=12,
                               mem=dict(model=2, G_s=0.06, C_skalak=10.0))])
    a = product_sim(case, graph=True)
    a.run(41)
    b = product_sim(case, graph=False)
    b.run(41)
    sa, sb = product_state(a), product_state(b)
    for k in ("rho", "ux", "uy", "x", "y"):
        assert np.array_equal(sa[k], sb[k]), k


def test_one_cell_gap_is_refused():
    # a block leaving one fluid row above the bottom wall, its face 0.3 cells
    # from that row: the links read the wall row's populations
    case = dict(nx=40, ny=10, tau=0.8, bfx=2e-5, coll=0, ibb=True,
                obstacles=[("rect", 10.0, 1.3, 30.0, 5.0)])
    with pytest.raises(Exception, match="one cell wide"):
        product_sim(case)


def test_batch_refuses_obstacles():
    from softflow_amd import softflow_core as sc

    plain = dict(nx=60, ny=24, tau=0.8, bfx=2e-5, coll=0, precision="fp32",
                 capsules=[dict(count=2, region=(5, 5, 55, 19), rmin=3.0, rmax=3.0, seed=4,
                                mem=dict(model=2, G_s=0.06, C_skalak=10.0))])
    with_ibb = dict(plain, ibb=True)
    ens = sc.Ensemble([product_sim(plain), product_sim(with_ibb)])
    with pytest.raises(RuntimeError, match="interpolated bounce-back"):
        ens.run(2)
