"""Host-visible diagnostics from FP64 device reductions.

- ``globalStats`` carries GlobalStatsWriter::writeTimestep's quantities
  (io/global_stats_writer.cpp:41-124) and equals them computed from the oracle's
  fields on the same seeded case (sums to 1e-12: the device adds in another
  order; minima and maxima exactly);
- ``checkStability`` (simulation.cpp:820-890) runs on the device reductions:
  it passes on a healthy run and fails on a velocity over the limit and on a
  non-finite density.
"""
import math

import numpy as np
import pytest

from tests.cases import oracle_sim, oracle_state, product_sim

pytestmark = pytest.mark.gpu

CASE = dict(nx=120, ny=40, tau=0.8, bfx=3e-5, coll=2, precision="fp64",
            capsules=[dict(count=8, region=(0, 0, 120, 40), rmin=3.0, rmax=4.0, seed=11,
                           mem=dict(model=2, G_s=0.06, C_skalak=10.0))],
            scalar=dict(D=0.05, blob=(60, 20, 6.0)))


def test_global_stats_match_oracle_fields():
    steps = 150
    o = oracle_sim(CASE)
    o.step(steps)
    ref = oracle_state(o)
    p = product_sim(CASE)
    p.run(steps)
    g = p.globalStats()
    fluid = np.ones((CASE["ny"], CASE["nx"]), bool)
    fluid[0, :] = fluid[-1, :] = False                  # the channel walls
    r, u, v = (np.asarray(ref[k]).reshape(fluid.shape)[fluid] for k in ("rho", "ux", "uy"))
    v2 = u * u + v * v
    n = fluid.sum()
    expect = {"mean_rho": r.sum() / n, "mean_velocity": np.sqrt(v2).sum() / n,
              "total_kinetic_energy": (0.5 * r * v2).sum()}
    for k, e in expect.items():
        assert abs(g[k] - e) <= 1e-12 * abs(e), (k, g[k], e)
    assert g["max_velocity"] == math.sqrt(v2.max())
    assert g["min_rho"] == r.min() and g["max_rho"] == r.max()
    assert g["timestep"] == steps and g["nonfinite_cells"] == 0
    caps = p.capsules()
    D = [caps[i].deformationIndex() for i in range(len(caps))]
    assert g["n_active_particles"] == len(caps)
    assert g["max_deformation"] == max(D)
    assert abs(g["mean_deformation"] - sum(D) / len(D)) <= 1e-15
    C = np.asarray(p.advectionDiffusion().concentration(0))
    assert abs(g["total_dissolved_mass"] - C.sum()) <= 1e-12 * C.sum()
    assert abs(g["total_dissolved_mass"] - np.asarray(ref["C"]).sum()) <= 1e-12 * C.sum()


def test_check_stability_on_device():
    sim = product_sim(CASE)
    sim.run(50)
    assert sim.checkStability()
    slow = product_sim(dict(CASE, params=dict(max_velocity=1e-6)))
    slow.run(50)
    assert not slow.checkStability()                     # |u| above the limit
    bad = product_sim(CASE)
    bad.run(5)
    # a non-finite density written through the field (no step is taken with it)
    bad.field().setRho(30, 20, float("nan"))
    g = bad.globalStats()
    assert g["nonfinite_cells"] >= 1
    assert not bad.checkStability()
