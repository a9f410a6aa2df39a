"""Workloads of the per-timestep hot path, built through the reference's
Simulation API (softflow_core), shared by bench.py and the GPU tests.

- ``large_domain_case`` / ``build_product``: the 4096 x 1024 single domain
  with 2000 bidisperse capsules and the scalar; ``pillar_case`` adds 200
  circular pillars with interpolated bounce-back.
- ``MDF_CHANNEL`` / ``build_mdf_channel``: the 2048 x 512 multi-direct-forcing
  channel of the learned warm start.
- the batched ensemble of independent channels is ``softflow_amd.ensemble``.

Capsules are placed by the reference's seeded rejection placement
(Simulation::addCapsuleRandom, simulation.cpp:220-274).
"""
from __future__ import annotations

import numpy as np

NX, NY = 4096, 1024


def large_domain_case():
    """4096 x 1024 channel, MRT (tau 0.7, s_e = s_eps = s_q = 1.4), body force
    1e-5, periodic x with half-way bounce-back walls in y; 2000 bidisperse
    Skalak capsules with Helfrich bending (k_b 1e-3): 800 of radius 6 with 38
    nodes and G_s 0.03 placed first (seed 101), then 1200 of radius 3 with 19
    nodes and G_s 0.06 (seed 202), C = 10, uniform and non-overlapping over the
    channel; an advection-diffusion scalar (D 0.05) from a unit Gaussian blob of
    sigma 20 at the centre."""
    big = dict(count=800, region=(0, 0, NX, NY), rmin=6.0, rmax=6.0, nodes=38, seed=101,
               mem=dict(model=2, G_s=0.03, C_skalak=10.0, use_helfrich=1, k_bend=1e-3))
    small = dict(count=1200, region=(0, 0, NX, NY), rmin=3.0, rmax=3.0, nodes=19, seed=202,
                 mem=dict(model=2, G_s=0.06, C_skalak=10.0, use_helfrich=1, k_bend=1e-3))
    return dict(nx=NX, ny=NY, tau=0.7, bfx=1e-5, coll=1, capsules=[big, small],
                scalar=dict(D=0.05, blob=(NX / 2, NY / 2, 20.0)))


def pillar_case():
    """4096x1024 channel through a staggered array of 200 circular pillars (radius 24,
    160 apart in x, 128 in y) with interpolated bounce-back, 1000 capsules placed
    around them, MRT, the scalar on: the large domain with obstacles."""
    pillars = []
    for row in range(8):
        y = 64.0 + 128.0 * row + 0.21
        x = (80.0 if row % 2 == 0 else 160.0) + 0.37
        while x < NX - 40:
            pillars.append(("circle", x, y, 24.0))
            x += 160.0
    big = dict(count=400, region=(0, 0, NX, NY), rmin=6.0, rmax=6.0, nodes=38, seed=303,
               mem=dict(model=2, G_s=0.03, C_skalak=10.0, use_helfrich=1, k_bend=1e-3))
    small = dict(count=600, region=(0, 0, NX, NY), rmin=3.0, rmax=3.0, nodes=19, seed=404,
                 mem=dict(model=2, G_s=0.06, C_skalak=10.0, use_helfrich=1, k_bend=1e-3))
    return dict(nx=NX, ny=NY, tau=0.7, bfx=1e-5, coll=1, capsules=[big, small], obstacles=pillars,
                ibb=True, scalar=dict(D=0.05, blob=(NX / 2, NY / 2, 20.0)))


def build_product(case, precision="fp32", graph=True):
    """One initialised Simulation of a channel case (large_domain_case,
    pillar_case): lattice storage and collide in FP32 (populations stored as
    f - w_q) or FP64; capsules and their forces in FP64 either way."""
    from softflow_amd import softflow_core as sc

    p = sc.SimulationParams()
    p.output_interval, p.probe_dump_every, p.stats_dump_every = 0, 0, 0   # no output files
    for k, v in case.get("output", {}).items():
        setattr(p, k, v)
    p.nx, p.ny = case["nx"], case["ny"]
    p.fluid.tau = case["tau"]
    p.fluid.boundary_type = sc.BoundaryType.PERIODIC
    p.fluid.body_force_x = case["bfx"]
    p.fluid.collision_model = [sc.CollisionModel.BGK, sc.CollisionModel.MRT,
                               sc.CollisionModel.REGULARIZED][case["coll"]]
    p.scalar.enabled = True
    p.scalar.diffusivity = [case["scalar"]["D"]]
    p.enable_stability_checks = False
    p.precision = sc.Precision.FP32 if precision == "fp32" else sc.Precision.FP64
    p.use_graph = graph
    p.fluid.use_interpolated_bb = bool(case.get("ibb", False))
    sim = sc.Simulation(p)
    b = sc.ChannelBuilder(p.nx, p.ny)
    b.addWalls()
    for ob in case.get("obstacles", []):
        b.addCirclePillar(*ob[1:])
    sim.setChannelBuilder(b)
    for spec in case["capsules"]:
        mp = sc.MembraneParams()
        m = spec["mem"]
        mp.model = sc.MembraneModel.SKALAK
        mp.G_s, mp.C_skalak = m["G_s"], m["C_skalak"]
        mp.use_helfrich_bending = bool(m["use_helfrich"])
        mp.k_bend = m["k_bend"]
        x0, y0, x1, y1 = spec["region"]
        sim.addCapsuleRandom(spec["count"], x0, y0, x1, y1, spec["rmin"], spec["rmax"],
                             spec["nodes"], mp, 0, spec["seed"], 1.0)
    sim.initialize()
    cx, cy, sg = case["scalar"]["blob"]
    yy, xx = np.mgrid[0:p.ny, 0:p.nx]
    blob = np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2.0 * sg * sg))
    ad = sim.advectionDiffusion()
    ad.initialize(0.0)
    ad.setField(blob.ravel(), 0)
    return sim




# Learned MDF warm start workload: 2048 x 512 channel, regularized tau = 0.8;
# 500 Skalak capsules (G_s = 0.06, C = 1) of radius 4 with 25 nodes seeded
# (seed 5) over the whole channel (2.4 % area fraction).  Body force 3e-6: at
# the 3e-5 of the original specification the reference's plain multi-direct
# forcing (five iterations) diverges within about 1750 steps as the plug
# accelerates, and so do C = 10 and a 30 % seeding band sooner (the stability
# record of bench.py --only mdf-stability); at 3e-6 it runs 20 000 steps.
MDF_CHANNEL = dict(nx=2048, ny=512, tau=0.8, bfx=3e-6, count=500, radius=4.0, nodes=25, seed=5,
                   region=(4, 8, 2044, 504), mem={"C_skalak": 1.0})


def build_mdf_channel(precision, ibm_iterations, graph=True, **changes):
    """The learned-MDF channel (MDF_CHANNEL, with ``changes`` applied: bfx,
    region, mem) with ``ibm_iterations`` fixed MDF iterations; returns
    (simulation, capsules placed)."""
    from softflow_amd import softflow_core as sc

    ch = dict(MDF_CHANNEL, **changes)
    p = sc.SimulationParams()
    p.output_interval, p.probe_dump_every, p.stats_dump_every = 0, 0, 0   # no output files
    p.nx, p.ny = ch["nx"], ch["ny"]
    p.fluid.tau = ch["tau"]
    p.fluid.boundary_type = sc.BoundaryType.PERIODIC
    p.fluid.body_force_x = ch["bfx"]
    p.fluid.collision_model = sc.CollisionModel.REGULARIZED
    p.ibm_iterations = ibm_iterations
    p.enable_stability_checks = False
    p.precision = sc.Precision.FP32 if precision == "fp32" else sc.Precision.FP64
    p.use_graph = graph
    sim = sc.Simulation(p)
    b = sc.ChannelBuilder(p.nx, p.ny)
    b.addWalls()
    sim.setChannelBuilder(b)
    mp = sc.MembraneParams()
    mp.model = sc.MembraneModel.SKALAK
    mp.G_s, mp.C_skalak = 0.06, 10.0
    for k, v in ch.get("mem", {}).items():
        setattr(mp, k, v)
    x0, y0, x1, y1 = ch["region"]
    placed = sim.addCapsuleRandom(ch["count"], x0, y0, x1, y1, ch["radius"], ch["radius"], ch["nodes"],
                                  mp, 0, ch["seed"], 1.0)
    sim.initialize()
    return sim, placed
