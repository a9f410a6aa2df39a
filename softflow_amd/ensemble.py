"""Ensembles of independent capsule channels for the batched stepper.

``ensemble_specs`` draws members of uneven size from one seed: lattice
nx x ny, capsule count, viscosity and body force per member.  ``build_member``
makes one initialised ``softflow_core.Simulation`` from a spec (x-periodic
channel with half-way bounce-back walls, BGK, Skalak capsules of random
radius placed by the reference's seeded rejection placement,
simulation.cpp:220-274, with the node count fixed by the radius,
capsule.h:14-17).  ``softflow_core.Ensemble(members)`` then advances them
together.
"""
from __future__ import annotations

import numpy as np


def ensemble_specs(count=256, seed=7, nx=(200, 800), ny=tuple(range(40, 161, 8)),
                   capsules=(5, 60), radius=(2.5, 6.0), tau=(0.6, 1.0), fx=(1e-5, 5e-5)):
    """Member specs: nx uniform in [nx0, nx1], ny one of ``ny``, capsule count
    uniform in [c0, c1], tau and fx uniform; each member gets its own
    placement seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        out.append(dict(
            nx=int(rng.integers(nx[0], nx[1] + 1)),
            ny=int(rng.choice(np.asarray(ny))),
            capsules=int(rng.integers(capsules[0], capsules[1] + 1)),
            rmin=float(radius[0]), rmax=float(radius[1]),
            tau=float(rng.uniform(*tau)),
            fx=float(rng.uniform(*fx)),
            seed=int(rng.integers(1, 2 ** 31 - 1)),
        ))
    return out


def build_member(spec, precision="fp32", graph=True):
    """One initialised Simulation for a member spec; returns (sim, placed)."""
    from softflow_amd import softflow_core as sc

    p = sc.SimulationParams()
    p.output_interval, p.probe_dump_every, p.stats_dump_every = 0, 0, 0   # no output files
    p.nx, p.ny = spec["nx"], spec["ny"]
    p.fluid.tau = spec["tau"]
    p.fluid.boundary_type = sc.BoundaryType.PERIODIC
    p.fluid.body_force_x = spec["fx"]
    p.fluid.collision_model = sc.CollisionModel.BGK
    p.enable_stability_checks = False
    p.precision = sc.Precision.FP32 if precision == "fp32" else sc.Precision.FP64
    p.use_graph = graph
    sim = sc.Simulation(p)
    b = sc.ChannelBuilder(p.nx, p.ny)
    b.addWalls()
    sim.setChannelBuilder(b)
    mp = sc.MembraneParams()
    mp.model = sc.MembraneModel.SKALAK
    placed = sim.addCapsuleRandom(spec["capsules"], 0, 0, p.nx, p.ny, spec["rmin"], spec["rmax"],
                                  0, mp, 0, spec["seed"], 1.0)
    sim.initialize()
    return sim, placed
