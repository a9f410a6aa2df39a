"""Shared case construction: the same configuration built in the product
(``softflow_core`` over the HIP kernels) and in the oracle (``oracle/libsfo.so``).

A case is a dict:
    nx, ny, tau, coll (0 BGK, 1 MRT, 2 regularized), bfx, mdf (ibm_iterations),
    boundary ('periodic' | 'closed' | 'inlet_outlet'), periodic_y, precision ('fp64' | 'fp32'),
    inlet: (inlet peak velocity, outlet density) for 'inlet_outlet' (optional),
    capsules: list of {count, region:(x0,y0,x1,y1), rmin, rmax, nodes, seed, min_gap, mem:{...}}
    scalar: {D, blob:(cx, cy, sigma)}      (optional)
    rep: {damping_normal, friction_coeff}  (optional)
    solids: list of ('rect', x0, y0, x1, y1) | ('circle', cx, cy, r): extra SOLID
            cells set after initialize (LatticeField::setCellType)   (optional)
    params: {name: value} further SimulationParams fields (product only)   (optional)
    walls: (top, bottom) x velocities of moving wall rows   (optional)
    obstacles: list of ('circle', cx, cy, r) | ('rect', x0, y0, x1, y1) |
               ('polygon', [(x, y), ...]): the channel builder's pillars   (optional)
    ibb: True for interpolated bounce-back towards them   (optional)
    fmax: max_lattice_force, the step's force cap (default 0.01)   (optional)
    capsule_periodic_x: True for capsule recirculation over a non-periodic fluid   (optional)
Membrane keys follow the oracle's struct (model, G_s, C_skalak, k_bend, use_helfrich, ...).
"""
from __future__ import annotations

import numpy as np

_MEM_RENAME = {"use_helfrich": "use_helfrich_bending"}


def _blob(case):
    s = case.get("scalar")
    if not s:
        return None
    cx, cy, sg = s["blob"]
    nx, ny = case["nx"], case["ny"]
    yy, xx = np.mgrid[0:ny, 0:nx]
    return np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2.0 * sg * sg))


def solid_mask(case):
    """Cells inside the case's extra solids (channel_builder.cpp:26-31 tests)."""
    nx, ny = case["nx"], case["ny"]
    yy, xx = np.mgrid[0:ny, 0:nx].astype(float)
    m = np.zeros((ny, nx), dtype=bool)
    for s in case.get("solids", []):
        if s[0] == "rect":
            x0, y0, x1, y1 = s[1:]
            m |= (xx >= x0) & (xx <= x1) & (yy >= y0) & (yy <= y1)
        else:
            cx, cy, r = s[1:]
            m |= (xx - cx) ** 2 + (yy - cy) ** 2 <= r * r
    return m


def oracle_sim(case, which="oracle"):
    from oracle import pyoracle as po

    s = case.get("scalar")
    rep = case.get("rep", {})
    p = po.default_params(
        nx=case["nx"], ny=case["ny"], tau=case.get("tau", 0.8),
        boundary_type={"closed": 2, "inlet_outlet": 0}.get(case.get("boundary"), 1),
        inlet_velocity=case.get("inlet", (0.02, 1.0))[0], outlet_density=case.get("inlet", (0.02, 1.0))[1],
        periodic_y=1 if case.get("periodic_y") else 0,
        body_force_x=case.get("bfx", 0.0), collision_model=case.get("coll", 0),
        ibm_iterations=case.get("mdf", 1), scalar_enabled=1 if s else 0,
        diffusivity=[s["D"]] if s else [0.01],
        rep_damping_normal=rep.get("damping_normal", 0.0),
        rep_friction_coeff=rep.get("friction_coeff", 0.0),
        top_wall_velocity=case.get("walls", (0.0, 0.0))[0],
        bottom_wall_velocity=case.get("walls", (0.0, 0.0))[1],
        use_interpolated_bb=1 if case.get("ibb") else 0,
        max_lattice_force=case.get("fmax", 0.01),
        capsule_periodic_x=1 if case.get("capsule_periodic_x") else 0,
    )
    sim = po.Sim(p, which)
    for ob in case.get("obstacles", []):
        sim.add_obstacle(ob)
    for spec in case.get("capsules", []):
        m = po.default_membrane(**spec.get("mem", {}))
        if "centers" in spec:
            for (cx, cy) in spec["centers"]:
                sim.add_capsule(cx, cy, spec["rmin"], spec.get("nodes", 0), m, spec.get("type", 0))
        else:
            x0, y0, x1, y1 = spec["region"]
            sim.add_capsule_random(spec["count"], x0, y0, x1, y1, spec["rmin"], spec["rmax"],
                                   spec.get("nodes", 0), m, spec.get("type", 0), spec.get("seed", 42),
                                   spec.get("min_gap", 1.0), 0)
    sim.initialize()
    if case.get("solids"):
        fl = sim.flags().copy()
        fl[solid_mask(case)] = 1
        sim.set_flags(fl)
    blob = _blob(case)
    if blob is not None:
        sim.scalar_initialize(0.0)
        for y in range(case["ny"]):
            for x in range(case["nx"]):
                sim.set_scalar_point(x, y, float(blob[y, x]), 0)
    return sim


def product_sim(case, graph=True):
    from softflow_amd import softflow_core as sc

    s = case.get("scalar")
    rep = case.get("rep", {})
    p = sc.SimulationParams()
    p.output_interval, p.probe_dump_every, p.stats_dump_every = 0, 0, 0   # no output files
    p.nx, p.ny = case["nx"], case["ny"]
    p.fluid.tau = case.get("tau", 0.8)
    p.fluid.boundary_type = {"closed": sc.BoundaryType.CLOSED,
                             "inlet_outlet": sc.BoundaryType.INLET_OUTLET}.get(case.get("boundary"),
                                                                              sc.BoundaryType.PERIODIC)
    p.fluid.inlet_velocity, p.fluid.outlet_density = case.get("inlet", (0.02, 1.0))
    p.fluid.periodic_y = bool(case.get("periodic_y", False))
    p.fluid.body_force_x = case.get("bfx", 0.0)
    p.fluid.collision_model = [sc.CollisionModel.BGK, sc.CollisionModel.MRT,
                               sc.CollisionModel.REGULARIZED][case.get("coll", 0)]
    p.ibm_iterations = case.get("mdf", 1)
    p.repulsion.damping_normal = rep.get("damping_normal", 0.0)
    p.repulsion.friction_coeff = rep.get("friction_coeff", 0.0)
    p.fluid.top_wall_velocity, p.fluid.bottom_wall_velocity = case.get("walls", (0.0, 0.0))
    p.fluid.use_interpolated_bb = bool(case.get("ibb", False))
    p.max_lattice_force = case.get("fmax", 0.01)
    p.fluid.capsule_periodic_x = bool(case.get("capsule_periodic_x", False))
    if s:
        p.scalar.enabled = True
        p.scalar.diffusivity = [s["D"]]
    p.enable_stability_checks = False
    p.precision = sc.Precision.FP32 if case.get("precision") == "fp32" else sc.Precision.FP64
    p.use_graph = graph
    for k, v in case.get("params", {}).items():     # further SimulationParams fields
        setattr(p, k, v)
    sim = sc.Simulation(p)
    b = sc.ChannelBuilder(p.nx, p.ny)
    b.addWalls()
    for ob in case.get("obstacles", []):
        if ob[0] == "circle":
            b.addCirclePillar(*ob[1:])
        elif ob[0] == "rect":
            b.addRectPillar(*ob[1:])
        else:
            b.addPolygonPillar([sc.Vec2d(float(x), float(y)) for x, y in ob[1]])
    sim.setChannelBuilder(b)
    models = [sc.MembraneModel.HOOKEAN, sc.MembraneModel.NEO_HOOKEAN, sc.MembraneModel.SKALAK,
              sc.MembraneModel.WLC]
    shapes = [sc.CapsuleShape.CIRCLE, sc.CapsuleShape.ELLIPSE, sc.CapsuleShape.BICONCAVE,
              sc.CapsuleShape.FIBER]
    for spec in case.get("capsules", []):
        mp = sc.MembraneParams()
        for k, v in spec.get("mem", {}).items():
            if k == "model":
                v = models[v]
            elif k == "shape":
                v = shapes[v]
            setattr(mp, _MEM_RENAME.get(k, k), v)
        if "centers" in spec:
            for (cx, cy) in spec["centers"]:
                sim.addCapsule(sc.Vec2d(cx, cy), spec["rmin"], spec.get("nodes", 0), mp, spec.get("type", 0))
        else:
            x0, y0, x1, y1 = spec["region"]
            sim.addCapsuleRandom(spec["count"], x0, y0, x1, y1, spec["rmin"], spec["rmax"],
                                 spec.get("nodes", 0), mp, spec.get("type", 0), spec.get("seed", 42),
                                 spec.get("min_gap", 1.0))
    sim.initialize()
    if case.get("solids"):
        fld = sim.field()
        ys, xs = np.nonzero(solid_mask(case))
        for y, x in zip(ys, xs):
            fld.setCellType(int(x), int(y), sc.CellType.SOLID)
    blob = _blob(case)
    if blob is not None:
        ad = sim.advectionDiffusion()
        ad.initialize(0.0)
        for y in range(case["ny"]):
            for x in range(case["nx"]):
                ad.setPoint(x, y, float(blob[y, x]), 0)
    return sim


def oracle_state(sim):
    out = dict(sim.field())
    n = sim.nodes()
    out.update(x=n["x"], y=n["y"], vx=n["vx"], vy=n["vy"], fx_node=n["fx"], fy_node=n["fy"])
    if sim.params.scalar_enabled:
        out["C"] = sim.scalar()
    out["capped"] = sim.last_capped()
    return out


def product_state(sim):
    f = sim.field()
    out = {"rho": np.array(f.density()), "ux": np.array(f.velocity_x()),
           "uy": np.array(f.velocity_y()), "fx": np.array(f.force_x()), "fy": np.array(f.force_y())}
    caps = sim.capsules()
    xs, ys, vxs, vys, fxs, fys = [], [], [], [], [], []
    for i in range(len(caps)):
        c = caps[i]
        for k in range(c.numNodes()):
            p, v, fo = c.nodePosition(k), c.nodeVelocity(k), c.nodeForce(k)
            xs.append(p.x); ys.append(p.y); vxs.append(v.x); vys.append(v.y)
            fxs.append(fo.x); fys.append(fo.y)
    out.update(x=np.array(xs), y=np.array(ys), vx=np.array(vxs), vy=np.array(vys),
               fx_node=np.array(fxs), fy_node=np.array(fys))
    ad = sim.advectionDiffusion()
    if ad is not None:
        out["C"] = np.array(ad.concentration(0))
    out["capped"] = sim.lastCappedNodes()
    return out


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    if a.size == 0:
        return 0.0
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    return float(num / den) if den > 0 else float(num)


# Channels with the builder's pillars (circle, rectangle, polygon), with and
# without interpolated bounce-back; run by the oracle-vs-reference and the
# device parity tests.
PILLARS = [("circle", 40.3, 20.6, 6.2), ("rect", 70.0, 9.5, 78.0, 17.2),
           ("polygon", [(100.0, 24.0), (112.0, 21.5), (108.0, 31.0)])]
OBSTACLE_CASES = dict(
    pillars_caps=dict(nx=130, ny=40, coll=0, obstacles=PILLARS,
                      capsules=[dict(count=10, region=(5, 5, 125, 35), rmin=3.0, rmax=4.0, seed=3,
                                     mem=dict(model=2, G_s=0.06, C_skalak=10.0))],
                      rep=dict(damping_normal=0.01, friction_coeff=0.1)),
    ibb_pillars_mrt_caps=dict(nx=130, ny=40, coll=1, obstacles=PILLARS, ibb=True,
                              capsules=[dict(count=8, region=(5, 5, 125, 35), rmin=3.0, rmax=4.0, seed=5,
                                             mem=dict(model=2, G_s=0.06, C_skalak=10.0))]),
    ibb_reg_mdf3_adr=dict(nx=130, ny=40, coll=2, mdf=3, obstacles=PILLARS[1:], ibb=True,
                          scalar=dict(D=0.05, blob=(30, 20, 5.0)),
                          capsules=[dict(count=6, region=(5, 5, 125, 35), rmin=3.0, rmax=4.0, seed=9,
                                         mem=dict(model=2, G_s=0.06, C_skalak=10.0))]),
    ibb_walls_only=dict(nx=60, ny=24, coll=0, ibb=True),
    ibb_periodic_seam=dict(nx=60, ny=24, coll=0, obstacles=[("circle", 3.5, 11.8, 3.2)], ibb=True),
    # a fluid run two cells long under a block: the links nearer than half a
    # cell read slots the opposite links of their upstream cells write
    ibb_two_cell_gap=dict(nx=40, ny=8, coll=1, obstacles=[("rect", 10.0, 2.3, 30.0, 5.0)], ibb=True),
    ibb_inlet_outlet=dict(nx=90, ny=30, coll=0, boundary="inlet_outlet", inlet=(0.03, 1.0),
                          obstacles=[("circle", 45.2, 14.7, 5.3)], ibb=True),
    # capsules placed against each pillar: contact with damping and friction
    obstacle_contact=dict(nx=130, ny=40, coll=0, obstacles=PILLARS,
                          capsules=[dict(centers=[(40.3, 31.4), (74.0, 5.2), (81.9, 13.0), (104.0, 19.3)],
                                         rmin=4.0, nodes=26, mem=dict(model=2, G_s=0.06, C_skalak=10.0))],
                          rep=dict(damping_normal=0.01, friction_coeff=0.1)),
    # capsules overlapping the pillars: nodes inside take the push-out branch
    obstacle_inside=dict(nx=130, ny=40, coll=0, obstacles=PILLARS, steps=60,
                         capsules=[dict(centers=[(40.3, 29.6), (74.0, 21.5), (109.5, 34.0)],
                                        rmin=4.0, nodes=26, mem=dict(model=2, G_s=0.06, C_skalak=10.0))],
                         rep=dict(damping_normal=0.01, friction_coeff=0.1)),
)


# Cases that drive the three operations of Simulation::step itself: the force
# cap fires in every step (capsules pressed into both walls), nodes are pushed
# past the wall lines and clamped, and capsules recirculate in x over a
# non-periodic fluid (inlet/outlet, closed box).  Run by the oracle-vs-
# reference, golden and device parity tests.
_SK = dict(model=2, G_s=0.06, C_skalak=10.0)
STEP_PIN_CASES = dict(
    walls_capped_clamped=dict(nx=48, ny=24, tau=0.8, bfx=3e-5, coll=0, steps=300,
                              capsules=[dict(centers=[(12.0, 2.0), (30.0, 23.0)], rmin=3.5, nodes=20,
                                             mem=_SK)]),
    walls_capped_mrt=dict(nx=48, ny=24, tau=0.7, bfx=1e-5, coll=1, fmax=2e-3, steps=200,
                          capsules=[dict(centers=[(10.0, 2.5), (26.0, 12.0), (40.0, 21.5)], rmin=3.5,
                                         nodes=20, mem=_SK)]),
    inlet_outlet_recirculate=dict(nx=40, ny=20, tau=0.8, bfx=0.0, coll=0, boundary="inlet_outlet",
                                  inlet=(0.05, 1.0), capsule_periodic_x=True, steps=100,
                                  capsules=[dict(centers=[(37.5, 10.0), (-0.05, 10.0), (20.0, 10.0)],
                                                 rmin=3.0, nodes=16, mem=_SK)]),
    closed_recirculate=dict(nx=40, ny=20, tau=0.8, bfx=3e-5, coll=2, boundary="closed",
                            capsule_periodic_x=True, steps=60,
                            capsules=[dict(centers=[(40.02, 10.0), (-0.3, 8.0)], rmin=3.0, nodes=16,
                                           mem=_SK)]),
)
