"""Host-side checks of the product that need no GPU.

- the C-ABI library loads and exports every entry point include/softflow.h
  declares (no compute calls are made);
- without a device, context creation fails cleanly with an error string;
- argument validation happens before any device call;
- the host engine (softflow_core) keeps the reference's parameter defaults,
  its seeded capsule placement and rest geometry, compared with the oracle;
- features outside the device path are refused at initialize().
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "softflow.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(sf_\w+)\s*\(", text, flags=re.M))


def test_header_declares_capi_exports():
    from softflow_amd import capi
    syms = header_symbols()
    assert len(syms) >= 25
    assert syms == set(capi.EXPORTS)


def test_library_exports_every_header_symbol():
    from softflow_amd import capi
    L = C.CDLL(capi.LIB_PATH)
    for s in header_symbols():
        assert hasattr(L, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sf_\w+)$", out, flags=re.M))
    assert header_symbols() <= exported


def test_argument_validation_before_device():
    from softflow_amd import capi
    L = capi.load()
    h = C.c_void_p()
    for kw, msg in ((dict(nx=2), "3x3"), (dict(tau=0.5), "tau"), (dict(boundary_type=0), "outlet density"), (dict(boundary_type=7), "boundary"),
                    (dict(collision_model=5), "collision"), (dict(precision=7), "precision"),
                    (dict(ibm_iterations=0), "ibm_iterations")):
        p = capi.make_params(nx=16, ny=8, **{k: v for k, v in kw.items() if k != "nx"})
        if "nx" in kw:
            p.nx = kw["nx"]
        rc = L.sf_create(C.byref(h), 0, C.byref(p))
        assert rc == capi.SF_E_ARG
        assert msg in L.sf_last_error().decode()
        assert not h.value
    assert L.sf_create(None, 0, None) == capi.SF_E_ARG
    assert L.sf_set_obstacles(None, 0, None, None, None) == capi.SF_E_ARG
    assert L.sf_set_capsule_launch(None, 0, 16) == capi.SF_E_ARG
    g = C.c_double()
    for nbytes, reps, out, msg in ((1 << 30, 10, None, "null"), (1000, 10, C.byref(g), "multiple of 16"),
                                   ((1 << 20) + 8, 10, C.byref(g), "multiple of 16"),
                                   (1 << 30, 0, C.byref(g), "reps")):
        assert L.sf_copy_ceiling(0, nbytes, reps, out) == capi.SF_E_ARG
        assert msg in L.sf_last_error().decode()


def test_batch_argument_validation_before_device():
    from softflow_amd import capi
    L = capi.load()
    b = C.c_void_p()
    assert L.sf_batch_create(C.byref(b), None, 1) == capi.SF_E_ARG
    arr = (C.c_void_p * 2)(None, None)
    assert L.sf_batch_create(C.byref(b), arr, 0) == capi.SF_E_ARG
    assert "at least one member" in L.sf_last_error().decode()
    assert L.sf_batch_create(C.byref(b), arr, 2) == capi.SF_E_ARG
    assert not b.value
    fake = (C.c_void_p * 2)(0x1000, 0x1000)     # never dereferenced: duplicates are refused first
    assert L.sf_batch_create(C.byref(b), fake, 2) == capi.SF_E_ARG
    assert "distinct" in L.sf_last_error().decode()
    for fn, args in (("sf_batch_step", (None, 1)), ("sf_batch_synchronize", (None,)),
                     ("sf_batch_set_graph_mode", (None, 1)), ("sf_batch_set_lbm_launch", (None, 64, 4, 1)),
                     ("sf_batch_info", (None, None))):
        assert getattr(L, fn)(*args) == capi.SF_E_ARG, fn
    assert L.sf_batch_destroy(None) == capi.SF_OK


@pytest.mark.skipif(os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"),
                    reason="a GPU may be present")
def test_create_fails_cleanly_without_device():
    from softflow_amd import capi
    assert capi.device_count() == 0
    with pytest.raises(capi.SoftFlowError):
        capi.Context(capi.make_params(nx=16, ny=8))


def test_engine_defaults_match_reference():
    from oracle import pyoracle as po
    from softflow_amd import softflow_core as sc
    p, o = sc.SimulationParams(), po.default_params()
    assert (p.nx, p.ny, p.dt, p.ibm_iterations) == (o.nx, o.ny, o.dt, o.ibm_iterations)
    f = p.fluid
    assert (f.tau, f.rho0, f.mrt_se, f.mrt_s_eps, f.mrt_sq) == (o.tau, o.rho0, o.mrt_se, o.mrt_s_eps, o.mrt_sq)
    assert p.max_lattice_force == o.max_lattice_force
    r = p.repulsion
    assert (r.epsilon, r.sigma, r.r_cut, r.power) == (o.rep_epsilon, o.rep_sigma, o.rep_r_cut, o.rep_power)
    m, om = sc.MembraneParams(), po.default_membrane()
    for k in ("k_stretch", "G_s", "C_skalak", "wlc_L_max_ratio", "k_bend", "k_area", "k_perimeter",
              "gamma_visc", "eta_membrane", "aspect_ratio", "indent_depth"):
        assert getattr(m, k) == getattr(om, k), k
    assert abs(p.kinematicViscosity() - (0.8 - 0.5) / 3) < 1e-16


@pytest.mark.parametrize("shape,nodes", [(0, 0), (1, 24), (2, 0)])
def test_random_placement_and_rest_geometry_match_oracle(shape, nodes):
    from oracle import pyoracle as po
    from softflow_amd import softflow_core as sc
    p = sc.SimulationParams()
    p.nx, p.ny = 160, 60
    p.fluid.boundary_type = sc.BoundaryType.PERIODIC
    sim = sc.Simulation(p)
    b = sc.ChannelBuilder(p.nx, p.ny)
    b.addWalls()
    sim.setChannelBuilder(b)        # the wall-distance test of the placement needs the walls
    mp = sc.MembraneParams()
    mp.model = sc.MembraneModel.SKALAK
    mp.shape = [sc.CapsuleShape.CIRCLE, sc.CapsuleShape.ELLIPSE, sc.CapsuleShape.BICONCAVE][shape]
    mp.aspect_ratio = 0.6
    placed = sim.addCapsuleRandom(20, 8, 6, 152, 54, 3.0, 5.0, nodes, mp, 0, 1234, 1.0)
    o = po.Sim(po.default_params(nx=160, ny=60, boundary_type=1))
    oplaced = o.add_capsule_random(20, 8, 6, 152, 54, 3.0, 5.0, nodes,
                                   po.default_membrane(model=2, shape=shape, aspect_ratio=0.6), 0, 1234, 1.0, 0)
    o.initialize()
    assert placed == oplaced > 5
    nd = o.nodes()
    off = o.offsets()
    caps = sim.capsules()
    assert len(caps) == placed
    for c in range(placed):
        cp = caps[c]
        assert cp.numNodes() == off[c + 1] - off[c]
        xy = np.array(cp.node_positions_array())
        assert np.array_equal(xy[:, 0], nd["x"][off[c]:off[c + 1]])
        assert np.array_equal(xy[:, 1], nd["y"][off[c]:off[c + 1]])
        assert cp.area() == o.geometry(c, 0)
        assert cp.perimeter() == o.geometry(c, 1)
        assert cp.deformationIndex() == o.geometry(c, 4)


def _pillar_builder(sc, nx, ny):
    from tests.cases import PILLARS
    b = sc.ChannelBuilder(nx, ny)
    b.addWalls()
    for ob in PILLARS:
        if ob[0] == "circle":
            b.addCirclePillar(*ob[1:])
        elif ob[0] == "rect":
            b.addRectPillar(*ob[1:])
        else:
            b.addPolygonPillar([sc.Vec2d(x, y) for x, y in ob[1]])
    return b


def test_obstacle_geometry_matches_oracle():
    # the engine's obstacles (the geometry the device path evaluates) against
    # the oracle's restatement, on a grid crossing every face, vertex and interior
    from softflow_amd import softflow_core as sc
    from tests.cases import PILLARS, oracle_sim
    obs = _pillar_builder(sc, 130, 40).getObstacles()
    o = oracle_sim(dict(nx=130, ny=40, obstacles=PILLARS))
    for i, ob in enumerate(obs):
        for x in np.linspace(0.0, 130.0, 263):
            for y in np.linspace(0.0, 40.0, 83):
                d, n = o.obstacle_distance(i, x, y)
                nn = ob.normalAt(x, y)
                assert ob.signedDistance(x, y) == d and (nn.x, nn.y) == (n[0], n[1]), (i, x, y)
                np_ = ob.nearestPoint(x, y)
                assert np.isfinite(np_.x) and np.isfinite(np_.y)


def test_capsule_placement_avoids_obstacles():
    # Simulation::addCapsuleRandom with a channel builder: the obstacle test
    # (signedDistance < r + min_gap, simulation.cpp:256-261) rejects the same
    # draws as the oracle's
    from softflow_amd import softflow_core as sc
    from tests.cases import PILLARS
    p = sc.SimulationParams()
    p.nx, p.ny = 130, 40
    p.fluid.boundary_type = sc.BoundaryType.PERIODIC
    sim = sc.Simulation(p)
    sim.setChannelBuilder(_pillar_builder(sc, 130, 40))
    mp = sc.MembraneParams()
    mp.model = sc.MembraneModel.SKALAK
    placed = sim.addCapsuleRandom(25, 2, 2, 128, 38, 3.0, 4.5, 0, mp, 0, 77, 1.0)
    from oracle import pyoracle as po
    o = po.Sim(po.default_params(nx=130, ny=40, boundary_type=1))
    for ob in PILLARS:
        o.add_obstacle(ob)
    oplaced = o.add_capsule_random(25, 2, 2, 128, 38, 3.0, 4.5, 0, po.default_membrane(model=2), 0, 77, 1.0, 0)
    assert placed == oplaced > 5
    nd = o.nodes()
    xs = np.concatenate([np.array(c.node_positions_array())[:, 0] for c in sim.capsules()])
    assert np.array_equal(xs, nd["x"])
    # no capsule centre within its radius + gap of a pillar
    caps = sim.capsules()
    for c in range(len(caps)):
        cx, cy = caps[c].centroid().x, caps[c].centroid().y
        for i in range(len(PILLARS)):
            assert o.obstacle_distance(i, cx, cy)[0] > 3.0


@pytest.mark.parametrize("what", ["gravity", "lubrication"])
def test_unsupported_features_refused(what):
    from softflow_amd import softflow_core as sc
    p = sc.SimulationParams()
    p.nx, p.ny = 32, 16
    p.fluid.boundary_type = sc.BoundaryType.PERIODIC
    if what == "gravity":
        p.fluid.gravity_y = -1e-6
    else:
        p.lubrication.enabled = True
    sim = sc.Simulation(p)
    with pytest.raises(ValueError, match="not implemented"):
        sim.initialize()
