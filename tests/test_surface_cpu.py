"""The Python surface the reference's front end uses, on the product module.

tests/golden/surface/setup_py_names.json lists every softflow_core name that
python/pysoftflow/setup.py reads, writes or calls (module classes and enum
members, Simulation methods, parameter fields, methods of the objects the
Simulation hands out), generated from that file by make_surface_names.py in
the same folder.  Each must exist on the product's softflow_core, except the
names of subsystems this build leaves out on purpose (EXPECTED_ABSENT), which
must stay absent so the list of gaps is exact.  No GPU is touched: classes
and parameter structs are inspected, not run.
"""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "surface", "setup_py_names.json")) as fh:
    NAMES = json.load(fh)["names"]

# Out of scope (not on the per-timestep hot path): particle chemistry (release,
# absorption, leaching, adsorption, per-capsule tracking), adhesion bonds, the
# free surface and the segregation metrics.  Their parameter fields exist and
# initialize() refuses them when enabled.
EXPECTED_ABSENT = {
    "AdhesionModel:getBonds": "adhesion bonds",
    "Simulation:adhesion": "adhesion bonds",
    "AdvectionDiffusion:getCapsuleAbsorbed": "particle chemistry",
    "AdvectionDiffusion:getCapsuleReleased": "particle chemistry",
    "AdvectionDiffusion:numTrackedCapsules": "particle chemistry",
    "Simulation:setAdsorptionParams": "particle chemistry",
    "Simulation:setLeachingParams": "particle chemistry",
    "Simulation:setParticleMass": "particle chemistry",
    "Simulation:setScalarAbsorptionRate": "particle chemistry",
    "Simulation:setScalarReleaseRate": "particle chemistry",
    "Simulation:enableFreeSurface": "free surface",
    "Simulation:setEmptyRegion": "free surface",
    "Simulation:lastSegregationResults": "segregation metrics",
}


def _resolve(sc, name):
    cls, path = name.split(":")
    if cls == "module":
        obj = sc
    elif cls in ("SimulationParams", "MembraneParams"):
        obj = getattr(sc, cls)()
    else:
        obj = getattr(sc, cls, None)
        if obj is None:
            return False
    for part in path.split("."):
        if not hasattr(obj, part):
            return False
        obj = getattr(obj, part)
    return True


def test_name_list_is_substantial():
    assert len(NAMES) > 100
    assert set(EXPECTED_ABSENT) <= set(NAMES)


@pytest.mark.parametrize("name", NAMES)
def test_setup_py_name_exists(name):
    from softflow_amd import softflow_core as sc
    present = _resolve(sc, name)
    if name in EXPECTED_ABSENT:
        assert not present, f"{name} now exists: take it off the expected-absent list"
    else:
        assert present, name


def test_out_of_scope_features_are_refused():
    from softflow_amd import softflow_core as sc
    for field in ("adhesion", "lubrication", "shan_chen"):
        p = sc.SimulationParams()
        p.nx, p.ny = 16, 8
        getattr(p, field).enabled = True
        sim = sc.Simulation(p)
        with pytest.raises(ValueError):
            sim.initialize()
