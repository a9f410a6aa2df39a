"""Capsule-side launch choices against the default launch.

The spread gather runs either one workgroup per lattice tile or a persistent
grid walking the touched tiles, and the interpolations (the step's node
velocities and the MDF corrections) take 16, 8 or 4 lanes per node.  Every
choice sums the same terms in the same order, so each case here must end bit
for bit where the default launch (one workgroup per tile, 16 lanes per node)
ends: capsules in a walled channel with the scalar, an MDF run (the correction
interpolation and the correction spreads), residual-controlled MDF, FP64, and a
seam crossing in x.  The default launch itself is checked against the oracle
by the parity tests.
"""
import numpy as np
import pytest

from tests.cases import product_sim, product_state

pytestmark = pytest.mark.gpu

SKALAK = dict(model=2, G_s=0.06, C_skalak=10.0)
CAPS = [dict(count=12, region=(5, 5, 125, 35), rmin=3.0, rmax=4.5, seed=11, mem=SKALAK)]
SCALAR = dict(D=0.05, blob=(60, 20, 6.0))

CASES = {
    "mrt_caps_scalar_fp32": dict(nx=130, ny=40, coll=1, bfx=2e-5, precision="fp32", capsules=CAPS,
                                 scalar=SCALAR),
    "reg_mdf3_caps_fp32": dict(nx=130, ny=40, coll=2, bfx=2e-5, mdf=3, precision="fp32", capsules=CAPS),
    "bgk_mdf4_caps_fp64": dict(nx=96, ny=40, coll=0, bfx=2e-5, mdf=4,
                               capsules=[dict(count=8, region=(5, 5, 90, 35), rmin=3.0, rmax=4.0, seed=5,
                                              mem=SKALAK)]),
    "seam_caps_fp32": dict(nx=64, ny=32, coll=1, bfx=3e-5, precision="fp32",
                           capsules=[dict(centers=[(62.5, 15.0), (20.0, 12.0)], rmin=4.0, nodes=24, mem=SKALAK)]),
}

# (spread grid, interpolation lanes per node); the first is the default
LAUNCHES = [(0, 16), (0, 8), (0, 4), (64, 16), (7, 4)]
STEPS = 60


def _run(case, launch, residual):
    sim = product_sim(case)
    sim.setCapsuleLaunch(*launch)
    if residual:
        sim.setMdfControl(1e-6, 6, False)
    sim.run(STEPS)
    return product_state(sim)


@pytest.mark.parametrize("name", sorted(CASES))
def test_capsule_launches_match_default(name):
    case = CASES[name]
    for residual in ((False, True) if case.get("mdf", 1) > 1 else (False,)):
        ref = None
        for launch in LAUNCHES:
            print(f"capsule launch {name}: {launch}, residual MDF {residual}, {STEPS} steps", flush=True)
            st = _run(case, launch, residual)
            if ref is None:
                ref = st
                assert np.all(np.isfinite(st["x"])) and np.all(np.isfinite(st["ux"]))
                continue
            assert sorted(st) == sorted(ref)
            for k in ref:
                if k == "capped":
                    assert st[k] == ref[k], (name, launch, k)
                    continue
                assert np.array_equal(np.asarray(st[k]), np.asarray(ref[k]), equal_nan=True), (name, launch, k)
