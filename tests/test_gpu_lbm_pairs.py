"""Two cells per lane in the collide-stream against one cell per lane.

The paired kernel takes cells (x, x+1) of FP32 lattices (x periodic without x
walls, no moving walls, nx even; generic cell types per cell) and evaluates
each cell with the single-cell kernel's expressions, so every case here must end
bit-identical either way: walls and periodic y, the three collision models,
capsules, the fused scalar with and without periodic y, an MDF run (the
partial shift force), a run whose populations were written from the host (the
stored shift), and generic cell types (pillars with interpolated bounce-back,
extra solids, where the two cells of a pair differ).  Launch shapes cover rows that end inside a
workgroup (nx = 130 gives 65 lanes per row), blocks taller than one row and a
persistent grid walking several tiles.  A configuration outside the paired
form (FP64) is checked to run the single-cell kernel unchanged.
"""
import numpy as np
import pytest

from tests.cases import PILLARS, product_sim, product_state

pytestmark = pytest.mark.gpu

CAPS = [dict(count=10, region=(5, 5, 125, 35), rmin=3.0, rmax=4.0, seed=3,
             mem=dict(model=2, G_s=0.06, C_skalak=10.0))]
SCALAR = dict(D=0.05, blob=(60, 20, 6.0))

CASES = {
    "mrt_caps_scalar": dict(nx=130, ny=40, coll=1, bfx=2e-5, precision="fp32", capsules=CAPS, scalar=SCALAR),
    "reg_caps": dict(nx=130, ny=40, coll=2, bfx=2e-5, precision="fp32", capsules=CAPS),
    "bgk_periodic_y_scalar": dict(nx=64, ny=32, coll=0, bfx=1e-5, periodic_y=True, precision="fp32",
                                  scalar=SCALAR | dict(blob=(30, 16, 5.0)),
                                  params=dict()),
    "reg_mdf3_caps": dict(nx=130, ny=40, coll=2, bfx=2e-5, mdf=3, precision="fp32", capsules=CAPS),
    "bgk_fp64_not_paired": dict(nx=130, ny=40, coll=0, bfx=2e-5, capsules=CAPS),
    # generic cell types: the two cells of a pair take their own types and masks
    "pillars_ibb_mrt_caps_scalar": dict(nx=130, ny=40, coll=1, bfx=2e-5, precision="fp32", obstacles=PILLARS,
                                        ibb=True, capsules=CAPS, scalar=SCALAR),
    "solids_bgk_periodic_y_scalar": dict(nx=64, ny=32, coll=0, bfx=1e-5, periodic_y=True, precision="fp32",
                                         solids=[("circle", 31.0, 16.0, 5.0), ("rect", 10, 4, 10, 27)],
                                         scalar=SCALAR | dict(blob=(45, 16, 4.0))),
    # the regularized and MDF (partial shift force) forms with per-cell types,
    # which the autotuner may select in production
    "pillars_ibb_reg_caps": dict(nx=130, ny=40, coll=2, bfx=2e-5, precision="fp32", obstacles=PILLARS,
                                 ibb=True, capsules=CAPS),
    "pillars_reg_mdf3_caps_scalar": dict(nx=130, ny=40, coll=2, bfx=2e-5, mdf=3, precision="fp32",
                                         obstacles=PILLARS[1:], capsules=CAPS, scalar=SCALAR),
}

# (bx, by, xcd_banded, persistent workgroups per XCD)
SHAPES = [(64, 4, True, 0), (256, 1, True, 0), (32, 8, False, 0), (64, 1, True, 2)]

STEPS = 40


def _run(case, shape, pairs, host_write):
    sim = product_sim(case)
    sim.setLbmLaunch(shape[0], shape[1], shape[2], shape[3])
    sim.setLbmPairs(pairs)
    sim.run(STEPS // 2)
    if host_write:
        # a host write of the populations: the next collision takes the stored macro state
        sim.field().initializeEquilibriumAt(7, 5)
    sim.run(STEPS - STEPS // 2)
    return product_state(sim)


@pytest.mark.parametrize("name", sorted(CASES))
def test_pairs_match_single_cells(name):
    case = CASES[name]
    for shape in SHAPES:
        host_write = shape == SHAPES[0]
        print(f"lbm pairs {name}: shape {shape}, {STEPS} steps, host write {host_write}", flush=True)
        a = _run(case, shape, False, host_write)
        b = _run(case, shape, True, host_write)
        assert sorted(a) == sorted(b)
        for k in a:
            if k == "capped":
                assert a[k] == b[k], (name, shape, k)
                continue
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k]), equal_nan=True), (name, shape, k)
        assert np.all(np.isfinite(a["ux"])), (name, shape)
