"""sf_field_stats_noalloc against sf_field_stats through the C-ABI.

The allocation-free variant keeps its partial sums in a fixed device array
shared by every call on a device, and computes each cell's rho and u inside
the reduction; its five values must equal the allocating entry point's bit for
bit.  Checked on calls in a row on different contexts, the second lattice
smaller than the first, in FP64 and FP32, with plain channel cells and with
generic cell types (inlet and outlet columns), and in the three macroscopic modes: a fresh
lattice (zero force), after steps (the last collision's force) and after a host
write of rho/u (the stored state).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def _context(nx, ny, precision, boundary, coll):
    from softflow_amd import capi
    p = capi.make_params(nx=nx, ny=ny, body_force_x=3e-5 if boundary == capi.SF_PERIODIC else 0.0,
                         collision_model=coll, precision=precision, boundary_type=boundary,
                         inlet_velocity=0.02, outlet_density=1.0)
    ctx = capi.Context(p)
    f = np.repeat(W[:, None], nx * ny, axis=1)
    ctx.set_populations(f)
    return ctx


def _same(ctx, tag):
    a = ctx.field_stats()
    b = ctx.field_stats_noalloc()
    print(tag, "field_stats", a.tolist(), flush=True)
    assert np.array_equal(a, b), (tag, a, b)
    assert np.isfinite(a).all() and a[0] > 0


def test_noalloc_stats_equal_allocating_entry_point():
    from softflow_amd import capi
    big = _context(160, 64, capi.SF_FP64, capi.SF_PERIODIC, 2)
    small = _context(48, 20, capi.SF_FP32, capi.SF_INLET_OUTLET, 1)
    _same(big, "big fresh")
    big.step(120)
    _same(big, "big after 120 steps")
    _same(small, "small fresh")
    small.step(40)
    _same(small, "small after 40 steps")        # a smaller lattice right after a larger one
    m = big.macroscopic()
    big.set_macroscopic(m["rho"] * 1.001, m["ux"] + 1e-3, m["uy"])
    _same(big, "big stored")
    big.step(7)
    _same(big, "big after 7 more steps")        # the larger one again: nothing stale is read
    small.step(5)
    _same(small, "small after 5 more steps")
