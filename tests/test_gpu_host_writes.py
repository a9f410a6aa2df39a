"""Host writes of the node state between steps, through the C-ABI.

The step keeps work ahead for the next step: the closing interpolation lists
the moved nodes in the next step's contact cells, and the spread's bins stay
open until the step closes.  A host write of the positions (sf_set_nodes) or
a new capsule set (sf_set_capsules) must drop what was prepared from the old
positions.  Each case is checked against a fresh context handed the same
state (populations, force, nodes, membrane state) before its first step, so
nothing was prepared there: both must end bit for bit alike.  The capsules
sit closer than the contact cut-off, so stale contact cells would change the
forces.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NY = 160, 48
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def _params(precision):
    from softflow_amd import capi
    return capi.make_params(nx=NX, ny=NY, tau=0.8, body_force_x=3e-5, collision_model=0, precision=precision)


def _capsules():
    """48 rings of 20 nodes (radius 3.5) on a 9-lattice-unit grid: neighbours
    2 units apart, inside the 3.5-unit contact cut-off."""
    from softflow_amd import capi
    n, r = 20, 3.5
    th = 2 * np.pi * np.arange(n) / n
    chord = 2 * r * np.sin(np.pi / n)
    xs, ys, off, desc = [], [], [0], []
    for cy in (12.0, 24.0, 36.0):
        for j in range(16):
            cx = 6.0 + 9.0 * j + (4.5 if cy == 24.0 else 0.0)
            xs.append(cx + r * np.cos(th))
            ys.append(cy + r * np.sin(th))
            off.append(off[-1] + n)
            d = capi.CapsuleDesc()
            d.model, d.G_s, d.C_skalak = 2, 0.06, 10.0
            d.k_bend, d.k_area, d.k_perimeter = 0.005, 0.5, 0.1
            d.rest_area = 0.5 * n * r * r * np.sin(2 * np.pi / n)
            d.rest_perimeter = n * chord
            desc.append(d)
    x = np.concatenate(xs) % NX
    y = np.concatenate(ys)
    return np.array(off, dtype=np.int32), desc, x, y, np.full(x.size, chord)


def _force_state(ctx):
    from softflow_amd import capi
    F = np.empty(2 * NX * NY)
    capi._check(ctx.L.sf_get_force_state(ctx.h, F.ctypes.data_as(C.POINTER(C.c_double))), "sf_get_force_state")
    return F


def _set_force_state(ctx, F):
    from softflow_amd import capi
    F = np.ascontiguousarray(F, dtype=np.float64)
    capi._check(ctx.L.sf_set_force_state(ctx.h, F.ctypes.data_as(C.POINTER(C.c_double))), "sf_set_force_state")


def _membrane_state(ctx, K, ncaps):
    from softflow_amd import capi
    pl = np.empty(K)
    hp = np.empty(ncaps, dtype=np.int32)
    capi._check(ctx.L.sf_get_membrane_state(ctx.h, pl.ctypes.data_as(C.POINTER(C.c_double)),
                                            hp.ctypes.data_as(C.POINTER(C.c_int32))), "sf_get_membrane_state")
    return pl, hp


def _state(ctx):
    # the wall rows are SOLID cells, whose populations are not state
    return ctx.populations()[:, 1:NY - 1, :], ctx.nodes()


def _same(a, b, what):
    fa, na = a
    fb, nb = b
    assert np.array_equal(fa, fb), f"{what}: populations differ"
    for k in ("x", "y", "vx", "vy", "fx", "fy"):
        assert np.array_equal(na[k], nb[k]), f"{what}: node {k} differs"


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_host_node_writes_drop_prepared_work(precision):
    from softflow_amd import capi
    prec = capi.SF_FP64 if precision == "fp64" else capi.SF_FP32
    off, desc, x0, y0, rest = _capsules()
    K, ncaps = int(off[-1]), len(off) - 1
    f0 = np.repeat(W[:, None], NX * NY, axis=1)
    print(f"host writes {precision}: {ncaps} capsules, {K} nodes", flush=True)

    a = capi.Context(_params(prec))
    a.set_populations(f0)
    a.set_capsules(off, desc, x0, y0, np.zeros(K), np.zeros(K), rest)
    a.step(12)
    f12 = a.populations()
    F12 = _force_state(a)
    n12 = a.nodes()
    pl, hp = _membrane_state(a, K, ncaps)
    xs = (n12["x"] + 0.6) % NX                    # every node into new contact cells
    a.set_nodes(x=xs)
    a.step(12)
    got = _state(a)

    # the same state handed to a fresh context before its first step
    b = capi.Context(_params(prec))
    b.set_populations(f12)
    b.set_capsules(off, desc, xs, n12["y"], n12["vx"], n12["vy"], rest, prev_len=pl, has_prev=hp)
    _set_force_state(b, F12)
    b.step(12)
    _same(got, _state(b), "sf_set_nodes after 12 steps")

    # a new capsule set on a context that has stepped
    c = capi.Context(_params(prec))
    c.set_populations(f0)
    c.set_capsules(off, desc, x0, y0, np.zeros(K), np.zeros(K), rest)
    c.step(12)
    c.set_populations(f12)
    c.set_capsules(off, desc, xs, n12["y"], n12["vx"], n12["vy"], rest, prev_len=pl, has_prev=hp)
    _set_force_state(c, F12)
    c.step(12)
    _same(_state(c), _state(b), "sf_set_capsules after 12 steps")
    for ctx in (a, b, c):
        ctx.close()
