"""Lifetime of the writable field views, on the CPU (the views exist before
initialize(), so no device is touched).

A view holds a lease on its array that is released when the view is freed.
The view must keep the field -- and the Simulation that owns it -- alive until
that release has run: when the view is the last owner, the release must not
touch a freed field (that was a use-after-free when the view held the field
only through a keep-alive weak reference, whose callback runs before the
array's base object is released).
"""
import gc
import subprocess
import sys
import textwrap

import numpy as np

SCRIPT = textwrap.dedent("""
    import gc, sys
    sys.path.insert(0, {root!r})
    from softflow_amd import softflow_core as sc
    p = sc.SimulationParams(); p.nx, p.ny = 256, 128
    for round_ in range(20):
        sim = sc.Simulation(p)
        f = sim.field()
        views = [f.density(), f.velocity_x(), f.distributions(), f.force_y()]
        views[0][1, 1] = 3.0
        views[2][4, 2, 3] = 0.5
        del sim, f              # the views are the last owners
        sub = views[1][2:5]     # an array derived from a view keeps the lease
        del views
        gc.collect()
        sub[0, 0] = 1.0
        del sub
        gc.collect()
    print("released")
""")


def test_last_view_outlives_its_simulation():
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, MALLOC_CHECK_="3")
    r = subprocess.run([sys.executable, "-c", SCRIPT.format(root=root)], capture_output=True, text=True,
                       env=env, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    assert "released" in r.stdout


def test_views_are_writable_and_alias_the_field():
    from softflow_amd import softflow_core as sc
    p = sc.SimulationParams()
    p.nx, p.ny = 12, 6
    sim = sc.Simulation(p)
    f = sim.field()
    a, b = f.density(), f.density()
    assert a.flags.writeable and a.shape == (6, 12)
    a[2, 3] = 7.5
    assert b[2, 3] == 7.5 and f.getRho(3, 2) == 7.5
    d = f.distributions()
    assert d.shape == (9, 6, 12) and d.flags.writeable
    d[5, 1, 1] = 0.25
    assert f.f(1, 1, 5) == 0.25
    del a, b, d
    gc.collect()
    assert np.array(f.density())[2, 3] == 7.5
