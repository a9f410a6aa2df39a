"""Writable field views (python/bindings.cpp:313-355, 435-440: zero-copy
numpy views of the live arrays in the reference).

A held view keeps showing the state after each step, and writes into it act
as the reference's do, checked in FP64 against the oracle given the same
write:
- density/velocity: the stored macroscopic state the next collision reads
  (the oracle's set_macro);
- distributions: the populations (the oracle's set_f);
- concentration: the scalar the next scalar collision relaxes towards, with
  the populations g unchanged (the oracle's set_concentration, which equals the
  reference's own array write bitwise, tests/test_oracle_ref.py);
- force: visible until the next step, which rebuilds the force field as the
  reference's clearForces does, so the run is unchanged.
"""
import numpy as np
import pytest

from tests.cases import oracle_sim, oracle_state, product_sim, product_state, rel_l2

pytestmark = pytest.mark.gpu

SKALAK = dict(model=2, G_s=0.06, C_skalak=10.0)
CASE = dict(nx=48, ny=24, tau=0.8, bfx=3e-5, coll=2, scalar=dict(D=0.05, blob=(24, 12, 4.0)),
            capsules=[dict(centers=[(14.0, 12.0), (33.0, 10.5)], rmin=3.5, nodes=20, mem=SKALAK)])
NX, NY = CASE["nx"], CASE["ny"]


def _check(got, ref, tag):
    for k in ("rho", "ux", "uy", "x", "y", "C"):
        err = rel_l2(got[k], ref[k])
        assert err <= 1e-12, f"{tag}: {k} rel L2 {err:.3e}"


def test_held_views_follow_the_steps():
    print("views: held across steps", flush=True)
    sim = product_sim(CASE)
    f = sim.field()
    rho, ux, d = f.density(), f.velocity_x(), f.distributions()
    C = sim.advectionDiffusion().concentration(0)
    assert rho.flags.writeable and d.flags.writeable and C.flags.writeable
    o = oracle_sim(CASE)
    for n in (3, 7):
        sim.run(n)
        o.step(n)
        ref = oracle_state(o)
        assert rel_l2(rho, ref["rho"]) <= 1e-12 and rel_l2(ux, ref["ux"]) <= 1e-12
        assert rel_l2(C, ref["C"]) <= 1e-12
        # populations of the fluid rows: SOLID cells' populations are not
        # state (the push stream never writes them, the reference's pull
        # stream copies neighbour values there, and no stage reads them)
        assert rel_l2(d[:, 1:-1], o.f().reshape(9, NY, NX)[:, 1:-1]) <= 1e-12


@pytest.mark.parametrize("what", ["velocity", "density", "distributions", "concentration"])
def test_view_writes_match_oracle(what):
    print(f"views: write {what}", flush=True)
    sim = product_sim(CASE)
    o = oracle_sim(CASE)
    sim.run(10)
    o.step(10)
    fld = o.field()
    if what == "velocity":
        v = sim.field().velocity_x()
        v[5:9, 10:30] += 0.01
        ux = fld["ux"].reshape(NY, NX).copy()
        ux[5:9, 10:30] += 0.01
        o.set_macro(fld["rho"], ux, fld["uy"])
        del v
    elif what == "density":
        r = sim.field().density()
        r[12, :] *= 1.001
        rho = fld["rho"].reshape(NY, NX).copy()
        rho[12, :] *= 1.001
        o.set_macro(rho, fld["ux"], fld["uy"])
        del r
    elif what == "distributions":
        d = sim.field().distributions()
        d[3, 10, 12] += 1e-3
        d[6, 4:7, 40] -= 2e-4
        f = o.f().reshape(9, NY, NX).copy()
        f[3, 10, 12] += 1e-3
        f[6, 4:7, 40] -= 2e-4
        o.set_f(f)
    else:
        c = sim.advectionDiffusion().concentration(0)
        c[5:9, 10:30] += 0.25
        Cw = o.scalar().reshape(NY, NX).copy()
        Cw[5:9, 10:30] += 0.25
        o.set_concentration(Cw)
    sim.run(15)
    o.step(15)
    _check(product_state(sim), oracle_state(o), what)


def test_force_view_writes_do_not_change_the_run():
    print("views: force write", flush=True)
    a, b = product_sim(CASE), product_sim(CASE)
    a.run(10)
    b.run(10)
    fx = a.field().force_x()
    fx[:, :] = 1.0
    assert np.all(np.array(a.field().force_x()) == 1.0)   # visible until the next step
    a.run(10)
    b.run(10)
    sa, sb = product_state(a), product_state(b)
    for k in ("rho", "ux", "uy", "fx", "fy", "x", "y", "C"):
        assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), k
