"""Batched ensembles on the HIP path.

A batch runs every member's kernels with that member's own arguments, so each
member must end bit for bit where the same simulation run alone ends: fields,
node positions, velocities and forces, in FP64 and in FP32, across the
member-by-member first step, eager steps, captured six-step graphs, a change of
launch shape and a member stepped on its own in between.  One member is also
held to the FP64 oracle bar (relative L2 <= 1e-12).  Members differ in lattice
size, viscosity, body force and capsule count; they mix the all-pairs and the
cell-list contact paths and capsules of more and fewer than 64 nodes; some
carry an advection-diffusion scalar, some apply multi-direct forcing.
"""
import numpy as np
import pytest

from tests.cases import oracle_sim, oracle_state, product_sim, product_state, rel_l2

pytestmark = pytest.mark.gpu

SKALAK = dict(model=2, G_s=0.06, C_skalak=10.0)
KEYS = ("rho", "ux", "uy", "x", "y", "vx", "vy", "fx_node", "fy_node")


def members(precision):
    base = [
        dict(nx=96, ny=34, tau=0.8, bfx=3e-5,          # 64 nodes: all-pairs contact
             capsules=[dict(count=4, region=(4, 4, 92, 30), rmin=3.0, rmax=3.0, nodes=16, seed=3, mem=SKALAK)]),
        dict(nx=150, ny=40, tau=0.7, bfx=2e-5,         # cell-list contact
             capsules=[dict(count=8, region=(0, 0, 150, 40), rmin=3.0, rmax=4.0, seed=11, mem=SKALAK)]),
        dict(nx=64, ny=48, tau=0.9, bfx=4e-5,          # one capsule of 76 nodes
             capsules=[dict(count=1, region=(20, 14, 44, 34), rmin=9.0, rmax=9.0, seed=5, mem=SKALAK)]),
        dict(nx=203, ny=56, tau=0.65, bfx=1e-5,
             capsules=[dict(count=12, region=(0, 0, 203, 56), rmin=2.5, rmax=4.0, seed=17, mem=SKALAK)]),
        dict(nx=128, ny=24, tau=1.0, bfx=5e-5,
             capsules=[dict(count=3, region=(0, 0, 128, 24), rmin=3.0, rmax=3.5, seed=23, mem=SKALAK)]),
    ]
    for i, c in enumerate(base):
        c.update(name=f"member{i}", coll=0, precision=precision)
    return base


def ensemble(sims):
    from softflow_amd import softflow_core as sc
    return sc.Ensemble(sims)


def assert_same(a, b, what):
    sa, sb = product_state(a), product_state(b)
    for k in KEYS:
        assert np.array_equal(sa[k], sb[k]), f"{what}: {k} differs"
    assert sa["capped"] == sb["capped"], what


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_members_match_single_runs(precision):
    cases = members(precision)
    print(f"batch parity {precision}: {len(cases)} members", flush=True)
    sims = [product_sim(c) for c in cases]
    twins = [product_sim(c) for c in cases]
    ens = ensemble(sims)
    schedule = [1, 3, 13, 8]
    for n in schedule:
        ens.run(n)
    for t in twins:
        t.run(sum(schedule))
    for i, (s, t) in enumerate(zip(sims, twins)):
        assert s.currentStep() == sum(schedule)
        assert_same(s, t, f"member {i} after {sum(schedule)} steps")
    # a member stepped on its own between batch steps, then a new launch shape
    sims[1].run(5)
    twins[1].run(5)
    ens.setLbmLaunch(128, 2, False)
    ens.run(14)
    for t in twins:
        t.run(14)
    for i, (s, t) in enumerate(zip(sims, twins)):
        assert_same(s, t, f"member {i} after the second batch run")
    # runs of the row-major cell index across row ends (rows narrower than
    # a workgroup, odd widths, the members' last partial workgroups)
    for shape in ((256, 1), (64, 1)):
        ens.setLbmLaunch(*shape, True)
        ens.setLbmFlat(True)
        ens.run(9)
        for t in twins:
            t.run(9)
        for i, (s, t) in enumerate(zip(sims, twins)):
            assert_same(s, t, f"member {i} after flat runs {shape}")


def test_batch_member_matches_oracle_fp64():
    cases = members("fp64")
    sims = [product_sim(c) for c in cases]
    ens = ensemble(sims)
    steps = 300
    ens.run(steps)
    for i in (1, 3):
        o = oracle_sim(cases[i])
        o.step(steps)
        ref = oracle_state(o)
        got = product_state(sims[i])
        for k in (("rho",), ("ux", "uy"), ("x", "y"), ("vx", "vy")):
            e = rel_l2(np.concatenate([np.ravel(got[p]) for p in k]),
                       np.concatenate([np.ravel(ref[p]) for p in k]))
            print(f"member {i} {k}: rel L2 {e:.3e}", flush=True)
            assert e <= 1e-12, (i, k, e)


def test_batch_autotune_keeps_results():
    cases = members("fp32")
    sims = [product_sim(c) for c in cases]
    twins = [product_sim(c) for c in cases]
    ens = ensemble(sims)
    ens.run(4)
    rep = ens.autotune(2)
    print(rep, flush=True)
    assert "chosen" in rep
    ens.run(20)
    for t in twins:
        t.run(24)
    for i, (s, t) in enumerate(zip(sims, twins)):
        assert_same(s, t, f"member {i} after autotune")
    info = ens.info()
    assert info[0] == len(cases)
    assert info[1] == sum(c["nx"] * c["ny"] for c in cases)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_members_with_scalar_match_single_runs(precision):
    # members with an advection-diffusion scalar beside members without one:
    # the batched collide-stream carries each scalar member's lagged scalar
    # step as one_step fuses it; a host write of the scalar between batch runs
    # sends the members through single steps until the fused state is back
    cases = members(precision)
    for i, c in enumerate(cases):
        if i % 2 == 0:
            c["scalar"] = dict(D=0.02 + 0.01 * i, blob=(c["nx"] / 2, c["ny"] / 2, 5.0))
    print(f"batch with scalar members {precision}", flush=True)
    sims = [product_sim(c) for c in cases]
    twins = [product_sim(c) for c in cases]
    ens = ensemble(sims)
    for n in (1, 2, 17):
        ens.run(n)
    for t in twins:
        t.run(20)
    for i, (s, t) in enumerate(zip(sims, twins)):
        assert_same(s, t, f"member {i} after 20 steps")
        if "scalar" in cases[i]:
            a = np.array(s.advectionDiffusion().concentration(0))
            b = np.array(t.advectionDiffusion().concentration(0))
            assert np.array_equal(a, b), f"member {i}: scalar differs"
    for s in (sims[2], twins[2]):
        s.advectionDiffusion().setPoint(10, 10, 2.0, 0)
    ens.run(13)
    for t in twins:
        t.run(13)
    for i, (s, t) in enumerate(zip(sims, twins)):
        assert_same(s, t, f"member {i} after the host write")
        if "scalar" in cases[i]:
            a = np.array(s.advectionDiffusion().concentration(0))
            b = np.array(t.advectionDiffusion().concentration(0))
            assert np.array_equal(a, b), f"member {i}: scalar differs after the host write"


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_members_with_mdf_match_single_runs(precision):
    # multi-direct forcing members beside plain ones: fixed 2 and 4
    # corrections, and residual control (each member its own tolerance and
    # cap, corrections stopping at different iterations in different steps)
    cases = members(precision)
    cases[0]["mdf"] = 3
    cases[1]["mdf"] = 5
    control = {2: (None, 6), 3: (None, 3)}      # member: (tolerance from its first residual, cap)
    print(f"batch with MDF members {precision}", flush=True)
    sims = [product_sim(c) for c in cases]
    twins = [product_sim(c) for c in cases]
    for i, (cap_tol, cap) in control.items():
        for s in (sims[i], twins[i]):
            s.run(20)
            s.setMdfControl(1e-300, 1, True)
            s.run(1)
        r1 = sims[i].mdfStats()["mean_residual"]
        twins[i].mdfStats()
        control[i] = (0.8 * r1, cap)
        for s in (sims[i], twins[i]):
            s.setMdfControl(control[i][0], cap, True)
    for i in range(len(cases)):
        if i not in control:
            for s in (sims[i], twins[i]):
                s.run(21)
    ens = ensemble(sims)
    for n in (1, 5, 12, 7):
        ens.run(n)
    for t in twins:
        t.run(25)
    for i, (s, t) in enumerate(zip(sims, twins)):
        assert_same(s, t, f"member {i} with MDF")
    for i in control:
        a, b = sims[i].mdfStats(), twins[i].mdfStats()
        print(f"member {i}: {a}", flush=True)
        assert a == b, (i, a, b)
        assert 0.0 < a["mean_corrections"] <= control[i][1]


def test_batch_rejects_unsupported_members():
    ok = members("fp32")[:2]
    bad_coll = dict(ok[0], name="mrt", coll=1)
    bad_mdf = dict(ok[0], name="mdf_scalar", mdf=3, scalar=dict(D=0.05, blob=(40, 17, 5.0)))
    bad_prec = dict(ok[0], name="fp64", precision="fp64")
    for bad, msg in ((bad_coll, "collision model"), (bad_mdf, "multi-direct forcing with a scalar"),
                     (bad_prec, "precision")):
        sims = [product_sim(ok[1]), product_sim(bad)]
        ens = ensemble(sims)
        with pytest.raises(RuntimeError, match=msg):
            ens.run(2)


def test_graph_replay_after_eager_remainder():
    # a remainder step run eagerly between two graph replays rotates the
    # buffer roles the captured graph was made with
    case = dict(name="graph_key", nx=80, ny=30, tau=0.8, bfx=3e-5, coll=0,
                capsules=[dict(count=3, region=(8, 6, 72, 24), rmin=3.0, rmax=3.0, nodes=16, seed=9,
                               mem=SKALAK)])
    a = product_sim(case, graph=True)
    for n in (7, 6, 11, 6):
        a.run(n)
    b = product_sim(case, graph=False)
    b.run(30)
    assert_same(a, b, "graph replays after remainders")
