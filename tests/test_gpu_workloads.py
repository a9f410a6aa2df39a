"""The benchmark workloads on the device at their full sizes.

The oracle finishes these sizes too slowly to compare field by field, so the
checks are the ones that hold at any size, beside the FP64 device mode (itself
held to the oracle at 1e-12 by the parity tests):

- the 4096 x 1024 channel with 2000 bidisperse capsules and the scalar, FP32
  lattice storage: fluid mass and scalar mass conserved to round-off, every
  field finite, the force-cap count reported; the same workload in FP64 over a
  shorter horizon, with the FP32 run's error against it on the velocity field,
  the capsule centroid trajectories and the Taylor deformation;
- the batched ensemble of 256 channels drawn with seed 7: every member finite,
  and members sampled across the batch (smallest, largest, and spread over
  the index range) bit-identical to the same channel stepped on its own;
- the 2048 x 512 multi-direct-forcing channel with 500 capsules in residual
  mode: plain FP64 MDF and the learned FP32 warm start (trained here on rows
  the plain solver collects) both hold the no-slip residual at or below the
  tolerance, and the learned run applies fewer corrections per step.
"""
import numpy as np
import pytest

from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


def _centroids(sim):
    caps = sim.capsules()
    c = np.empty((len(caps), 2))
    for i in range(len(caps)):
        p = caps[i].centroid()
        c[i] = (p.x, p.y)
    return c


def _deformation(sim):
    caps = sim.capsules()
    return np.array([caps[i].deformationIndex() for i in range(len(caps))])


def _velocity(sim):
    f = sim.field()
    return np.r_[np.ravel(f.velocity_x()), np.ravel(f.velocity_y())]


def _fluid_mass(sim):
    # rows 1 .. ny-2 are the fluid between the half-way bounce-back walls
    return float(np.array(sim.field().density())[1:-1].sum(dtype=np.float64))


def test_large_domain_fp32_conserves_and_stays_finite():
    from softflow_amd.workloads import build_product, large_domain_case
    print("large domain FP32: build", flush=True)
    sim = build_product(large_domain_case(), "fp32")
    m0 = _fluid_mass(sim)
    c0 = float(np.sum(sim.advectionDiffusion().concentration(0), dtype=np.float64))
    steps = 600
    sim.run(steps)
    print("large domain FP32: ran", steps, "steps", flush=True)
    m1 = _fluid_mass(sim)
    c1 = float(np.sum(sim.advectionDiffusion().concentration(0), dtype=np.float64))
    stats = sim.fieldStats()
    capped = sim.lastCappedNodes()
    print(f"fluid mass {m0:.10e} -> {m1:.10e} (rel {abs(m1 - m0) / m0:.2e}); scalar mass {c0:.10e} -> "
          f"{c1:.10e} (rel {abs(c1 - c0) / c0:.2e}); max |u| {stats[3]:.4e}; non-finite cells "
          f"{stats[4]:.0f}; capped cells at the last step {capped}", flush=True)
    assert abs(m1 - m0) <= 1e-6 * m0
    assert abs(c1 - c0) <= 1e-5 * c0
    assert stats[4] == 0
    assert 0.0 < stats[3] < 0.05
    for k in ("density", "velocity_x", "velocity_y", "force_x", "force_y"):
        assert np.all(np.isfinite(np.array(getattr(sim.field(), k)()))), k
    caps = sim.capsules()
    assert len(caps) == 2000 and caps.totalNodes() == 800 * 38 + 1200 * 19
    assert np.all(np.isfinite(_centroids(sim)))
    assert capped >= 0


def test_large_domain_fp32_against_fp64():
    """FP32 deviation storage against the FP64 mode on the same seeded
    4096 x 1024 workload: the errors the FP32 path carries, reported and
    bounded (velocity field relative L2, the largest centroid trajectory
    error over the run, the largest Taylor deformation error)."""
    from softflow_amd.workloads import build_product, large_domain_case
    case = large_domain_case()
    runs = {}
    for prec in ("fp64", "fp32"):
        print(f"large domain {prec}: build", flush=True)
        runs[prec] = build_product(case, prec)
    steps, every = 1000, 100
    traj = {k: [] for k in runs}
    for s in range(0, steps, every):
        for k, sim in runs.items():
            sim.run(every)
            traj[k].append(_centroids(sim))
        print(f"large domain: step {s + every}", flush=True)
    du = rel_l2(_velocity(runs["fp32"]), _velocity(runs["fp64"]))
    dc = max(float(np.max(np.hypot(*(a - b).T))) for a, b in zip(traj["fp32"], traj["fp64"]))
    dD = float(np.max(np.abs(_deformation(runs["fp32"]) - _deformation(runs["fp64"]))))
    umax = runs["fp64"].fieldStats()[3]
    print(f"FP32 vs FP64 after {steps} steps (max |u| {umax:.3e}): velocity rel L2 {du:.3e}, "
          f"centroid trajectory max error {dc:.3e} lu, Taylor D max error {dD:.3e}", flush=True)
    assert du < 1e-3
    assert dc < 0.02
    assert dD < 1e-3


def test_ensemble_256_members_bitwise_to_solo():
    from softflow_amd import softflow_core as sc
    from softflow_amd.ensemble import build_member, ensemble_specs
    specs = ensemble_specs(256, 7)
    print("ensemble: build 256 members", flush=True)
    sims = [build_member(sp, "fp32")[0] for sp in specs]
    ens = sc.Ensemble(sims)
    steps = 300
    ens.autotune(3)
    ens.run(steps)
    ens.synchronize()
    print(f"ensemble: ran {steps} steps", flush=True)
    bad = [i for i, s in enumerate(sims) if s.fieldStats()[4] > 0]
    assert not bad, bad
    cells = [sp["nx"] * sp["ny"] for sp in specs]
    pick = sorted({int(np.argmin(cells)), int(np.argmax(cells)), 0, 37, 101, 164, 200, 255})
    for i in pick:
        solo, _ = build_member(specs[i], "fp32")
        solo.run(steps)
        a, b = sims[i], solo
        fa, fb = a.field(), b.field()
        for k in ("density", "velocity_x", "velocity_y", "force_x", "force_y"):
            assert np.array_equal(np.array(getattr(fa, k)()), np.array(getattr(fb, k)())), (i, k)
        ca, cb = a.capsules(), b.capsules()
        assert len(ca) == len(cb)
        for c in range(len(ca)):
            for n in range(ca[c].numNodes()):
                pa, pb = ca[c].nodePosition(n), cb[c].nodePosition(n)
                assert pa.x == pb.x and pa.y == pb.y, (i, c, n)
        print(f"member {i} ({specs[i]['nx']}x{specs[i]['ny']}, {len(ca)} capsules): bitwise", flush=True)


def test_mdf_channel_learned_against_plain():
    from softflow_amd import softflow_core as sc
    from softflow_amd.workloads import build_mdf_channel
    tol, max_corr, develop, collect, timed = 1e-5, 16, 1500, 60, 100
    print(f"mdf channel: plain FP64, {develop} steps of 5-iteration MDF", flush=True)
    plain, placed = build_mdf_channel("fp64", 5)
    assert placed == 500
    plain.run(develop)
    plain.setMdfControl(tol, max_corr, True)
    plain.run(6)
    plain.mdfStats()
    plain.run(timed)
    sp = plain.mdfStats()
    print(f"mdf channel: learned FP32, develop, collect {collect} steps, train", flush=True)
    learned, _ = build_mdf_channel("fp32", 5)
    learned.run(develop)
    learned.setMdfControl(1e-300, 4, True)
    learned.collectMdfSamples(collect * learned.capsules().totalNodes())
    learned.run(collect)
    model = sc.MdfModel(0, 7)
    loss = model.train(learned, epochs=4, batch=4096, lr=2e-3, seed=11)
    learned.collectMdfSamples(0)
    learned.setMdfModel(model, True)
    learned.setMdfControl(tol, max_corr, True)
    learned.run(6)
    learned.mdfStats()
    learned.run(timed)
    sl = learned.mdfStats()
    print(f"mdf channel over {timed} steps at tolerance {tol:g}: plain FP64 {sp['mean_corrections']:.3f} "
          f"corrections/step, residual max {sp['max_residual']:.3e}; learned FP32 {sl['mean_corrections']:.3f}, "
          f"residual max {sl['max_residual']:.3e}; training loss {loss[0]:.3f} -> {loss[-1]:.3f}", flush=True)
    assert sp["max_residual"] <= tol and sl["max_residual"] <= tol
    assert sl["mean_corrections"] < sp["mean_corrections"]
    assert loss[-1] < loss[0]
