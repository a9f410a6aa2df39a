"""Benchmark of the per-timestep hot path on one MI355X.

Workload (the large single-domain configuration): 4096 x 1024 D2Q9 channel,
MRT collision (tau = 0.7, s_e = s_eps = s_q = 1.4), body force 1e-5, periodic
x with half-way bounce-back walls in y; 2000 bidisperse Skalak capsules
(800 of radius 6 with 38 nodes and G_s = 0.03 placed first with seed 101,
then 1200 of radius 3 with 19 nodes and G_s = 0.06 with seed 202; C = 10,
Helfrich bending k_b = 1e-3); an advection-diffusion scalar (D = 0.05)
starting from a unit Gaussian blob of sigma = 20 at the centre.  Lattice
storage and collide in FP32 (populations stored as f - w_q); capsules FP64.

A step is one Simulation::step: membrane + contact forces, IBM spread, LBM
collide-stream, scalar step, interpolation + node move.  ``value`` is whole
step MLUPS over the timed region (HIP-graph replay, inputs resident in HBM).
``roofline`` is for the dominant kernel (the fused collide-stream): its mean
duration comes from HIP events around each launch in a second timed region
of the same length, run kernel by kernel.

The line also carries the learned-vs-plain MDF comparison (the second half of
the metric): steps/s of the plain 5-iteration MDF in FP64 against the FP32
path whose first correction comes from the trained network, at the same
no-slip residual, on a 2048 x 512 channel with 500 capsules.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-mdf]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from softflow_amd.workloads import (MDF_CHANNEL, NX, NY, build_mdf_channel,  # noqa: E402
                                   build_product, large_domain_case, pillar_case)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per lattice update of the fused collide-stream in FP32:
# 9 populations read + 9 written (72 B), the step's raw force (8 B) and the
# previous step's raw force that sets the collision velocity (8 B)
LBM_BYTES_PER_LUP = 88
# the scalar step carried by the same kernel: 9 g read + 9 g written (FP32)
ADR_BYTES_PER_LUP = 72


def host_cpu():
    """CPU model and logical core count of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_threads():
    """OpenMP threads for the CPU baseline: every core this process is given
    (OMP_NUM_THREADS when set, which a shared GPU host uses to allot its
    share), else every core of the host."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)
    return max(1, min(n, os.cpu_count() or 1))


def cpu_baseline(case, steps, threads):
    """The reference's own hot-path components (oracle/_ref: its lbm, coupling
    and membrane translation units compiled in place under its Release flags,
    driven in Simulation::step order) with ``threads`` OpenMP threads on this
    host: ``steps`` full steps of the same workload.  Where the Release build is
    absent, the strict (parity) build; where neither exists, the
    single-threaded FP64 oracle restatement."""
    from oracle import pyoracle as po

    if os.path.exists(po.REF_RELEASE_SO):
        kind, which = "reference", "ref-release"
    elif os.path.exists(po.REF_SO):
        kind, which = "reference", "ref"
    else:
        kind, which = "port", "oracle"
    p = po.default_params(nx=case["nx"], ny=case["ny"], tau=case["tau"], boundary_type=1,
                          body_force_x=case["bfx"], collision_model=case["coll"],
                          scalar_enabled=1, diffusivity=[case["scalar"]["D"]],
                          use_interpolated_bb=1 if case.get("ibb") else 0)
    s = po.Sim(p, which)
    for ob in case.get("obstacles", []):
        s.add_obstacle(ob)
    if kind == "reference":
        s.set_threads(threads)
    else:
        threads = 1
    for spec in case["capsules"]:
        m = po.default_membrane(**spec["mem"])
        x0, y0, x1, y1 = spec["region"]
        s.add_capsule_random(spec["count"], x0, y0, x1, y1, spec["rmin"], spec["rmax"],
                             spec["nodes"], m, 0, spec["seed"], 1.0, 0)
    s.initialize()
    s.step(1)   # first step outside the timing (page-in of the arrays)
    t = time.perf_counter()
    s.step(steps)
    dt = time.perf_counter() - t
    flags = {"ref-release": po.REF_RELEASE_FLAGS, "ref": "-O2 -ffp-contract=off -fno-fast-math -fopenmp",
             "oracle": "-O2 -ffp-contract=off (single thread)"}[which]
    return case["nx"] * case["ny"] * steps / dt / 1e6, dt, kind, threads, flags


USE_GRAPH = True   # HIP-graph replay of the steps (off: every launch eager, for tracing)
GRAPH_PERIOD = 6   # the step's buffer roles repeat every six steps


def prime_graphs(sim, steps):
    """Capture, outside any timed region, every graph that ``sim.run(steps)``
    will replay from the current buffer phase.  A run of ``steps`` from phase P
    replays the six-step graph of phase P and a (steps mod 6)-step graph of
    phase P: six steps capture the first and end at phase P again, r = steps
    mod 6 more capture the second, and 6 - r more bring the roles back to
    phase P (12 steps at most)."""
    sim.run(min(steps, GRAPH_PERIOD))
    r = steps % GRAPH_PERIOD
    if r and steps > GRAPH_PERIOD:
        sim.run(r)
        sim.run(GRAPH_PERIOD - r)
    elif r:
        sim.run(GRAPH_PERIOD - r)
    sim.synchronize()


def timed_run(sim, steps):
    """Seconds for ``sim.run(steps)`` (which returns once the device is done),
    with its graphs primed first, and the number of graphs captured inside the
    timed region (0 when every step was a replay)."""
    prime_graphs(sim, steps)
    cap0 = sim.graphCaptures()
    t = time.perf_counter()
    sim.run(steps)
    dt = time.perf_counter() - t
    return dt, sim.graphCaptures() - cap0


def timed_steps(sim, steps):
    dt, _ = timed_run(sim, steps)
    return steps / dt


MDF_TOL = 1e-5        # the no-slip residual both MDF schemes are held to (<= 1e-4)
MDF_MAX_CORR = 16     # corrections allowed per step before a step gives up on the tolerance
MDF_DEVELOP = 3000    # steps of the reference's 5-iteration MDF from rest before timing


def _mdf_stats(st):
    return {"corrections_per_step": round(st["mean_corrections"], 3), "residual_mean": st["mean_residual"],
            "residual_max": st["max_residual"]}


def _residual_curve(sim, kmax=6, steps=60):
    """Mean and max no-slip residual after k = 1 .. kmax corrections per step
    (the learned correction counts as the first when a model is set)."""
    out = []
    for k in range(1, kmax + 1):
        sim.setMdfControl(1e-300, k, True)
        sim.run(6)
        sim.mdfStats()
        sim.run(steps)
        st = sim.mdfStats()
        out.append({"corrections": k, "residual_mean": st["mean_residual"], "residual_max": st["max_residual"]})
    return out


def mdf_learned_bench(steps, train_steps, epochs, curve=True, develop=None):
    """The learned MDF warm start against plain MDF on the MDF channel
    (MDF_CHANNEL), both held to the same no-slip residual MDF_TOL.

    Protocol, the same for both schemes: ``develop`` (default MDF_DEVELOP) steps of the reference's
    5-iteration MDF from rest, then ``steps`` timed steps.
    - plain, FP64: plain corrections until the residual r = max_k |V_k - U(X_k)|
      is <= MDF_TOL (at most MDF_MAX_CORR per step); the reference's fixed
      5-iteration MDF (4 corrections) is timed beside it;
    - learned, FP32: the network's correction first, then plain corrections
      until r <= MDF_TOL (same cap).  The network is trained here, by the
      device's own kernels, on (state, converged correction) rows the plain
      FP32 solver collects over ``train_steps`` steps after its development.
    Reported: steps/s, corrections per step, residuals, the velocity field of
    the learned FP32 run against the plain FP64 run after the same number of
    steps, and the residual after k corrections for both schemes."""
    from softflow_amd import softflow_core as sc

    develop = MDF_DEVELOP if develop is None else develop
    ch = MDF_CHANNEL
    out = {"workload": f"{ch['nx']}x{ch['ny']} regularized tau={ch['tau']} fx={ch['bfx']:g}, {ch['count']} Skalak "
                       f"(G_s=0.06, C={ch['mem']['C_skalak']:g}) r={ch['radius']:g} ({ch['nodes']} nodes) seeded "
                       "over the channel (2.4% area fraction)",
           "residual_tolerance": MDF_TOL, "max_corrections": MDF_MAX_CORR, "develop_steps": develop,
           "timed_steps": steps}
    # plain FP64: develop, the reference's fixed scheme, then residual control
    plain, placed = build_mdf_channel("fp64", 5, USE_GRAPH)
    out["capsules"] = placed
    plain.run(develop)
    plain.setMdfControl(1e-300, 4, True)        # 4 corrections, residual measured
    plain.run(6)
    plain.mdfStats()
    dt, _ = timed_run(plain, steps)
    out["plain_fp64_5_iterations"] = dict(steps_per_s=round(steps / dt, 1), **_mdf_stats(plain.mdfStats()))
    plain.setMdfControl(MDF_TOL, MDF_MAX_CORR, True)
    plain.run(6)
    plain.mdfStats()
    dt, _ = timed_run(plain, steps)
    sp = plain.mdfStats()
    out["plain_fp64_steps_per_s"] = round(steps / dt, 1)
    out["plain"] = _mdf_stats(sp)
    # learned FP32: develop, collect rows from the plain solver, train
    t0 = time.perf_counter()
    learned, _ = build_mdf_channel("fp32", 5, USE_GRAPH)
    learned.run(develop)
    learned.setMdfControl(1e-300, 4, True)
    K = learned.capsules().totalNodes()
    learned.collectMdfSamples(train_steps * K)
    learned.run(train_steps)
    rows = learned.mdfSampleCount()
    model = sc.MdfModel(0, 7)
    loss = model.train(learned, epochs=epochs, batch=4096, lr=2e-3, seed=11)
    learned.collectMdfSamples(0)
    train_s = time.perf_counter() - t0
    # the same step count as the plain run before its residual-controlled segment
    gap = (steps + len_prime(steps) + 6) - train_steps
    if gap > 0:
        learned.run(gap)
    learned.setMdfModel(model, True)
    learned.setMdfControl(MDF_TOL, MDF_MAX_CORR, True)
    learned.run(6)
    learned.mdfStats()
    dt, _ = timed_run(learned, steps)
    sl = learned.mdfStats()
    out["learned_fp32_steps_per_s"] = round(steps / dt, 1)
    out["learned"] = _mdf_stats(sl)
    out["speedup"] = round(out["learned_fp32_steps_per_s"] / out["plain_fp64_steps_per_s"], 3)
    out["speedup_vs_5_iterations"] = round(out["learned_fp32_steps_per_s"] /
                                           out["plain_fp64_5_iterations"]["steps_per_s"], 3)
    out["mlups_plain_fp64"] = round(out["plain_fp64_steps_per_s"] * ch["nx"] * ch["ny"] / 1e6, 1)
    out["mlups_learned_fp32"] = round(out["learned_fp32_steps_per_s"] * ch["nx"] * ch["ny"] / 1e6, 1)

    def vel(sim):
        f = sim.field()
        return np.r_[np.ravel(f.velocity_x()), np.ravel(f.velocity_y())]
    ua = vel(plain)
    out["velocity_rel_l2_vs_plain_fp64"] = float(np.linalg.norm(vel(learned) - ua) / np.linalg.norm(ua))
    out["training"] = {"rows": int(rows), "epochs": epochs, "seconds": round(train_s, 2),
                       "loss_first_epoch": loss[0], "loss_last_epoch": loss[-1]}
    if curve:
        out["residual_after_k_corrections"] = {"plain_fp64": _residual_curve(plain),
                                               "learned_fp32": _residual_curve(learned)}
    return out


def len_prime(steps):
    """Steps prime_graphs runs before a timed run of ``steps``."""
    r = steps % GRAPH_PERIOD
    return min(steps, GRAPH_PERIOD) + (GRAPH_PERIOD if r and steps > GRAPH_PERIOD else (GRAPH_PERIOD - r if r else 0))


def mdf_iteration_costs(steps=1200, develop=300):
    """The parts of an MDF step on the MDF channel, for the cost model
    T = B + n A + (max - n) I: A, one applied correction (interpolation +
    correction launch), from an unreachable tolerance with 4 against 1
    corrections; I, one iteration after convergence (both launches exit at
    once), from a tolerance every step meets at once with 16 against 1; and
    the step with one applied correction.  FP64 and FP32, each developed
    ``develop`` steps with the reference's 5-iteration MDF, graph replay."""
    out = {}
    for prec in ("fp64", "fp32"):
        sim, _ = build_mdf_channel(prec, 5, USE_GRAPH)
        sim.run(develop)

        def per_step(tol, mx):
            sim.setMdfControl(tol, mx, True)
            sim.run(6)
            dt, _ = timed_run(sim, steps)
            return dt / steps * 1e6
        t1, t16 = per_step(1.0, 1), per_step(1.0, 16)
        a1, a4 = per_step(1e-300, 1), per_step(1e-300, 4)
        out[prec] = {"applied_correction_us": round((a4 - a1) / 3, 2),
                     "iteration_after_convergence_us": round((t16 - t1) / 15, 2),
                     "step_with_one_correction_us": round(a1, 1)}
    return out


def mdf_stability_bench(steps=4000, chunk=250):
    """The reference's plain 5-iteration MDF (FP64, the oracle's arithmetic) on
    the MDF channel at the original specification's body force 3e-5, with
    C = 10, and with the capsules seeded into a 30 % band: the step at which
    each first shows non-finite cells or a speed above 0.5, or None when it
    ran ``steps`` steps.  The benchmark channel (body force 3e-6) is run the
    same way for comparison."""
    cases = {"benchmark channel (fx 3e-6, C 1, whole channel)": {},
             "fx 3e-5": {"bfx": 3e-5},
             "C 10": {"mem": {"C_skalak": 10.0}},
             "30% band (164 x 504)": {"region": (942, 4, 1106, 508)}}
    out = {}
    for name, changes in cases.items():
        sim, placed = build_mdf_channel("fp64", 5, USE_GRAPH, **changes)
        at, umax = None, 0.0
        try:
            for s in range(0, steps, chunk):
                sim.run(chunk)
                fs = sim.fieldStats()
                umax = fs[3]
                if fs[4] > 0 or not np.isfinite(fs[3]) or fs[3] > 0.5:
                    at = s + chunk
                    break
        except RuntimeError as e:   # a capacity overflow as the capsules collapse
            at = f"{s + chunk} ({e})"
        out[name] = {"capsules": placed, "diverged_by_step": at, "max_speed": umax, "steps": steps}
        del sim
    return out


def ensemble_bench(steps, count=256, seed=7, seq_steps=60):
    """Batched ensemble: ``count`` channels of uneven size drawn from one seed
    (nx 200-800, ny 40-160 in steps of 8, 5-60 Skalak capsules of radius
    2.5-6 with the node count set by the radius, tau 0.6-1.0, fx 1e-5 to
    5e-5; BGK, FP32 lattice storage), advanced by one launch per kernel of the
    step over all members with six-step graph replay.  Reported: whole-ensemble
    MLUPS, the same members stepped one after another (each with its own graph
    replay), and the batched collide-stream against the HBM roofline (its
    launch time from the autotuner's HIP-event timing of the chosen shape)."""
    from softflow_amd import softflow_core as sc
    from softflow_amd.ensemble import build_member, ensemble_specs

    specs = ensemble_specs(count, seed)
    sims, placed = [], 0
    for sp in specs:
        sim, n = build_member(sp, "fp32")
        sims.append(sim)
        placed += n
    cells = sum(sp["nx"] * sp["ny"] for sp in specs)
    ens = sc.Ensemble(sims)
    info = ens.info()
    ens.run(7)
    tune = ens.autotune(10)
    chosen = tune.strip().splitlines()[-1]
    want = chosen[len("chosen "):]
    lbm_ms = None
    for line in tune.strip().splitlines()[:-1]:
        desc, ms, _ = line.rsplit(" ", 2)          # "<launch description> <ms> ms"
        if desc == want:
            lbm_ms = float(ms)
    ens.run(12)
    ens.synchronize()
    t = time.perf_counter()
    ens.run(steps)
    ens.synchronize()
    dt = time.perf_counter() - t
    dt_seq = None
    if seq_steps > 0:
        for sim in sims:
            sim.run(6)
        t = time.perf_counter()
        for sim in sims:
            sim.run(seq_steps)      # returns after the device is done
        dt_seq = time.perf_counter() - t
    worst = max(sim.fieldStats()[3] for sim in sims)
    bad = sum(1 for sim in sims if sim.fieldStats()[4] > 0)
    out = {"workload": f"{count} x-periodic BGK channels drawn with seed {seed}: nx 200-800, ny 40-160, "
                       "5-60 Skalak capsules r 2.5-6 (auto nodes), tau 0.6-1.0, fx 1e-5..5e-5; FP32",
           "members": count, "cells": cells, "capsules": placed, "membrane_nodes": int(info[2]),
           "steps": steps, "ms_per_step": round(dt * 1e3 / steps, 5),
           "mlups": round(cells * steps / dt / 1e6, 1),
           "sequential_mlups": round(cells * seq_steps / dt_seq / 1e6, 1) if dt_seq else None,
           "lbm_launch": chosen, "max_speed": worst, "nonfinite_members": bad}
    if dt_seq:
        out["speedup_vs_sequential"] = round(out["mlups"] / out["sequential_mlups"], 2)
    if lbm_ms:
        achieved = LBM_BYTES_PER_LUP * cells / (lbm_ms * 1e-3) / 1e9
        out["lbm_roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                               "kernel": "kb_lbm<BGK,fp32> (batched collide + push-stream)",
                               "kernel_ms": lbm_ms, "bytes_per_lup": LBM_BYTES_PER_LUP}
    return out


def output_bench(case, steps):
    """The writers on the 4096x1024 workload.  (1) The step rate with the
    reference's default statistics (every 100 steps) and probe (every 10 steps,
    4 probes) output against the same workload without output, each measured
    twice, interleaved, in this process.  (2) One VTK + CSV dump: wall time
    from the snapshot to the files closed, bytes written.  (3) The reference's
    own VTK writer (oracle/_ref, one thread) on a 1024x256 field of the same
    kind, where that build exists, as the CPU rate."""
    import shutil
    import tempfile
    from softflow_amd import softflow_core as sc

    def rate(sim):
        dt, _ = timed_run(sim, steps)
        return dt * 1e3 / steps

    out = {}
    tmp = tempfile.mkdtemp(prefix="sf_out_")
    try:
        plain = build_product(case, "fp32", graph=USE_GRAPH)
        plain.run(20)
        # default statistics and probe intervals of SimulationParams
        p2 = dict(output_dir=tmp, stats_dump_every=100, probe_dump_every=10)
        sim = build_product(dict(case, output=p2), "fp32", graph=USE_GRAPH)
        for k, (i, j) in enumerate([(100, 100), (2048, 512), (3000, 900), (4000, 20)]):
            sim.addFluidProbe(i, j, f"p{k}")
        sim.run(20)
        base = [rate(plain)]
        ms = [rate(sim)]
        base.append(rate(plain))
        ms.append(rate(sim))
        del plain
        base_ms, ms = sum(base) / 2, sum(ms) / 2
        out["base_ms_per_step"] = round(base_ms, 5)
        out["stats_probes_ms_per_step"] = round(ms, 5)
        out["stats_probes_overhead"] = round(ms / base_ms - 1.0, 4)
        # one VTK + CSV dump at the next step (the writer thread formats on 16 threads)
        sim.setVTKOutput(1, "ascii", sc.FluidOutputFields(), sc.ParticleVTKFields())
        sim.setCSVOutput(1, "csv", sc.ParticleOutputFields(), sc.ParticleFilter(), True)
        t = time.perf_counter()
        sim.run(1)                       # finalize(): waits for the files
        dump_s = time.perf_counter() - t
        nbytes = 0
        for root, _, files in os.walk(tmp):
            for fn in files:
                if fn.endswith((".vti", ".vtp")) or fn == "particle_data.csv":
                    nbytes += os.path.getsize(os.path.join(root, fn))
        out["vtk_csv_dump_s"] = round(dump_s, 3)
        out["vtk_csv_dump_mb"] = round(nbytes / 1e6, 1)
        out["vtk_csv_dump_mb_per_s"] = round(nbytes / 1e6 / dump_s, 1)
        del sim
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    try:
        from oracle import pyoracle as po
        if os.path.exists(po.REF_SO):
            rtmp = tempfile.mkdtemp(prefix="sf_ref_out_")
            try:
                r = po.Sim(po.default_params(nx=1024, ny=256, tau=0.7, boundary_type=1, body_force_x=1e-5,
                                             collision_model=1, scalar_enabled=1, diffusivity=[0.05]), "ref")
                r.initialize()
                r.step(2)
                r.output_open(rtmp)
                t = time.perf_counter()
                r.output_write(0, 0.0, vtk=True)
                rdt = time.perf_counter() - t
                r.output_close()
                rb = sum(os.path.getsize(os.path.join(a, f)) for a, _, fs in os.walk(rtmp) for f in fs
                         if f.endswith(".vti"))
                out["reference_vtk_writer"] = {"mb_per_s": round(rb / 1e6 / rdt, 1), "cores": 1,
                                               "sample": f"VTKWriter::writeFluidField on 1024x256 with the "
                                                         f"scalar ({rb / 1e6:.1f} MB in {rdt:.2f} s)"}
            finally:
                shutil.rmtree(rtmp, ignore_errors=True)
    except Exception as e:  # the reference build is optional on this host
        out["reference_vtk_writer"] = {"error": str(e)}
    return out


def geometry_bench(steps, cpu_steps, threads, with_cpu=True):
    """The pillar array (pillar_case): whole-step rate with obstacles, their contact
    forces and interpolated bounce-back, and the reference's own components on the
    same workload on the host cores (where oracle/_ref exists)."""
    case = pillar_case()
    sim = build_product(case, "fp32", graph=USE_GRAPH)
    sim.run(10)
    tune = sim.autotune(10).strip().splitlines()[-1].strip()   # as the 4096x1024 channel is tuned
    sim.run(20)
    dt, _ = timed_run(sim, steps)
    out = {"workload": "4096x1024 MRT tau=0.7 fx=1e-5 through 200 staggered circular pillars (r 24) "
                       "with interpolated bounce-back, 1000 Skalak+Helfrich capsules, ADR D=0.05; FP32",
           "pillars": len(case["obstacles"]), "capsules": sum(c["count"] for c in case["capsules"]),
           "ms_per_step": round(dt * 1e3 / steps, 5), "mlups": round(NX * NY * steps / dt / 1e6, 1),
           "lbm_launch": tune}
    del sim
    if with_cpu:
        try:
            v, cdt, kind, cores, flags = cpu_baseline(case, cpu_steps, threads)
            out["cpu_baseline"] = {"value": round(v, 3), "unit": "MLUPS", "cores": cores, "kind": kind,
                                   "flags": flags,
                                   "sample": f"{cpu_steps} full steps of the pillar workload ({cdt:.1f} s)"}
        except Exception as e:
            out["cpu_baseline"] = {"value": None, "sample": f"unavailable: {e}"}
    return out


def calibrate_copy(mib):
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}))


def traffic_from_profile():
    """HBM bytes per launch of the collide-stream kernel, as measured with the
    FETCH_SIZE and WRITE_SIZE counters and scaled by the copy-kernel calibration
    (profiles/k_lbm_traffic_4096x1024.txt); None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_lbm_traffic_4096x1024.txt")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing-pass", action="store_true")
    ap.add_argument("--lbm-launch", default="auto",
                    help="'auto' (autotune) or BXxBY[xl|xpN] for a fixed collide-stream launch shape "
                         "(l: linear workgroup order, pN: persistent grid of N workgroups per XCD)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (the known-size "
                         "reference for the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--no-mdf", action="store_true",
                    help="skip the learned-vs-plain MDF comparison")
    ap.add_argument("--mdf-steps", type=int, default=20000)
    ap.add_argument("--mdf-train-steps", type=int, default=160)
    ap.add_argument("--mdf-develop", type=int, default=MDF_DEVELOP,
                    help="steps of the reference's 5-iteration MDF before the timed segments")
    ap.add_argument("--no-mdf-curve", action="store_true",
                    help="skip the residual-after-k-corrections curves of the MDF comparison")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched-ensemble measurement")
    ap.add_argument("--ensemble-steps", type=int, default=2000)
    ap.add_argument("--ensemble-sequential-steps", type=int, default=60,
                    help="steps of the members run one after another for comparison (0: skip)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel eagerly instead of replaying HIP graphs (for tracing)")
    ap.add_argument("--no-output", action="store_true",
                    help="skip the writer measurement")
    ap.add_argument("--output-steps", type=int, default=500)
    ap.add_argument("--no-geometry", action="store_true",
                    help="skip the pillar-array measurement")
    ap.add_argument("--only", choices=["ensemble", "mdf", "mdf-costs", "mdf-stability", "output", "geometry"],
                    default=None,
                    help="run only the batched-ensemble or the learned-vs-plain MDF measurement "
                         "and print its object as the JSON line (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    global USE_GRAPH
    USE_GRAPH = not args.no_graph
    # progress lines of the native engine go to stderr; stdout carries the JSON line only
    json_fd = os.dup(1)
    os.dup2(2, 1)
    if args.only:
        if args.only == "ensemble":
            part = ensemble_bench(args.ensemble_steps, seq_steps=args.ensemble_sequential_steps)
        elif args.only == "mdf":
            part = mdf_learned_bench(args.mdf_steps, args.mdf_train_steps, 6, not args.no_mdf_curve,
                                     args.mdf_develop)
        elif args.only == "mdf-costs":
            part = mdf_iteration_costs()
        elif args.only == "mdf-stability":
            part = mdf_stability_bench()
        elif args.only == "geometry":
            part = geometry_bench(args.steps, 5, cpu_threads(), not args.no_cpu)
        else:
            part = output_bench(large_domain_case(), args.output_steps)
        sys.stdout.flush()
        os.write(json_fd, (json.dumps(part) + "\n").encode())
        return

    case = large_domain_case()
    t0 = time.perf_counter()
    sim = build_product(case, "fp32", graph=USE_GRAPH)
    print(f"# setup {time.perf_counter() - t0:.1f}s, {sum(c['count'] for c in case['capsules'])} capsules",
          file=sys.stderr, flush=True)
    sim.run(min(args.warmup, 10))
    if args.lbm_launch == "auto":
        tune = sim.autotune(10)
        print("# autotune:\n" + "".join("#   " + l + "\n" for l in tune.strip().splitlines()),
              file=sys.stderr, flush=True)
    else:
        parts = args.lbm_launch.split("x")
        bx, by = int(parts[0]), int(parts[1])
        banded = not (len(parts) > 2 and parts[2] == "l")
        persist = int(parts[2][1:]) if len(parts) > 2 and parts[2].startswith("p") else 0
        sim.setLbmLaunch(bx, by, banded, persist)
        tune = ("fixed %dx%d persistent %d/xcd" % (bx, by, persist) if persist else
                "fixed %dx%d %s" % (bx, by, "xcd-banded" if banded else "linear"))
        # this workload (FP32, x-periodic channel, nx even) takes the paired kernel by default
        tune += " 2 cells/lane"
    sim.run(args.warmup)
    # the timed region replays graphs only: those it needs are captured first
    dt, captures = timed_run(sim, args.steps)
    if captures:
        raise RuntimeError(f"{captures} step graphs were captured inside the timed region")
    ms_per_step = dt * 1e3 / args.steps
    mlups = NX * NY * args.steps / dt / 1e6
    print(f"# graph replay: {ms_per_step:.4f} ms/step, {mlups:.0f} MLUPS", file=sys.stderr, flush=True)

    roof = None
    kt = {}
    if not args.no_timing_pass:
        sim.setTiming(True)
        t = time.perf_counter()
        sim.run(args.steps)
        dt_eager = time.perf_counter() - t
        kt = dict(sim.kernelTimes())
        sim.setTiming(False)
        # with the scalar on, the collide-stream kernel also carries the scalar
        # step: 9 scalar distributions read + 9 written per cell on top
        fused = "lbm+adr" in kt
        paired = "2 cells/lane" in tune.strip().splitlines()[-1]
        lbm_ms = kt.get("lbm+adr", kt.get("lbm"))
        bpl = LBM_BYTES_PER_LUP + (ADR_BYTES_PER_LUP if fused else 0)
        if lbm_ms:
            achieved = bpl * NX * NY / (lbm_ms * 1e-3) / 1e9
            roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic_from_profile(),
                    "kernel": ("k_lbm_pair<MRT,scalar> (FP32, two cells per lane)" if paired else
                               "k_lbm<MRT,fp32,channel,scalar>" if fused else "k_lbm<MRT,fp32>")
                              + (": fused collide + push-stream + lagged scalar step" if fused
                                 else ": fused collide + push-stream"),
                    "kernel_ms": round(lbm_ms, 5), "bytes_per_lup": bpl}
        print(f"# event-timed pass: {dt_eager * 1e3 / args.steps:.4f} ms/step; per-kernel ms: "
              + ", ".join(f"{k}={v:.4f}" for k, v in sorted(kt.items())), file=sys.stderr, flush=True)

    stats = sim.fieldStats()
    del sim
    # the achievable streaming rate on this device, beside the 8 TB/s peak
    copy_ref = None
    try:
        from softflow_amd import capi
        gbs = capi.copy_ceiling(1 << 30, 10)
        copy_ref = {"value": round(gbs, 1), "unit": "GB/s",
                    "kernel": "nontemporal 16-byte streaming copy, 1 GiB each way (sf_copy_ceiling)",
                    "collide_stream_frac": round(roof["achieved"] / gbs, 4) if roof and gbs > 0 else None}
    except Exception as e:   # reported, so the main line is still printed
        copy_ref = {"value": None, "error": str(e)}
    print("# copy ceiling: " + json.dumps(copy_ref), file=sys.stderr, flush=True)
    outb = None
    if not args.no_output:
        try:
            outb = output_bench(case, args.output_steps)
        except RuntimeError as e:
            outb = {"error": str(e)}
        print("# writers: " + json.dumps(outb), file=sys.stderr, flush=True)
    geo = None
    if not args.no_geometry:
        try:
            geo = geometry_bench(min(args.steps, 1000), 5, cpu_threads(), not args.no_cpu)
        except RuntimeError as e:
            geo = {"error": str(e)}
        print("# pillar array: " + json.dumps(geo), file=sys.stderr, flush=True)
    mdf = None
    if not args.no_mdf:
        try:
            mdf = mdf_learned_bench(args.mdf_steps, args.mdf_train_steps, 6, not args.no_mdf_curve,
                                    args.mdf_develop)
        except RuntimeError as e:   # reported, so the main line is still printed
            mdf = {"error": str(e)}
        print("# learned vs plain MDF: " + json.dumps(mdf), file=sys.stderr, flush=True)
    ens = None
    if not args.no_ensemble:
        try:
            ens = ensemble_bench(args.ensemble_steps, seq_steps=args.ensemble_sequential_steps)
        except RuntimeError as e:
            ens = {"error": str(e)}
        print("# batched ensemble: " + json.dumps(ens), file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        try:
            threads = cpu_threads()
            v, cdt, kind, cores, flags = cpu_baseline(case, args.cpu_steps, threads)
            what = ("the reference's own components (OpenMP, FP64)" if kind == "reference"
                    else "the FP64 oracle restatement")
            model, nproc = host_cpu()
            cpu = {"value": round(v, 3), "unit": "MLUPS", "cores": cores, "kind": kind,
                   "flags": flags, "host_cpu": model, "host_nproc": nproc,
                   "sample": f"{args.cpu_steps} full steps of the same 4096x1024 + 2000-capsule "
                             f"workload in {what} ({cdt:.1f} s)"}
        except Exception as e:  # the oracle library is test infrastructure; report its absence
            cpu = {"value": None, "unit": "MLUPS", "cores": 1, "kind": "port", "sample": f"unavailable: {e}"}

    out = {
        "metric": "MLUPS + % HBM roofline, 4096x1024 + 2000 capsules; learned vs plain MDF steps/s",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "large single domain: 4096x1024 MRT tau=0.7 fx=1e-5, 2000 bidisperse "
                               "Skalak+Helfrich capsules, ADR D=0.05; FP32 lattice storage",
                   "nx": NX, "ny": NY, "capsules": 2000, "membrane_nodes": 800 * 38 + 1200 * 19},
        "roofline": roof,
        "copy_ceiling": copy_ref,
        "cpu_baseline": cpu,
        "mdf_learned_vs_plain": mdf,
        "batched_ensemble": ens,
        "writers": outb,
        "pillar_array": geo,
        "kernel_ms": {k: round(v, 5) for k, v in sorted(kt.items())},
        "lbm_launch": tune.strip().splitlines()[-1],
        "graph_captures_in_timed_region": captures,
        "mass": stats[0],
        "max_speed": stats[3],
    }
    sys.stdout.flush()
    os.write(json_fd, (json.dumps(out) + "\n").encode())


if __name__ == "__main__":
    main()
