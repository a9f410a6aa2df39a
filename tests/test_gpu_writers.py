"""The engine's writers against the reference's, file for file.

tests/golden/writers/<case>/ holds what the reference's own writers
(io/vtk_writer.cpp, csv_writer.cpp, fluid_probe_writer.cpp,
global_stats_writer.cpp) wrote on reference-component runs of each case
(tests/golden/make_golden.py).  The engine runs the same case in FP64 (whose
state equals the reference's bit for bit on these cases) with the same output
configuration; its files, written by the writer thread from asynchronous
device snapshots, must equal the reference's byte for byte.  The statistics
file is the exception: its sums come from device reductions whose order
differs from the reference's serial loop, so each value is held to one unit in
the last printed digit.
"""
import gzip
import json
import os

import numpy as np
import pytest

from tests.cases import product_sim

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "writers")
with open(os.path.join(HERE, "cases.json")) as fh:
    CASES = json.load(fh)


def _golden_files(name):
    root = os.path.join(HERE, name)
    out = {}
    for r, _, files in os.walk(root):
        for fn in files:
            if fn.endswith(".gz"):
                rel = os.path.relpath(os.path.join(r, fn), root)[:-3]
                with gzip.open(os.path.join(r, fn), "rb") as g:
                    out[rel] = g.read()
    return out


def _run_product(w, outdir):
    from softflow_amd import softflow_core as sc

    every = w["every"]
    params = dict(output_dir=str(outdir), output_interval=every, probe_dump_every=every,
                  stats_dump_every=every, vtk_format="legacy" if w["legacy"] else "ascii")
    sim = product_sim(dict(w["case"], params=params))
    if w["extra_fluid"]:
        ff = sc.FluidOutputFields()
        ff.vorticity = ff.strain_rate = ff.ibm_force = True
        sim.setVTKOutput(every, "legacy" if w["legacy"] else "ascii", ff, sc.ParticleVTKFields())
    if w["all_csv"]:
        pf = sc.ParticleOutputFields()
        pf.force = pf.diameter = pf.group = pf.deformation = pf.area_volume = True
        pf.orientation = pf.angular_vel = True
        sim.setCSVOutput(every, "csv", pf, sc.ParticleFilter(), True)
    for k, (i, j) in enumerate(w["probes"]):
        sim.addFluidProbe(i, j, f"p{k}")
    sim.run(w["steps"])          # ends with finalize(): .pvd files written, writers closed
    return sim


def _stats_close(a, b):
    la, lb = a.decode().splitlines(), b.decode().splitlines()
    assert len(la) == len(lb) and la[0] == lb[0]
    for ra, rb in zip(la[1:], lb[1:]):
        va, vb = ra.split(","), rb.split(",")
        assert len(va) == len(vb)
        for x, y in zip(va, vb):
            if x == y:
                continue
            fx, fy = float(x), float(y)
            assert abs(fx - fy) <= 1.0000001e-6 * max(abs(fx), abs(fy)), (ra, rb)


@pytest.mark.parametrize("name", sorted(CASES))
def test_writers_match_reference(name, tmp_path):
    w = CASES[name]
    print(f"writers {name}: {w['steps']} steps, output every {w['every']}", flush=True)
    _run_product(w, tmp_path)
    ref = _golden_files(name)
    got = {}
    for r, _, files in os.walk(tmp_path):
        for fn in files:
            rel = os.path.relpath(os.path.join(r, fn), tmp_path)
            with open(os.path.join(r, fn), "rb") as fh:
                got[rel] = fh.read()
    assert sorted(got) == sorted(ref)
    for rel, data in ref.items():
        if rel == "global_stats.csv":
            _stats_close(got[rel], data)
            continue
        if got[rel] != data:
            ga, ra = got[rel].decode().splitlines(), data.decode().splitlines()
            first = next((i for i, (x, y) in enumerate(zip(ga, ra)) if x != y), min(len(ga), len(ra)))
            pytest.fail(f"{rel}: first difference at line {first + 1}: "
                        f"{ga[first] if first < len(ga) else '<eof>'!r} != {ra[first] if first < len(ra) else '<eof>'!r}")


def test_output_does_not_change_the_run(tmp_path):
    # writers on (every step) and off: the same state after the same steps
    w = CASES["periodic_caps_scalar"]
    a = _run_product(dict(w, every=1, steps=23), tmp_path)
    b = product_sim(w["case"])
    b.run(23)
    fa, fb = a.field(), b.field()
    assert np.array_equal(np.asarray(fa.velocity_x()), np.asarray(fb.velocity_x()))
    assert np.array_equal(np.asarray(a.advectionDiffusion().concentration(0)),
                          np.asarray(b.advectionDiffusion().concentration(0)))
    ca, cb = a.capsules(), b.capsules()
    for i in range(len(ca)):
        assert ca[i].nodePosition(3).x == cb[i].nodePosition(3).x
    n = len([f for f in os.listdir(os.path.join(tmp_path, "fluid")) if f.endswith(".vti")])
    assert n == 23
