"""Write the golden vectors under tests/golden/ from the reference build.

Runs only where the reference tree exists: it loads oracle/_ref/libsfref.so,
the reference's own hot-path components (compiled from its sources by
``make -C oracle ref``) driven in Simulation::step order at one OpenMP thread,
and stores inputs (the case dict) and outputs (fields, node state) as .npz.

    python tests/golden/make_golden.py [--force]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.cases import STEP_PIN_CASES, oracle_sim, oracle_state  # noqa: E402

SKALAK = dict(model=2, G_s=0.06, C_skalak=10.0)

CASES = {
    "channel_bgk": (dict(nx=40, ny=18, tau=0.8, bfx=3e-5, coll=0), 800),
    "channel_mrt": (dict(nx=40, ny=18, tau=0.7, bfx=1e-5, coll=1), 800),
    "channel_reg": (dict(nx=40, ny=18, tau=1.2, bfx=2e-5, coll=2), 800),
    "closed_box_bgk": (dict(nx=24, ny=20, tau=0.9, bfx=2e-5, coll=0, boundary="closed"), 400),
    "capsules_reg_mdf3": (dict(nx=80, ny=40, tau=0.8, bfx=3e-5, coll=2, mdf=3,
                               capsules=[dict(count=4, region=(10, 8, 70, 32), rmin=3.0, rmax=3.0,
                                              nodes=20, seed=42, mem=SKALAK)]), 400),
    "capsules_bgk_celllist": (dict(nx=80, ny=40, tau=0.8, bfx=3e-5, coll=0,
                                   capsules=[dict(count=7, region=(8, 8, 72, 32), rmin=3.0, rmax=4.0,
                                                  nodes=0, seed=7, mem=SKALAK)]), 400),
    "capsules_helfrich_mrt_scalar": (dict(nx=64, ny=32, tau=0.7, bfx=1e-5, coll=1,
                                          scalar=dict(D=0.05, blob=(32, 16, 5.0)),
                                          capsules=[dict(count=3, region=(8, 8, 56, 24), rmin=3.0,
                                                         rmax=4.0, nodes=0, seed=5,
                                                         mem=dict(model=2, G_s=0.06, C_skalak=10.0,
                                                                  use_helfrich=1, k_bend=1e-3))]), 300),
    "capsules_wlc_neo_hooke": (dict(nx=64, ny=32, tau=0.8, bfx=3e-5, coll=0,
                                    capsules=[dict(count=2, region=(8, 8, 56, 24), rmin=3.5, rmax=3.5,
                                                   nodes=0, seed=11, mem=dict(model=3, G_s=0.05)),
                                              dict(count=2, region=(8, 8, 56, 24), rmin=3.0, rmax=3.0,
                                                   nodes=0, seed=12, mem=dict(model=1, G_s=0.05,
                                                                              shape=1, aspect_ratio=0.6)),
                                              dict(count=1, region=(8, 8, 56, 24), rmin=3.0, rmax=3.0,
                                                   nodes=16, seed=13, mem=dict(model=0, k_stretch=0.05,
                                                                               shape=2))]), 300),
    # the channel builder's pillars with interpolated bounce-back; capsules
    # against them (contact with damping and friction)
    "pillars_ibb_mrt_caps": (dict(nx=96, ny=36, tau=0.8, bfx=3e-5, coll=1, ibb=True,
                                  obstacles=[("circle", 30.3, 18.6, 5.2), ("rect", 55.0, 8.5, 62.0, 15.2),
                                             ("polygon", [(75.0, 22.0), (86.0, 19.5), (82.0, 29.0)])],
                                  capsules=[dict(centers=[(30.3, 28.2), (58.5, 4.6), (80.0, 16.0)],
                                                 rmin=3.5, nodes=22, mem=SKALAK)],
                                  rep=dict(damping_normal=0.01, friction_coeff=0.1)), 300),
    # a fluid run two cells long between a wall and a block (cells whose links
    # read slots another cell's links write), periodic seam pillar
    "ibb_gap_seam": (dict(nx=48, ny=10, tau=0.8, bfx=2e-5, coll=0, ibb=True,
                          obstacles=[("rect", 10.0, 2.3, 30.0, 6.0), ("circle", 40.5, 5.0, 1.6),
                                     ("circle", 1.5, 4.5, 1.2)]), 400),
}
# the step's own operations: force cap firing at the last step, nodes clamped
# at the wall lines, capsules recirculating over a non-periodic fluid
for _name, _c in STEP_PIN_CASES.items():
    _c = dict(_c)
    CASES[_name] = (_c, _c.pop("steps"))


# The reference's writers (io/*_writer.cpp) on reference-component runs: each
# case's files, gzipped, under writers/<case>/, written at every output step of
# Simulation::step's schedule (simulation.cpp:736-767) and closed as
# Simulation::finalize does.
WRITER_CASES = {
    "periodic_caps_scalar": dict(
        case=dict(nx=32, ny=14, tau=0.8, bfx=3e-5, coll=0, scalar=dict(D=0.05, blob=(14, 7, 3.0)),
                  capsules=[dict(centers=[(2.5, 7.0), (19.0, 6.5)], rmin=3.0, nodes=16, mem=SKALAK)]),
        steps=11, every=10, legacy=False, all_csv=True, extra_fluid=True, probes=[(4, 4), (16, 7), (40, -3)]),
    "closed_legacy_mrt": dict(
        case=dict(nx=30, ny=16, tau=0.9, bfx=2e-5, coll=1, boundary="closed",
                  capsules=[dict(centers=[(12.0, 8.0)], rmin=3.5, nodes=18, mem=SKALAK)]),
        steps=11, every=5, legacy=True, all_csv=False, extra_fluid=False, probes=[(10, 8)]),
}


def write_writer_cases(force):
    import gzip
    import shutil
    import tempfile

    for name, w in WRITER_CASES.items():
        dest = os.path.join(HERE, "writers", name)
        if os.path.isdir(dest) and not force:
            continue
        tmp = tempfile.mkdtemp()
        sim = oracle_sim(w["case"], which="ref")
        sim.output_open(tmp, legacy=w["legacy"], all_csv=w["all_csv"], extra_fluid=w["extra_fluid"],
                        probes=w["probes"])
        for st in range(w["steps"]):
            sim.step(1)
            if st % w["every"] == 0:
                sim.output_write(st, float(st), vtk=True, csv=True, probes=True, stats=True)
        sim.output_close()
        shutil.rmtree(dest, ignore_errors=True)
        for root, _, files in os.walk(tmp):
            for fn in files:
                rel = os.path.relpath(os.path.join(root, fn), tmp)
                out = os.path.join(dest, rel + ".gz")
                os.makedirs(os.path.dirname(out), exist_ok=True)
                with open(os.path.join(root, fn), "rb") as src, open(out, "wb") as raw:
                    with gzip.GzipFile(filename="", mode="wb", fileobj=raw, mtime=0) as gz:
                        gz.write(src.read())
        shutil.rmtree(tmp)
        print("writers", name, "saved", flush=True)
    with open(os.path.join(HERE, "writers", "cases.json"), "w") as f:
        json.dump(WRITER_CASES, f, indent=1)


def main():
    from oracle import pyoracle

    if not pyoracle.ref_available():
        sys.exit("oracle/_ref/libsfref.so is not built (make -C oracle ref)")
    force = "--force" in sys.argv
    index = {}
    for name, (case, steps) in CASES.items():
        index[name] = {"case": case, "steps": steps}
        path = os.path.join(HERE, f"{name}.npz")
        if os.path.exists(path) and not force:
            continue            # vectors already written (--force rewrites them)
        sim = oracle_sim(case, which="ref")
        sim.step(steps)
        st = oracle_state(sim)
        arrays = {k: np.asarray(v) for k, v in st.items() if k != "capped"}
        arrays["capped"] = np.array(st["capped"])
        np.savez_compressed(path, **arrays)
        print(name, "saved", flush=True)
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(index, f, indent=1)
    write_writer_cases(force)


if __name__ == "__main__":
    main()
