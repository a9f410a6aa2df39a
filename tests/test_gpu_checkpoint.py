"""Checkpoints in the reference's .sfck layout (io/checkpoint.cpp), version 4.

- a run saved, reloaded into a fresh simulation and continued ends bit for
  bit where the uninterrupted run ends (FP64 and FP32; capsules with
  Kelvin-Voigt membranes, fixed MDF, the scalar), through the extension block;
- the file's layout is the reference's, field by field, against the state it
  was written from;
- without the extension block the restart is the reference's (the next
  collision takes rho and u from the file), still at the FP64 parity bar;
- mismatched or truncated files are refused and change nothing;
- checkpoint_interval writes checkpoint_<step>.sfck after the steps whose
  index is a positive multiple of it (simulation.cpp:773-779).
"""
import os

import numpy as np
import pytest

from tests.cases import product_sim, product_state, rel_l2

pytestmark = pytest.mark.gpu

KV = dict(model=2, G_s=0.06, C_skalak=10.0, eta_membrane=0.5)
CASE = dict(nx=96, ny=40, tau=0.8, bfx=3e-5, coll=1,
            capsules=[dict(count=6, region=(4, 4, 92, 36), rmin=3.0, rmax=4.0, seed=7, mem=KV)],
            scalar=dict(D=0.05, blob=(48, 20, 6.0)))
KEYS = ("rho", "ux", "uy", "x", "y", "vx", "vy", "fx_node", "fy_node", "C")


def _same(a, b, what):
    for k in KEYS:
        if k in a or k in b:
            assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"


@pytest.mark.parametrize("variant", ["fp64", "fp32", "mdf3_fp64"])
def test_resume_is_bit_exact(tmp_path, variant):
    case = dict(CASE, precision="fp32" if variant == "fp32" else "fp64")
    if variant.startswith("mdf3"):
        case = dict(case, coll=2, mdf=3)
        case.pop("scalar")
    path = str(tmp_path / "run.sfck")
    print(f"checkpoint resume {variant}", flush=True)
    a = product_sim(case)
    a.run(37)
    assert a.saveCheckpoint(path)
    a.run(25)
    ref = product_state(a)
    b = product_sim(case)
    b.run(3)                           # a different state, overwritten by the load
    assert b.loadCheckpoint(path)
    assert b.currentStep() == 37
    b.run(25)
    got = product_state(b)
    _same(got, ref, variant)


def _read_sfck(path):
    """Parse a version-4 file as io/checkpoint.cpp:10-58 lays it out."""
    buf = open(path, "rb").read()
    o = 0

    def take(fmt, n=1):
        nonlocal o
        size = np.dtype(fmt).itemsize * n
        v = np.frombuffer(buf, dtype=np.dtype(fmt), count=n, offset=o)
        o += size
        return v

    assert buf[:4] == b"SFCK"
    o = 4
    version, step, nx, ny, ncaps = take("<i4", 5)
    N = nx * ny
    out = dict(version=version, step=step, nx=nx, ny=ny, ncaps=ncaps)
    out["f"] = take("<f8", 9 * N)
    out["rho"], out["ux"], out["uy"] = take("<f8", N), take("<f8", N), take("<f8", N)
    out["flags"] = take("<i4", N)
    caps = []
    for _ in range(ncaps):
        cid, tid, nn = take("<i4", 3)
        pos = take("<f8", 2 * nn).reshape(nn, 2)
        vel = take("<f8", 2 * nn).reshape(nn, 2)
        area, perim = take("<f8", 2)
        pnx = take("<i4")[0]
        caps.append(dict(id=cid, type=tid, pos=pos, vel=vel, area=area, perim=perim, pnx=pnx))
    out["caps"] = caps
    out["nbonds"] = take("<i4")[0]
    out["has_scalar"] = take("<i4")[0]
    if out["has_scalar"]:
        nsp = take("<i4")[0]
        out["g"] = [None] * nsp
        out["C"] = [None] * nsp
        for s in range(nsp):
            out["g"][s] = take("<f8", 9 * N)
            out["C"][s] = take("<f8", N)
        out["trackers"] = (take("<i4")[0], take("<i4")[0], take("<i4")[0])
    out["ext_offset"] = o
    out["ext_magic"] = buf[o:o + 4]
    o += 4
    if out["ext_magic"] == b"SFXT":
        out["ext_version"] = take("<i4")[0]
        out["force"] = take("<f8", 2 * N)
        for c in caps:
            c["has_prev"] = take("<i4")[0]
            c["prev_len"] = take("<f8", len(c["pos"]))
    out["end"] = o
    out["size"] = len(buf)
    return out


def test_file_layout_is_the_references(tmp_path):
    path = str(tmp_path / "layout.sfck")
    a = product_sim(dict(CASE, precision="fp64"))
    a.run(21)
    assert a.saveCheckpoint(path)
    d = _read_sfck(path)
    assert (d["version"], d["step"], d["nx"], d["ny"]) == (4, 21, 96, 40)
    fl = a.field()
    assert np.array_equal(d["f"], np.asarray(fl.distributions()).ravel())
    for k, v in (("rho", fl.density()), ("ux", fl.velocity_x()), ("uy", fl.velocity_y())):
        assert np.array_equal(d[k], np.asarray(v).ravel()), k
    assert d["flags"][0] == 1 and d["flags"][40 * 96 // 2] == 0     # bottom wall, interior
    caps = a.capsules()
    assert d["ncaps"] == len(caps)
    for i, c in enumerate(d["caps"]):
        cp = caps[i]
        assert c["id"] == cp.getId() and c["pnx"] == 96
        assert np.array_equal(c["pos"], np.array([[cp.nodePosition(k).x, cp.nodePosition(k).y]
                                                  for k in range(cp.numNodes())]))
        assert c["area"] == cp.area() and c["perim"] == cp.perimeter()
        assert c["has_prev"] == 1 and np.all(c["prev_len"] > 0)
    assert d["nbonds"] == 0 and d["has_scalar"] == 1 and d["trackers"] == (0, 0, 0)
    assert np.array_equal(d["C"][0], np.asarray(a.advectionDiffusion().concentration(0)).ravel())
    assert d["ext_magic"] == b"SFXT" and d["ext_version"] == 1
    assert d["end"] == d["size"]


def test_restart_without_extension_is_the_references(tmp_path):
    # the version-4 content alone: rho and u from the file drive the next
    # collision.  It has no Kelvin-Voigt state (the reference's restart of
    # viscoelastic membranes is not exact either), so the membranes here are
    # purely elastic.
    elastic = dict(CASE["capsules"][0], mem=dict(KV, eta_membrane=0.0))
    case = dict(CASE, precision="fp64", capsules=[elastic])
    path = str(tmp_path / "full.sfck")
    a = product_sim(case)
    a.run(30)
    assert a.saveCheckpoint(path)
    d = _read_sfck(path)
    plain = str(tmp_path / "v4only.sfck")
    with open(path, "rb") as src, open(plain, "wb") as dst:
        dst.write(src.read()[:d["ext_offset"]])
    a.run(40)
    ref = product_state(a)
    b = product_sim(case)
    assert b.loadCheckpoint(plain)
    b.run(40)
    got = product_state(b)
    for k in ("rho", "ux", "uy", "x", "y", "vx", "vy"):
        e = rel_l2(got[k], ref[k])
        print(f"v4-only restart {k}: rel L2 {e:.3e}", flush=True)
        assert e <= 1e-12, k


def test_bad_files_are_refused(tmp_path):
    path = str(tmp_path / "a.sfck")
    a = product_sim(dict(CASE, precision="fp64"))
    a.run(5)
    assert a.saveCheckpoint(path)
    other = product_sim(dict(nx=64, ny=40, tau=0.8, bfx=3e-5, coll=1))
    assert not other.loadCheckpoint(path)                       # domain mismatch
    short = str(tmp_path / "short.sfck")
    data = open(path, "rb").read()
    with open(short, "wb") as fh:
        fh.write(data[:len(data) // 3])
    b = product_sim(dict(CASE, precision="fp64"))
    b.run(2)
    before = product_state(b)
    assert not b.loadCheckpoint(short)                          # truncated: nothing changes
    assert b.currentStep() == 2
    _same(product_state(b), before, "after a refused load")
    bogus = str(tmp_path / "bogus.sfck")
    with open(bogus, "wb") as fh:
        fh.write(b"XXXX" + data[4:])
    assert not b.loadCheckpoint(bogus)
    assert not b.loadCheckpoint(str(tmp_path / "missing.sfck"))


def test_checkpoint_interval_files(tmp_path):
    case = dict(CASE, precision="fp32", params=dict(output_dir=str(tmp_path), checkpoint_interval=10))
    case.pop("scalar")
    sim = product_sim(case)
    sim.run(25)
    names = sorted(f for f in os.listdir(tmp_path) if f.endswith(".sfck"))
    assert names == ["checkpoint_10.sfck", "checkpoint_20.sfck"], names
    for n, step in (("checkpoint_10.sfck", 10), ("checkpoint_20.sfck", 20)):
        d = _read_sfck(os.path.join(tmp_path, n))
        assert d["step"] == step and d["end"] == d["size"]
    # the file written after step index 20 holds the state after 21 steps
    plain = dict(case, params={})
    twin = product_sim(plain)
    twin.run(21)
    b = product_sim(plain)
    assert b.loadCheckpoint(os.path.join(tmp_path, "checkpoint_20.sfck"))
    sb, st = product_state(b), product_state(twin)
    for k in ("x", "y", "vx", "vy"):
        assert np.array_equal(sb[k], st[k]), k
    assert np.array_equal(np.asarray(b.field().distributions()), np.asarray(twin.field().distributions()))


def _expected_layout(nx, ny, node_counts, nsp):
    """Byte offsets the .sfck format fixes (io/checkpoint.cpp:10-58, 76-196):
    the end of the version-4 content (where the extension block starts) and the
    end of the file with the extension block, from the sizes alone."""
    N = nx * ny
    size = 4 + 5 * 4                      # magic; version, step, nx, ny, ncaps
    size += 9 * N * 8 + 3 * N * 8 + N * 4  # f (SoA), rho, ux, uy, cell types
    for n in node_counts:                 # id, type, nodes; positions, velocities; area, perimeter; periodic nx
        size += 3 * 4 + 2 * (2 * n * 8) + 2 * 8 + 4
    size += 4 + 4                         # adhesion bond count (0), has_scalar
    if nsp:
        size += 4 + nsp * (9 * N * 8 + N * 8) + 3 * 4   # species; g, C per species; 3 tracker counts
    ext = 4 + 4 + 2 * N * 8 + sum(4 + n * 8 for n in node_counts)   # SFXT, version, force, KV state
    return size, size + ext


@pytest.mark.parametrize("scalar", [True, False])
def test_file_size_follows_from_the_sizes(tmp_path, scalar):
    case = dict(CASE, precision="fp64")
    if not scalar:
        case.pop("scalar")
    a = product_sim(case)
    a.run(9)
    path = str(tmp_path / "sz.sfck")
    assert a.saveCheckpoint(path)
    caps = a.capsules()
    counts = [caps[i].numNodes() for i in range(len(caps))]
    v4_end, total = _expected_layout(96, 40, counts, 1 if scalar else 0)
    d = _read_sfck(path)
    assert d["ext_offset"] == v4_end and d["size"] == total == os.path.getsize(path)
    assert d["has_scalar"] == (1 if scalar else 0)


def test_versions_2_and_3_are_accepted_and_others_refused(tmp_path):
    # a version-3 file is the version-4 content without the two particle
    # counts after the tracker count; a version-2 file ends after the bond
    # count (io/checkpoint.cpp:213-222, 317-356)
    case = dict(CASE, precision="fp64")
    a = product_sim(case)
    a.run(12)
    path = str(tmp_path / "v4.sfck")
    assert a.saveCheckpoint(path)
    d = _read_sfck(path)
    raw = open(path, "rb").read()
    v4 = raw[:d["ext_offset"]]
    ver = np.frombuffer(v4, dtype="<i4", count=1, offset=4)[0]
    assert ver == 4

    def with_version(body, v):
        return body[:4] + np.array([v], dtype="<i4").tobytes() + body[8:]
    v3 = with_version(v4[:-8], 3)
    bonds_end = d["ext_offset"] - (4 + 4 + 9 * 96 * 40 * 8 + 96 * 40 * 8 + 12)   # before has_scalar
    v2 = with_version(v4[:bonds_end], 2)
    fl = np.asarray(a.field().distributions()).copy()
    for name, body in (("v3", v3), ("v2", v2)):
        p = str(tmp_path / f"{name}.sfck")
        open(p, "wb").write(body)
        b = product_sim(case)
        b.run(3)
        assert b.loadCheckpoint(p), name
        assert b.currentStep() == 12
        assert np.array_equal(np.asarray(b.field().distributions()), fl), name
        if name == "v3":
            assert np.array_equal(np.asarray(b.advectionDiffusion().concentration(0)),
                                  np.asarray(a.advectionDiffusion().concentration(0)))
    for bad in (1, 5):
        p = str(tmp_path / f"v{bad}.sfck")
        open(p, "wb").write(with_version(v4, bad))
        b = product_sim(case)
        assert not b.loadCheckpoint(p), bad
