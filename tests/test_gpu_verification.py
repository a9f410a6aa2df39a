"""The reference's own verification problems run on the HIP path.

The reference checks its engine against analytic solutions in
tests/verification/*.cpp; the same checks run here on the device, in FP64 and
in the FP32 deviation-storage mode, with the reference's tolerances:
Poiseuille profile and flow rate (test_poiseuille_collision_operators.cpp,
2 %), translation invariance (same file, 1e-10), Taylor-Green decay
(test_taylor_green.cpp, 2 %), pure diffusion (test_pure_diffusion.cpp, 2 %)
and Taylor dispersion (test_taylor_dispersion.cpp, 5 %); plus the start-up
series of the 400 x 80 Poiseuille channel at 10^4 steps and exact mass
conservation.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make(nx, ny, tau=0.8, coll=0, G=0.0, periodic_y=False, precision="fp64", D=None, walls=(0.0, 0.0)):
    from softflow_amd import softflow_core as sc
    p = sc.SimulationParams()
    p.output_interval, p.probe_dump_every, p.stats_dump_every = 0, 0, 0   # no output files
    p.nx, p.ny = nx, ny
    p.fluid.tau = tau
    p.fluid.boundary_type = sc.BoundaryType.PERIODIC
    p.fluid.periodic_y = periodic_y
    p.fluid.body_force_x = G
    p.fluid.top_wall_velocity, p.fluid.bottom_wall_velocity = walls
    p.fluid.collision_model = [sc.CollisionModel.BGK, sc.CollisionModel.MRT,
                               sc.CollisionModel.REGULARIZED][coll]
    if D is not None:
        p.scalar.enabled = True
        p.scalar.n_species = 1
        p.scalar.diffusivity = [D]
        p.scalar.periodic_y = periodic_y
    p.enable_stability_checks = False
    p.precision = sc.Precision.FP32 if precision == "fp32" else sc.Precision.FP64
    sim = sc.Simulation(p)
    sim.initialize()
    return sim


def ux(sim):
    return np.array(sim.field().velocity_x())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("coll", [0, 1, 2])
@pytest.mark.parametrize("tau", [0.8, 1.2])
def test_poiseuille_profile_and_flow_rate(precision, coll, tau):
    nx, ny, G = 8, 34, 1e-6
    nu = (tau - 0.5) / 3
    H = ny - 2
    sim = make(nx, ny, tau, coll, G, precision=precision)
    steps = int(5.0 * H * H / nu)
    print(f"poiseuille {precision} coll={coll} tau={tau} steps={steps}", flush=True)
    sim.run(steps)
    u = ux(sim)[1:H + 1, nx // 2]
    y = np.arange(H) + 0.5
    ua = G / (2 * nu) * y * (H - y)
    l2 = math.sqrt(np.sum((u - ua) ** 2) / np.sum(ua ** 2))
    q = u.sum() / (G * H ** 3 / (12 * nu))
    assert l2 < 0.02 and abs(q - 1) < 0.02, (l2, q)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("coll", [0, 1, 2])
def test_couette_profile(precision, coll):
    # moving wall rows at Ub and Ut (lbm/moving_wall.cpp); the steady profile
    # between the half-way walls is linear
    nx, ny, tau, Ut, Ub = 6, 22, 0.8, 0.02, -0.01
    H = ny - 2
    sim = make(nx, ny, tau, coll, 0.0, precision=precision, walls=(Ut, Ub))
    sim.run(int(8 * H * H / ((tau - 0.5) / 3)))
    u = ux(sim)[1:H + 1, nx // 2]
    ua = Ub + (Ut - Ub) * (np.arange(H) + 0.5) / H
    err = np.max(np.abs(u - ua)) / abs(Ut - Ub)
    print(f"couette {precision} coll={coll}: max rel err {err:.2e}", flush=True)
    assert err < (1e-6 if precision == "fp64" else 1e-4)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_translation_invariance(precision):
    out = {}
    for nx in (4, 32):
        sim = make(nx, 34, 0.8, 0, 1e-6, precision=precision)
        sim.run(20000)
        u = ux(sim)
        out[nx] = (np.max(np.ptp(u[1:33], axis=1)), u[1:33, nx // 2])
    tol = 1e-10 if precision == "fp64" else 1e-7
    assert out[4][0] < tol and out[32][0] < tol
    assert np.max(np.abs(out[4][1] - out[32][1])) < (1e-8 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_poiseuille_startup_series_400x80(precision):
    # the plain Poiseuille channel: BGK tau 0.8 on 400 x 80, fx = 3e-5,
    # 10^4 steps, against the start-up series (the run is far from steady)
    ny, G, tau, T = 80, 3e-5, 0.8, 10000
    nu = (tau - 0.5) / 3
    H = ny - 2
    sim = make(400, ny, tau, 0, G, precision=precision)
    sim.run(T)
    u = ux(sim)
    assert np.max(np.ptp(u[1:H + 1], axis=1)) < 1e-9 * np.max(u)
    y = np.arange(H) + 0.5
    ua = G / (2 * nu) * y * (H - y)
    for n in range(200):
        k = 2 * n + 1
        ua -= 4 * G * H * H / (nu * math.pi ** 3 * k ** 3) * np.sin(k * math.pi * y / H) * \
            math.exp(-k * k * math.pi ** 2 * nu * T / (H * H))
    err = np.max(np.abs(u[1:H + 1, 200] - ua)) / np.max(ua)
    print(f"startup series {precision}: max rel err {err:.3e}", flush=True)
    assert err < 0.01


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_taylor_green_decay(precision):
    from softflow_amd import softflow_core as sc  # noqa: F401
    n, tau, U0 = 64, 0.8, 0.04
    nu = (tau - 0.5) / 3
    k = 2 * math.pi / n
    sim = make(n, n, tau, 0, 0.0, periodic_y=True, precision=precision)
    f = sim.field()
    for y in range(n):
        for x in range(n):
            f.setRho(x, y, 1.0 - U0 * U0 * 0.75 * (math.cos(2 * k * x) + math.cos(2 * k * y)))
            f.setUx(x, y, -U0 * math.cos(k * x) * math.sin(k * y))
            f.setUy(x, y, U0 * math.sin(k * x) * math.cos(k * y))
            f.initializeEquilibriumAt(x, y)

    def energy():
        fl = sim.field()
        a, b = np.array(fl.velocity_x()), np.array(fl.velocity_y())
        return np.sum(a * a + b * b)
    E0 = energy()
    T = int(1.0 / (nu * 2 * k * k))
    sim.run(T)
    ratio = energy() / E0
    assert abs(ratio / math.exp(-2.0) - 1) < 0.02, ratio


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pure_diffusion(precision):
    nx, ny, D, s0, T = 256, 8, 0.05, 8.0, 5000
    sim = make(nx, ny, 0.8, 0, 0.0, periodic_y=True, precision=precision, D=D)
    x = np.arange(nx)
    c = np.exp(-((x - nx / 2) ** 2) / (2 * s0 * s0))
    ad = sim.advectionDiffusion()
    ad.setField(np.tile(c, (ny, 1)))
    sim.run(T)
    C = np.array(ad.concentration(0)).mean(axis=0)
    m = C.sum()
    mu = (x * C).sum() / m
    var = ((x - mu) ** 2 * C).sum() / m
    assert abs(var / (s0 * s0 + 2 * D * T) - 1) < 0.02
    assert abs(m - c.sum()) < (1e-9 if precision == "fp64" else 1e-5) * m


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_taylor_dispersion(precision):
    nx, ny, D = 1024, 24, 0.02
    sim = make(nx, ny, 0.6, 0, 1e-5, precision=precision, D=D)
    sim.run(30000)
    u_mean = float(ux(sim)[1:ny - 1].mean())
    H = ny - 2
    Pe = u_mean * H / D
    expect = D + Pe * Pe * D / 210.0
    ad = sim.advectionDiffusion()
    xx = np.arange(nx)
    C0 = np.zeros((ny, nx))
    C0[1:ny - 1] = np.exp(-(xx - nx / 4.0) ** 2 / (2 * 36.0))
    ad.setField(C0)

    def moments():
        Cb = np.array(ad.concentration(0))[1:ny - 1].mean(axis=0)
        th = 2 * math.pi * xx / nx
        m = Cb.sum()
        xb = math.atan2((Cb * np.sin(th)).sum() / m, (Cb * np.cos(th)).sum() / m) * nx / (2 * math.pi)
        xb = xb + nx if xb < 0 else xb
        d = xx - xb
        d = np.where(d > nx / 2, d - nx, np.where(d < -nx / 2, d + nx, d))
        return (Cb * d * d).sum() / m
    t_mix = int(H * H / (4 * D))
    sim.run(6 * t_mix)
    v1 = moments()
    sim.run(6 * t_mix)
    v2 = moments()
    meas = (v2 - v1) / (2 * 6 * t_mix)
    print(f"taylor dispersion {precision}: Pe={Pe:.2f} D_eff {meas:.5f} vs {expect:.5f}", flush=True)
    assert abs(meas - expect) / expect < 0.05


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mass_conservation(precision):
    sim = make(64, 40, 0.8, 2, 3e-5, precision=precision)
    rho0 = np.array(sim.field().density())[1:-1].sum()
    sim.run(3000)
    rho = np.array(sim.field().density())[1:-1].sum()
    assert abs(rho - rho0) < (1e-10 if precision == "fp64" else 1e-6) * rho0
