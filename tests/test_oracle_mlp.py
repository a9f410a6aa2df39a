"""Pins of the oracle's restatement of the MDF control and of the warm-start
network (neither has a counterpart in the reference).

- network forward: a hand-built network whose output is a closed form in
  erf (checks the parameter layout W1[w][NI], b1, W2[w][w], b2, W3[2][w], b3
  and the normalisation);
- network gradient: central finite differences of the oracle's own loss;
- residual-mode MDF: with an unreachable tolerance it is the reference's
  fixed-iteration MDF bit for bit; with a reachable one it stops at the
  first iteration whose residual is within tolerance, and the residual
  after k corrections falls with k (MDF converges).
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.cases import oracle_sim, oracle_state

W = 64
NI = 17
NP = W * NI + W + W * W + W + 2 * W + 2


def offsets(w=W):
    o = {}
    o["W1"] = 0
    o["b1"] = o["W1"] + w * NI
    o["W2"] = o["b1"] + w
    o["b2"] = o["W2"] + w * w
    o["W3"] = o["b2"] + w
    o["b3"] = o["W3"] + 2 * w
    return o


def ident_norm():
    return np.r_[np.zeros(NI), np.ones(NI), np.zeros(2), np.ones(2)]


def gelu(z):
    return 0.5 * z * (1 + math.erf(z / math.sqrt(2)))


def test_forward_closed_form():
    o = offsets()
    P = np.zeros(NP)
    P[o["W1"] + 3 * NI + 7] = 2.0        # hidden unit 3 reads input 7
    P[o["b1"] + 3] = -0.5
    P[o["W2"] + 10 * W + 3] = 1.5        # unit 10 of layer 2 reads unit 3
    P[o["W3"] + 1 * W + 10] = -3.0       # output 1 reads unit 10
    P[o["b3"] + 0] = 0.25
    nrm = ident_norm()
    nrm[7], nrm[NI + 7] = 1.0, 4.0       # x^7 = (x7 - 1) / 4
    nrm[2 * NI + 1], nrm[2 * NI + 2 + 1] = 0.1, 2.0   # y1 = 0.1 + 2 out1
    rng = np.random.default_rng(0)
    x = rng.normal(size=(20, NI))
    y = po.mlp_predict(P, nrm, x)
    g0 = gelu(0.0)
    for s in range(20):
        h3 = gelu(2.0 * (x[s, 7] - 1.0) / 4.0 - 0.5)
        z10 = 1.5 * h3
        assert abs(y[s, 0] - (0.25 + 0.0 * g0)) < 1e-15
        assert abs(y[s, 1] - (0.1 + 2.0 * (-3.0 * gelu(z10)))) < 1e-14


def test_gradient_finite_differences():
    rng = np.random.default_rng(1)
    P = rng.normal(scale=0.3, size=NP)
    nrm = np.r_[rng.normal(size=NI), rng.uniform(0.5, 2, NI), rng.normal(size=2), rng.uniform(0.5, 2, 2)]
    x = rng.normal(size=(7, NI))
    t = rng.normal(size=(7, 2))
    L, G = po.mlp_loss_grad(P, nrm, x, t)
    assert L > 0
    idx = np.r_[rng.choice(NP, 60, replace=False), [0, NP - 1, NP - 2, offsets()["b1"], offsets()["W3"]]]
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (po.mlp_loss_grad(Pp, nrm, x, t)[0] - po.mlp_loss_grad(Pm, nrm, x, t)[0]) / (2 * h)
        assert abs(fd - G[i]) < 1e-7 * max(1.0, abs(G[i])), (i, fd, G[i])


CASE = dict(nx=80, ny=40, tau=0.8, bfx=3e-5, coll=2,
            capsules=[dict(count=5, region=(8, 8, 72, 32), rmin=3.0, rmax=3.5, nodes=20, seed=42,
                           mem=dict(model=2, G_s=0.06, C_skalak=10.0))])


def test_residual_mode_unreachable_tolerance_is_fixed_mdf():
    a = oracle_sim(dict(CASE, mdf=5))
    b = oracle_sim(dict(CASE, mdf=1))
    b.set_mdf_control(1e-300, 4)
    a.step(60)
    b.step(60)
    sa, sb = oracle_state(a), oracle_state(b)
    for k in ("rho", "ux", "uy", "x", "y", "vx", "vy"):
        assert np.array_equal(sa[k], sb[k]), k
    assert b.mdf_stats()[0] == 4.0


def test_residual_mode_stops_within_tolerance():
    s = oracle_sim(dict(CASE, mdf=1))
    s.step(40)
    s.set_mdf_control(1e-300, 1, probe=True)
    s.step(1)
    r1 = s.mdf_stats()[1]               # residual after one correction
    s.set_mdf_control(1e-300, 6, probe=True)
    s.step(1)
    r6 = s.mdf_stats()[1]
    assert r6 < 0.5 * r1                 # corrections reduce the residual
    tol = 0.5 * (r1 + r6)
    s.set_mdf_control(tol, 12, probe=True)
    s.step(5)
    st = s.mdf_stats()
    assert st[2] <= tol                  # every step ended within tolerance
    assert 1.0 <= st[0] < 12.0
