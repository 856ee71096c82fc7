"""CPU pins of the shared random stream and of the oracle's stochastic cell
rescaling.

Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) is pinned by the
known-answer vectors Random123 publishes for it (counter, key -> output), and
the keyed draws every stream uses by their counter layout.  The barostat
restatement (md/integrators.py:657-811, Bernetti and Bussi 2020) is pinned by
closed forms that a dropped term, a wrong sign or a misplaced factor breaks:
beta = 0 leaves everything unchanged, the deterministic part is
exp(dt beta (P - P0) / (3 tau)) with P = 2K/3V + tr W/3V, the noise-only part
has Var(ln mu) = 2 kT beta dt / (9 tau V) and a normal law, the clip bounds,
and the reference's own cases (test/md/test_integrators.py:373-426):
volume = det of the scaled cell, positions scaled with the box, expansion
under high pressure.
"""

import math

import numpy as np
import pytest
from scipy import stats

import oracle

KAT = [
    # (counter, key, output): Random123 kat_vectors, philox4x32 10 rounds
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("case", range(len(KAT)))
def test_philox_known_answers(case):
    ctr, key, out = KAT[case]
    np.testing.assert_array_equal(oracle.philox_raw(ctr, key), np.array(out, np.uint32))


def test_keyed_draw_counter_layout():
    seed, a, t, tag = 0x0123456789ABCDEF, 77, (5 << 32) + 9, 0x47434D41
    want = oracle.philox_raw([a, t & 0xFFFFFFFF, t >> 32, tag],
                             [seed & 0xFFFFFFFF, seed >> 32])
    np.testing.assert_array_equal(oracle.philox(seed, a, t, tag), want)


# ------------------------------------------------ stochastic cell rescaling --

def _system(n=8, L=20.0, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-L / 2, L / 2, (n, 3))
    mom = rng.normal(0, 1.0, (n, 3))
    mass = np.full(n, 2.0)
    tril = np.array([[L, 0.0, L, 0.0, 0.0, L]])
    return pos, mom, mass, tril


def _scr(pos, mom, mass, tril, kT, dt, baro, W, t=3, seed=11):
    pos, mom, tril = pos.copy(), mom.copy(), tril.copy()
    mu = oracle.scr_scale(seed, t, mom, pos, tril, mass, [0, len(mass)], [kT], [dt], baro,
                          np.asarray(W, float).reshape(1, 3, 3))
    return mu[0], pos, tril


def _pressure(mom, mass, W, V):
    K = 0.5 * np.sum(mom * mom / mass[:, None])
    return 2 * K / (3 * V) + np.trace(W) / (3 * V)


def test_scr_zero_compressibility_is_identity():
    pos, mom, mass, tril = _system()
    W = np.diag([3.0, -1.0, 2.0])
    mu, p2, t2 = _scr(pos, mom, mass, tril, 0.02, 0.5, [[1e-4, 100.0, 0.0, 0.9]], W)
    assert mu == 1.0
    np.testing.assert_array_equal(p2, pos)
    np.testing.assert_array_equal(t2, tril)


def test_scr_deterministic_part_closed_form():
    pos, mom, mass, tril = _system(seed=1)
    W = np.array([[5.0, 0.3, 0.0], [0.3, -2.0, 0.1], [0.0, 0.1, 4.0]])
    V = tril[0, 0] * tril[0, 2] * tril[0, 5]
    P0, tau, beta, dt = 1e-4, 50.0, 3.0, 0.2
    mu, p2, t2 = _scr(pos, mom, mass, tril, 0.0, dt, [[P0, tau, beta, 0.5]], W)
    P = _pressure(mom, mass, W, V)
    assert mu == pytest.approx(math.exp(dt / tau * beta * (P - P0) / 3.0), rel=1e-14)
    # the cell and the positions are scaled by mu; the volume by mu^3
    np.testing.assert_allclose(t2, tril * mu, rtol=1e-15)
    np.testing.assert_allclose(p2, pos * mu, rtol=1e-15)
    assert t2[0, 0] * t2[0, 2] * t2[0, 5] == pytest.approx(V * mu**3, rel=1e-14)


def test_scr_expands_under_high_pressure_and_clips():
    """test_integrators.py:404-426: a box under pressure above the target
    grows (mu > 1), below it shrinks; a huge excess is clipped to 1/mu_min."""
    pos, mom, mass, tril = _system(seed=2)
    V = tril[0, 0] * tril[0, 2] * tril[0, 5]
    up, _, _ = _scr(pos, mom, mass, tril, 0.0, 0.5, [[0.0, 10.0, 1.0, 0.5]], np.eye(3) * 30.0)
    down, _, _ = _scr(pos, mom, mass, tril, 0.0, 0.5, [[10.0, 10.0, 1.0, 0.5]], np.zeros((3, 3)))
    assert up > 1.0 > down
    big, _, _ = _scr(pos, mom, mass, tril, 0.0, 1.0, [[0.0, 1.0, 1e9, 0.9]], np.eye(3) * 1e6)
    assert big == pytest.approx(1 / 0.9, rel=1e-15)
    small, _, _ = _scr(pos, mom, mass, tril, 0.0, 1.0, [[1e6, 1.0, 1e9, 0.9]], np.zeros((3, 3)))
    assert small == pytest.approx(0.9, rel=1e-15)


def test_scr_noise_law():
    """With P = P0 only the noise remains: 3 ln mu / sqrt(2 kT beta dt / (tau V))
    over many steps is standard normal (Kolmogorov-Smirnov), one draw per step
    shared by every system of the batch."""
    pos, mom, mass, tril = _system(seed=3)
    mom[:] = 0.0
    V = tril[0, 0] * tril[0, 2] * tril[0, 5]
    P0 = 2e-3
    W = np.eye(3) * P0 * V  # tr W / 3V = P0
    kT, dt, tau, beta = 0.025, 0.3, 40.0, 2.0
    scale = math.sqrt(2 * kT * beta * dt / (tau * V))
    z = []
    for t in range(3000):
        mu, _, _ = _scr(pos, mom, mass, tril, kT, dt, [[P0, tau, beta, 0.5]], W, t=t, seed=5)
        z.append(3.0 * math.log(mu) / scale)
    z = np.array(z)
    assert stats.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 4 / math.sqrt(len(z))
    assert np.var(z) == pytest.approx(1.0, abs=0.1)
    # two systems at one step draw the same R (ratio of their log factors is
    # the ratio of their noise scales)
    pos2 = np.concatenate([pos, pos])
    mom2 = np.zeros_like(pos2)
    mass2 = np.concatenate([mass, mass])
    tril2 = np.array([[20.0, 0, 20.0, 0, 0, 20.0], [30.0, 0, 30.0, 0, 0, 30.0]])
    V2 = tril2[:, 0] * tril2[:, 2] * tril2[:, 5]
    W2 = np.stack([np.eye(3) * P0 * v for v in V2])
    mu2 = oracle.scr_scale(9, 4, mom2, pos2.copy(), tril2.copy(), mass2, [0, 8, 16], [kT, kT],
                           [dt, dt], [[P0, tau, beta, 0.5]] * 2, W2)
    r = np.log(mu2) / np.sqrt(2 * kT * beta * dt / (tau * V2))
    assert r[0] == pytest.approx(r[1], rel=1e-12)
