"""CPU pins of the GCMC pieces: Peng-Robinson fugacity (host code), the
oracle's acceptance rules (exact log ratios, ideal-gas limit) and its random
stream."""

import math

import numpy as np
import pytest

import oracle
from kups_amd.core.constants import BAR, BOLTZMANN_CONSTANT, PASCAL
from kups_amd.mcmc.fugacity import peng_robinson_log_fugacity

PC, TC, OMEGA = 7.84e6 * PASCAL, 303.75, 0.22394


def _pr_pure(p, T):
    """Textbook pure-component PR (independent of the module's mixture code)."""
    m = 0.37464 + 1.54226 * OMEGA - 0.26992 * OMEGA**2
    alpha = (1 + m * (1 - math.sqrt(T / TC))) ** 2
    A = 0.45724 * alpha * (p / PC) / (T / TC) ** 2
    B = 0.0778 * (p / PC) / (T / TC)
    roots = np.roots([1, B - 1, A - 3 * B**2 - 2 * B, -(A * B - B**2 - B**3)])
    Z = roots[np.abs(roots.imag) < 1e-12].real.max()
    s2 = math.sqrt(2)
    ln_phi = Z - 1 - math.log(Z - B) - A / (2 * s2 * B) * math.log((Z + (1 + s2) * B) / (Z + (1 - s2) * B))
    return ln_phi + math.log(p), Z


def test_pr_low_pressure_ideal_limit():
    # test/mcmc/test_fugacity.py:57-69
    p = 1e-2 * BAR
    lnf, Z = peng_robinson_log_fugacity(p, 300.0, PC, TC, OMEGA)
    assert Z == pytest.approx(1.0, abs=1e-4)
    assert lnf[0] == pytest.approx(math.log(p), abs=1e-4)


@pytest.mark.parametrize("p_bar", [1.0, 10.0, 50.0])
def test_pr_against_numpy_roots(p_bar):
    # test/mcmc/test_fugacity.py:72-97
    p = p_bar * BAR
    lnf, Z = peng_robinson_log_fugacity(p, 300.0, PC, TC, OMEGA)
    lnf_ref, Z_ref = _pr_pure(p, 300.0)
    assert Z == pytest.approx(Z_ref, rel=1e-10)
    assert lnf[0] == pytest.approx(lnf_ref, rel=1e-10, abs=1e-10)


def test_pr_binary_ideal_limit():
    p = 1e-2 * BAR
    y = np.array([0.25, 0.75])
    lnf, Z = peng_robinson_log_fugacity(p, 300.0, [PC, PC], [TC, TC], [OMEGA, OMEGA], y)
    np.testing.assert_allclose(lnf, np.log(y * p), atol=2e-4)


def _empty_box_state(aV: float, L: float = 25.0, gmax: int = 256):
    """One chain, no host, a 3-site guest without charges or LJ: an ideal gas."""
    kT = 298.0 * BOLTZMANN_CONSTANT
    V = L**3
    return dict(
        host_off=np.array([0, 0]), host_pos=np.zeros((1, 3)), host_q=np.zeros(1),
        host_spec=np.zeros(1), tril=np.array([[L, 0, L, 0, 0, L]]),
        motif_pos=np.array([[0, 0, 0], [-1.16, 0, 0], [1.16, 0, 0]]), motif_q=np.zeros(3),
        motif_spec=np.zeros(3), sigma=np.array([[3.0]]), epsilon=np.array([[0.0]]),
        alpha=np.array([0.3]), kmax=np.array([-1.0]), tail=np.zeros((1, 1, 1)),
        mol_const=np.zeros(1), kT=np.array([kT]), log_activity=np.array([math.log(aV / V)]),
        guest_pos=np.zeros((1, gmax, 3, 3)), n_guest=np.zeros(1, np.int32),
        counts=np.zeros((1, 1), np.int32), sk=np.zeros((1, 1, 2)), energy=np.zeros(1),
        width=np.full((1, 4), 0.1, np.float32), hist=np.zeros((1, 4, 4), np.uint32),
        hist_idx=np.zeros((1, 4), np.int32), stats=np.zeros((1, 8), np.int64),
        ins_acc=np.zeros(1, np.int64), del_acc=np.zeros(1, np.int64),
        rc_lj=12.0, rc_cl=12.0, grid=(0, 0, 0), w_move=(1.0, 1.0, 1.0, 3.0),
        upper=(12.5, 1e30, 1e30, 1e30))


def test_oracle_exchange_log_ratio_exact():
    # probability.py:116-173: insertion into an empty box: ln(aV) - ln(1);
    # deletion of the only molecule: -ln(aV) + ln(1)
    st = _empty_box_state(aV=5.0)
    tr = oracle.gcmc_run(st, 400, 0, seed=3, n_trace=1)[0]
    ins = tr[(tr[:, 0] == 6)]
    for move, dE, logp, acc in tr:
        assert dE == 0.0
    # the first insertion attempt from N = 0
    n = 0
    for move, dE, logp, acc in tr:
        if move == 6:
            assert logp == pytest.approx(math.log(5.0) - math.log(n + 1), rel=1e-12)
        elif move == 7 and n > 0:
            assert logp == pytest.approx(-math.log(5.0) + math.log(n), rel=1e-12)
        if acc and move == 6:
            n += 1
        elif acc and move == 7 and n > 0:
            n -= 1
    assert n == st["n_guest"][0]
    assert len(ins) > 50


def test_oracle_ideal_gas_loading():
    """Empty host, no interactions: <N> = a V = beta f V (Poisson)."""
    st = _empty_box_state(aV=6.0)
    st["w_move"] = (0.0, 0.0, 0.0, 1.0)
    ns = []
    step = 0
    for _ in range(200):
        oracle.gcmc_run(st, 100, step, seed=11)
        step += 100
        ns.append(int(st["n_guest"][0]))
    ns = np.array(ns[20:])
    assert ns.mean() == pytest.approx(6.0, abs=0.6)
    assert ns.var() == pytest.approx(6.0, rel=0.35)


def test_batch_shared_move_type():
    # core/propagator.py:263-268: one move type for every chain per step
    words = [oracle.philox(5, 0xFFFFFFFF, t, 0x42415443) for t in range(3)]
    assert len({tuple(w) for w in words}) == 3
    # counter-based: the same (seed, counter) always gives the same words
    np.testing.assert_array_equal(oracle.philox(5, 7, 9, 1), oracle.philox(5, 7, 9, 1))


def test_oracle_mlp_gradient_finite_differences():
    """The oracle's MLP gradient (used to check the GPU backward pass) against
    central finite differences of its loss."""
    from kups_amd.mcmc.proposal import init_params

    p = init_params(1).astype(np.float64)
    rng = np.random.default_rng(0)
    X = rng.normal(size=(7, 32))
    y = rng.normal(size=7)
    _, g = oracle.mlp_loss_and_grad(p, X, y)
    h = 1e-6
    for idx in [0, 4100, 9224, 20000, 30000, 37300, 37376]:
        q = p.copy(); q[idx] += h
        lp = oracle.mlp_loss_and_grad(q, X, y)[0]
        q[idx] -= 2 * h
        lm = oracle.mlp_loss_and_grad(q, X, y)[0]
        assert (lp - lm) / (2 * h) == pytest.approx(g[idx], rel=1e-5, abs=1e-9)


def test_csvr_chi_square_sampler_matches_scipy():
    """The Marsaglia-Tsang Gamma draws of the CSVR step: 2 Gamma(k/2) is
    chi^2(k) (Kolmogorov-Smirnov against scipy), for large and small k."""
    from scipy import stats
    for k in (1.0, 3.0, 40.0, 12285.0):
        x = np.array([2.0 * oracle.gamma_mt(0.5 * k, 11, s, t)
                      for s in range(40) for t in range(100)])
        assert stats.kstest(x, stats.chi2(k).cdf).pvalue > 1e-3, k


def test_csvr_keeps_the_canonical_kinetic_energy_distribution():
    """Force-free CSVR steps (the thermostat alone) sample K from the canonical
    Gamma(N_dof / 2, kT) law whatever tau is (the exactness property of Bussi,
    Donadio and Parrinello): mean N_dof kT / 2, variance N_dof (kT)^2 / 2, from
    a start far from equilibrium."""
    rng = np.random.default_rng(3)
    n, kT = 20, 0.025
    mass = np.full(n, 39.948)
    mom = rng.normal(0, 5.0, (n, 3))  # ~40x too hot
    mom -= mom.mean(axis=0)
    ndof = 3 * n - 3
    ks = []
    for t in range(6000):
        oracle.csvr_scale(5, t, mom, mass, [0, n], [kT], [1.0], [7.0])
        if t >= 1000:
            ks.append(0.5 * np.sum(mom ** 2 / mass[:, None]))
    ks = np.array(ks)
    # effective sample size from the coupling: correlation time tau / dt = 7
    assert abs(ks.mean() / (0.5 * ndof * kT) - 1) < 0.03
    assert abs(ks.var() / (0.5 * ndof * kT ** 2) - 1) < 0.15


def test_csvr_shares_r1_and_needs_dof():
    """One R1 for the whole batch; a one-atom system (N_dof = 0) is left as it is."""
    mom = np.ones((3, 3))
    a = oracle.csvr_scale(1, 0, mom, np.ones(3), [0, 1, 3], [0.01, 0.01], [1.0, 1.0], [2.0, 2.0])
    assert a[0] == 1.0 and a[1] != 1.0
