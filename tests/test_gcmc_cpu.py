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


def _empty_box_state(aV, L: float = 25.0, gmax: int = 256):
    """One chain, no host, guests without charges or LJ: an ideal gas.  One
    3-site species, or with ``aV`` a sequence, a mixture whose species t has
    3 - t sites (1..3) and activity aV[t] / V."""
    kT = 298.0 * BOLTZMANN_CONSTANT
    V = L**3
    aV = np.atleast_1d(np.asarray(aV, float))
    nmot = len(aV)
    mpos = np.zeros((nmot, 3, 3))
    for t in range(nmot):
        mpos[t, :3 - t, 0] = np.linspace(-1.16, 1.16, 3 - t) if t < 2 else 0.0
    return dict(
        host_off=np.array([0, 0]), host_pos=np.zeros((1, 3)), host_q=np.zeros(1),
        host_spec=np.zeros(1), tril=np.array([[L, 0, L, 0, 0, L]]),
        motif_pos=mpos, motif_q=np.zeros((nmot, 3)), motif_spec=np.zeros((nmot, 3)),
        motif_nsite=np.array([3 - t for t in range(nmot)]),
        sigma=np.array([[3.0]]), epsilon=np.array([[0.0]]),
        alpha=np.array([0.3]), kmax=np.array([-1.0]), tail=np.zeros((1, 1, 1)),
        mol_const=np.zeros((1, nmot)), kT=np.array([kT]), log_activity=np.log(aV / V)[None],
        guest_pos=np.zeros((1, gmax, 3, 3)), n_guest=np.zeros(1, np.int32),
        guest_motif=np.zeros((1, gmax), np.int32),
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


def test_oracle_mixture_exchange_log_ratios_exact():
    """Two species (activities 5.5 / V and pi / V, 3 and 2 sites): every
    exchange trial carries the fugacity term of ONE species with that
    species' own count, ln(a_t V) - ln(N_t + 1) or -ln(a_t V) + ln N_t
    (probability.py:116-173, the species drawn uniformly, moves.py:600-602,
    711-712); a deletion of a species with no molecule present changes
    nothing.  The ratio a_0 / a_1 is irrational, so the trace's log p names
    the species; the tracked counts must end where the kernel's do."""
    aV = (5.5, math.pi)
    st = _empty_box_state(aV=aV)
    st["w_move"] = (0.0, 0.0, 0.0, 1.0)
    tr = oracle.gcmc_run(st, 3000, 0, seed=9, n_trace=1)[0]
    n = [0, 0]
    picked = [0, 0]
    noop = 0
    for code, dE, logp, acc in tr:
        assert dE == 0.0
        if code == 6:
            cand = [math.log(aV[t]) - math.log(n[t] + 1) for t in (0, 1)]
        else:
            cand = [-math.log(aV[t]) + math.log(n[t]) if n[t] else 0.0 for t in (0, 1)]
        hit = [t for t in (0, 1) if abs(logp - cand[t]) < 1e-12]
        assert hit, (code, logp, cand, n)
        t = hit[0]
        if code == 7 and n[t] == 0:
            noop += 1
            assert logp == 0.0
            continue
        picked[t] += 1
        if acc:
            n[t] += 1 if code == 6 else -1
    gm = st["guest_motif"][0, :st["n_guest"][0]]
    assert n == [int((gm == 0).sum()), int((gm == 1).sum())]
    assert min(picked) > 600 and noop > 0


def test_oracle_mixture_ideal_gas_loading():
    """Ideal mixture in an empty box: each species is an independent Poisson
    gas, <N_t> = a_t V, whatever its site count."""
    st = _empty_box_state(aV=(4.0, 2.0, 1.0))
    st["w_move"] = (0.0, 0.0, 0.0, 1.0)
    ns = []
    step = 0
    for _ in range(300):
        oracle.gcmc_run(st, 100, step, seed=13)
        step += 100
        gm = st["guest_motif"][0, :st["n_guest"][0]]
        ns.append(np.bincount(gm, minlength=3))
    ns = np.array(ns[30:], float)
    np.testing.assert_allclose(ns.mean(0), [4.0, 2.0, 1.0], rtol=0.15, atol=0.2)
    np.testing.assert_allclose(ns.var(0), [4.0, 2.0, 1.0], rtol=0.4)


def test_oracle_motif_padding_never_read():
    """Interacting guests (LJ between sites): a run whose motif table and
    guest buffer are padded to 4 sites, the padding site holding a charge and
    a position that would change every energy if read, gives the same trace,
    bit for bit, as the unpadded 3-site tables."""
    a = _empty_box_state(aV=40.0)
    a["epsilon"] = np.array([[0.005]])
    b = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in a.items()}
    b["motif_pos"] = np.concatenate([a["motif_pos"], [[[0.3, 0.2, 0.1]]]], 1)
    b["motif_q"] = np.concatenate([a["motif_q"], [[5.0]]], 1)
    b["motif_spec"] = np.concatenate([a["motif_spec"], [[0]]], 1)
    b["guest_pos"] = np.zeros((1, a["guest_pos"].shape[1], 4, 3))
    tra = oracle.gcmc_run(a, 800, 0, seed=4, n_trace=1)
    trb = oracle.gcmc_run(b, 800, 0, seed=4, n_trace=1)
    np.testing.assert_array_equal(tra, trb)
    assert a["n_guest"][0] > 5 and np.any(tra[0, :, 1] != 0)


def test_mixture_species_table():
    """Each LJ label appears once; every site of every motif maps to its own
    label's row (a site table pointing past the LJ table would read
    another species' parameters)."""
    from kups_amd import workloads
    from kups_amd.mcmc.gcmc import mixture_species

    motifs, lj, gs = mixture_species(("CO2", "N2", "CH4"))
    assert len(set(lj.labels)) == len(lj.labels) == 10
    for m_, sp in zip(motifs, gs):
        assert [lj.labels[i] for i in sp] == list(m_.species)
    # single CO2: the table the earlier single-species path built
    _, lj1, gs1 = mixture_species(("CO2",))
    assert lj1.labels == tuple(workloads.HOST_TYPES) + ("C_co2", "O_co2") and gs1 == [(5, 6, 6)]


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
