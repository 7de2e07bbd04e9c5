"""Host logic of the learned lambda schedule (q_amd/schedule.py): BAR and
overlap against closed forms, features, and the greedy ladder with a
predictor whose answer is known."""
import numpy as np

from q_amd import schedule as S


def test_bar_recovers_gaussian_free_energy():
    # forward work ~ N(mu, s2) (kT units) gives df = mu - s2/2; the reverse
    # work sampled in B is ~ N(-(mu - s2), s2)
    rng = np.random.default_rng(0)
    mu, s2 = 3.0, 2.0
    wF = rng.normal(mu, np.sqrt(s2), 50000)
    wR = rng.normal(-(mu - s2), np.sqrt(s2), 50000)
    assert abs(S.bar_df(wF, wR) - (mu - s2 / 2)) < 0.03


def test_overlap_limits():
    rng = np.random.default_rng(1)
    w = rng.normal(0.0, 0.01, 2000)
    assert S.bar_overlap(w, -w) > 0.99          # identical states
    far = S.bar_overlap(rng.normal(40, 1, 2000), rng.normal(40, 1, 2000))
    assert far < 1e-3
    # overlap falls as the states move apart
    o = [S.bar_overlap(rng.normal(m, np.sqrt(2 * m), 4000) if m > 0 else np.zeros(4000),
                       rng.normal(m, np.sqrt(2 * m), 4000) if m > 0 else np.zeros(4000)) for m in (0.5, 2, 8)]
    assert o[0] > o[1] > o[2]


def test_features_layout():
    dE = np.linspace(-5, 5, 200) + 40.0
    f = S.features(dE, 0.3, 10.0)
    assert f.shape == (S.N_IN,) and f.dtype == np.float32
    assert abs(f[:S.NBINS].sum() - 1.0) < 1e-6
    assert np.allclose(f[S.NBINS:], [0.3, 1.0])
    # a constant shift of the energy gap does not change the features
    assert np.array_equal(f, S.features(dE - 40.0, 0.3, 10.0))


def test_gaussian_overlap_matches_sampled_bar():
    # the closed form against bar_overlap on Gaussian works of known spread
    rng = np.random.default_rng(4)
    for sw in (0.3, 1.0, 2.5):
        df = 1.7
        wF = rng.normal(df + sw ** 2 / 2, sw, 200000)
        wR = rng.normal(-df + sw ** 2 / 2, sw, 200000)
        assert abs(S.gaussian_overlap(sw)[0] - S.bar_overlap(wF, wR)) < 0.01
    assert abs(S.gaussian_overlap(0.0)[0] - 1.0) < 1e-12
    o = S.gaussian_overlap(np.linspace(0, 10, 50))
    assert np.all(np.diff(o) < 0)
    sw = np.array([0.05, 0.7, 3.0, 8.0])
    assert np.allclose(S.work_spread(S.gaussian_overlap(sw)), sw, rtol=1e-6)


def test_equal_length_ladder():
    lams = S.PILOT_LAMBDAS
    # equal overlaps: the ladder is uniform
    s, implied = S.ladder_from_overlaps(lams, np.full(10, 0.8), 16, 22)
    assert np.allclose(s, np.linspace(0, 1, len(s)))
    # ten times the work spread on the last interval: windows crowd there,
    # and every step carries the same spread
    ovl = S.gaussian_overlap(np.array([0.5] * 9 + [5.0]))
    s, implied = S.ladder_from_overlaps(lams, ovl, 16, 22)
    assert 16 <= len(s) <= 22 and s[0] == 0.0 and s[-1] == 1.0 and np.all(np.diff(s) > 0)
    rate = np.array([5.0] * 9 + [50.0])
    spread = [np.sum(rate * np.clip(np.minimum(b, lams[1:]) - np.maximum(a, lams[:-1]), 0, None))
              for a, b in zip(s[:-1], s[1:])]
    assert np.allclose(spread, spread[0], rtol=1e-9)
    assert np.mean(s > 0.9) >= 0.45     # the last interval holds 5 of the 9.5 units of length
    assert abs(implied - S.gaussian_overlap(spread[0])[0]) < 1e-9
    # the window count follows the target overlap within the bounds
    assert len(S.ladder_from_overlaps(lams, np.full(10, 0.99), 16, 22)[0]) == 16
    assert len(S.ladder_from_overlaps(lams, np.full(10, 0.05), 16, 22)[0]) == 22
    assert len(S.ladder_from_overlaps(lams, np.full(10, 0.05), n=19)[0]) == 19


def test_learned_ladder_follows_the_predictor():
    # a predictor whose overlap falls with the spread of the centred work
    # histogram: the ladder crowds where the pilot energy gaps are wide
    lams = list(S.PILOT_LAMBDAS)
    rng = np.random.default_rng(2)
    dEs = [rng.normal(0, 1 + 9 * lam ** 4, 400) for lam in lams]
    centres = np.linspace(S.U_LO, S.U_HI, S.NBINS + 1)[:-1] + (S.U_HI - S.U_LO) / S.NBINS / 2

    def predict(F):
        h = F[:, :S.NBINS]
        sd = np.sqrt(np.maximum((h * centres ** 2).sum(1) - (h * centres).sum(1) ** 2, 0))
        return S.gaussian_overlap(sd)

    s = S.propose_n(predict, lams, dEs, 10.0, 16, 22)
    assert 16 <= len(s) <= 22
    steps = np.diff(s)
    assert steps[:3].mean() > 2 * steps[-3:].mean()


def test_fallback_to_uniform_when_the_network_disagrees_with_the_pilots():
    # pilots whose measured neighbour overlaps the predictor reproduces keep
    # the learned spacing; a predictor far from them (overlap 0.3 where the
    # pilots measure far more) makes the schedule fall back to the uniform
    # ladder with as many windows, and the decision is recorded
    lams = list(S.PILOT_LAMBDAS)
    rng = np.random.default_rng(2)
    dEs = [rng.normal(0, 1 + 9 * lam ** 4, 400) for lam in lams]
    meas = S.measured_overlaps(lams, dEs)

    def faithful(F):
        # per window: the mean of the measured overlaps of its intervals
        m = np.concatenate([[meas[0]], meas])
        m2 = np.concatenate([meas, [meas[-1]]])
        return 0.5 * (m + m2)[:len(F)]

    s, dec = S.learned_schedule(faithful, lams, dEs, 10.0, 16, 22)
    assert not dec["fallback"] and dec["disagreement"] < S.FALLBACK_DISAGREEMENT
    assert not np.allclose(np.diff(s), np.diff(s)[0])
    s2, dec2 = S.learned_schedule(lambda F: np.full(len(F), 0.3), lams, dEs, 10.0, 16, 22)
    assert dec2["fallback"] and dec2["disagreement"] > S.FALLBACK_DISAGREEMENT
    assert np.allclose(s2, np.linspace(0.0, 1.0, len(s2)))


def test_bar_variance_matches_replica_spread():
    # Bennett's variance against the spread of BAR over independent replicas
    rng = np.random.default_rng(3)
    mu, s2, n = 2.0, 3.0, 300
    est, var = [], []
    for _ in range(400):
        wF = rng.normal(mu, np.sqrt(s2), n)
        wR = rng.normal(-(mu - s2), np.sqrt(s2), n)
        est.append(S.bar_df(wF, wR))
        var.append(S.bar_var(wF, wR))
    ratio = np.var(est) / np.mean(var)
    assert 0.8 < ratio < 1.25, ratio
    # identical states: no uncertainty
    w = np.zeros(100)
    assert S.bar_var(w, w) < 1e-12


def test_simulated_fraction_bound():
    # pilots included: the largest ladder within half the 51-window time
    assert S.simulated_fraction(22, 4000, 1000) <= 0.5 < S.simulated_fraction(23, 4000, 1000)


def test_reperturb_redraws_only_the_fep_states():
    from q_amd import synth
    base = S.build_system(S.Perturbation(1001, "ligwat", 6))
    a = synth.reperturb(base, 11, 3)
    b = synth.reperturb(base, 11, 3)
    c = synth.reperturb(base, 12, 9)
    assert a.qcharges == b.qcharges and a.qtypes == b.qtypes and a.qbonds == b.qbonds
    assert np.array_equal(np.array(a.x), np.array(base.x)) and a.crg == base.crg and a.iac == base.iac and a.bonds == base.bonds
    assert a.qatoms == base.qatoms
    for s, n in ((a, 3), (c, 9)):
        changed = [k for k, (t, q) in enumerate(zip(s.qtypes, s.qcharges)) if t[0] != t[1] or q[0] != q[1]]
        assert 1 <= len(changed) <= n
        assert sum(1 for al in s.qalpha if al[0] > 0) == n
        # both states keep the ligand's net charge
        assert abs(sum(q for q, _ in s.qcharges) - sum(q for _, q in s.qcharges)) < 1e-9
    # state 1 is the topology
    assert [q for q, _ in a.qcharges] == base.crg[:len(base.qatoms)]
