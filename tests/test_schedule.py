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
    dE = np.linspace(-5, 5, 200)
    f = S.features(dE, 0.3, 0.1, 10.0)
    assert f.shape == (S.N_IN,) and f.dtype == np.float32
    assert abs(f[:S.NBINS].sum() - 1.0) < 1e-6
    assert np.allclose(f[S.NBINS:], [0.3, 0.1, 1.0])


def test_greedy_ladder_follows_the_predictor():
    # predicted overlap exp(-k |delta| sd(dE)), with the energy gap spread
    # ten times wider at the lambda = 1 end: steps must shrink there
    lams = list(S.PILOT_LAMBDAS)
    rng = np.random.default_rng(2)
    dEs = [rng.normal(0, 1 + 9 * lam ** 4, 400) for lam in lams]
    spread = {round(l, 6): d.std() for l, d in zip(lams, dEs)}

    def predict(F):
        out = []
        for f in F:
            lam, delta = float(f[S.NBINS]), float(f[S.NBINS + 1])
            k = int(np.argmin(np.abs(np.array(lams) - lam)))
            out.append(np.exp(-3.0 * abs(delta) * spread[round(lams[k], 6)]))
        return np.array(out)

    s = S.propose_n(predict, lams, dEs, 10.0, 16, 24)
    assert 16 <= len(s) <= 24
    assert s[0] == 0.0 and s[-1] == 1.0 and np.all(np.diff(s) > 0)
    steps = np.diff(s)
    assert steps[:3].mean() > 2 * steps[-3:].mean()


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
    assert S.simulated_fraction(22, 2000, 500) <= 0.5 < S.simulated_fraction(23, 2000, 500)


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
