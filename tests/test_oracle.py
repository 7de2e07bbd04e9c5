"""The CPU oracle (oracle/qdyn_np.py) pinned three ways, on the CPU:

1. against the golden vectors of tests/golden/ (the reference Q6 routines,
   step-0 energies, gradient and pair lists of five seeded windows, one of
   them under the arithmetic LJ combination rule, one with a 3-atom solvent);
2. against closed forms: the LRF Taylor expansion against direct Coulomb
   sums, a single water pair against Coulomb + Lennard-Jones by hand;
3. its gradient against central finite differences of its own energy
   (catches a force term whose sign, index or factor disagrees with its
   energy term).
"""
import os

import numpy as np
import pytest

from oracle import qdyn_np as O
from tests.golden.make_golden import E_NAMES, canonical_lists, file_sha, input_files, sha

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# seeded synthetic windows (the protein leg under the arithmetic LJ rule,
# the ligand in a 3-atom water with LJ on the hydrogens, the ligand with the
# further FEP sections -- el_scale, soft pairs, excluded pairs, a Morse
# q-bond, FEP SHAKE constraints --, with library charges for q-q pairs under
# either rule, with sequence, position, distance, angle and wall restraints,
# and with a q_atom cut-off that selects the Q-atom lists) and the
# reference's own test system
CASES = [("water", 0), ("ligwat", 0), ("protein", 2), ("kat", 0), ("protarith", 2), ("ligtip3", 0), ("ligfepx", 0),
         ("liglibq", 0), ("ligarithlibq", 0), ("ligrestr", 0), ("ligqcut", 0), ("ligshake", 0)]


def load_gold(cfg, seed):
    return dict(np.load(os.path.join(GOLD, f"{cfg}_s{seed}.npz")))


@pytest.fixture(scope="module")
def oracle_eval(golden_inp):
    cache = {}

    def get(cfg, seed):
        if (cfg, seed) not in cache:
            g = load_gold(cfg, seed)
            inp = golden_inp(cfg, seed)
            files = input_files(cfg, seed, os.path.dirname(inp))
            assert file_sha(files) == str(g["input_sha"]), "input files changed: regenerate golden"
            E, EQ, dgrad, L = O.evaluate(inp, g["X0"], float(g["temp0"][1]))
            cache[(cfg, seed)] = (g, E, EQ, dgrad, L, inp)
        return cache[(cfg, seed)]

    return get


@pytest.mark.parametrize("cfg,seed", CASES)
def test_oracle_energies_vs_golden(oracle_eval, cfg, seed):
    g, E, EQ, _, _, _ = oracle_eval(cfg, seed)
    for k, name in enumerate(E_NAMES):
        ref = g["E0"][k]
        assert abs(E[name] - ref) <= 1e-9 * max(abs(ref), 1.0), name
    if g["EQ0"].size:
        rel = np.abs(EQ - g["EQ0"]) / np.maximum(np.abs(g["EQ0"]), 1.0)
        assert rel.max() < 1e-9


@pytest.mark.parametrize("cfg,seed", CASES)
def test_oracle_gradient_vs_golden(oracle_eval, cfg, seed):
    g, _, _, dgrad, _, _ = oracle_eval(cfg, seed)
    ref = g["D0"]
    assert np.abs(dgrad - ref).max() < 1e-6 * max(1.0, np.sqrt((ref ** 2).mean()))


@pytest.mark.parametrize("cfg,seed", CASES)
def test_oracle_pairlists_vs_golden(oracle_eval, cfg, seed):
    g, _, _, _, L, _ = oracle_eval(cfg, seed)
    pp = [(i + 1, j + 1, 3 if (i, j) in L["pp14"] else 1) for i, j in L["pp"]]
    pw = [(i + 1, j + 1, 1) for i, j in L["pw"]]
    ww = [(i + 1, j + 1) for i, j in L["ww"]]
    qp = [(q + 1, j + 1, c) for q, j, c in L["qp"]]
    qw = [o + 1 for o in L["qw"]]
    C = canonical_lists(pp, pw, ww, qp, qw)
    for k, a in C.items():
        assert len(a) == int(g[f"n_{k}"]), k
        assert sha(a) == str(g[f"sha_{k}"]), k


def test_lrf_expansion_matches_coulomb():
    """Potential to 2nd and field to 3rd order of the expansion against the
    direct sums for charges 12-16 A away evaluated 1 A off the centre."""
    rng = np.random.default_rng(5)
    n = 40
    u = rng.normal(size=(n, 3))
    src = u / np.linalg.norm(u, axis=1)[:, None] * rng.uniform(12.0, 16.0, n)[:, None]
    q = rng.normal(size=n)
    P = O.lrf_expand(src, q)
    for _ in range(20):
        x = rng.normal(size=3)
        x *= 1.0 / np.linalg.norm(x)
        v, fld = O.lrf_eval(P, -x)             # r = centre - x with centre at 0
        dr = x - src
        rr = np.linalg.norm(dr, axis=1)
        v_ref = (q / rr).sum()
        fld_ref = (q[:, None] * dr / rr[:, None] ** 3).sum(0)   # E = -grad V
        scale = np.abs(q).sum() / 12.0
        assert abs(v - v_ref) < 2e-3 * scale        # 3rd-order remainder (1/12)^3
        assert np.abs(fld - fld_ref).max() < 2e-4 * scale / 12.0
    # the expansion's own order: halving the offset shrinks the error ~8x
    errs = []
    for h in (1.0, 0.5):
        x = np.array([0.6, -0.3, 0.74]) * h
        v, _ = O.lrf_eval(P, -x)
        errs.append(abs(v - (q / np.linalg.norm(x - src, axis=1)).sum()))
    assert errs[1] < errs[0] / 5.0


def test_single_water_pair_by_hand(systems):
    """Two SPC waters alone: ww energies equal Coulomb + LJ written out from
    the topology charges and O-O parameters."""
    inp = systems("water", seed=0, steps=10)
    w = O.Window(inp)
    t = w.t
    ns = w.ns
    # keep the first two waters, put the rest far outside every cutoff
    x = w.x.copy()
    far = np.repeat(np.arange(1, w.nwat - 1), 3)
    x[ns + 6:, 0] += 1.0e3 * far
    x[ns + 3: ns + 6] = x[ns: ns + 3] + np.array([2.9, 0.4, -0.3])
    w.x = x
    w.pair_lists()
    assert w.lists["ww"] == {(ns, ns + 3)}
    E, _, _ = w.pot_energy(0.0)
    kc = t["kc"]
    q = np.array([t["crg"][ns], t["crg"][ns + 1], t["crg"][ns + 2]], float)
    el = sum(kc * q[a] * q[b] / np.linalg.norm(x[ns + a] - x[ns + 3 + b]) for a in range(3) for b in range(3))
    o = t["iac"][ns] - 1
    r = np.linalg.norm(x[ns] - x[ns + 3])
    vdw = t["avdw"][o, 0] ** 2 / r ** 12 - t["bvdw"][o, 0] ** 2 / r ** 6
    # charges are single precision in the reference (topo.f90 crg real(4))
    assert abs(E["ww.el"] - el) < 1e-5 * abs(el)
    assert abs(E["ww.vdw"] - vdw) < 1e-12 * max(1.0, abs(vdw))


def _energy(w, x, tfree):
    w.x = x
    E, EQ, d = w.pot_energy(tfree)
    return E["potential"], d


@pytest.mark.parametrize("cfg,seed,npick", [("ligwat", 0, 4), ("protein", 2, 2), ("kat", 0, 3), ("protarith", 2, 2),
                                                ("ligtip3", 0, 3)])
def test_gradient_is_derivative_of_energy(golden_inp, cfg, seed, npick):
    """Central differences of the potential (lists frozen, LRF off: the LRF
    force is a Taylor field, not the derivative of its energy) against the
    analytic gradient at q-atoms, solute atoms and water atoms."""
    g = load_gold(cfg, seed)
    inp = golden_inp(cfg, seed)
    w = O.Window(inp, g["X0"])
    w.p["lrf"] = False
    w.pair_lists()
    x0 = w.x.copy()
    _, d = _energy(w, x0.copy(), 300.0)
    rng = np.random.default_rng(1)
    picks = list(w.f["iqseq"][:npick]) + list(rng.choice(np.arange(w.ns), npick, replace=False)) + \
        list(w.ns + rng.choice(np.arange(3 * w.nwat), npick, replace=False))
    h = 1e-5
    for a in picks:
        for c in range(3):
            xp, xm = x0.copy(), x0.copy()
            xp[a, c] += h
            xm[a, c] -= h
            ep, _ = _energy(w, xp, 300.0)
            em, _ = _energy(w, xm, 300.0)
            fd = (ep - em) / (2 * h)
            assert abs(fd - d[a, c]) < 2e-4 * max(1.0, abs(d[a, c])), (a, c, fd, d[a, c])


def test_evb_offdiagonal_records_golden():
    """The reference's EVB off-diagonal records (offdiag, md.f90:12911-12937;
    put_ene, nrgy.f90:94-119): one (i, j, Hij, r_kl) per [off_diagonals]
    line in input order with Hij = A exp(-mu r_kl), and the functions leave
    the dynamics alone (state records equal to the same window without
    them)."""
    from q_amd import synth
    g = load_gold("ligevb", 0)
    od = g["en_offd"]
    funcs = synth.CONFIGS["ligevb"][1]["offdiag"]
    assert od.shape == (len(g["en"]), len(funcs), 4)
    for m, (i, j, _k, _l, a, mu) in enumerate(funcs):
        assert np.all(od[:, m, 0] == i) and np.all(od[:, m, 1] == j)
        assert np.allclose(od[:, m, 2], a * np.exp(-mu * od[:, m, 3]), rtol=1e-13, atol=0)
        assert np.all((od[:, m, 3] > 0.5) & (od[:, m, 3] < 20.0))
    assert np.array_equal(g["en"], load_gold("ligwat", 0)["en"])
