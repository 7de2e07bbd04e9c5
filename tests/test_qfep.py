"""The qfep judge (q_amd/qfep.py) against the reference qfep
(src/q6/qfep.f90) on two five-window lambda ladders: tests/golden/qfep holds
the energy files written by the reference qdyn and the tables the reference
qfep printed for them (tests/golden/make_golden.py, make_qfep);
tests/golden/qfep_evb the same for a window with EVB off-diagonal functions,
analysed with a qfep off-diagonal function of r_kl (ground-state mixing,
c1**2, c2**2, <r_xy>).  Every printed number of Parts 0-6 must agree to the
printed digits."""
import io
import os
import sys

import numpy as np
import pytest

from q_amd import qfep as QF

GOLDS = {k: os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", k) for k in ("qfep", "qfep_evb")}


def _rows(txt, title):
    """numeric rows of the table that follows `title`"""
    lines = txt[txt.index(title):].split("\n")[1:]
    out = []
    for ln in lines:
        s = ln.strip()
        if not s or s.startswith("-->"):
            if out and not s:
                break
            continue
        if s.startswith("#"):
            if out:
                break
            continue
        out.append(s.split())
    return out


@pytest.fixture(scope="module", params=list(GOLDS))
def both(request):
    gold = GOLDS[request.param]
    ref = open(os.path.join(gold, "qfep.out")).read()
    cwd = os.getcwd()
    os.chdir(gold)
    try:
        stdin, stdout = sys.stdin, sys.stdout
        sys.stdin = open("qfep.inp")
        sys.stdout = io.StringIO()
        QF.main()
        mine = sys.stdout.getvalue()
    finally:
        sys.stdin.close()
        sys.stdin, sys.stdout = stdin, stdout
        os.chdir(cwd)
    return ref, mine


@pytest.mark.parametrize("part", ["# Part 1", "# Part 2", "# Part 3", "# Part 4", "# Part 5", "# Part 6"])
def test_tables_match_reference(both, part):
    ref, mine = both
    a, b = _rows(ref, part), _rows(mine, part)
    assert len(a) == len(b) > 0
    for ra, rb in zip(a, b):
        assert [float(v) for v in ra] == [float(v) for v in rb], (ra, rb)


def test_part0_averages(both):
    ref, mine = both
    a = [r for r in _rows(ref, "# Part 0") if r[0].endswith(".en")]
    b = [r for r in _rows(mine, "# Part 0") if r[0].endswith(".en")]
    assert len(a) == len(b) == 10
    for ra, rb in zip(a, b):
        assert ra[1:4] == rb[1:4]
        assert np.allclose([float(v) for v in ra[4:]], [float(v) for v in rb[4:]], atol=0.006)


def test_bar_equals_zwanzig_on_identical_states(tmp_path):
    """Closed form: when both states have the same energy in every frame the
    free energy differences vanish in every estimator."""
    rec = np.zeros((30, 2, 15))
    rng = np.random.default_rng(0)
    e = rng.normal(-500, 3, 30)
    paths = []
    for k, lam in enumerate([1.0, 0.5, 0.0]):
        r = rec.copy()
        r[:, 0, 0], r[:, 1, 0] = lam, 1 - lam
        r[:, 0, 1] = r[:, 1, 1] = e + k
        p = tmp_path / f"w{k}.en"
        with open(p, "wb") as f:
            for t in range(len(r)):
                for s in range(2):
                    f.write(np.int32(124).tobytes() + np.int32(s + 1).tobytes() + r[t, s].tobytes() +
                            np.int32(124).tobytes())
                f.write(np.int32(0).tobytes() + np.int32(0).tobytes())
        paths.append(str(p))
    R = QF.qfep(paths, 2, 0.596, 0, 10, 1)
    assert abs(R.dG_bar) < 1e-12 and abs(R.dG_fep) < 1e-12
    assert abs(R.dglusum[3]) < 1e-12 and abs(R.dgtisum[3]) < 1e-12


def test_bar_matches_exact_gaussian_free_energy(tmp_path):
    """Closed form: with energy gaps drawn so that dU is Gaussian with
    variance s2 in the forward direction and consistent in reverse
    (mean s2/(2kT) apart), the exact free energy is mu - s2/(2 kT); BAR
    recovers it to sampling accuracy."""
    kT, mu, s2 = 0.596, 2.0, 0.5
    rng = np.random.default_rng(3)
    n = 20000
    dU0 = rng.normal(mu, np.sqrt(s2), n)            # sampled in state A
    dU1 = rng.normal(mu - s2 / kT, np.sqrt(s2), n)  # sampled in state B
    paths = []
    for k, (lam, dU) in enumerate([(1.0, dU0), (0.0, dU1)]):
        p = tmp_path / f"g{k}.en"
        with open(p, "wb") as f:
            for t in range(n):
                EA, EB = 0.0, dU[t]
                for s, (l, E) in enumerate([(lam, EA), (1 - lam, EB)]):
                    row = np.zeros(15)
                    row[0], row[1] = l, E
                    f.write(np.int32(124).tobytes() + np.int32(s + 1).tobytes() + row.tobytes() +
                            np.int32(124).tobytes())
                f.write(np.int32(0).tobytes() + np.int32(0).tobytes())
        paths.append(str(p))
    R = QF.qfep(paths, 2, kT, 0, 10, 1)
    exact = mu - s2 / (2 * kT)
    assert abs(R.dG_bar - exact) < 0.02
