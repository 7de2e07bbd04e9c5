"""Seeded-ensemble free energy on the GPU: a five-window ladder (lambda 0,
0.25, 0.5, 0.75, 1) of a two-atom perturbation of the ligand in water from
eight Maxwell seeds, 4000 steps per window, run by the engine as one batch
of 40 windows, against the same ladder run by the reference qdyn
(tests/golden/ensemble_ligwat.npz).  Per seed, the qfep judge (q_amd.qfep,
the restatement of qfep.f90) gives BAR and Zwanzig free energies from each
side's .en records, the first 100 steps skipped.  Within a few hundred steps
the GPU and FP64 trajectories decorrelate (chaos amplifies the FP32 force
rounding), so the two ensembles are independent samples of the same
free-energy distribution: their means must agree within three standard
errors of the difference, and within 0.1 kcal/mol when that is larger; the
spreads are reported.  The reference records are also run through the
judge, which must give the reference qfep's printed totals."""
import os

import numpy as np
import pytest

from q_amd import qfep as QF
from tests.golden.make_golden import (ENS_SEEDS, ENS_SKIP, ENS_STEPS, QFEP_LADDER, ensemble_system, ensemble_window,
                                      file_sha)

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def write_en(path, recs):
    """A 2-state .en file (nrgy.f90:94-119) from the lambda and total of its
    state records (records x states x 2; the other fields zero), with empty
    off-diagonal records."""
    with open(path, "wb") as f:
        for block in recs:
            for s, r in enumerate(block):
                full = np.zeros(15)
                full[:2] = r
                f.write(np.int32(124).tobytes() + np.int32(s + 1).tobytes() + full.tobytes() +
                        np.int32(124).tobytes())
            f.write(np.int32(0).tobytes() * 2)


def judge(paths):
    R = QF.qfep(paths, 2, 0.596, ENS_SKIP, 20, 5)
    return R.dG_bar, R.dG_fep


def test_seeded_ensemble_free_energy(tmp_path):
    from q_amd.engine import Engine
    g = dict(np.load(os.path.join(GOLD, "ensemble_ligwat.npz")))
    d = str(tmp_path)
    system = ensemble_system()
    inps = []
    for s in ENS_SEEDS:
        for k in range(QFEP_LADDER):
            inps.append(ensemble_window(d, s, k, system if not inps else None))
    assert file_sha([os.path.join(d, "ligwat_s0.top"), os.path.join(d, "ligwat_s0.fep")]) == str(g["input_sha"])
    e = Engine(0)
    for p in inps:
        e.add_window(p)
    e.prepare()
    print(f"[gpu] ensemble: {len(inps)} windows x {ENS_STEPS} steps", flush=True)
    e.run(ENS_STEPS, finalize=True)
    e.close()
    mine, ref = [], []
    for si, s in enumerate(ENS_SEEDS):
        mine.append(judge([os.path.join(d, f"s{s}_w{k}.en") for k in range(QFEP_LADDER)]))
        rp = []
        for k in range(QFEP_LADDER):
            p = os.path.join(d, f"ref_s{s}_w{k}.en")
            write_en(p, g["en"][si, k])
            rp.append(p)
        ref.append(judge(rp))
    mine, ref = np.array(mine), np.array(ref)
    # the judge reproduces the reference qfep's totals on the reference records
    assert np.allclose(ref[:, 0], g["dG_bar"], atol=1.5e-3) and np.allclose(ref[:, 1], g["dG_fep"], atol=1.5e-3)
    for k, name in enumerate(("BAR", "Zwanzig")):
        print(f"{name}: GPU per seed {np.round(mine[:, k], 3)} mean {mine[:, k].mean():.3f} sd {mine[:, k].std():.3f}; "
              f"reference {np.round(ref[:, k], 3)} mean {ref[:, k].mean():.3f} sd {ref[:, k].std():.3f}; "
              f"difference of means {mine[:, k].mean() - ref[:, k].mean():+.3f} kcal/mol")
    n = len(ENS_SEEDS)
    for k in range(2):
        se = np.sqrt((mine[:, k].var(ddof=1) + ref[:, k].var(ddof=1)) / n)
        print(f"  standard error of the difference {se:.3f} kcal/mol")
        assert abs(mine[:, k].mean() - ref[:, k].mean()) < max(0.1, 3.0 * se)
