"""The learned lambda schedule end to end on a held-out perturbation
(ligand in water): pilots on the engine, the shipped overlap network (run by
the library's HIP kernels) proposes a ladder, and the ladder's BAR free
energy is compared with the uniform 51-window ladder.

Tolerance: the ladders' free energies carry sampling noise far above the
0.3 kcal/mol of a converged comparison.  Measured on four held-out
ligand-in-water legs with 150 ps windows and 4 Maxwell seeds each
(q_amd/data/schedule_eval_150ps.json), the paired difference of one seed's
learned and 51-window ladders scatters with a standard deviation of 0.3-1.2
kcal/mol per leg, 0.8 kcal/mol RMS over the legs; more sampled time did not
shrink it (the 51-window ladder's own replica spread is 0.3-0.7 kcal/mol
there, 0.4 with 30 ps analysed per window).  This test's 4 ps windows scatter
more.  So agreement is asserted at 3 kcal/mol, under 4 times that RMS, which
a ladder with a gap in its overlap misses by far; the time budget and the
window count are exact."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def result(tmp_path_factory):
    from q_amd import schedule as S
    from q_amd import synth
    net = S.load_predictor()
    p = S.Perturbation(9001, "ligwat", 6)
    system = synth.reperturb(S.build_system(p), p.seed, p.n_pert, p.alpha)
    work = str(tmp_path_factory.mktemp("sched"))
    print(f"schedule case {p}", flush=True)
    row = S.evaluate(p, system, net.forward, work, pilot_steps=1000, prod_steps=4000, pilot_skip=40, prod_skip=200)
    net.close()
    print(json.dumps({k: v for k, v in row.items()}, indent=1))
    return row


def test_ladder_size_and_time_budget(result):
    assert 16 <= result["windows"] <= 22
    assert result["simulated_fraction"] <= 0.5
    s = np.array(result["schedule"])
    assert s[0] == 0.0 and s[-1] == 1.0 and np.all(np.diff(s) > 0)


def test_learned_ladder_free_energy(result):
    st = result["stats"]
    assert abs(st["learned"]["dG"] - st["ref51"]["dG"]) <= 3.0
    assert st["learned"]["min_overlap"] > 0.3


def test_predictions_follow_pilot_overlaps(result):
    # the network's interval overlaps, against the pilot windows' own BAR
    # overlaps: closer than the constant training-set mean
    pred = np.array(result["predicted_overlap"])
    meas = np.array(result["pilot_overlap"])
    z = np.load(__import__("os").path.join(__import__("os").path.dirname(__import__("q_amd").__file__), "data",
                                           "overlap_training_rows.npz"))
    base = float(np.mean(z["Y"]))
    assert np.mean(np.abs(pred - meas)) < np.mean(np.abs(base - meas))
