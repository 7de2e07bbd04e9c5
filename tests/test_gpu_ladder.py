"""The 64-window batched lambda ladder at its stated shape: 32 ligand-in-water
windows and 32 protein-leg windows (lambda_k = k/31 per leg, two Maxwell
seeds per leg) advanced together by one engine, as the benchmark runs them.

Checks: every window's step-0 energies against the reference's golden
vectors (the windows share the golden systems; at step 0 the per-state Q
energies and the lambda-independent terms do not depend on lambda or on the
velocity seed, because the initial SHAKE of the coordinates does not use the
velocities, md.f90:2506-2525); 75 steps with three pair-list updates and
energy steps, after which every window is still running, and the first and
last window of each leg follow bitwise the trajectory they have alone."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import E_NAMES

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N_LEG = 32
STEPS = 75
# terms that depend neither on lambda (the Q terms enter E weighted) nor on
# the velocities at step 0; the radial solvent restraint does, through the
# free temperature in its radius shift (restrain_solvent, md.f90:14910-14957),
# and is compared for the windows with the golden's velocity seed
LAMBDA_FREE = ["LRF", "p.bond", "p.angle", "p.torsion", "p.improper", "w.bond", "w.angle", "pp.el", "pp.vdw",
               "pw.el", "pw.vdw", "ww.el", "ww.vdw", "restraint.fix", "restraint.shell", "restraint.water_pol"]
GOLDEN_MD_SEED = 1


@pytest.fixture(scope="module")
def ladder(systems):
    from q_amd import synth
    inps = []
    for cfg, seed, md0 in (("ligwat", 0, 1), ("protein", 2, 3)):
        # window 0 writes the system's files; the others share them
        first = systems(cfg, seed=seed, steps=STEPS + 5, lam=0.0, md_seed=md0, energy=10, output=5,
                        tag=f"{cfg}_w00")
        d = os.path.dirname(first)
        inps.append(first)
        for k in range(1, N_LEG):
            inps.append(synth.make_window(cfg, d, seed=seed, steps=STEPS + 5, lam=k / (N_LEG - 1),
                                          md_seed=md0 + (k % 2), energy=10, output=5, tag=f"{cfg}_w{k:02d}"))
    return inps


def test_ladder_64_windows(ladder):
    from q_amd.engine import Engine
    e = Engine(0)
    for p in ladder:
        e.add_window(p)
    e.prepare()
    print(f"[gpu] ladder: {len(ladder)} windows, step-0 evaluation", flush=True)
    e.evaluate(rebuild_lists=True)
    seeds = [1 + (k % 2) for k in range(N_LEG)] + [3 + (k % 2) for k in range(N_LEG)]
    for w, (cfg, seed) in enumerate([("ligwat", 0)] * N_LEG + [("protein", 2)] * N_LEG):
        g = np.load(os.path.join(GOLD, f"{cfg}_s{seed}.npz"))
        eq = e.q_energies(w)
        rel = np.abs(eq[:, 1:] - g["EQ0"][:, 1:]) / np.maximum(np.abs(g["EQ0"][:, 1:]), 1.0)
        assert rel.max() < 1e-6, (w, rel.max())
        E = e.energies(w)
        for name in LAMBDA_FREE:
            ref = float(g["E0"][E_NAMES.index(name)])
            # energy terms are FP64 on energy steps (the pair kernels too)
            tol = 1e-6 * abs(ref) + 1e-6
            assert abs(E[name] - ref) < tol, (w, name, E[name], ref)
        if seeds[w] == GOLDEN_MD_SEED:
            ref = float(g["E0"][E_NAMES.index("restraint.solvent_radial")])
            assert abs(E["restraint.solvent_radial"] - ref) < 1e-6 * abs(ref) + 1e-6, w
    print(f"[gpu] ladder: {STEPS} steps", flush=True)
    e.run(STEPS, finalize=False)
    for w in range(len(ladder)):
        assert e.info(w)["status"] == 0, w
    xb = {w: e.coords(w) for w in (0, N_LEG - 1, N_LEG, 2 * N_LEG - 1)}
    e.close()
    for w, x in xb.items():
        es = Engine(0)
        es.add_window(ladder[w])
        es.prepare()
        print(f"[gpu] ladder: window {w} alone", flush=True)
        es.run(STEPS, finalize=False)
        assert np.array_equal(es.coords(0), x), w
        es.close()
