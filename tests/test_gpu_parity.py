"""GPU parity of the HIP engine against the reference Q6 routines.

The reference side is the golden vectors of tests/golden/ (written by
tests/golden/make_golden.py from the unmodified Q6 modules, src/q6/md.f90
make_pair_lists + pot_energy at step 0) for the same seeded input files.
Step-0 energies (per-state Q energies and every term of the potential)
must agree within 1e-6 relative, the FP32 gradient within 1e-3 of its rms,
pair lists exactly (as sets of unordered atom pairs).
"""
import os

import numpy as np
import pytest

from tests.golden.make_golden import E_NAMES, canonical_lists, sha

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# three seeded synthetic windows, the reference's own test system (KAT
# ligand in water, one state, sequence restraint) and the protein leg under
# the arithmetic LJ combination rule (nonbon2_*, md.f90:8972-10227), and
# the ligand in water with soft-core maximum potentials under either rule
# (sc_lookup, qatom.f90:1600-1664), the ligand in a 3-atom water with LJ
# on the hydrogens (nonbond_3atomsolvent, nonbond_qw_3atom), the ligand with
# el_scale, soft pairs, excluded pairs, a Morse q-bond and FEP SHAKE
# constraints, with library charges for q-q pairs under either LJ rule
# (nonbond_qq_lib_charges, nonbon2_qq_lib_charges), and with sequence,
# position, distance, angle and wall restraints (p_restrain), and a
# QligFEP-style dual-topology perturbation (qligfep.py:311-362), and both
# legs with Q-atom lists rebuilt at list updates (q_atom cut-off inside the
# exclusion radius, nbqplist / nbqwlist)
CASES = {"water": 0, "ligwat": 0, "protein": 2, "kat": 0, "protarith": 2, "ligmaxpot": 0, "ligarithmax": 0,
         "ligtip3": 0, "ligfepx": 0, "liglibq": 0, "ligarithlibq": 0, "ligrestr": 0, "ligdual": 0,
         "ligqcut": 0, "protqcut": 2, "ligshake": 0}


def _rel(a, b, floor=1.0):
    return abs(a - b) / max(abs(b), floor)


@pytest.fixture(scope="module")
def evaluated(golden_inp):
    from q_amd.engine import Engine
    out = {}
    for cfg, seed in CASES.items():
        inp = golden_inp(cfg, seed)
        g = dict(np.load(os.path.join(GOLD, f"{cfg}_s{seed}.npz")))
        ref = dict(zip(E_NAMES, g["E0"].tolist()))
        dump = {"E0": ref, "EQ0": g["EQ0"], "D0": g["D0"], "gold": g}
        e = Engine(0)
        w = e.add_window(inp)
        e.prepare()
        print(f"[gpu] evaluate {cfg}", flush=True)
        e.evaluate(rebuild_lists=True)
        out[cfg] = (e, w, dump)
    return out


@pytest.mark.parametrize("cfg", list(CASES))
def test_step0_energies(evaluated, cfg):
    e, w, dump = evaluated[cfg]
    E = e.energies(w)
    ref = dump["E0"]
    print(cfg)
    for k in ["pp.el", "pp.vdw", "pw.el", "pw.vdw", "ww.el", "ww.vdw", "LRF", "p.bond", "p.angle", "p.torsion",
              "p.improper", "w.bond", "w.angle", "q.bond", "q.angle", "q.torsion", "q.improper", "qx.el", "qx.vdw",
              "restraint.total", "restraint.fix", "restraint.shell", "restraint.solvent_radial",
              "restraint.water_pol", "potential"]:
        print(f"  {k:18s} gpu {E[k]: .10f} ref {ref[k]: .10f} rel {_rel(E[k], ref[k]):.2e}")
    # every energy term: on energy steps the pair kernels evaluate the
    # energy terms in FP64 from the FP64 coordinates (the forces stay FP32),
    # so the bulk nonbonded sums are held to the FP64 gate as well.  The
    # totals cancel to far less than the sum of the terms' magnitudes (e.g.
    # ww.el 152 kcal/mol of 7e6 for the ligand in water), so a relative gate
    # on the total is strict; the 1e-6 kcal/mol floor covers the fixed-point
    # quantum (2^-36 kcal/mol per wave sum)
    for k in ["ww.el", "ww.vdw", "pw.el", "pw.vdw", "pp.el", "pp.vdw", "LRF", "p.bond", "p.angle", "p.torsion",
              "p.improper", "q.bond", "q.angle", "q.torsion", "q.improper", "restraint.total",
              "restraint.solvent_radial", "restraint.water_pol", "restraint.shell", "restraint.protein",
              "potential"]:
        assert abs(E[k] - ref[k]) < 1e-6 * abs(ref[k]) + 1e-6, (k, E[k], ref[k])


@pytest.mark.parametrize("cfg", ["ligwat", "protein", "kat", "protarith", "ligmaxpot", "ligarithmax", "ligtip3",
                                 "ligfepx", "liglibq", "ligarithlibq", "ligrestr", "ligdual", "ligqcut",
                                 "protqcut"])
def test_step0_q_energies(evaluated, cfg):
    e, w, dump = evaluated[cfg]
    EQ = e.q_energies(w)
    ref = dump["EQ0"]
    names = ["lambda", "total", "q.bond", "q.angle", "q.torsion", "q.improper", "qx.el", "qx.vdw", "qq.el",
             "qq.vdw", "qp.el", "qp.vdw", "qw.el", "qw.vdw", "restraint"]
    for s in range(EQ.shape[0]):
        for k, n in enumerate(names):
            print(f"  s{s} {n:10s} gpu {EQ[s, k]: .12f} ref {ref[s, k]: .12f} rel {_rel(EQ[s, k], ref[s, k]):.2e}")
    for s in range(EQ.shape[0]):
        for k in range(EQ.shape[1]):
            assert _rel(EQ[s, k], ref[s, k]) < 1e-6, (s, names[k])


@pytest.mark.parametrize("cfg", list(CASES))
def test_step0_gradient(evaluated, cfg):
    e, w, dump = evaluated[cfg]
    d = e.forces(w).reshape(-1, 3)
    ref = dump["D0"].reshape(-1, 3)
    err = np.linalg.norm(d - ref, axis=1)
    scale = np.linalg.norm(ref, axis=1)
    print(cfg, "max |d-ref|", err.max(), "at", int(err.argmax()), "|ref| there", scale[err.argmax()],
          "rms ref", np.sqrt((scale ** 2).mean()))
    assert err.max() < 1e-3 * max(1.0, np.sqrt((scale ** 2).mean()))


@pytest.mark.parametrize("cfg", list(CASES))
def test_pairlists(evaluated, cfg):
    from q_amd import engine as EN
    e, w, dump = evaluated[cfg]
    g = dump["gold"]
    i, j, lj = e.pairlist(w, EN.LIST_PP)
    pp = np.stack([i, j, lj], 1) if len(i) else np.zeros((0, 3))
    pi, pj, plj = e.pairlist(w, EN.LIST_PW)
    pw = np.stack([pi, pj, plj], 1) if len(pi) else np.zeros((0, 3))
    wi, wj, _ = e.pairlist(w, EN.LIST_WW)
    ww = np.stack([wi, wj], 1) if len(wi) else np.zeros((0, 2))
    qi, qj, qlj = e.pairlist(w, EN.LIST_QP)
    qp = np.stack([qi, qj, qlj], 1) if len(qi) else np.zeros((0, 3))
    _, ow, _ = e.pairlist(w, EN.LIST_QW)
    C = canonical_lists(pp, pw, ww, qp, ow)
    for k, a in C.items():
        print(f"  {cfg} {k}: gpu {len(a)} ref {int(g['n_' + k])}")
        assert len(a) == int(g[f"n_{k}"]), k
        assert len(np.unique(a, axis=0)) == len(a), k
        assert sha(a) == str(g[f"sha_{k}"]), k
