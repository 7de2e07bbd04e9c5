"""GPU dynamics: 100 MD steps against the reference md_run on the same
inputs (src/q6/md.f90:3917-4263; golden vectors of tests/golden/), the .en
file layout (nrgy.f90:94-119), and batching invariance (a window gives
bitwise the same trajectory alone or batched with others)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def read_en(path, nstates):
    """Records of a Fortran unformatted .en file: per write, nstates records
    of (int32 state, 15 f64) and one off-diagonal record."""
    raw = open(path, "rb").read()
    out, k = [], 0
    while k < len(raw):
        block = []
        for s in range(nstates):
            n = np.frombuffer(raw, np.int32, 1, k)[0]
            assert n == 124
            st = np.frombuffer(raw, np.int32, 1, k + 4)[0]
            assert st == s + 1
            block.append(np.frombuffer(raw, np.float64, 15, k + 8).copy())
            assert np.frombuffer(raw, np.int32, 1, k + 4 + n)[0] == n
            k += 8 + n
        n = np.frombuffer(raw, np.int32, 1, k)[0]
        assert n == 0  # no off-diagonal elements
        k += 8
        out.append(np.array(block))
    return out


def rel_err(a, b):
    """Relative difference of .en fields (lambda, total, q bond, angle,
    torsion, improper, qx el, qx vdw, qq el, qq vdw, qp el, qp vdw, qw el,
    qw vdw, restraint): each field against its own magnitude, except the
    sums -- qx (= qq + qp + qw) and total -- which are taken against the sum
    of their parts' magnitudes, as a sum of parts of opposite sign can be far
    smaller than the parts whose errors it carries."""
    b = np.asarray(b)
    scale = np.maximum(np.abs(b), 1.0)
    scale[..., 6] = np.maximum(np.abs(b[..., [8, 10, 12]]).sum(-1), 1.0)
    scale[..., 7] = np.maximum(np.abs(b[..., [9, 11, 13]]).sum(-1), 1.0)
    scale[..., 1] = np.maximum(np.abs(b[..., [2, 3, 4, 5, 8, 9, 10, 11, 12, 13, 14]]).sum(-1), 1.0)
    return np.abs(np.asarray(a) - b) / scale


@pytest.mark.parametrize("cfg,seed", [("ligwat", 0), ("protein", 2), ("kat", 0), ("protarith", 2), ("ligtip3", 0),
                                      ("ligfepx", 0), ("liglibq", 0), ("ligarithlibq", 0), ("ligrestr", 0),
                                      ("ligdual", 0), ("ligqcut", 0), ("protqcut", 2), ("ligshake", 0)])
def test_100_steps(golden_inp, cfg, seed):
    from q_amd.engine import Engine
    g = dict(np.load(os.path.join(GOLD, f"{cfg}_s{seed}.npz")))
    inp = golden_inp(cfg, seed)
    d = os.path.dirname(inp)
    tag = os.path.splitext(os.path.basename(inp))[0]
    e = Engine(0)
    w = e.add_window(inp)
    e.prepare()
    print(f"[gpu] engine run {cfg}", flush=True)
    e.run(100, finalize=True)
    x = e.coords(w)
    dx = np.abs(x - g["XN"]).max()
    print(cfg, "max |x - x_ref| after 100 steps", dx)
    assert dx < 0.05
    eq = e.q_energies(w)
    ref = g["EQN"]
    rel = rel_err(eq, ref)
    print(cfg, "final EQ max rel", rel.max(), "per field", rel.max(axis=0))
    assert rel.max() < 1e-4
    S = eq.shape[0]
    en_ref = g["en"]
    en = read_en(os.path.join(d, tag + ".en"), S)
    assert len(en) == len(en_ref) == 9     # steps 10, 20, ..., 90
    for k, (a, b) in enumerate(zip(en, en_ref)):
        r = rel_err(a, b)
        print(f"  .en record {k} max rel {r.max():.2e}")
        assert r.max() < 1e-4


def test_evb_offdiagonal_records(golden_inp):
    """EVB off-diagonal functions (md.f90:12911-12937): the .en off-diagonal
    records (i, j, Hij, r_kl) of a 100-step run against the reference's, and
    the log's H(i,j) lines."""
    from q_amd.engine import Engine
    g = dict(np.load(os.path.join(GOLD, "ligevb_s0.npz")))
    inp = golden_inp("ligevb", 0)
    d = os.path.dirname(inp)
    tag = os.path.splitext(os.path.basename(inp))[0]
    e = Engine(0)
    w = e.add_window(inp)
    e.prepare()
    print("[gpu] engine run ligevb", flush=True)
    e.run(100, finalize=True)
    from tests.golden.make_golden import read_en as read_en_offd
    en, od = read_en_offd(os.path.join(d, tag + ".en"), 2, offd=True)
    ref = g["en_offd"]
    assert od.shape == ref.shape
    assert np.array_equal(od[:, :, :2], ref[:, :, :2])
    rel = np.abs(od[:, :, 2:] - ref[:, :, 2:]) / np.abs(ref[:, :, 2:])
    print("offd max rel", rel.max())
    assert rel.max() < 1e-4
    r = np.abs(en - g["en"]) / np.maximum(np.abs(g["en"]), 1.0)
    assert r.max() < 1e-4
    log = open(os.path.join(d, tag + ".log")).read()
    assert log.count("H( 1, 2) =") >= 1 and log.count("dist. between Q-atoms   3   7 =") >= 1


def test_batched_equals_single(systems):
    from q_amd.engine import Engine
    # the two LJ combination rules and both solvent models in one batch
    inps = [systems("ligwat", seed=0, steps=40, lam=0.25), systems("protein", seed=2, steps=40, lam=0.75),
            systems("protarith", seed=2, steps=40, lam=0.5), systems("ligtip3", seed=0, steps=40, lam=0.5)]
    eb = Engine(0)
    for p in inps:
        eb.add_window(p)
    eb.prepare()
    eb.run(30, finalize=False)
    xb = [eb.coords(w) for w in range(len(inps))]
    for w, p in enumerate(inps):
        # the single windows run on the default launch shapes, the batch on
        # the tuned ones: neither the batch nor the tuner may change a result
        es = Engine(0)
        es.set_autotune(False)
        es.add_window(p)
        es.prepare()
        es.run(30, finalize=False)
        assert np.array_equal(es.coords(0), xb[w]), p


def test_deterministic(systems):
    from q_amd.engine import Engine
    inp = systems("ligwat", seed=0, steps=40, lam=0.5)
    xs = []
    for _ in range(2):
        e = Engine(0)
        e.add_window(inp)
        e.prepare()
        e.run(25, finalize=False)
        xs.append(e.coords(0))
    assert np.array_equal(xs[0], xs[1])
