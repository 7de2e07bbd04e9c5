"""The drop-in boundary on the GPU: restart files (write_xfin,
src/q6/md.f90:15452-15470), restart-chained windows (a window started from
the previous window's restart, as QligFEP chains its lambda steps,
qligfep.py:773-801, and the reference's restart-consistency check,
test/q6/test1/eval_test.sh:77-120), the qmi-qdyn command and its log
(write_out, md.f90:15292-15438), the backup restart every 1000 steps
(md.f90:4165-4167), die on a SHAKE failure (md.f90:400-455, 14531-14550) and
runs split over several qmi_run calls.

The reference side is tests/golden/restart_ligwat.npz: the ligand-in-water
golden window run by the reference qdyn for 100 steps (final restart file,
log) and a second 100-step window started from that restart (.en records,
log, final restart)."""
import os
import subprocess

import numpy as np
import pytest

from tests.golden.make_golden import RESTART_STEPS, file_sha, read_en

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
QMI_QDYN = os.path.join(ROOT, "q_amd", "qmi-qdyn")


def golden():
    return dict(np.load(os.path.join(GOLD, "restart_ligwat.npz")))


def parse_re(raw: bytes):
    """Records of a restart file: x (f64), v (f64), and the water
    polarisation theta_corr (f32) when present."""
    recs, k = [], 0
    while k < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, k)[0])
        recs.append(raw[k + 4:k + 4 + n])
        assert int(np.frombuffer(raw, np.int32, 1, k + 4 + n)[0]) == n
        k += 8 + n
    x = np.frombuffer(recs[0], np.float64, offset=4)
    v = np.frombuffer(recs[1], np.float64, offset=4)
    th = np.frombuffer(recs[2], np.float32, offset=4) if len(recs) > 2 else None
    assert int(np.frombuffer(recs[0], np.int32, 1)[0]) == x.size
    return x, v, th


def log_blocks(text: str) -> dict:
    """The Energy-summary and Q-atom blocks of a qdyn log:
    {(kind, step): {line key: values}}, step -1 for the FINAL blocks."""
    out, cur = {}, None
    for line in text.splitlines():
        if "Energy summary" in line or "Q-atom energies" in line:
            kind = "E" if "Energy summary" in line else "Q"
            step = -1 if "FINAL" in line else int(line.split("step")[1].split()[0])
            cur = out.setdefault((kind, step), {})
            continue
        if line.startswith("=====") or not line.strip():
            if line.startswith("====="):
                cur = None
            continue
        if cur is None:
            continue
        t = line.split()
        try:
            vals = [float(v) for v in t[1:]]
        except ValueError:
            continue  # column headers
        key = t[0]
        if kind == "Q" or key in ("Q-Q", "Q-prot", "Q-wat", "Q-surr.", "Q-any", "Q-SUM"):
            key, vals = (t[0], t[1]), vals[1:]
        cur[key] = vals
    return out


def compare_logs(ours: str, ref: str, tol_rel: float, label: str):
    """Every value of the reference's blocks against ours: equal at the
    printed %10.2f (0.01) within tol_rel relative."""
    a, b = log_blocks(ours), log_blocks(ref)
    assert set(b) <= set(a), (label, sorted(set(b) - set(a)))
    worst = 0.0
    for k, lines in b.items():
        for key, vals in lines.items():
            assert key in a[k], (label, k, key)
            got = a[k][key]
            assert len(got) == len(vals), (label, k, key)
            for g, r in zip(got, vals):
                err = abs(g - r)
                worst = max(worst, err / max(abs(r), 1.0))
                assert err <= 0.0100001 + tol_rel * abs(r), (label, k, key, g, r)
    print(f"{label}: {len(b)} blocks, worst relative difference {worst:.2e}")


def window_a(d):
    from q_amd import synth
    inp = synth.make_window("ligwat", d, seed=0, steps=RESTART_STEPS, energy=10, output=10)
    assert file_sha([os.path.join(d, "ligwat_s0.top"), os.path.join(d, "ligwat_s0.fep")]) == \
        str(golden()["input_sha"]), "generator drift"
    return inp


def test_qdyn_command_log_and_restart(tmp_path):
    """qmi-qdyn A.inp > A.log as qdyn A.inp > A.log: exit status 0, the
    termination line, the log's energy blocks at every output step and the
    final restart file against the reference's."""
    g = golden()
    d = str(tmp_path)
    inp = window_a(d)
    print("[gpu] qmi-qdyn ligwat", flush=True)
    r = subprocess.run([QMI_QDYN, os.path.basename(inp)], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "terminated normally" in r.stdout
    # FP32 bulk pair sums differ from the reference's FP64 ones by up to
    # ~5e-5 relative (solvent el), and after 100 steps the trajectories by
    # ~1e-5 relative
    compare_logs(r.stdout, str(g["a_log"]), 1e-4, "ligwat log")
    x, v, th = parse_re(open(os.path.join(d, "ligwat_s0.re"), "rb").read())
    xr, vr, thr = parse_re(g["a_re"].tobytes())
    print("restart: max |dx|", np.abs(x - xr).max(), "max |dv|", np.abs(v - vr).max(), "rms v", np.sqrt((vr ** 2).mean()),
          "theta_corr", th, thr)
    assert x.shape == xr.shape and v.shape == vr.shape
    assert np.abs(x - xr).max() < 1e-3
    assert np.abs(v - vr).max() < 1e-2 * np.sqrt((vr ** 2).mean()) * 10
    assert th is not None and thr is not None and th.shape == thr.shape
    assert np.abs(th.astype(np.float64) - thr).max() < 1e-4


def test_restart_chain(tmp_path):
    """A window started from a restart file: (1) from the engine's own restart
    of a 100-step window, the step-0 potential of the new window equals the
    final potential of the previous one as printed (eval_test.sh:77-120);
    (2) from the reference's restart file, its .en records, log and final
    restart match the reference's chained window."""
    from q_amd import synth
    from q_amd.engine import Engine
    g = golden()
    d = str(tmp_path / "own")
    inp_a = window_a(d)
    e = Engine(0)
    e.add_window(inp_a)
    e.prepare()
    print("[gpu] chain: window A", flush=True)
    e.run(RESTART_STEPS, finalize=True)
    e.close()
    inp_b = synth.make_window("ligwat", d, seed=0, steps=RESTART_STEPS, energy=10, output=10, tag="chain",
                              restart="ligwat_s0.re")
    e = Engine(0)
    e.add_window(inp_b)
    e.prepare()
    print("[gpu] chain: window B from the engine's restart", flush=True)
    e.run(RESTART_STEPS, finalize=True)
    e.close()
    la = log_blocks(open(os.path.join(d, "ligwat_s0.log")).read())
    lb = log_blocks(open(os.path.join(d, "chain.log")).read())
    assert la[("E", -1)]["SUM"][1] == lb[("E", 0)]["SUM"][1]

    d2 = str(tmp_path / "ref")
    inp_b2 = synth.make_window("ligwat", d2, seed=0, steps=RESTART_STEPS, energy=10, output=10, tag="chain",
                               restart="ligwat_s0.re")
    open(os.path.join(d2, "ligwat_s0.re"), "wb").write(g["a_re"].tobytes())
    e = Engine(0)
    w = e.add_window(inp_b2)
    e.prepare()
    print("[gpu] chain: window B from the reference restart", flush=True)
    e.run(RESTART_STEPS, finalize=True)
    en = read_en(os.path.join(d2, "chain.en"), 2)
    assert en.shape == g["b_en"].shape
    rel = np.abs(en - g["b_en"]) / np.maximum(np.abs(g["b_en"]), 1.0)
    print("chained .en max rel", rel.max())
    assert rel.max() < 1e-4
    compare_logs(open(os.path.join(d2, "chain.log")).read(), str(g["b_log"]), 1e-4, "chained log")
    x, v, th = parse_re(open(os.path.join(d2, "chain.re"), "rb").read())
    xr, vr, thr = parse_re(g["b_re"].tobytes())
    print("chained restart: max |dx|", np.abs(x - xr).max(), "theta_corr", th, thr)
    assert np.abs(x - xr).max() < 1e-3
    assert np.abs(th.astype(np.float64) - thr).max() < 1e-4
    e.close()


def test_backup_restart_every_1000_steps(tmp_path):
    """The restart file is rewritten after steps 0, 1000, 2000, ... with the
    state after that step."""
    from q_amd import synth
    from q_amd.engine import Engine
    inp = synth.make_window("ligwat", str(tmp_path), seed=0, steps=1005, energy=10, output=500)
    re_path = os.path.join(str(tmp_path), "ligwat_s0.re")
    e = Engine(0)
    w = e.add_window(inp)
    e.prepare()
    print("[gpu] backup restart: step 0", flush=True)
    e.run(1, finalize=False)
    x, v, _ = parse_re(open(re_path, "rb").read())
    assert np.array_equal(x, e.coords(w)) and np.array_equal(v, e.velocities(w))
    print("[gpu] backup restart: steps 1-1001", flush=True)
    e.run(1001, finalize=False)
    x, v, _ = parse_re(open(re_path, "rb").read())
    assert not np.array_equal(x, e.coords(w))          # written after step 1000, one step back
    e2 = Engine(0)
    w2 = e2.add_window(inp)
    e2.prepare()
    e2.run(1001, finalize=False)
    assert np.array_equal(x, e2.coords(w2)) and np.array_equal(v, e2.velocities(w2))
    e.close()
    e2.close()


def test_shake_failure_dies_alone(tmp_path):
    """A window whose SHAKE cannot converge (a water hydrogen placed on its
    oxygen from the host: the constraint direction vanishes, and the kernel's
    1000-sweep limit stops it) ends as die ends qdyn -- its log gets the
    termination record and its restart file is written -- while the other
    window of the batch runs to its end, gets its final summary and restart,
    and follows bitwise the trajectory it has alone."""
    from q_amd import synth
    from q_amd.engine import Engine, QmiError
    d = str(tmp_path)
    good = synth.make_window("ligwat", d, seed=0, steps=20, energy=10, output=10, tag="good", lam=0.25)
    bad = synth.make_window("ligwat", d, seed=0, steps=20, energy=10, output=10, tag="bad", lam=0.75)
    e = Engine(0)
    wg = e.add_window(good)
    wb = e.add_window(bad)
    e.prepare()
    info = e.info(wb)
    x = e.coords(wb).reshape(-1, 3).copy()
    o = info["nat_solute"]              # first water oxygen
    x[o + 1] = x[o]
    e.set_state(wb, x=x.ravel())
    print("[gpu] shake failure batch", flush=True)
    with pytest.raises(QmiError, match="shake failure in window"):
        e.run(20, finalize=True)
    assert e.info(wb)["status"] == 1 and e.info(wg)["status"] == 0
    log_bad = open(os.path.join(d, "bad.log")).read()
    assert "ABNORMAL TERMINATION of qdyn" in log_bad and "Terminating due to shake failure" in log_bad
    assert "restart file written at step" in log_bad
    assert os.path.getsize(os.path.join(d, "bad.re")) > 0
    log_good = open(os.path.join(d, "good.log")).read()
    assert "FINAL  Energy summary" in log_good and "ABNORMAL" not in log_good
    xg, _, _ = parse_re(open(os.path.join(d, "good.re"), "rb").read())
    e.close()
    d2 = str(tmp_path / "alone")
    good2 = synth.make_window("ligwat", d2, seed=0, steps=20, energy=10, output=10, tag="good", lam=0.25)
    e2 = Engine(0)
    e2.add_window(good2)
    e2.prepare()
    e2.run(20, finalize=True)
    assert np.array_equal(e2.coords(0), xg)
    e2.close()


def test_run_in_several_calls(tmp_path):
    """Windows with different [MD] steps advanced over two qmi_run calls: each
    window's .en records and log summaries are those of its own run alone
    (every step once, in order)."""
    from q_amd import synth
    from q_amd.engine import Engine
    d = str(tmp_path)
    inps = [synth.make_window("ligwat", d, seed=0, steps=40, energy=10, output=10, tag="w40", lam=0.3),
            synth.make_window("ligwat", d, seed=0, steps=70, energy=10, output=10, tag="w70", lam=0.6)]
    e = Engine(0)
    for p in inps:
        e.add_window(p)
    e.prepare()
    print("[gpu] two calls", flush=True)
    e.run(30, finalize=False)
    e.run(100, finalize=True)
    e.close()
    for tag, n in (("w40", 40), ("w70", 70)):
        en = read_en(os.path.join(d, tag + ".en"), 2)
        assert len(en) == n // 10 - 1
        steps = sorted(k[1] for k in log_blocks(open(os.path.join(d, tag + ".log")).read()) if k[0] == "E")
        assert steps == [-1] + list(range(0, n, 10)), (tag, steps)
        d1 = os.path.join(d, "alone_" + tag)
        alone = synth.make_window("ligwat", d1, seed=0, steps=n, energy=10, output=10, tag=tag,
                                  lam=0.3 if n == 40 else 0.6)
        e1 = Engine(0)
        e1.add_window(alone)
        e1.prepare()
        e1.run(n, finalize=True)
        e1.close()
        assert np.array_equal(en, read_en(os.path.join(d1, tag + ".en"), 2)), tag
