"""DCD trajectory output (trj.f90:184-262, md.f90:4151-4154, 4199-4203)
against the reference qdyn's file for the same input: a ligand-in-water
window written every 5 steps for 20 steps with a three-row atom mask.  The
header records (CORD block, title with topology and mask rows, atom count)
must be byte-identical; the frames must hold the same atoms at the same
steps, within 1e-3 A (the trajectories agree to ~1e-5 A after 20 steps,
see the dynamics tests, and the file stores real(4))."""
import os
import struct
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


def records(path):
    b = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(b):
        n = struct.unpack("<i", b[pos:pos + 4])[0]
        assert struct.unpack("<i", b[pos + 4 + n:pos + 8 + n])[0] == n
        out.append(b[pos + 4:pos + 4 + n])
        pos += 8 + n
    return out


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    """The engine's trajectory for the golden input (directory, path)."""
    import make_golden as M
    from q_amd.engine import Engine
    d = str(tmp_path_factory.mktemp("trj"))
    inp = M.trj_window(d)
    sha = M.file_sha([inp, os.path.join(d, "ligwat_s0.top"), os.path.join(d, "ligwat_s0.fep")])
    assert sha == open(os.path.join(HERE, "golden", "trj", "input_sha.txt")).read().strip()
    e = Engine(0)
    e.add_window(inp)
    e.set_log(0, None)
    e.prepare()
    e.run(M.TRJ_STEPS, finalize=True)
    e.close()
    return d, os.path.join(d, "trj.dcd")


def test_dcd_matches_reference(written):
    import make_golden as M
    d, path = written
    got = records(path)
    ref = records(os.path.join(HERE, "golden", "trj", "trj.dcd"))
    assert len(got) == len(ref) == 3 + 3 * (M.TRJ_STEPS // M.TRJ_EVERY)
    for k in range(3):
        assert got[k] == ref[k], f"header record {k}"
    worst = 0.0
    for k in range(3, len(ref)):
        a, b = np.frombuffer(got[k], "<f4"), np.frombuffer(ref[k], "<f4")
        assert a.shape == b.shape
        worst = max(worst, float(np.abs(a - b).max()))
    print("frames max |dx|", worst)
    assert worst < 1e-3


QCALC = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "qcalc")


def _qcalc_rows(d, dcd):
    """Result rows (frame, values) of the reference qcalc on a trajectory:
    distances, an angle and a torsion (with their energies) over atoms of
    the mask."""
    import subprocess
    cmds = "ligwat_s0.top\ndist\n1 5\ndist\n3 54\nangle\n1 2 3\ntorsion\n1 2 3 4\ngo\n" + dcd + "\n.\n"
    r = subprocess.run([QCALC], cwd=d, input=cmds, capture_output=True, text=True, timeout=60)
    rows = [ln.split() for ln in r.stdout.splitlines() if ln.startswith(dcd)]
    return [(int(t[1]), [float(v) for v in t[2:]]) for t in rows]


@pytest.mark.skipif(not os.path.exists(QCALC), reason="the reference qcalc is not built in oracle/_ref")
def test_qcalc_reads_engine_trajectory(written):
    """qcalc compatibility: the reference qcalc (src/q6/qcalc.f90, built in
    oracle/_ref) reads the engine's DCD and prints the same geometry as for
    the reference qdyn's file."""
    import shutil
    d, path = written
    shutil.copy(os.path.join(HERE, "golden", "trj", "trj.dcd"), os.path.join(d, "ref.dcd"))
    mine, ref = _qcalc_rows(d, "trj.dcd"), _qcalc_rows(d, "ref.dcd")
    print("qcalc rows", mine)
    assert len(mine) == len(ref) == 4 and all(len(v) == 6 for _, v in mine)
    for (fa, va), (fb, vb) in zip(mine, ref):
        assert fa == fb
        assert np.allclose(va, vb, atol=0.011), (va, vb)
