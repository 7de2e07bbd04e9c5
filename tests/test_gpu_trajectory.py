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


def test_dcd_matches_reference(tmp_path):
    import make_golden as M
    from q_amd.engine import Engine
    d = str(tmp_path)
    inp = M.trj_window(d)
    sha = M.file_sha([inp, os.path.join(d, "ligwat_s0.top"), os.path.join(d, "ligwat_s0.fep")])
    assert sha == open(os.path.join(HERE, "golden", "trj", "input_sha.txt")).read().strip()
    e = Engine(0)
    e.add_window(inp)
    e.set_log(0, None)
    e.prepare()
    e.run(M.TRJ_STEPS, finalize=True)
    e.close()
    got = records(os.path.join(d, "trj.dcd"))
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
