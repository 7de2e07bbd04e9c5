"""The C-ABI library on the CPU: it loads, exports every entry point that
include/qmi.h declares, and does host-side work (reading and setting up a
window) without touching a GPU.  The host set-up must reproduce the
reference's initial state bit for bit: Maxwell velocities from the
reference RNG (randm/gauss/maxwell, src/q6/md.f90:14408-14442, 1453-1468),
initial SHAKE and centre-of-mass removal (2506-2525, 14670-14702), checked
against the golden vectors of tests/golden/."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tests.golden.make_golden import file_sha, input_files

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_symbols():
    src = open(os.path.join(ROOT, "include", "qmi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qmi_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from q_amd import engine as EN
    lib = EN.load_library()
    syms = header_symbols()
    assert len(syms) >= 19
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(EN.EXPORTED) == syms
    assert lib.qmi_abi_version() == 1


def test_errors_are_reported_not_fatal(tmp_path):
    from q_amd import engine as EN
    lib = EN.load_library()
    h = C.c_void_p()
    assert lib.qmi_engine_create(-1, C.byref(h)) < 0
    assert b"device" in lib.qmi_last_error()
    e = EN.Engine(0)
    with pytest.raises(EN.QmiError, match="no windows"):
        e.prepare()
    with pytest.raises(EN.QmiError):
        e.add_window(str(tmp_path / "missing.inp"))
    bad = tmp_path / "pbc.inp"
    bad.write_text("[MD]\nsteps 10\nstepsize 1\ntemperature 300\n[PBC]\nbox_centre 0 0 0\n")
    with pytest.raises(EN.QmiError):
        e.add_window(str(bad))
    with pytest.raises(EN.QmiError, match="not prepared"):
        e.evaluate()
    e.close()


@pytest.mark.parametrize("cfg,seed", [("water", 0), ("ligwat", 0), ("protein", 2), ("kat", 0), ("ligfepx", 0),
                                      ("liglibq", 0), ("ligarithlibq", 0), ("ligrestr", 0),
                                      ("ligdual", 0), ("ligqcut", 0)])
def test_host_setup_matches_reference_start(golden_inp, cfg, seed):
    from q_amd.engine import Engine
    g = dict(np.load(os.path.join(GOLD, f"{cfg}_s{seed}.npz")))
    inp = golden_inp(cfg, seed)
    assert file_sha(input_files(cfg, seed, os.path.dirname(inp))) == str(g["input_sha"])
    e = Engine(0)
    w = e.add_window(inp)
    info = e.info(w)
    assert [info[k] for k in ("natom", "nat_solute", "nwat", "nstates", "nqat")] == g["sizes"].tolist()
    assert np.array_equal(e.coords(w), g["X0"])
    assert np.array_equal(e.velocities(w), g["V0"])
    e.close()
