"""Reader and runner for the reference golden-vector harness (oracle/qprobe.f90).

Test infrastructure only: runs ``oracle/_ref/qprobe`` (the reference Q6
modules compiled unmodified plus this project's driver) on a qdyn input and
parses its full-precision dump: step-0 energies (type energies/q_energies,
src/q6/nrgy.f90:43-58), gradient d(:), coordinates and velocities after the
reference's Maxwell start (md.f90:3787, 2506, 14670), its pair lists
(md.f90:5175-8556) and the final state after md_run (md.f90:3917-4263).
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
QPROBE = os.path.join(REF_DIR, "qprobe")
QDYN = os.path.join(REF_DIR, "qdyn")
QFEP = os.path.join(REF_DIR, "qfep")

E_FIELDS = {
    "pot_kin_lrf": ["potential", "kinetic", "LRF"],
    "p": ["p.bond", "p.angle", "p.torsion", "p.improper"],
    "w": ["w.bond", "w.angle", "w.torsion", "w.improper"],
    "q": ["q.bond", "q.angle", "q.torsion", "q.improper"],
    "pp": ["pp.el", "pp.vdw"], "pw": ["pw.el", "pw.vdw"], "ww": ["ww.el", "ww.vdw"], "qx": ["qx.el", "qx.vdw"],
    "rst": ["restraint.total", "restraint.fix", "restraint.shell", "restraint.protein",
            "restraint.solvent_radial", "restraint.water_pol"],
}


def available() -> bool:
    return os.path.exists(QPROBE)


def run_qprobe(inp_path: str, out_name: str = "qprobe.dump", timeout: float = 600) -> dict:
    """Run the reference harness in the directory of ``inp_path``."""
    d = os.path.dirname(os.path.abspath(inp_path))
    env = dict(os.environ)
    env["QPROBE_OUT"] = out_name
    with open(os.path.join(d, out_name + ".log"), "w") as log:
        subprocess.run([QPROBE, os.path.basename(inp_path)], cwd=d, env=env, stdout=log, stderr=subprocess.STDOUT,
                       check=True, timeout=timeout)
    return read_dump(os.path.join(d, out_name))


def read_dump(path: str) -> dict:
    with open(path) as f:
        lines = f.read().split("\n")
    out: dict = {}
    k = 0

    def vec(n):
        nonlocal k
        rows = (n + 2) // 3
        a = np.array(" ".join(lines[k:k + rows]).split(), dtype=float)
        k += rows
        return a

    def energies():
        nonlocal k
        e = {}
        eq = []
        while k < len(lines):
            t = lines[k].split()
            if not t:
                k += 1
                continue
            if t[0] in E_FIELDS:
                for name, v in zip(E_FIELDS[t[0]], t[1:]):
                    e[name] = float(v)
                k += 1
            elif t[0] == "EQ":
                eq.append([float(v) for v in t[2:]])
                k += 1
            else:
                break
        return e, np.array(eq)

    while k < len(lines):
        t = lines[k].split()
        k += 1
        if not t or t[0].startswith("#"):
            continue
        tag = t[0]
        if tag == "SIZES":
            v = list(map(int, t[1:]))
            out["sizes"] = dict(zip(["natom", "nat_solute", "nwat", "nstates", "nqat", "ncgp", "ncgp_solute",
                                     "shake_constraints", "shake_molecules"], v))
        elif tag == "NDEGF":
            out["ndegf"] = list(map(int, t[1:]))
        elif tag == "DT":
            out["dt"], out["tau_t"], out["temp0"] = map(float, t[1:4])
        elif tag in ("X0", "V0", "D0", "XN", "VN", "DN"):
            out[tag] = vec(int(t[1]))
        elif tag == "TEMP0":
            out["temp0_state"] = list(map(float, t[1:]))
        elif tag in ("E0", "EN"):
            out[tag], out[tag.replace("E", "EQ")] = energies()
        elif tag in ("NBPP", "NBPW", "NBQP"):
            n = int(t[1])
            a = np.array(" ".join(lines[k:k + n]).split(), dtype=np.int64).reshape(n, -1) if n else np.zeros((0, 3), np.int64)
            k += n
            out[tag] = a
        elif tag == "NBWW":
            n = int(t[1])
            a = np.array(" ".join(lines[k:k + n]).split(), dtype=np.int64).reshape(n, 2) if n else np.zeros((0, 2), np.int64)
            k += n
            out[tag] = a
        elif tag == "NBQW":
            n = int(t[1])
            a = np.array(" ".join(lines[k:k + n]).split(), dtype=np.int64) if n else np.zeros(0, np.int64)
            k += n
            out[tag] = a
        elif tag == "NBQQ":
            s, n = int(t[1]), int(t[2])
            rows = [lines[k + r].split() for r in range(n)]
            k += n
            out.setdefault("NBQQ", {})[s] = rows
        elif tag == "LRF":
            n = int(t[1])
            a = np.array(" ".join(lines[k:k + n]).split(), dtype=float).reshape(n, 44)
            k += n
            out["LRF"] = a
        elif tag == "END":
            break
    return out


def unordered(i, j) -> set:
    return {(min(a, b), max(a, b)) for a, b in zip(np.asarray(i).tolist(), np.asarray(j).tolist())}
