"""ctypes binding of libqmi.so (include/qmi.h).

This mirrors the process interface of the reference `qdyn <input.inp>`
(src/q6/md.f90:2568-2577): a window is one qdyn input file; the library reads
the files named in it and writes the .en / .re / log outputs at the
intervals it asks for.  Errors raise :class:`QmiError` with the library's
message (the reference stops with `die`, md.f90:400-455).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqmi.so")

N_ENERGIES = 29
N_Q_ENERGIES = 15
LIST_PP, LIST_PW, LIST_WW, LIST_QP, LIST_QW = range(5)

# field names of qmi_get_energies (type energies, nrgy.f90:43-49)
ENERGY_FIELDS = [
    "potential", "kinetic", "LRF",
    "p.bond", "p.angle", "p.torsion", "p.improper",
    "w.bond", "w.angle", "w.torsion", "w.improper",
    "q.bond", "q.angle", "q.torsion", "q.improper",
    "pp.el", "pp.vdw", "pw.el", "pw.vdw", "ww.el", "ww.vdw", "qx.el", "qx.vdw",
    "restraint.total", "restraint.fix", "restraint.shell", "restraint.protein",
    "restraint.solvent_radial", "restraint.water_pol",
]
# per-state fields (type q_energies, nrgy.f90:51-58) -- one .en record
Q_FIELDS = ["lambda", "total", "q.bond", "q.angle", "q.torsion", "q.improper", "qx.el", "qx.vdw",
            "qq.el", "qq.vdw", "qp.el", "qp.vdw", "qw.el", "qw.vdw", "restraint"]


class QmiError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libqmi.so; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise QmiError(f"{path} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    p = C.c_void_p
    i32, i64, dp = C.c_int, C.c_longlong, C.POINTER(C.c_double)
    sig = {
        "qmi_abi_version": (i32, []),
        "qmi_engine_create": (i32, [i32, C.POINTER(p)]),
        "qmi_window_add": (i32, [p, C.c_char_p, i64, C.POINTER(i32)]),
        "qmi_prepare": (i32, [p]),
        "qmi_window_set_log": (i32, [p, i32, C.c_char_p]),
        "qmi_run": (i32, [p, i64, i32]),
        "qmi_evaluate": (i32, [p, i32]),
        "qmi_get_energies": (i32, [p, i32, dp, i32]),
        "qmi_get_q_energies": (i32, [p, i32, dp, i32]),
        "qmi_get_coords": (i32, [p, i32, dp, i64]),
        "qmi_get_velocities": (i32, [p, i32, dp, i64]),
        "qmi_get_forces": (i32, [p, i32, dp, i64]),
        "qmi_set_state": (i32, [p, i32, dp, dp, i64]),
        "qmi_get_pairlist": (i32, [p, i32, i32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int16), i64, C.POINTER(i64)]),
        "qmi_window_info": (i32, [p, i32, C.POINTER(C.c_int32), i32]),
        "qmi_write_restart": (i32, [p, i32, C.c_char_p]),
        "qmi_timings": (i32, [p, dp, i32]),
        "qmi_set_timing": (i32, [p, i32]),
        "qmi_time_kernel": (i32, [p, i32, i32, dp]),
        "qmi_set_autotune": (i32, [p, i32]),
        "qmi_tuning": (i32, [p, dp, i32]),
        "qmi_last_error": (C.c_char_p, []),
        "qmi_engine_destroy": (None, [p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


EXPORTED = [
    "qmi_abi_version", "qmi_engine_create", "qmi_window_add", "qmi_window_set_log", "qmi_prepare", "qmi_run",
    "qmi_evaluate",
    "qmi_get_energies", "qmi_get_q_energies", "qmi_get_coords", "qmi_get_velocities", "qmi_get_forces",
    "qmi_set_state", "qmi_get_pairlist", "qmi_window_info", "qmi_write_restart", "qmi_timings",
    "qmi_set_timing", "qmi_time_kernel", "qmi_set_autotune", "qmi_tuning", "qmi_last_error", "qmi_engine_destroy",
    "qmi_mlp_create", "qmi_mlp_num_params", "qmi_mlp_get_params", "qmi_mlp_set_params", "qmi_mlp_forward",
    "qmi_mlp_grad", "qmi_mlp_train", "qmi_mlp_last_error", "qmi_mlp_destroy", "qmi_mlp_forward_noalloc",
]


def _check(rc: int):
    if rc is not None and rc < 0:
        raise QmiError(_lib.qmi_last_error().decode() or f"qmi error {rc}")
    return rc


class Engine:
    """One engine = one GPU stream holding any number of windows."""

    def __init__(self, device: int = 0):
        lib = load_library()
        self._h = C.c_void_p()
        _check(lib.qmi_engine_create(device, C.byref(self._h)))
        self.lib = lib
        self.windows = []

    def close(self):
        if self._h:
            self.lib.qmi_engine_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_window(self, inp_path: str, seed: int = -1) -> int:
        wid = C.c_int()
        _check(self.lib.qmi_window_add(self._h, os.fspath(inp_path).encode(), seed, C.byref(wid)))
        self.windows.append(inp_path)
        return wid.value

    def set_log(self, w: int, path):
        """Log destination of window w: a path, "-" for stdout, None for none."""
        _check(self.lib.qmi_window_set_log(self._h, w, None if path is None else os.fspath(path).encode()))

    def prepare(self):
        _check(self.lib.qmi_prepare(self._h))

    def run(self, nsteps: int, finalize: bool = True):
        _check(self.lib.qmi_run(self._h, nsteps, 1 if finalize else 0))

    def evaluate(self, rebuild_lists: bool = True):
        _check(self.lib.qmi_evaluate(self._h, 1 if rebuild_lists else 0))

    def info(self, w: int) -> dict:
        a = (C.c_int32 * 6)()
        _check(self.lib.qmi_window_info(self._h, w, a, 6))
        return dict(natom=a[0], nat_solute=a[1], nwat=a[2], nstates=a[3], nqat=a[4], status=a[5])

    def energies(self, w: int) -> dict:
        a = np.zeros(N_ENERGIES)
        _check(self.lib.qmi_get_energies(self._h, w, a.ctypes.data_as(C.POINTER(C.c_double)), N_ENERGIES))
        return dict(zip(ENERGY_FIELDS, a.tolist()))

    def q_energies(self, w: int) -> np.ndarray:
        s = self.info(w)["nstates"]
        a = np.zeros(max(1, s) * N_Q_ENERGIES)
        _check(self.lib.qmi_get_q_energies(self._h, w, a.ctypes.data_as(C.POINTER(C.c_double)), a.size))
        return a[: s * N_Q_ENERGIES].reshape(s, N_Q_ENERGIES)

    def _vec(self, fn, w: int) -> np.ndarray:
        n = 3 * self.info(w)["natom"]
        a = np.zeros(n)
        _check(fn(self._h, w, a.ctypes.data_as(C.POINTER(C.c_double)), n))
        return a

    def coords(self, w: int) -> np.ndarray:
        return self._vec(self.lib.qmi_get_coords, w)

    def velocities(self, w: int) -> np.ndarray:
        return self._vec(self.lib.qmi_get_velocities, w)

    def forces(self, w: int) -> np.ndarray:
        """Gradient d(:) (kcal/mol/A), as md.f90 accumulates it."""
        return self._vec(self.lib.qmi_get_forces, w)

    def set_state(self, w: int, x=None, v=None):
        n = 3 * self.info(w)["natom"]
        xp = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        vp = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        _check(self.lib.qmi_set_state(self._h, w,
                                      None if xp is None else xp.ctypes.data_as(C.POINTER(C.c_double)),
                                      None if vp is None else vp.ctypes.data_as(C.POINTER(C.c_double)), n))

    def pairlist(self, w: int, kind: int):
        n = C.c_longlong()
        _check(self.lib.qmi_get_pairlist(self._h, w, kind, None, None, None, 0, C.byref(n)))
        cap = n.value
        i = np.zeros(max(cap, 1), np.int32)
        j = np.zeros(max(cap, 1), np.int32)
        l = np.zeros(max(cap, 1), np.int16)
        _check(self.lib.qmi_get_pairlist(self._h, w, kind, i.ctypes.data_as(C.POINTER(C.c_int32)),
                                         j.ctypes.data_as(C.POINTER(C.c_int32)),
                                         l.ctypes.data_as(C.POINTER(C.c_int16)), cap, C.byref(n)))
        return i[:cap], j[:cap], l[:cap]

    def write_restart(self, w: int, path: str):
        _check(self.lib.qmi_write_restart(self._h, w, os.fspath(path).encode()))

    def set_autotune(self, on: bool):
        _check(self.lib.qmi_set_autotune(self._h, 1 if on else 0))

    def tuning(self) -> dict:
        """launch shapes chosen by the autotuner and their measured times (us)"""
        a = np.zeros(13)
        _check(self.lib.qmi_tuning(self._h, a.ctypes.data_as(C.POINTER(C.c_double)), 13))
        return dict(q_tiles_per_wave=int(a[0]), rows_block=int(a[1]), lrf_block=int(a[2]), bonded_block=int(a[3]),
                    xcd_grouped=int(a[4]), bonded_mode=int(a[5]), lrf_source_parts=int(a[6]),
                    us=dict(q_tiles=a[7], rows=a[8], lrf_build=a[9], bonded=a[10], xcd=a[11], bonded_mode=a[12]))

    def set_timing(self, on: bool):
        _check(self.lib.qmi_set_timing(self._h, 1 if on else 0))

    def time_kernel(self, which: int = 0, reps: int = 50) -> float:
        """device microseconds per launch of a step kernel (0: water-water)"""
        us = C.c_double()
        _check(self.lib.qmi_time_kernel(self._h, which, reps, C.byref(us)))
        return us.value

    def timings(self) -> list:
        a = np.zeros(8)
        n = _check(self.lib.qmi_timings(self._h, a.ctypes.data_as(C.POINTER(C.c_double)), 8))
        return a[:n].tolist()
