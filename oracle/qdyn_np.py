"""CPU restatement (numpy, FP64) of the reference qdyn step -- TEST INFRASTRUCTURE.

Only tests, the smoke check and the bench's CPU leg may use this module, and
only as a checker: the product path never imports it.  It re-derives, with
its own readers, what the reference computes at one step of md_run for a
spherical-boundary, SPC-water window with the geometric or the arithmetic
LJ combination rule:

  reading     initialize (src/q6/md.f90:2528-3262), topo_read (topo.f90:529-1082),
              qatom_load_atoms / qatom_load_fep (qatom.f90:253-548, 874-1745)
  set-up      topology (md.f90:14705-14790), prep_coord/make_shell (13596-13721),
              get_fep (1472-1652), prep_sim (13787-13892), wat_sphere/wat_shells
              (14961-15118), make_qconn/nbqqlist (7312-7530)
  lists       nbpplist_lrf (5448-5700), nbpwlist_lrf (6864-7068),
              nbwwlist_lrf (8370-8556), nbqplist (7875-7964), nbqwlist (8066-8118)
  forces      nonbond_pp (10387-10491), nonbond_pw (10631-10682), nonbond_ww_spc
              (11823-12047), nonbond_qp_qvdw / nonbond_qp (11138-11204, 10980-11045),
              nonbond_qw_spc (11302-11411), nonbond_qq (10763-10868); arithmetic
              rule: nonbon2_pp/pw/qp/qw/qq/ww (8972-9089, 9242-9297, 9603-9689,
              9798-9936, 9385-9491, 10099-10227); lrf_taylor
              (3646-3712), bond/angle/torsion/improper2 (954-997, 814-901,
              14793-14907, 1789-1897), qbond/qangle/qtorsion/qimproper
              (14060-14114, 13896-13988, 14262-14403, 14117-14258), fix_shell
              (1414-1450), restrain_solvent (14910-14957), watpol (15122-15288),
              pot_energy totals (13278-13473)

The starting coordinates are taken as given (the restart, or the reference's
Maxwell start supplied by the caller); dynamics is pinned against the
reference binary directly.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

F32 = np.float32
PI = float(F32(4.0) * F32(math.atan(F32(1.0))))       # md.f90:383 pi = 4.0*atan(1.0) in real(4)
DEG2RAD = PI / float(F32(180.0))
BOLTZ = float(F32(0.001986))
RHO_WAT = float(F32(0.0335))


# ------------------------------------------------------------------ files --
def read_prm(path):
    """Sections of a qdyn-style file (prmfile.f90:873-1000)."""
    secs, cur = {}, None
    for raw in open(path):
        line = raw.rstrip("\n")
        st = line.lstrip(" \t")
        if not st or st[0] in "!#*":
            continue
        if st[0] == "[":
            cur = st[1:st.index("]")].upper()
            secs.setdefault(cur, [])
            continue
        cut = len(st)
        for k, ch in enumerate(st):
            if ch == "!" or (ch in "#*" and k > 0 and st[k - 1] in " \t"):
                cut = k
                break
        secs[cur].append(st[:cut].strip())
    return secs


def _kv(sec, key, conv=float, default=None):
    for l in sec or []:
        t = l.split()
        if len(t) >= 2 and t[0] == key:
            try:
                return conv(t[1])
            except ValueError:
                continue
    return default


def _onoff(sec, key, default):
    v = _kv(sec, key, str)
    if v is None:
        return default
    return v in ("on", "ON", "1")


def read_inp(path):
    s = read_prm(path)
    md, co, sp, so, iv, fi = (s.get(k) for k in ("MD", "CUT-OFFS", "SPHERE", "SOLVENT", "INTERVALS", "FILES"))
    p = dict(dir=os.path.dirname(os.path.abspath(path)))
    p["nsteps"] = _kv(md, "steps", int)
    p["dt"] = float(F32(0.020462)) * _kv(md, "stepsize")
    p["temp0"] = _kv(md, "temperature")
    p["tau_t"] = float(F32(0.020462)) * _kv(md, "bath_coupling", float, 10.0)
    p["separate_scaling"] = _onoff(md, "separate_scaling", False)
    p["lrf"] = _onoff(md, "lrf", False)
    p["rcpp"] = _kv(co, "solute_solute", float, 10.0)
    p["rcww"] = _kv(co, "solvent_solvent", float, 10.0)
    p["rcpw"] = _kv(co, "solute_solvent", float, 10.0)
    p["rcq"] = _kv(co, "q_atom", float, 99.0)
    p["rclrf"] = _kv(co, "lrf", float, 99.0)
    p["rexcl_i"] = _kv(sp, "shell_radius", float, 0.85)
    p["fk_pshell"] = _kv(sp, "shell_force", float, 10.0)
    p["fk_wsphere"] = _kv(so, "radial_force", float, -1.0)
    p["fkwpol"] = _kv(so, "polarization_force", float, -1.0)
    p["wpol"] = _onoff(so, "polarization", True)
    p["wpol_born"] = _onoff(so, "charge_correction", p["wpol"])
    p["rwat_in"] = _kv(so, "radius", float, 0.0)
    p["shake_solvent"] = _onoff(md, "shake_solvent", True)
    p["shake_solute"] = _onoff(md, "shake_solute", False)
    p["shake_hydrogens"] = _onoff(md, "shake_hydrogens", False)
    p["top"] = _kv(fi, "topology", str)
    p["fep"] = _kv(fi, "fep", str)
    lam = []
    for l in s.get("LAMBDAS", []):
        lam += [float(v) for v in l.split()]
    if not lam and p["fep"]:
        lam = [1.0]          # one state at lambda 1 without [lambdas] (md.f90:3093-3098)
    p["lambda"] = np.array(lam)
    # restraint sections (md.f90:3119-3259); atoms 0-based
    p["rst_seq"] = []
    for l in s.get("SEQUENCE_RESTRAINTS", []):
        v = (l.split() + ["0", "0", "0", "0", "0"])[:5]
        p["rst_seq"].append((int(v[0]) - 1, int(v[1]) - 1, float(v[2]), int(v[3]), int(v[4])))
    p["rst_pos"] = [(int(v[0]) - 1, [float(t) for t in v[1:4]], [float(t) for t in v[4:7]], int(v[7]))
                    for v in (l.split() for l in s.get("ATOM_RESTRAINTS", []))]
    p["rst_dis"] = [(int(v[0]) - 1, int(v[1]) - 1, float(v[2]), float(v[3]), float(v[4]), int(v[5]))
                    for v in (l.split() for l in s.get("DISTANCE_RESTRAINTS", []))]
    p["rst_ang"] = [(int(v[0]) - 1, int(v[1]) - 1, int(v[2]) - 1, float(v[3]), float(v[4]), int(v[5]))
                    for v in (l.split() for l in s.get("ANGLE_RESTRAINTS", []))]
    p["rst_wal"] = [(int(v[0]) - 1, int(v[1]) - 1, float(v[2]), float(v[3]), float(v[4]), float(v[5]), int(v[6]))
                    for v in (l.split() for l in s.get("WALL_RESTRAINTS", []))]
    return p


def _nums(tokens):
    out = []
    for t in tokens:
        try:
            out.append(float(t))
        except ValueError:
            break
    return out


class _Lines:
    def __init__(self, path):
        self.l = open(path).read().split("\n")
        self.k = 0

    def line(self):
        self.k += 1
        return self.l[self.k - 1]

    def values(self, n):
        out = []
        while len(out) < n:
            for t in self.l[self.k].replace(",", " ").split():
                if len(out) < n:
                    out.append(float(t))
            self.k += 1
        return np.array(out)


def read_top(path):
    """Q topology (topo.f90:529-1082); indices returned 0-based."""
    L = _Lines(path)
    t = {}
    if L.line().strip() == "Q topology file":
        while True:
            toks = L.line().split()
            if toks and toks[0].upper() == "END":
                break
            if toks and toks[0].upper() == "FF_TYPE":
                t["ff_type"] = int(toks[1])
    h = _nums(L.line().split())
    n = int(h[0])
    t["nat"] = n
    t["nsol"] = int(h[1]) if len(h) > 1 else n
    t["xtop"] = L.values(3 * n).reshape(n, 3)
    L.line()
    t["iac"] = L.values(n).astype(int)
    h = _nums(L.line().split())
    nb = int(h[0])
    t["nb_sol"] = int(h[1]) if len(h) > 1 else nb
    t["bnd"] = L.values(3 * nb).astype(int).reshape(nb, 3) - [1, 1, 0]
    nbc = int(L.values(1)[0])
    t["bondlib"] = {}
    for _ in range(nbc):
        v = _nums(L.line().split())
        t["bondlib"][int(v[0])] = (v[1], v[2])
    h = _nums(L.line().split())
    na = int(h[0])
    t["na_sol"] = int(h[1]) if len(h) > 1 else na
    t["ang"] = L.values(4 * na).astype(int).reshape(na, 4) - [1, 1, 1, 0]
    nac = int(L.values(1)[0])
    t["anglib"] = {}
    for _ in range(nac):
        v = _nums(L.line().split())
        t["anglib"][int(v[0])] = (v[1], v[2])
    h = _nums(L.line().split())
    nt = int(h[0])
    t["tor"] = L.values(5 * nt).astype(int).reshape(nt, 5) - [1, 1, 1, 1, 0]
    ntc = int(L.values(1)[0])
    t["torlib"] = {}
    for _ in range(ntc):
        v = L.values(5)
        t["torlib"][int(v[0])] = tuple(F32(x) for x in v[1:5])
    h = _nums(L.line().split())
    ni = int(h[0])
    t["imp"] = L.values(5 * ni).astype(int).reshape(ni, 5) - [1, 1, 1, 1, 0]
    h = _nums(L.line().split())
    nic = int(h[0])
    t["implib"] = {}
    for _ in range(nic):
        v = L.values(3)
        t["implib"][int(v[0])] = (F32(v[1]), F32(v[2]))
    L.line()
    t["crg"] = L.values(n).astype(F32)
    h = _nums(L.line().split())
    ncgp = int(h[0])
    t["ncgp_sol"] = int(h[1]) if len(h) > 1 else ncgp
    cg = []
    for _ in range(ncgp):
        hh = L.values(2).astype(int)
        cg.append((int(hh[1]) - 1, (L.values(int(hh[0])).astype(int) - 1).tolist()))
    t["cgp"] = cg
    nty = int(L.values(1)[0])
    t["vdw_rule"] = int(L.values(1)[0])
    v = _nums(L.line().split())
    t["el14"], t["kc"] = v[0], (v[1] if len(v) > 1 and v[1] > 0 else 332.0)
    L.line()
    t["mass"] = L.values(nty)
    A, B = np.zeros((nty, 3)), np.zeros((nty, 3))
    for j in range(3):
        L.line()
        A[:, j] = L.values(nty)
        L.line()
        B[:, j] = L.values(nty)
    if t["vdw_rule"] not in (1, 2):
        raise ValueError("unknown vdW combination rule")
    if t["vdw_rule"] == 2:
        # epsilons enter as their square roots (topology, md.f90:14747-14752)
        B = np.sqrt(np.abs(B))
    t["avdw"], t["bvdw"] = A, B
    nlj2 = int(L.values(1)[0])
    if nlj2:
        raise ValueError("type-2 LJ pairs are not restated")

    def bitmap():
        need = 25 * t["nsol"]
        bits = []
        while len(bits) < need:
            line = L.line()
            bits += [c == "1" for c in line[:80].ljust(80)][: need - len(bits)]
        return np.array(bits, bool).reshape(t["nsol"], 25) if need else np.zeros((0, 25), bool)

    L.values(1)
    t["list14"] = bitmap()
    n14l = int(L.values(1)[0])
    t["l14long"] = [tuple(L.values(2).astype(int) - 1) for _ in range(n14l)]
    L.values(1)
    t["listex"] = bitmap()
    nexl = int(L.values(1)[0])
    t["exlong"] = [tuple(L.values(2).astype(int) - 1) for _ in range(nexl)]
    line = L.line()
    if not line.strip():
        line = L.line()
    nres = int(_nums(line.split())[0])
    L.values(nres)
    L.line()
    L.k += (nres + 15) // 16
    nmol = int(L.values(1)[0])
    t["istart_mol"] = (L.values(nmol).astype(int) - 1).tolist() + [n]
    L.line()
    L.k += (nty + 7) // 8
    L.line()
    L.k += (nty + 12) // 13
    t["solvent_type"] = int(L.values(1)[0])
    if t["solvent_type"] not in (0, 1):
        raise ValueError("only SPC (0) and 3-atom (1) solvent are restated")
    v = _nums(L.line().split())
    t["rexcl_o"], t["rwat"] = v[0], v[1]
    t["xpcent"] = L.values(3)
    t["xwcent"] = L.values(3)
    L.values(2)
    fl = []
    while len(fl) < n:
        fl += [c in "Tt" for c in L.line()[:80]]
    t["excl"] = np.array(fl[:n], bool)
    t["heavy"] = t["mass"][t["iac"] - 1] >= 4.0
    t.setdefault("ff_type", 1)
    return t


def lj_ab(ai, bi, aj, bj, rule):
    """A, B of a pair's V = A/r^12 - B/r^6 from the two atoms' LJ columns:
    products under the geometric rule (nonbond_pp, md.f90:10432-10439);
    under the arithmetic rule R* = a_i + a_j and eps = b_i b_j (b being the
    square root of epsilon), A = eps R*^12, B = 2 eps R*^6 (nonbon2_pp,
    md.f90:8993-9000, 9056-9058)."""
    if rule == 1:
        return ai * aj, bi * bj
    s = ai + aj
    s6 = (s * s) ** 3
    e = bi * bj
    return e * s6 * s6, 2.0 * e * s6


def read_fep(path, nq_states, nat_solute, natyps, vdw_rule=1):
    s = read_prm(path)
    S = nq_states
    f = dict(S=S)
    f["iqseq"] = [int(l.split()[1]) - 1 for l in s["ATOMS"]]
    nq = len(f["iqseq"])
    f["nq"] = nq
    f["qcrg"] = None
    f["cc"] = [[float(v) for v in l.split()[1:1 + S]] for l in s.get("CHANGE_CHARGES", [])]
    names, qa, qb = [], [], []
    for l in s.get("ATOM_TYPES", []):
        t = l.split()
        names.append(t[0])
        qa.append([float(t[1]), float(t[3]), float(t[5])])
        qb.append([float(t[2]), float(t[4]), float(t[6])])
    f["qavdw"], f["qbvdw"] = np.array(qa), np.array(qb)
    ca = s.get("CHANGE_ATOMS", [])
    f["qvdw"] = len(ca) > 0
    f["qiac"] = np.array([[names.index(x) for x in l.split()[1:1 + S]] for l in ca]) if ca else None
    alpha = np.zeros((nq, S))
    for l in s.get("SOFTCORE", []):
        t = l.split()
        alpha[int(t[0]) - 1] = [float(v) for v in t[1:1 + S]]
    f["alpha"] = alpha
    fep = s.get("FEP", [])
    if _onoff(fep, "softcore_use_max_potential", False):
        raise ValueError("softcore_use_max_potential is not restated")
    # q-q pairs with topology charges (qatom.f90:297; md.f90:13387-13397)
    f["libq"] = _onoff(fep, "qq_use_library_charges", False)
    # soft pairs, electrostatic scaling and special exclusions among q-atoms
    # (qatom.f90:1049-1135); q-atom numbers and topology atoms 0-based
    f["soft"] = {tuple(sorted((int(v[0]) - 1, int(v[1]) - 1))) for v in (l.split() for l in s.get("SOFT_PAIRS", []))}
    f["el_scale"] = {}
    for l in s.get("EL_SCALE", []):
        v = l.split()
        f["el_scale"][tuple(sorted((int(v[0]) - 1, int(v[1]) - 1)))] = [float(t) for t in v[2:2 + S]]
    f["exspec"] = [(int(v[0]) - 1, int(v[1]) - 1, [int(t) == 1 for t in v[2:2 + S]])
                   for v in (l.split() for l in s.get("EXCLUDED_PAIRS", []))]
    # extra SHAKE constraints (qatom.f90:1498-1512), topology atoms 0-based
    f["shake"] = [(int(v[0]) - 1, int(v[1]) - 1) for v in (l.split() for l in s.get("SHAKE_CONSTRAINTS", []))]
    if vdw_rule == 2 and f["qvdw"]:
        # square roots of the q-type epsilons, normal and 1-4 columns only
        # (get_fep, md.f90:1518-1523)
        f["qbvdw"][:, 0] = np.sqrt(f["qbvdw"][:, 0])
        f["qbvdw"][:, 2] = np.sqrt(f["qbvdw"][:, 2])
    for sec in ("OFF_DIAGONALS", "COUPLINGS"):
        if s.get(sec):
            raise ValueError(f"[{sec.lower()}] is not restated")

    def lib(sec, ncol):
        out = {}
        for l in s.get(sec, []):
            v = [float(x) for x in l.split()]
            out[int(v[0])] = v[1:1 + ncol]
        return out

    def changes(sec, nat):
        return [[int(float(v)) for v in l.split()[:nat + S]] for l in s.get(sec, [])]

    qbl = {}
    for l in s.get("BOND_TYPES", []):
        v = [float(x) for x in l.split()]
        qbl[int(v[0])] = (v[1], v[2], v[3], 0.0) if len(v) >= 4 else (0.0, 0.0, v[2], v[1])
    f["qbondlib"] = qbl
    f["qbnd"] = changes("CHANGE_BONDS", 2)
    f["qanglib"] = {k: (v[0], DEG2RAD * v[1]) for k, v in lib("ANGLE_TYPES", 2).items()}
    f["qang"] = changes("CHANGE_ANGLES", 3)
    f["qtorlib"] = {k: (v[0], v[1], DEG2RAD * v[2]) for k, v in lib("TORSION_TYPES", 3).items()}
    f["qtor"] = changes("CHANGE_TORSIONS", 4)
    f["qimplib"] = {k: (v[0], DEG2RAD * v[1]) for k, v in lib("IMPROPER_TYPES", 2).items()}
    f["qimp"] = changes("CHANGE_IMPROPERS", 4)
    return f


# ------------------------------------------------------------------- LRF --
_T3_IDX = [(0, 0, 0), (0, 0, 1), (0, 0, 2), (0, 1, 1), (0, 1, 2), (0, 2, 2), (1, 1, 1), (1, 1, 2), (1, 2, 2),
           (2, 2, 2)]


def lrf_expand(dr, q):
    """The 20 unique Taylor components (phi0, phi1[3], phi2[6], phi3[10]) that
    source charges q at offsets dr (n,3) from a group centre add to that
    centre's expansion (nbpplist_lrf accumulates the same sums at
    md.f90:5589-5700, storing phi2/phi3 as full 3x3 / 3x3x3)."""
    r2 = (dr ** 2).sum(1)
    f0 = q / (r2 * np.sqrt(r2))
    f1 = 3. * f0 / r2
    f2 = -f1 / r2
    X, Y, Z = dr[:, 0], dr[:, 1], dr[:, 2]
    return np.array([
        (f0 * r2).sum(), -(f0 * X).sum(), -(f0 * Y).sum(), -(f0 * Z).sum(),
        (f1 * X * X - f0).sum(), (f1 * X * Y).sum(), (f1 * X * Z).sum(),
        (f1 * Y * Y - f0).sum(), (f1 * Y * Z).sum(), (f1 * Z * Z - f0).sum(),
        (f2 * (5 * X * X * X - r2 * 3 * X)).sum(), (f2 * (5 * X * X * Y - r2 * Y)).sum(),
        (f2 * (5 * X * X * Z - r2 * Z)).sum(), (f2 * (5 * X * Y * Y - r2 * X)).sum(),
        (f2 * (5 * X * Y * Z)).sum(), (f2 * (5 * X * Z * Z - r2 * X)).sum(),
        (f2 * (5 * Y * Y * Y - r2 * 3 * Y)).sum(), (f2 * (5 * Y * Y * Z - r2 * Z)).sum(),
        (f2 * (5 * Y * Z * Z - r2 * Y)).sum(), (f2 * (5 * Z * Z * Z - r2 * 3 * Z)).sum()])


def lrf_eval(P, r):
    """Potential and field of expansions P (..., 20) at r = centre - x (..., 3)
    (lrf_taylor, md.f90:3646-3712): Vij to second order, the field to third."""
    P = np.asarray(P, float)
    r = np.asarray(r, float)
    T2 = P[..., [4, 5, 6, 5, 7, 8, 6, 8, 9]].reshape(P.shape[:-1] + (3, 3))
    T3 = np.zeros(P.shape[:-1] + (3, 3, 3))
    for n, (a, b, c) in enumerate(_T3_IDX):
        for perm in {(a, b, c), (a, c, b), (b, a, c), (b, c, a), (c, a, b), (c, b, a)}:
            T3[(Ellipsis,) + perm] = P[..., 10 + n]
    T2r = np.einsum("...ij,...j->...i", T2, r)
    vij = P[..., 0] + (P[..., 1:4] * r).sum(-1) + 0.5 * (r * T2r).sum(-1)
    fld = P[..., 1:4] + T2r + 0.5 * np.einsum("...kij,...i,...j->...k", T3, r, r)
    return vij, fld


# ---------------------------------------------------------------- system --
class Window:
    """Everything pot_energy needs, derived as the reference derives it."""

    def __init__(self, inp_path, x=None):
        p = read_inp(inp_path)
        t = read_top(os.path.join(p["dir"], p["top"]))
        self.p, self.t = p, t
        n, ns = t["nat"], t["nsol"]
        self.n, self.ns, self.nwat = n, ns, (n - ns) // 3
        self.x = t["xtop"].copy() if x is None else np.asarray(x, float).reshape(n, 3).copy()
        # topology(): unit conversions (real(4) libraries stay real(4))
        self.anglib = {k: (v[0], DEG2RAD * v[1]) for k, v in t["anglib"].items()}
        self.torlib = {k: (v[0], v[1], F32(DEG2RAD * float(v[2])), F32(F32(1.0) / v[3])) for k, v in t["torlib"].items()}
        self.implib = {k: (v[0], F32(DEG2RAD * float(v[1]))) for k, v in t["implib"].items()}
        # prep_coord(): restrained shell from topology coordinates (make_shell)
        ri = p["rexcl_i"]
        if ri <= 1.0:
            ri = ri * t["rexcl_o"]
        self.shell = np.zeros(n, bool)
        for sw, atoms in t["cgp"][: t["ncgp_sol"]]:
            if t["excl"][sw] or not t["heavy"][sw]:
                continue
            if ((t["xtop"][sw] - t["xpcent"]) ** 2).sum() > ri * ri:
                self.shell[atoms] = True
        # get_fep()
        S = len(p["lambda"])
        self.S = S
        self.iqatom = np.zeros(n, int)
        self.bnd_cod = t["bnd"][:, 2].copy()
        self.ang_cod = t["ang"][:, 3].copy()
        self.tor_cod = t["tor"][:, 4].copy()
        self.imp_cod = t["imp"][:, 4].copy()
        if S:
            f = read_fep(os.path.join(p["dir"], p["fep"]), S, ns, len(t["mass"]), t["vdw_rule"])
            self.f = f
            for q, a in enumerate(f["iqseq"]):
                self.iqatom[a] = q + 1
            qcrg = np.array([[t["crg"][a]] * S for a in f["iqseq"]], F32)
            for k, row in enumerate(f["cc"]):
                qcrg[k] = row
            self.qcrg = qcrg.astype(F32)
            for qb in f["qbnd"]:
                m = ((t["bnd"][:, 0] == qb[0] - 1) & (t["bnd"][:, 1] == qb[1] - 1)) | \
                    ((t["bnd"][:, 0] == qb[1] - 1) & (t["bnd"][:, 1] == qb[0] - 1))
                self.bnd_cod[m] = 0
            for qa in f["qang"]:
                i, j, k = qa[0] - 1, qa[1] - 1, qa[2] - 1
                a = t["ang"]
                m = ((a[:, 0] == i) & (a[:, 1] == j) & (a[:, 2] == k)) | ((a[:, 0] == k) & (a[:, 1] == j) & (a[:, 2] == i))
                self.ang_cod[m] = 0
            for qt in f["qtor"]:
                i, j, k, l = (v - 1 for v in qt[:4])
                a = t["tor"]
                m = (((a[:, 0] == i) & (a[:, 1] == j) & (a[:, 2] == k) & (a[:, 3] == l)) |
                     ((a[:, 0] == l) & (a[:, 1] == k) & (a[:, 2] == j) & (a[:, 3] == i)))
                self.tor_cod[m] = 0
            for qi in f["qimp"]:
                j, k = qi[1] - 1, qi[2] - 1
                a = t["imp"]
                m = ((a[:, 1] == j) & (a[:, 2] == k)) | ((a[:, 1] == k) & (a[:, 2] == j))
                self.imp_cod[m] = 0
            # soft-core lookup with explicit alphas (qatom.f90:1559-1650)
            nty = len(t["mass"])
            sc = np.zeros((f["nq"], nty + f["nq"], S))
            al = f["alpha"]
            for s in range(S):
                sc[:, :nty, s] = al[:, s][:, None]
                for i in range(f["nq"]):
                    for j in range(f["nq"]):
                        if al[i, s] > float(F32(1e-6)) or al[j, s] > float(F32(1e-6)):
                            sc[i, nty + j, s] = max(al[i, s], al[j, s])
            self.sc = sc
        # prep_sim(): water charges, sphere and shells, scaled charges
        sq = math.sqrt(t["kc"])
        self.crg_ow = F32(t["crg"][ns]) if self.nwat else F32(0)
        self.crg_hw = F32(-self.crg_ow / F32(2.0))
        lam = p["lambda"]
        self.rwat = p["rwat_in"] if p["rwat_in"] > 0 else t["rwat"]
        crgq = 0.0
        if S:
            for q in range(self.f["nq"]):
                for s in range(S):
                    crgq += float(self.qcrg[q, s]) * lam[s]
        self.crgqtot = crgq if p["wpol_born"] else 0.0
        self.fk_ws = 60.0 if p["fk_wsphere"] == -1 else p["fk_wsphere"]
        self.fkwpol = 20.0 if p["fkwpol"] == -1 else p["fkwpol"]
        self.dwmz = float(F32(0.26)) * math.exp(-float(F32(0.19)) * (self.rwat - 15.0)) + float(F32(0.74))
        self.awmz = float(F32(0.2)) / (1.0 + math.exp(float(F32(0.4)) * (self.rwat - 25.0))) + float(F32(0.3))
        self.shells = []
        if p["wpol"] and self.nwat:
            b = t["bnd"][-1]
            a = t["ang"][-1]
            mu = F32(-float(self.crg_ow) * t["bondlib"][b[2]][1] * math.cos(self.anglib[a[3]][1] / 2))
            drs = float(F32(3.0001) / F32(0.5))
            nsh = int(-0.5 + math.sqrt(2 * drs + 0.25))
            rout = self.rwat
            for k in range(nsh):
                dr = float(F32(0.5) * F32(k + 1))
                rin = rout - dr
                rsh = (0.5 * (rout ** 3 + rin ** 3)) ** float(F32(1.0) / F32(3.0))
                # rho_wat*mu_w*4. multiplies two real(4) values (md.f90:15110)
                cstb = F32(self.crgqtot * float(F32(0.98750)) / (float(F32(RHO_WAT) * mu * F32(4.0)) * PI * rsh ** 2))
                self.shells.append(dict(rout=F32(rout), dr=F32(dr), cstb=cstb, theta_corr=F32(0.0)))
                rout = rout - dr
        self.crg = (t["crg"].astype(np.float64) * sq).astype(F32)
        self.crg_ow = F32(float(self.crg_ow) * sq)
        self.crg_hw = F32(float(self.crg_hw) * sq)
        # the three solvent charges: O, H, H of the SPC routines, or the
        # topology charges of the first solvent molecule for a 3-atom
        # solvent (nonbond_3atomsolvent, md.f90:12335-12340)
        self.wat3 = t["solvent_type"] == 1 and self.nwat > 0
        self.crg_w = ([self.crg[ns], self.crg[ns + 1], self.crg[ns + 2]] if self.wat3
                      else [self.crg_ow, self.crg_hw, self.crg_hw])
        if S:
            self.qcrg = (self.qcrg.astype(np.float64) * sq).astype(F32)
        self.cgp_of = np.zeros(n, int)
        for g, (_sw, atoms) in enumerate(t["cgp"]):
            self.cgp_of[atoms] = g
        # special pairs (exclusions first, then 1-4)
        self.excl_pairs, self.pairs14 = set(), set()
        for i in range(ns):
            for k in range(25):
                if i + k + 1 < ns:
                    if t["listex"][i, k]:
                        self.excl_pairs.add((i, i + k + 1))
                    if t["list14"][i, k]:
                        self.pairs14.add((i, i + k + 1))
        for a, b in t["exlong"]:
            self.excl_pairs.add((min(a, b), max(a, b)))
        for a, b in t["l14long"]:
            self.pairs14.add((min(a, b), max(a, b)))
        # init_shake (md.f90:2326-2503): constrained bonds leave the bond
        # terms; for each constraint the first angle whose end atoms are the
        # constrained pair leaves the angle terms
        self.shake_pairs = []
        for k, (i, j, _c) in enumerate(t["bnd"]):
            if self.bnd_cod[k] == 0:
                continue
            if ((p["shake_hydrogens"] and (not t["heavy"][i] or not t["heavy"][j])) or
                    (p["shake_solute"] and i < ns) or (p["shake_solvent"] and i >= ns)):
                self.shake_pairs.append((i, j))
                self.bnd_cod[k] = -1
        # the .fep file's constraints join them (a pair already constrained
        # gets the new distance; its bond term stays), md.f90:2416-2441
        if S:
            for (i, j) in f["shake"]:
                if (i, j) not in self.shake_pairs and (j, i) not in self.shake_pairs:
                    self.shake_pairs.append((i, j))
        for (i, j) in self.shake_pairs:
            a = t["ang"]
            hit = np.nonzero(((a[:, 0] == i) & (a[:, 2] == j)) | ((a[:, 0] == j) & (a[:, 2] == i)))[0]
            if len(hit):
                self.ang_cod[hit[0]] = 0
        if S:
            self._qconn()

    # --------------------------------------------------------------------
    def _qconn(self):
        """make_qconn / find_bonded (md.f90:7312-7411): bond distance + 1, up to 4."""
        t, f, S, ns = self.t, self.f, self.S, self.ns
        nq = f["nq"]
        qc = np.full((S, ns, nq), 9, int)
        adj = [[] for _ in range(ns)]
        for (i, j, _c), cod in zip(t["bnd"][: t["nb_sol"]], self.bnd_cod[: t["nb_sol"]]):
            if cod == 0:
                continue
            adj[i].append(j)
            adj[j].append(i)
        for s in range(S):
            ad = [list(a) for a in adj]
            for qb in f["qbnd"]:
                if qb[2 + s] > 0:
                    ad[qb[0] - 1].append(qb[1] - 1)
                    ad[qb[1] - 1].append(qb[0] - 1)
            for q, a0 in enumerate(f["iqseq"]):
                dist = {a0: 1}
                front = [a0]
                for lev in (2, 3, 4):
                    nxt = []
                    for a in front:
                        for b in ad[a]:
                            if b not in dist:
                                dist[b] = lev
                                nxt.append(b)
                    front = nxt
                for a, lv in dist.items():
                    qc[s, a, q] = lv
        # special exclusions among q-atoms, per state (make_qconn,
        # md.f90:7330-7348); with a non-q atom they change the topology
        # exclusions instead, which is not restated
        iq_of = {a: q for q, a in enumerate(f["iqseq"])}
        for (i, j, flags) in f["exspec"]:
            if i not in iq_of or j not in iq_of:
                raise ValueError("excluded pairs with non-q atoms are not restated")
            for s in range(S):
                if flags[s]:
                    qc[s, j, iq_of[i]] = 0
                    qc[s, i, iq_of[j]] = 0
        self.qconn = qc

    # ---------------------------------------------------------- pair lists --
    def pair_lists(self):
        """Unordered atom-pair sets of nbpp / nbpw, the O-O pairs of nbww, nbqp
        (iq, j, LJcod) and nbqw, plus the LRF far-field group pairs."""
        t, x, p = self.t, self.x, self.p
        ns, n = self.ns, self.n
        cg = t["cgp"]
        sw = np.array([c[0] for c in cg])
        nsg = t["ncgp_sol"]
        wat_o = ns + 3 * np.arange(self.nwat)
        # group switch positions: solute groups use their switching atom,
        # waters their oxygen (nbpwlist_lrf / nbwwlist_lrf)
        gsw = np.concatenate([sw[:nsg], wat_o])
        ok = ~t["excl"][gsw]
        P = x[gsw]
        d = P[:, None, :] - P[None, :, :]
        r2 = (d[..., 0] ** 2 + d[..., 1] ** 2) + d[..., 2] ** 2
        ng = len(gsw)
        gw = np.arange(ng) >= nsg
        rc2 = np.where(gw[:, None] & gw[None, :], p["rcww"] ** 2,
                       np.where(gw[:, None] | gw[None, :], p["rcpw"] ** 2, p["rcpp"] ** 2))
        valid = ok[:, None] & ok[None, :]
        inside = valid & (r2 <= rc2)
        self.far = valid & (r2 > rc2) & (r2 <= p["rclrf"] ** 2) & ~np.eye(ng, dtype=bool)
        self.gsw = gsw
        iq = self.iqatom
        pp, pp14, pw, ww = set(), set(), set(), set()
        ljc = 1  # no type-2 pairs
        crg, A, B, iac = self.crg, t["avdw"], t["bvdw"], t["iac"]

        def nonzero(i, j, c):
            return not ((F32(crg[i]) * F32(crg[j]) == 0) and A[iac[i] - 1, c - 1] * A[iac[j] - 1, c - 1] == 0
                        and B[iac[i] - 1, c - 1] * B[iac[j] - 1, c - 1] == 0)

        gi, gj = np.nonzero(np.triu(inside[:nsg, :nsg]))
        for a, b in zip(gi, gj):
            for i in cg[a][1]:
                if iq[i]:
                    continue
                for j in cg[b][1]:
                    if iq[j] or (a == b and i >= j) or i == j:
                        continue
                    key = (min(i, j), max(i, j))
                    if key in self.excl_pairs or not nonzero(i, j, ljc):
                        continue
                    if key in self.pairs14:
                        pp14.add(key)
                    pp.add(key)
        gi, gj = np.nonzero(inside[:nsg, nsg:])
        for a, b in zip(gi, gj):
            o = wat_o[b]
            for i in cg[a][1]:
                if iq[i]:
                    continue
                for j in (o, o + 1, o + 2):
                    if nonzero(i, j, ljc):
                        pw.add((i, j))
        gi, gj = np.nonzero(np.triu(inside[nsg:, nsg:], 1))
        for a, b in zip(gi, gj):
            ww.add((wat_o[a], wat_o[b]))
        qp, qw = [], []
        if self.S:
            rcq2 = p["rcq"] ** 2
            conn = (self.qconn <= 3).any(axis=(0, 2))
            for g in range(nsg):
                s_ = sw[g]
                if t["excl"][s_] or ((x[s_] - t["xpcent"]) ** 2).sum() > rcq2:
                    continue
                for j in cg[g][1]:
                    if conn[j]:
                        continue
                    for q in range(self.f["nq"]):
                        c = 3 if self.qconn[0, j, q] == 4 else 1
                        qp.append((q, j, c))
            for o in wat_o:
                if not t["excl"][o] and ((x[o] - t["xpcent"]) ** 2).sum() <= rcq2:
                    qw.append(o)
        self.lists = dict(pp=pp, pp14=pp14, pw=pw, ww=ww, qp=qp, qw=qw)
        return self.lists

    # ------------------------------------------------------------ energies --
    def pot_energy(self, tfree):
        """Energies (dict) and gradient d (n,3) at the current coordinates."""
        if not hasattr(self, "lists"):
            self.pair_lists()
        x, t, p = self.x, self.t, self.p
        E = {}
        d = np.zeros((self.n, 3))
        crg, A, B, iac = self.crg.astype(np.float64), t["avdw"], t["bvdw"], t["iac"] - 1
        rule = t["vdw_rule"]
        el14 = t["el14"]
        L = self.lists
        # solute-solute (nonbond_pp)
        pr = np.array(sorted(L["pp"]), int).reshape(-1, 2)
        is14 = np.array([tuple(v) in L["pp14"] for v in pr], bool)
        if len(pr):
            i, j = pr[:, 0], pr[:, 1]
            dx = x[j] - x[i]
            r2 = 1.0 / (dx ** 2).sum(1)
            r = np.sqrt(r2)
            r6 = r2 ** 3
            qq = (self.crg[i] * self.crg[j]).astype(np.float64)
            vel = qq * r * np.where(is14, el14, 1.0)
            c = np.where(is14, 2, 0)
            aa, bb = lj_ab(A[iac[i], c], B[iac[i], c], A[iac[j], c], B[iac[j], c], rule)
            va = r6 * r6 * aa
            vb = r6 * bb
            dv = r2 * (-vel - 12. * va + 6. * vb)
            np.add.at(d, i, -dv[:, None] * dx)
            np.add.at(d, j, dv[:, None] * dx)
            E["pp.el"], E["pp.vdw"] = vel.sum(), (va - vb).sum()
        else:
            E["pp.el"] = E["pp.vdw"] = 0.0
        # solute-water (nonbond_pw)
        pr = np.array(sorted(L["pw"]), int).reshape(-1, 2)
        if len(pr):
            i, j = pr[:, 0], pr[:, 1]
            dx = x[j] - x[i]
            r2 = 1.0 / (dx ** 2).sum(1)
            r = np.sqrt(r2)
            r6 = r2 ** 3
            vel = (self.crg[i] * self.crg[j]).astype(np.float64) * r
            aa, bb = lj_ab(A[iac[i], 0], B[iac[i], 0], A[iac[j], 0], B[iac[j], 0], rule)
            va = aa * r6 * r6
            vb = bb * r6
            dv = r2 * (-vel - 12. * va + 6. * vb)
            np.add.at(d, i, -dv[:, None] * dx)
            np.add.at(d, j, dv[:, None] * dx)
            E["pw.el"], E["pw.vdw"] = vel.sum(), (va - vb).sum()
        else:
            E["pw.el"] = E["pw.vdw"] = 0.0
        # water-water (nonbond_ww_spc; 3-atom solvent: nonbond_3atomsolvent,
        # md.f90:12317-12591, LJ on all nine atom pairs)
        E["ww.el"] = E["ww.vdw"] = 0.0
        if L["ww"]:
            pr = np.array(sorted(L["ww"]), int)
            qw = self.crg_w
            wt = [iac[self.ns + k] for k in range(3)]
            for a in range(3):
                for b in range(3):
                    i, j = pr[:, 0] + a, pr[:, 1] + b
                    dx = x[j] - x[i]
                    rinv = np.sqrt(1.0 / (dx ** 2).sum(1))
                    r2 = rinv * rinv
                    qq = float(F32(qw[a] * qw[b]))
                    vel = qq * rinv
                    if (a == 0 and b == 0) or self.wat3:
                        AOO, BOO = lj_ab(A[wt[a], 0], B[wt[a], 0], A[wt[b], 0], B[wt[b], 0], rule)
                        va = AOO * r2 ** 6
                        vb = BOO * r2 ** 3
                        dv = r2 * (-vel - 12. * va + 6. * vb)
                        E["ww.vdw"] += (va - vb).sum()
                    else:
                        dv = r2 * (-vel)
                    E["ww.el"] += vel.sum()
                    np.add.at(d, i, -dv[:, None] * dx)
                    np.add.at(d, j, dv[:, None] * dx)
        # Q-atom terms
        S = self.S
        EQ = np.zeros((S, 15))
        if S:
            self._q_terms(d, EQ)
            EQ[:, 0] = p["lambda"]
        # LRF
        E["LRF"] = self._lrf(d) if p["lrf"] else 0.0
        # bonded
        self._bonded(d, E, EQ)
        # restraints
        self._restraints(d, E, tfree)
        self._p_restrain(d, E, EQ)
        # totals (pot_energy, md.f90:13442-13471)
        for s in range(S):
            EQ[s, 6] = EQ[s, 8] + EQ[s, 10] + EQ[s, 12]
            EQ[s, 7] = EQ[s, 9] + EQ[s, 11] + EQ[s, 13]
            EQ[s, 1] = EQ[s, 2:8].sum() + EQ[s, 14]
        lam = p["lambda"] if S else np.zeros(0)
        for k, nm in ((2, "q.bond"), (3, "q.angle"), (4, "q.torsion"), (5, "q.improper"), (6, "qx.el"), (7, "qx.vdw")):
            E[nm] = float((EQ[:, k] * lam).sum()) if S else 0.0
        E["restraint.protein"] = E["p_restrain"] + (float((EQ[:, 14] * lam).sum()) if S else 0.0)
        E["restraint.total"] = (E["restraint.fix"] + E["restraint.shell"] + E["restraint.protein"] +
                                E["restraint.solvent_radial"] + E["restraint.water_pol"])
        E["potential"] = sum(E[k] for k in ("p.bond", "w.bond", "p.angle", "w.angle", "p.torsion", "p.improper",
                                            "pp.el", "pp.vdw", "pw.el", "pw.vdw", "ww.el", "ww.vdw", "q.bond",
                                            "q.angle", "q.torsion", "q.improper", "qx.el", "qx.vdw",
                                            "restraint.total", "LRF"))
        return E, EQ, d

    def _q_terms(self, d, EQ):
        f, t, x, S = self.f, self.t, self.x, self.S
        lam = self.p["lambda"]
        A, B, iac = t["avdw"], t["bvdw"], t["iac"] - 1
        rule = t["vdw_rule"]
        nty = len(t["mass"])
        el14 = t["el14"]
        iq_at = np.array(f["iqseq"])
        qv = f["qvdw"]
        # q-solute (nonbond_qp_qvdw)
        qp = np.array(self.lists["qp"], int).reshape(-1, 3)
        if len(qp):
            q, j, c = qp[:, 0], qp[:, 1], qp[:, 2]
            i = iq_at[q]
            dx = x[j] - x[i]
            r2h = (dx ** 2).sum(1)
            r6 = r2h ** 3
            r2 = 1.0 / r2h
            r = np.sqrt(r2)
            col = c - 1
            dv = np.zeros(len(q))
            for s in range(S):
                if qv:
                    qt = f["qiac"][q, s]
                    aLJ, bLJ = lj_ab(f["qavdw"][qt, col], f["qbvdw"][qt, col], A[iac[j], col], B[iac[j], col], rule)
                    r6sc = r6 + self.sc[q, iac[j], s]
                else:
                    aLJ, bLJ = lj_ab(A[iac[i], col], B[iac[i], col], A[iac[j], col], B[iac[j], col], rule)
                    r6sc = r6
                vel = (self.qcrg[q, s] * self.crg[j]).astype(np.float64) * r * np.where(c == 3, el14, 1.0)
                va = aLJ / r6sc ** 2
                vb = bLJ / r6sc
                dv += r2 * (-vel - (12. * va - 6. * vb) * (r6 / r6sc)) * lam[s]
                EQ[s, 10] += vel.sum()
                EQ[s, 11] += (va - vb).sum()
            np.add.at(d, i, -dv[:, None] * dx)
            np.add.at(d, j, dv[:, None] * dx)
        # q-water (nonbond_qw_spc; 3-atom solvent: nonbond_qw_3atom,
        # md.f90:11545-11670, LJ and soft-core on every solvent atom)
        qw = np.array(self.lists["qw"], int)
        if len(qw):
            wt = [iac[self.ns + k] for k in range(3)]
            nlj = 3 if self.wat3 else 1
            for q, i in enumerate(iq_at):
                dk = [x[qw + k] - x[i] for k in range(3)]
                r2h = [(v ** 2).sum(1) for v in dk]
                rk = [np.sqrt(1.0 / r) for r in r2h]
                r6hc = [r ** 3 for r in r2h]
                dv = [np.zeros(len(qw)) for _ in range(3)]
                for s in range(S):
                    for k in range(3):
                        qq = float(F32(self.crg_w[k] * self.qcrg[q, s]))
                        vel = qq * rk[k]
                        va = vb = 0.0
                        r6sc = r6hc[k]
                        if k < nlj:
                            if qv:
                                qt = f["qiac"][q, s]
                                aLJ, bLJ = lj_ab(f["qavdw"][qt, 0], f["qbvdw"][qt, 0], A[wt[k], 0], B[wt[k], 0], rule)
                                r6sc = r6hc[k] + self.sc[q, wt[k], s]
                            else:
                                aLJ, bLJ = lj_ab(A[iac[i], 0], B[iac[i], 0], A[wt[k], 0], B[wt[k], 0], rule)
                            va = aLJ / r6sc ** 2
                            vb = bLJ / r6sc
                        dv[k] += (1.0 / r2h[k]) * (-vel - (12. * va - 6. * vb) * (r6hc[k] / r6sc)) * lam[s]
                        EQ[s, 12] += vel.sum()
                        EQ[s, 13] += np.sum(va - vb)
                d[i] -= sum(dv[k][:, None] * dk[k] for k in range(3)).sum(0)
                for k in range(3):
                    np.add.at(d, qw + k, dv[k][:, None] * dk[k])
        # q-q and q-connected pairs (nbqqlist + nonbond_qq)
        qc = self.qconn
        nq = f["nq"]
        for s in range(S):
            pairs = []
            for a in range(nq - 1):
                for b in range(a + 1, nq):
                    lv = qc[s, iq_at[b], a]
                    if lv > 3:
                        # LJ code 2 marks a soft pair (nbqqlist, md.f90:7468-7477)
                        lj = 3 if lv == 4 else (2 if qv and (a, b) in f["soft"] else 1)
                        pairs.append((a, iq_at[b], b + 1, lj))
            conn = (qc <= 3).any(axis=(0, 2))
            for j in range(self.ns):
                if self.iqatom[j] or not conn[j]:
                    continue
                for a in range(nq):
                    if qc[s, j, a] >= 4:
                        pairs.append((a, j, 0, 3 if qc[s, j, a] == 4 else 1))
            for (a, j, jq, lj) in pairs:
                i = iq_at[a]
                col = lj - 1
                qi = float(self.qcrg[a, s])
                qj = float(self.qcrg[jq - 1, s]) if jq else float(self.crg[j])
                if f["libq"] and jq:
                    # nonbond_qq_lib_charges (md.f90:10896-10900, 10916)
                    qi, qj = float(self.crg[i]), float(self.crg[j])
                es = f["el_scale"].get((a, jq - 1), [1.0] * S)[s] if jq else 1.0
                expo = qv and jq and lj == 2
                if not qv:
                    # topology types for both atoms (nonbond_qq, md.f90:10786-10796)
                    aLJ, bLJ = lj_ab(A[iac[i], col], B[iac[i], col], A[iac[j], col], B[iac[j], col], rule)
                else:
                    qt = f["qiac"][a, s]
                    if jq:
                        qtj = f["qiac"][jq - 1, s]
                        # the exponential repulsion takes plain products
                        # under either rule (nonbon2_qq, md.f90:9424-9429)
                        aLJ, bLJ = lj_ab(f["qavdw"][qt, col], f["qbvdw"][qt, col], f["qavdw"][qtj, col],
                                         f["qbvdw"][qtj, col], 1 if expo else rule)
                    else:
                        aLJ, bLJ = lj_ab(f["qavdw"][qt, col], f["qbvdw"][qt, col], A[iac[j], col],
                                         B[iac[j], col], rule)
                dx = x[j] - x[i]
                r2 = (dx ** 2).sum()
                r6hc = r2 ** 3
                r6 = 1.0 / (r6hc + self.sc[a, nty + jq - 1, s])
                r2i = 1.0 / r2
                r = math.sqrt(r2i)
                vel = qi * qj * r * es
                if lj == 3:
                    vel *= el14
                if expo:
                    # soft pair: V = A exp(-B r) (nonbond_qq, md.f90:10838-10842)
                    va, vb = aLJ * math.exp(-bLJ / r), 0.0
                    dv = r2i * (-vel - bLJ * va / r) * lam[s]
                else:
                    va, vb = aLJ * r6 * r6, bLJ * r6
                    dv = r2i * (-vel - (12. * va - 6. * vb) * r6hc * r6) * lam[s]
                d[i] -= dv * dx
                d[j] += dv * dx
                EQ[s, 8 if jq else 10] += vel
                EQ[s, 9 if jq else 11] += va - vb

    def _lrf(self, d):
        """Far-field Taylor expansions per target group, then lrf_taylor."""
        t, x = self.t, self.x
        cg = t["cgp"]
        nsg = t["ncgp_sol"]
        groups = [c[1] for c in cg[:nsg]] + [[o, o + 1, o + 2] for o in self.ns + 3 * np.arange(self.nwat)]
        cent = np.array([x[g].sum(0) / float(len(g)) for g in groups])
        src_atoms = [np.array([a for a in g if not self.iqatom[a]], int) for g in groups]
        crg = self.crg.astype(np.float64)
        phi = np.zeros((len(groups), 20))
        for g in range(len(groups)):
            far = np.nonzero(self.far[g])[0]
            if not len(far):
                continue
            at = np.concatenate([src_atoms[h] for h in far])
            if not len(at):
                continue
            phi[g] = lrf_expand(x[at] - cent[g], crg[at])
        self.phi, self.cent = phi, cent
        gof = np.zeros(self.n, int)
        for g, mem in enumerate(groups):
            gof[mem] = g
        at = np.nonzero(~t["excl"] & (self.iqatom == 0))[0]
        g = gof[at]
        vij, fld = lrf_eval(phi[g], cent[g] - x[at])
        d[at] -= crg[at, None] * fld
        return float((0.5 * crg[at] * vij).sum())

    def _bonded(self, d, E, EQ):
        t, x = self.t, self.x
        lam = self.p["lambda"]
        # topology terms after shrink_topology (solute / solvent split by index)
        for nm in ("bond", "angle", "torsion", "improper"):
            E["p." + nm] = E["w." + nm] = 0.0
        for k, (i, j, c) in enumerate(t["bnd"]):
            c = self.bnd_cod[k]
            if c <= 0:
                continue
            fk, b0 = t["bondlib"][c]
            r = x[j] - x[i]
            b = math.sqrt(r @ r)
            db = b - b0
            dv = fk * db / b
            d[j] += r * dv
            d[i] -= r * dv
            E[("p." if k < t["nb_sol"] else "w.") + "bond"] += 0.5 * fk * db * db
        for k, (i, j, kk, c) in enumerate(t["ang"]):
            c = self.ang_cod[k]
            if c == 0:
                continue
            fk, a0 = self.anglib[c]
            e = self._angle(d, i, j, kk, fk, a0, 1.0)
            E[("p." if k < t["na_sol"] else "w.") + "angle"] += e
        # torsions and impropers, vectorised over the terms
        for key, cod, lib in (("tor", self.tor_cod, self.torlib), ("imp", self.imp_cod, self.implib)):
            a = t[key]
            m = cod != 0
            if not m.any():
                continue
            I, J, K, L = (a[m, c] for c in range(4))
            par = np.array([[float(v) for v in lib[c]] for c in cod[m]])
            phi, dp = self._dihedrals(I, J, K, L)
            if key == "tor":
                fk, rm, dl, paths = par.T
                arg = rm * phi - dl
                E["p.torsion"] += float((fk * (1.0 + np.cos(arg)) * paths).sum())
                dv = -rm * fk * np.sin(arg) * paths
            elif t["ff_type"] == 2:
                fk, i0 = par.T
                arg = 2 * phi - i0
                E["p.improper"] += float((fk * (1 + np.cos(arg))).sum())
                dv = -2 * fk * np.sin(arg)
            else:
                fk, i0 = par.T
                arg = phi - i0
                arg = arg - 2. * PI * np.round(arg / (2. * PI))
                dv = fk * arg
                E["p.improper"] += float((0.5 * dv * arg).sum())
            for q, at in enumerate((I, J, K, L)):
                np.add.at(d, at, dv[:, None] * dp[:, 3 * q:3 * q + 3])
        if not self.S:
            return
        f = self.f
        for s in range(self.S):
            for row in f["qbnd"]:
                c = row[2 + s]
                if c <= 0:
                    continue
                i, j = row[0] - 1, row[1] - 1
                Dmz, amz, r0, fk = f["qbondlib"][c]
                r = x[j] - x[i]
                b = math.sqrt(r @ r)
                db = b - r0
                fe = math.exp(-amz * db)
                EQ[s, 2] += Dmz * (fe * fe - 2. * fe) + 0.5 * fk * db * db
                dv = (2. * Dmz * amz * (fe - fe * fe) + fk * db) * lam[s] / b
                d[i] -= dv * r
                d[j] += dv * r
            for row in f["qang"]:
                c = row[3 + s]
                if c <= 0:
                    continue
                fk, a0 = f["qanglib"][c]
                EQ[s, 3] += self._angle(d, row[0] - 1, row[1] - 1, row[2] - 1, fk, a0, lam[s])
            for row in f["qtor"]:
                c = row[4 + s]
                if c <= 0:
                    continue
                fk, rm, dl = f["qtorlib"][c]
                phi, dp = self._dihedral(*(v - 1 for v in row[:4]))
                arg = rm * phi - dl
                EQ[s, 4] += fk * (1.0 + math.cos(arg))
                dv = -rm * fk * math.sin(arg) * lam[s]
                for a, at in enumerate(row[:4]):
                    d[at - 1] += dv * dp[3 * a:3 * a + 3]
            for row in f["qimp"]:
                c = row[4 + s]
                if c <= 0:
                    continue
                fk, i0 = f["qimplib"][c]
                phi, dp = self._dihedral(*(v - 1 for v in row[:4]))
                arg = phi - i0
                arg -= 2. * PI * round(arg / (2. * PI))
                dv = fk * arg
                EQ[s, 5] += 0.5 * dv * arg
                for a, at in enumerate(row[:4]):
                    d[at - 1] += dv * lam[s] * dp[3 * a:3 * a + 3]

    def _angle(self, d, i, j, k, fk, a0, lam):
        x = self.x
        rji, rjk = x[i] - x[j], x[k] - x[j]
        bji2, bjk2 = 1.0 / (rji @ rji), 1.0 / (rjk @ rjk)
        bji, bjk = math.sqrt(bji2), math.sqrt(bjk2)
        scp = min(1.0, max(-1.0, (rji @ rjk) * bji * bjk))
        ang = math.acos(scp)
        da = ang - a0
        dv = fk * da * lam
        f1 = math.sin(ang)
        f1 = -1.e12 if abs(f1) < 1.e-12 else -1.0 / f1
        di = f1 * (rjk * bji * bjk - scp * rji * bji2)
        dk = f1 * (rji * bji * bjk - scp * rjk * bjk2)
        d[i] += dv * di
        d[k] += dv * dk
        d[j] -= dv * (di + dk)
        return 0.5 * fk * da * da

    def _dihedral(self, i, j, k, l):
        phi, dp = self._dihedrals(np.array([i]), np.array([j]), np.array([k]), np.array([l]))
        return float(phi[0]), dp[0]

    def _dihedrals(self, I, J, K, L):
        """Dihedral angles and their 12 derivative components
        (torsion, md.f90:14793-14907)."""
        x = self.x
        rji, rjk, rkl = x[I] - x[J], x[K] - x[J], x[L] - x[K]
        rnj = np.cross(rji, rjk)
        rnk = -np.cross(rjk, rkl)
        bj2, bk2 = 1.0 / (rnj * rnj).sum(1), 1.0 / (rnk * rnk).sum(1)
        bj, bk = np.sqrt(bj2), np.sqrt(bk2)
        scp = np.clip((rnj * rnk).sum(1) * (bj * bk), -1.0, 1.0)
        phi = np.arccos(scp)
        phi = np.where((rjk * np.cross(rnj, rnk)).sum(1) < 0, -phi, phi)
        f1 = np.sin(phi)
        f1 = np.where(np.abs(f1) < 1.e-12, 1.e-12, f1)
        f1 = -1.0 / f1
        di = f1[:, None] * (rnk * (bj * bk)[:, None] - (scp * bj2)[:, None] * rnj)
        dl = f1[:, None] * (rnj * (bj * bk)[:, None] - (scp * bk2)[:, None] * rnk)
        rki, rlj = rji - rjk, -rjk - rkl
        dp = np.concatenate([np.cross(rjk, di), np.cross(rki, di) + np.cross(rkl, dl),
                             np.cross(rlj, dl) - np.cross(rji, di), np.cross(rjk, dl)], axis=1)
        return phi, dp

    def _p_restrain(self, d, E, EQ):
        """Sequence, position, distance, angle and wall restraints
        (p_restrain, md.f90:12940-13274)."""
        t, x, p = self.t, self.x, self.p
        mass = t["mass"][t["iac"] - 1]
        lam = p["lambda"]
        e_prot = 0.0

        def take(a, ih):
            return t["heavy"][a] or ih == 1

        for (i0, j0, fk, ih, tc) in p["rst_seq"]:
            at = [a for a in range(i0, j0 + 1) if take(a, ih)]
            if tc in (1, 2):
                if not at:
                    continue
                m = mass[at] if tc == 2 else np.ones(len(at))
                dr = ((x[at] - t["xtop"][at]) * m[:, None]).sum(0) / m.sum()
                e_prot += 0.5 * fk * (dr @ dr)
                for a in at:
                    d[a] += fk * dr * (mass[a] / float(F32(12.010)) if tc == 1 else 1.0)
            else:
                for a in at:
                    dr = x[a] - t["xtop"][a]
                    e_prot += 0.5 * fk * (dr @ dr)
                    d[a] += fk * dr

        def state_energy(ipsi, e):
            nonlocal e_prot
            if ipsi == 0:
                EQ[:, 14] += e
                if self.S == 0:
                    e_prot += e
            else:
                EQ[ipsi - 1, 14] += e

        for (i, x0, fk, ipsi) in p["rst_pos"]:
            wgt = lam[ipsi - 1] if ipsi else 1.0
            dr = x[i] - np.array(x0)
            d[i] += np.array(fk) * dr * wgt
            state_energy(ipsi, float((0.5 * np.array(fk) * dr * dr).sum()))
        for (i, j, d1, d2, fk, ipsi) in p["rst_dis"]:
            wgt = lam[ipsi - 1] if ipsi else 1.0
            dr = x[j] - x[i]
            b = math.sqrt(dr @ dr)
            if d1 <= b <= d2:
                continue
            db = b - (d1 if b < d1 else d2)
            dv = wgt * fk * db / b
            d[j] += dr * dv
            d[i] -= dr * dv
            state_energy(ipsi, 0.5 * fk * db * db)
        for (i, j, k, ang, fk, ipsi) in p["rst_ang"]:
            wgt = lam[ipsi - 1] if ipsi else 1.0
            e = self._angle(d, i, j, k, fk, ang * DEG2RAD, wgt)
            state_energy(ipsi, e)
        for (i0, j0, dw, fk, aM, dM, ih) in p["rst_wal"]:
            for a in range(i0, j0 + 1):
                if not take(a, ih):
                    continue
                dr = x[a] - t["xwcent"]
                b = math.sqrt(dr @ dr)
                db = b - dw
                if db > 0:
                    e_prot += 0.5 * fk * db * db - dM
                    dv = fk * db / b
                else:
                    fe = math.exp(aM * db)
                    e_prot += dM * (fe * fe - 2. * fe)
                    dv = -2. * dM * aM * (fe - fe * fe) / b
                d[a] += dv * dr
        E["p_restrain"] = e_prot

    def _restraints(self, d, E, tfree):
        t, x, p = self.t, self.x, self.p
        E["restraint.fix"] = E["restraint.shell"] = 0.0
        for i in np.nonzero(t["excl"] | self.shell)[0]:
            fk = 200.0 if t["excl"][i] else p["fk_pshell"]
            dr = x[i] - t["xtop"][i]
            e = 0.5 * fk * (dr @ dr)
            if t["excl"][i]:
                E["restraint.fix"] += e
            if self.shell[i]:
                E["restraint.shell"] += e
            d[i] += fk * dr
        E["restraint.solvent_radial"] = 0.0
        E["restraint.water_pol"] = 0.0
        if not self.nwat:
            return
        shift = math.sqrt(BOLTZ * tfree / self.fk_ws) if self.fk_ws != 0 else 0.0
        for w in range(self.nwat):
            i = self.ns + 3 * w
            if t["excl"][i]:
                continue
            dr = x[i] - t["xwcent"]
            b = math.sqrt(dr @ dr)
            db = b - (self.rwat - shift)
            if db > 0:
                e, dv = 0.5 * self.fk_ws * db * db - self.dwmz, self.fk_ws * db / b
            else:
                fe = math.exp(self.awmz * db)
                e, dv = self.dwmz * (fe * fe - 2. * fe), -2. * self.dwmz * self.awmz * (fe - fe * fe) / b
            E["restraint.solvent_radial"] += e
            d[i] += dv * dr
        if not self.shells:
            return
        sh = self.shells
        nsh = len(sh)
        theta = np.zeros(self.nwat)
        members = [[] for _ in range(nsh)]
        for w in range(self.nwat):
            i = self.ns + 3 * w
            if t["excl"][i]:
                continue
            mu = x[i + 1] + x[i + 2] - 2. * x[i]
            mu = mu / math.sqrt(mu @ mu)
            rc_v = x[i] - t["xwcent"]
            rc = math.sqrt(rc_v @ rc_v)
            scp = min(1.0, max(-1.0, mu @ (rc_v / rc)))
            theta[w] = math.acos(scp)
            if rc > float(F32(sh[-1]["rout"]) - F32(sh[-1]["dr"])):   # real(4) difference
                s = nsh - 1
                while s >= 1 and not rc <= float(sh[s]["rout"]):
                    s -= 1
                members[s].append(w)
        for s in range(nsh):
            m = members[s]
            nin = len(m)
            order = sorted(m, key=lambda w: (theta[w], w))
            for il, w in enumerate(order, 1):
                arg = float(F32(1.0) + (F32(1.0) - F32(2.0) * F32(il)) / F32(nin))
                th0 = math.acos(arg)
                th0 = th0 - 3. * math.sin(th0) * float(sh[s]["cstb"]) / 2.
                th0 = min(max(th0, 0.0), PI)
                dth = theta[w] - th0 + float(sh[s]["theta_corr"])
                E["restraint.water_pol"] += 0.5 * self.fkwpol * dth * dth
                dv = self.fkwpol * dth
                i = self.ns + 3 * w
                mu = x[i + 1] + x[i + 2] - 2. * x[i]
                rm = math.sqrt(mu @ mu)
                mu = mu / rm
                rc_v = x[i] - t["xwcent"]
                rc = math.sqrt(rc_v @ rc_v)
                rcu = rc_v / rc
                scp = min(1.0, max(-1.0, mu @ rcu))
                f0 = math.sin(math.acos(scp))
                f0 = dv * (-1.0 / max(f0, 1.e-12))
                f1 = (rcu - mu * scp) / rm
                f2 = (mu - rcu * scp) / rc
                d[i] += f0 * (-2. * f1 + f2)
                d[i + 1] += f0 * f1
                d[i + 2] += f0 * f1


def evaluate(inp_path, x0, tfree):
    """Energies, per-state energies and gradient of the window at x0."""
    w = Window(inp_path, x0)
    lists = w.pair_lists()
    E, EQ, d = w.pot_energy(tfree)
    return E, EQ, d.reshape(-1), lists
