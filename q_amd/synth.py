"""Seeded synthetic FEP systems written as Q topology (.top), .fep and qdyn .inp files.

The files follow the text formats that Q's readers accept:
  * topology: ``topo_read`` (src/q6/topo.f90:529-1082), version-5 header, sphere boundary;
  * fep file: ``qatom_load_atoms``/``qatom_load_fep`` (src/q6/qatom.f90:253-548, 874-1745);
  * input file: ``initialize`` (src/q6/md.f90:2528-3262).

Workloads (all seeded, all spherical boundary, SPC water, geometric LJ rule, FF_TYPE 2):
  * ``water_droplet``  -- 15 A sphere of SPC water on a 3.1 A lattice (1410 atoms);
  * ``ligand_in_water`` -- 50-atom two-state ligand (30 heavy + 20 H) in a 20 A SPC sphere;
  * ``protein_leg``     -- 160 residue-like 16-atom units + 40-atom two-state ligand in a 25 A sphere.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------
# force-field tables (OPLS-like; sqrt(A), sqrt(B) convention of the Q topology)
# ---------------------------------------------------------------------------
# name, sigma (A), epsilon (kcal/mol), mass
LJ_TYPES = [
    ("CT", 3.50, 0.066, 12.011),
    ("CA", 3.55, 0.070, 12.011),
    ("NT", 3.25, 0.170, 14.007),
    ("OS", 2.96, 0.210, 15.999),
    ("SH", 3.55, 0.250, 32.060),
    ("HC", 2.50, 0.030, 1.008),
    ("HN", 0.00, 0.000, 1.008),
    ("OW", 3.1656, 0.1554, 15.9994),
    ("HW", 0.00, 0.000, 1.008),
    ("DU", 0.00, 0.000, 1.008),
]
T_CT, T_CA, T_NT, T_OS, T_SH, T_HC, T_HN, T_OW, T_HW, T_DU = range(1, 11)
HEAVY_TYPES = [T_CT, T_CA, T_NT, T_OS, T_SH]
SPC_QO = -0.82
# a TIP3P-like 3-site water with Lennard-Jones on the hydrogens (topology
# solvent type 1, "3-atom"), SPC geometry; its two types follow LJ_TYPES
TIP3_TYPES = [("OT", 3.1507, 0.1521, 15.9994), ("HT", 0.4000, 0.0460, 1.008)]
TIP3_QO = -0.834
SPC_ROH = 1.0
SPC_HOH = 109.47


def lj_sqrt(sigma: float, eps: float) -> tuple[float, float]:
    """sqrt(A), sqrt(B) of 4 eps sigma^12 / 4 eps sigma^6."""
    if sigma == 0.0 or eps == 0.0:
        return 0.0, 0.0
    a = 4.0 * eps * sigma ** 12
    b = 4.0 * eps * sigma ** 6
    return math.sqrt(a), math.sqrt(b)


def lj_columns(sigma: float, eps: float, rule: int) -> tuple:
    """The (a, b) pairs of the three LJ columns (normal, polar, 1-4) of one
    atom type as a Q topology or FEP file stores them.  Geometric rule (1):
    sqrt(A), sqrt(B), with the 1-4 column at half the A and B.  Arithmetic
    rule (2): R*_i = r_min/2 = 2^(1/6) sigma / 2 and epsilon_i, the pair
    taking R* = R*_i + R*_j and eps = sqrt(eps_i eps_j) (topo.f90:1236-1270,
    md.f90:14747-14752); the 1-4 column halves epsilon."""
    if rule == 1:
        a, b = lj_sqrt(sigma, eps)
        r2 = math.sqrt(2.0)
        return (a, b), (a, b), (a / r2, b / r2)
    if sigma == 0.0 or eps == 0.0:
        return (0.0, 0.0), (0.0, 0.0), (0.0, 0.0)
    r = 2.0 ** (1.0 / 6.0) * sigma / 2.0
    return (r, eps), (r, eps), (r, eps / 2.0)


# bond codes: (fk, r0); angle codes: (fk, theta0 deg); torsion codes (fk, mult, delta deg, paths)
BOND_LIB = [(620.0, 1.46), (680.0, 1.09), (1106.0, SPC_ROH), (0.0, 2 * SPC_ROH * math.sin(math.radians(SPC_HOH / 2)))]
B_HEAVY, B_XH, B_OW_H, B_HW_HW = 1, 2, 3, 4
ANGLE_LIB = [(80.0, 112.0), (70.0, 109.5), (200.0, SPC_HOH)]
A_HEAVY, A_XH, A_WAT = 1, 2, 3
TORSION_LIB = [(0.70, 1.0, 0.0, 1.0), (0.25, 2.0, 180.0, 1.0), (0.35, 3.0, 0.0, 1.0)]
IMPROPER_LIB = [(1.10, 180.0)]


@dataclass
class System:
    """A Q topology in memory (1-based atom numbers in all index arrays)."""

    name: str
    x: list = field(default_factory=list)          # (n,3) coordinates
    iac: list = field(default_factory=list)        # atom type per atom (1-based)
    crg: list = field(default_factory=list)        # charges
    bonds: list = field(default_factory=list)      # (i, j, code)
    angles: list = field(default_factory=list)     # (i, j, k, code)
    torsions: list = field(default_factory=list)   # (i, j, k, l, code)
    impropers: list = field(default_factory=list)  # (i, j, k, l, code)
    cgps: list = field(default_factory=list)       # (switch atom, [atoms])
    residues: list = field(default_factory=list)   # (start atom, name)
    molecules: list = field(default_factory=list)  # start atoms
    nat_solute: int = 0
    nbonds_solute: int = 0
    nangles_solute: int = 0
    ncgp_solute: int = 0
    radius: float = 20.0
    vdw_rule: int = 1                              # 1 geometric, 2 arithmetic LJ combination
    sc_max_potential: bool = False                 # [softcore] values are maximum potentials, not alphas
    offdiag: list = field(default_factory=list)    # EVB (state i, state j, q-atom k, q-atom l, A, mu)
    water: str = "spc"                             # "spc" (solvent type 0) or "tip3" (3-atom, type 1)
    # FEP description (q-atoms are topology atoms, 1-based)
    qatoms: list = field(default_factory=list)
    qcharges: list = field(default_factory=list)   # per q-atom (state1, state2)
    qtypes: list = field(default_factory=list)     # per q-atom (type name s1, s2)
    qalpha: list = field(default_factory=list)     # per q-atom (alpha s1, s2)
    qbonds: list = field(default_factory=list)     # (i, j, code1, code2)
    qangles: list = field(default_factory=list)    # (i, j, k, code1, code2)
    qtorsions: list = field(default_factory=list)  # (i, j, k, l, code1, code2)
    qimpropers: list = field(default_factory=list)
    bond_lib: list = field(default_factory=lambda: list(BOND_LIB))
    angle_lib: list = field(default_factory=lambda: list(ANGLE_LIB))
    qbond_lib: list = field(default_factory=list)    # (fk, r0) harmonic or (Dmz, amz, r0) Morse per q-bond type
    qangle_lib: list = field(default_factory=list)   # (fk, theta0 deg) per q-angle type
    # further .fep sections (qatom.f90:297, 1049-1135, 1498-1512):
    # qq_use_library_charges (bool), el_scale [(iq, jq, s1, s2)], soft_pairs
    # [(iq, jq)] with exp_repulsion (A, B) columns for the q types,
    # excluded_pairs [(i, j, f1, f2)], shake [(i, j, d1, d2)]
    fep_extra: dict = field(default_factory=dict)
    inp_extra: list = field(default_factory=list)    # restraint sections appended to the .inp

    @property
    def natom(self) -> int:
        return len(self.x)


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------

def _rand_rot(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d],
    ])


def _place(a: np.ndarray, b: np.ndarray, c: np.ndarray, bond: float, angle_deg: float,
           tors_deg: float) -> np.ndarray:
    """Position of atom d bonded to c with |cd|=bond, angle bcd, torsion abcd."""
    th = math.radians(angle_deg)
    ph = math.radians(tors_deg)
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    nn = np.linalg.norm(n)
    if nn < 1e-8:
        n = np.cross(bc, np.array([1.0, 0.0, 0.0]))
        if np.linalg.norm(n) < 1e-8:
            n = np.cross(bc, np.array([0.0, 1.0, 0.0]))
        nn = np.linalg.norm(n)
    n /= nn
    m = np.cross(n, bc)
    d2 = np.array([-bond * math.cos(th), bond * math.sin(th) * math.cos(ph), bond * math.sin(th) * math.sin(ph)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


class _Grid:
    """Cell grid for clash checks while growing molecules."""

    def __init__(self, cell: float = 3.0):
        self.cell = cell
        self.cells: dict = {}

    def _key(self, p):
        return (int(math.floor(p[0] / self.cell)), int(math.floor(p[1] / self.cell)),
                int(math.floor(p[2] / self.cell)))

    def add(self, p, tag):
        self.cells.setdefault(self._key(p), []).append((np.asarray(p, float), tag))

    def min_dist(self, p, skip=()):
        k = self._key(p)
        best = 1e9
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q, tag in self.cells.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if tag in skip:
                            continue
                        d = float(np.linalg.norm(q - p))
                        if d < best:
                            best = d
        return best


def _tetra_dirs(p, nbrs, m, rng):
    """Unit vectors for ``m`` hydrogens completing a tetrahedral centre at
    ``p`` that already has neighbour positions ``nbrs``."""
    us = [(q - p) / np.linalg.norm(q - p) for q in nbrs]
    k = len(us)
    if k == 0:
        d = rng.normal(size=3)
        us = [d / np.linalg.norm(d)]
        k = 1
        m -= 0
    if k == 1:
        u = us[0]
        t = np.cross(u, rng.normal(size=3))
        t /= np.linalg.norm(t)
        t2 = np.cross(u, t)
        ang = math.radians(109.47)
        phi0 = rng.uniform(0, 2 * math.pi)
        out = []
        for j in range(3):
            ph = phi0 + 2 * math.pi * j / 3
            out.append(math.cos(ang) * u + math.sin(ang) * (math.cos(ph) * t + math.sin(ph) * t2))
        return out[:m]
    if k == 2:
        b = -(us[0] + us[1])
        b /= np.linalg.norm(b)
        n = np.cross(us[0], us[1])
        n /= np.linalg.norm(n)
        h = math.radians(54.74)
        return [math.cos(h) * b + math.sin(h) * n, math.cos(h) * b - math.sin(h) * n][:m]
    d = -(us[0] + us[1] + us[2])
    nd = np.linalg.norm(d)
    if nd < 1e-6:
        d = np.cross(us[0], us[1])
        nd = np.linalg.norm(d)
    return [d / nd][:m]


# ---------------------------------------------------------------------------
# molecule builders
# ---------------------------------------------------------------------------

def _grow_ligand(rng, n_heavy: int, n_h: int, center, grid: _Grid, first_atom: int):
    """Random-walk ligand: a tree of heavy atoms (bonds 1.40-1.52 A, angles
    109-120 deg) decorated with hydrogens.  Returns coordinates, types,
    neighbour lists (0-based within the ligand) and the heavy/H flags."""
    for _attempt in range(200):
        pos = [np.asarray(center, float)]
        nbr = [[]]
        ok = True
        while len(pos) < n_heavy and ok:
            cand = [i for i in range(len(pos)) if len(nbr[i]) < 3]
            placed = False
            for _try in range(60):
                p = int(rng.choice(cand))
                bl = rng.uniform(1.40, 1.52)
                ang = rng.uniform(109.0, 120.0)
                if nbr[p]:
                    q = nbr[p][0]
                    a = pos[nbr[q][0]] if nbr[q] and nbr[q][0] != p else pos[q] + rng.normal(size=3)
                    newp = _place(a, pos[q], pos[p], bl, ang, rng.uniform(0, 360))
                else:
                    d = rng.normal(size=3)
                    newp = pos[p] + bl * d / np.linalg.norm(d)
                dmin = min(float(np.linalg.norm(newp - r)) for r in pos)
                if dmin < 1.38 or any(np.linalg.norm(newp - pos[k]) < 2.3 for k in range(len(pos)) if k != p
                                      and k not in nbr[p]):
                    continue
                if grid.min_dist(newp) < 3.0:
                    continue
                nbr[p].append(len(pos))
                nbr.append([p])
                pos.append(newp)
                placed = True
                break
            ok = placed
        if not ok:
            continue
        # hydrogens on heavy atoms with free valence (max 4 neighbours),
        # placed on the tetrahedral directions left free by the heavy bonds
        heavy_n = len(pos)
        hpos, hpar = [], []
        order = [int(v) for v in rng.permutation(heavy_n)]
        want = [0] * heavy_n
        left = n_h
        while left > 0:
            progressed = False
            for p in order:
                if left == 0:
                    break
                if len(nbr[p]) + want[p] < 4:
                    want[p] += 1
                    left -= 1
                    progressed = True
            if not progressed:
                break
        clash = False
        for p in range(heavy_n):
            if want[p] == 0:
                continue
            near = set(nbr[p]) | {p}
            best, best_d = None, -1.0
            for _tt in range(24):
                cand = [pos[p] + 1.09 * d for d in _tetra_dirs(pos[p], [pos[q] for q in nbr[p]], want[p], rng)]
                others = [r for k, r in enumerate(pos) if k not in near] + hpos
                dmin = min((float(np.linalg.norm(c - r)) for c in cand for r in others), default=9.0)
                dmin = min(dmin, min((grid.min_dist(c) for c in cand), default=9.0))
                if dmin > best_d:
                    best, best_d = cand, dmin
                if dmin >= 2.1:
                    break
            if best_d < 1.9:
                clash = True
                break
            for c in best:
                hpos.append(c)
                hpar.append(p)
        if clash:
            continue
        if len(hpos) < n_h:
            continue
        xyz = pos + hpos
        heavy = [True] * heavy_n + [False] * n_h
        nb = [list(v) for v in nbr] + [[p] for p in hpar]
        for hi, p in enumerate(hpar):
            nb[p].append(heavy_n + hi)
        for k, r in enumerate(xyz):
            grid.add(r, ("lig", first_atom + k))
        types = [int(rng.choice(HEAVY_TYPES)) for _ in range(heavy_n)] + [T_HC] * n_h
        return xyz, types, nb, heavy
    raise RuntimeError("could not grow ligand")


def _topology_from_graph(nb: list, offset: int, heavy: list):
    """Bonds, angles, torsions, impropers from a neighbour graph (0-based)."""
    n = len(nb)
    bonds, angles, tors, imps = [], [], [], []
    for i in range(n):
        for j in nb[i]:
            if i < j:
                code = B_HEAVY if (heavy[i] and heavy[j]) else B_XH
                bonds.append((i + offset, j + offset, code))
    for j in range(n):
        for a in range(len(nb[j])):
            for b in range(a + 1, len(nb[j])):
                i, k = nb[j][a], nb[j][b]
                code = A_HEAVY if (heavy[i] and heavy[k]) else A_XH
                angles.append((i + offset, j + offset, k + offset, code))
    for j in range(n):
        for k in nb[j]:
            if j < k:
                for i in nb[j]:
                    if i == k:
                        continue
                    for l in nb[k]:
                        if l == j or l == i:
                            continue
                        for code in (1, 2, 3):
                            tors.append((i + offset, j + offset, k + offset, l + offset, code))
    for j in range(n):
        if heavy[j] and len(nb[j]) == 3:
            i, k, l = nb[j]
            imps.append((i + offset, j + offset, k + offset, l + offset, 1))
    return bonds, angles, tors, imps


def _add_water_lattice(rng, sysm: System, radius: float, grid: _Grid | None, spacing: float = 3.1,
                       heavy_grid: _Grid | None = None):
    """Waters (SPC, or the TIP3P-like model when ``sysm.water`` is "tip3")
    on a cubic lattice inside ``radius``, random orientations, skipping
    lattice points that clash with solute atoms."""
    if sysm.water == "tip3":
        t_o, t_h, q_o = len(LJ_TYPES) + 1, len(LJ_TYPES) + 2, TIP3_QO
    else:
        t_o, t_h, q_o = T_OW, T_HW, SPC_QO
    n = int(math.ceil(radius / spacing)) + 1
    pts = []
    for i in range(-n, n + 1):
        for j in range(-n, n + 1):
            for k in range(-n, n + 1):
                p = np.array([i, j, k], float) * spacing
                if np.dot(p, p) <= radius * radius:
                    pts.append(p)
    half = math.radians(SPC_HOH / 2)
    h1 = np.array([SPC_ROH * math.sin(half), SPC_ROH * math.cos(half), 0.0])
    h2 = np.array([-SPC_ROH * math.sin(half), SPC_ROH * math.cos(half), 0.0])
    qh = -q_o / 2
    for p in pts:
        if heavy_grid is not None:
            if heavy_grid.min_dist(p) < 2.7 or grid.min_dist(p) < 2.0:
                continue
        elif grid is not None and grid.min_dist(p) < 2.5:
            continue
        rot = _rand_rot(rng)
        base = sysm.natom + 1
        sysm.x += [p, p + rot @ h1, p + rot @ h2]
        sysm.iac += [t_o, t_h, t_h]
        sysm.crg += [q_o, qh, qh]
        sysm.bonds += [(base, base + 1, B_OW_H), (base, base + 2, B_OW_H), (base + 1, base + 2, B_HW_HW)]
        sysm.angles += [(base + 1, base, base + 2, A_WAT)]
        sysm.cgps.append((base, [base, base + 1, base + 2]))
        sysm.residues.append((base, "HOH"))
        sysm.molecules.append(base)


def _fill_water_gaps(rng, sysm: System, radius: float, grid: _Grid, heavy_grid: _Grid, first_water: int,
                     step: float = 0.9, oo_min: float = 2.55, hv_min: float = 2.35, any_min: float = 1.75,
                     h_min: float = 1.35) -> None:
    """Add waters in the voids the lattice left next to the solute: points of
    a finer grid, visited in random order, take a water when its oxygen is at
    least ``oo_min`` from every other water oxygen, ``hv_min`` from solute
    heavy atoms and ``any_min`` from any solute atom, and its (randomly
    oriented) hydrogens at least ``h_min`` from every other atom (close
    contacts, as hydrogen bonds have them)."""
    t_o, t_h, q_o = T_OW, T_HW, SPC_QO
    half = math.radians(SPC_HOH / 2)
    h1 = np.array([SPC_ROH * math.sin(half), SPC_ROH * math.cos(half), 0.0])
    h2 = np.array([-SPC_ROH * math.sin(half), SPC_ROH * math.cos(half), 0.0])
    qh = -q_o / 2
    wo = _Grid()
    wall = _Grid()
    for a in range(first_water - 1, sysm.natom, 3):
        wo.add(sysm.x[a], a)
        for k in range(3):
            wall.add(sysm.x[a + k], a + k)
    n = int(math.ceil(radius / step))
    pts = []
    r_in = radius - 0.5
    for i in range(-n, n + 1):
        for j in range(-n, n + 1):
            for k in range(-n, n + 1):
                p = (np.array([i, j, k], float) + 0.5) * step
                if np.dot(p, p) <= r_in * r_in:
                    pts.append(p)
    order = rng.permutation(len(pts))
    for idx in order:
        p = pts[idx]
        if wo.min_dist(p) < oo_min or heavy_grid.min_dist(p) < hv_min or grid.min_dist(p) < any_min:
            continue
        rot = _rand_rot(rng)
        hs = [p + rot @ h1, p + rot @ h2]
        if any(grid.min_dist(h) < h_min or wall.min_dist(h) < h_min for h in hs):
            continue
        base = sysm.natom + 1
        sysm.x += [p, hs[0], hs[1]]
        sysm.iac += [t_o, t_h, t_h]
        sysm.crg += [q_o, qh, qh]
        sysm.bonds += [(base, base + 1, B_OW_H), (base, base + 2, B_OW_H), (base + 1, base + 2, B_HW_HW)]
        sysm.angles += [(base + 1, base, base + 2, A_WAT)]
        sysm.cgps.append((base, [base, base + 1, base + 2]))
        sysm.residues.append((base, "HOH"))
        sysm.molecules.append(base)
        wo.add(p, base - 1)
        for k, q in enumerate([p] + hs):
            wall.add(q, base - 1 + k)


def _strain_free(s: System) -> None:
    """Give every solute bond and angle its own equilibrium parameter taken
    from the built geometry, so the bonded terms start unstrained."""
    x = np.asarray(s.x, float)
    nb = []
    for (i, j, c) in s.bonds[: s.nbonds_solute]:
        r = float(np.linalg.norm(x[j - 1] - x[i - 1]))
        s.bond_lib.append((BOND_LIB[c - 1][0], round(r, 5)))
        nb.append((i, j, len(s.bond_lib)))
    s.bonds = nb + s.bonds[s.nbonds_solute:]
    na = []
    for (i, j, k, c) in s.angles[: s.nangles_solute]:
        u = x[i - 1] - x[j - 1]
        w = x[k - 1] - x[j - 1]
        th = math.degrees(math.acos(max(-1.0, min(1.0, float(np.dot(u, w) / np.linalg.norm(u) / np.linalg.norm(w))))))
        s.angle_lib.append((ANGLE_LIB[c - 1][0], round(th, 3)))
        na.append((i, j, k, len(s.angle_lib)))
    s.angles = na + s.angles[s.nangles_solute:]


def _neutral_charges(rng, n: int, sd: float = 0.3) -> list:
    q = rng.normal(0.0, sd, size=n)
    q = np.round(q - q.mean(), 4)
    q[-1] = round(-float(np.sum(q[:-1])), 4)
    return [float(v) for v in q]


def _ligand_fep(rng, sysm: System, first: int, n_lig: int, types: list, n_pert: int, alpha: float,
                nb: list, heavy: list, spread: bool = False):
    """All ligand atoms become q-atoms; ``n_pert`` of them change charge and
    LJ type between state 1 and state 2 and get soft-core alpha.  State 2
    is kept neutral by moving the residual charge onto the last perturbed
    heavy atom, or, with ``spread``, in equal shares onto all of them."""
    pert = sorted(int(v) for v in rng.choice(n_lig, size=n_pert, replace=False))
    # at least one perturbed heavy atom carries the neutrality correction
    if not any(heavy[k] for k in pert):
        pert[-1] = int(rng.choice([k for k in range(n_lig) if heavy[k] and k not in pert]))
        pert = sorted(pert)
    q1 = [sysm.crg[first - 1 + k] for k in range(n_lig)]
    q2 = list(q1)
    for k in pert:
        # hydrogens vanish (dummy type, no LJ) in state 2: their charge goes to 0
        q2[k] = round(q1[k] + float(rng.normal(0.0, 0.3)), 4) if heavy[k] else 0.0
    # keep state 2 neutral: put the residual on a perturbed heavy atom
    resid = round(sum(q2) - sum(q1), 4)
    if spread:
        hs = [k for k in pert if heavy[k]]
        share = round(resid / len(hs), 4)
        for k in hs[:-1]:
            q2[k] = round(q2[k] - share, 4)
        resid = round(sum(q2) - sum(q1), 4)
    kh = [k for k in pert if heavy[k]][-1]
    q2[kh] = round(q2[kh] - resid, 4)
    names = {t: LJ_TYPES[t - 1][0] for t in range(1, len(LJ_TYPES) + 1)}
    sysm.qatoms = [first + k for k in range(n_lig)]
    sysm.qcharges = [(q1[k], q2[k]) for k in range(n_lig)]
    sysm.qtypes = []
    sysm.qalpha = []
    for k in range(n_lig):
        t1 = types[k]
        if k in pert:
            t2 = T_DU if not heavy[k] else int(rng.choice([t for t in HEAVY_TYPES if t != t1]))
            sysm.qtypes.append((names[t1], names[t2]))
            sysm.qalpha.append((alpha, alpha))
        else:
            sysm.qtypes.append((names[t1], names[t1]))
            sysm.qalpha.append((0.0, 0.0))
    # redefine the bonded terms touching perturbed atoms as state-dependent
    # q-terms: state 1 keeps the topology parameters, state 2 shifts them
    pset = {first + k for k in pert}
    for (i, j, c) in [b for b in sysm.bonds if b[0] in pset or b[1] in pset]:
        if i in sysm.qatoms and j in sysm.qatoms:
            fk, r0 = sysm.bond_lib[c - 1]
            sysm.qbond_lib += [(fk, r0), (fk, round(r0 + 0.03, 5))]
            n = len(sysm.qbond_lib)
            sysm.qbonds.append((i, j, n - 1, n))
    for (i, j, k, c) in [a for a in sysm.angles if a[1] in pset]:
        fk, t0 = sysm.angle_lib[c - 1]
        sysm.qangle_lib += [(fk, t0), (fk, round(t0 + 2.0, 3))]
        n = len(sysm.qangle_lib)
        sysm.qangles.append((i, j, k, n - 1, n))
    seen = set()
    for (i, j, k, l, c) in sysm.torsions:
        if (j in pset or k in pset) and (i, j, k, l) not in seen and c == 1:
            seen.add((i, j, k, l))
            sysm.qtorsions.append((i, j, k, l, 1, 2))
    for (i, j, k, l, c) in sysm.impropers:
        if j in pset:
            sysm.qimpropers.append((i, j, k, l, 1, 1))


def reperturb(base: System, seed: int, n_pert: int, alpha: float = 10.0, spread: bool = True) -> System:
    """A copy of ``base`` whose two FEP states are redrawn: the same ligand
    (the first ``len(base.qatoms)`` topology atoms) and environment, a new
    seeded choice of ``n_pert`` perturbed atoms with new charges and types
    (neutrality residual spread over the perturbed heavy atoms unless
    ``spread`` is false).  Used to make many ligand perturbations in one
    costly environment."""
    import copy
    if not base.qatoms or base.qatoms[0] != 1:
        raise ValueError("reperturb needs a system whose ligand atoms come first")
    n_lig = len(base.qatoms)
    s = copy.deepcopy(base)
    for name in ("qatoms", "qcharges", "qtypes", "qalpha", "qbonds", "qangles", "qtorsions", "qimpropers",
                 "qbond_lib", "qangle_lib"):
        setattr(s, name, [])
    types = [int(t) for t in s.iac[:n_lig]]
    heavy = [t in HEAVY_TYPES for t in types]
    nb: list = [[] for _ in range(n_lig)]
    for (i, j, _c) in s.bonds:
        if i <= n_lig and j <= n_lig:
            nb[i - 1].append(j - 1)
            nb[j - 1].append(i - 1)
    _ligand_fep(np.random.default_rng(seed), s, 1, n_lig, types, n_pert, alpha, nb, heavy, spread=spread)
    return s


def water_droplet(radius: float = 15.0, seed: int = 0) -> System:
    rng = np.random.default_rng(seed)
    s = System(name=f"water{radius:g}", radius=radius)
    _add_water_lattice(rng, s, radius, None)
    return s


def _graph_distances(nb: list) -> np.ndarray:
    """Bond-count distances between all atoms of a molecule graph."""
    n = len(nb)
    dist = np.full((n, n), 99, int)
    for a in range(n):
        dist[a, a] = 0
        front = [a]
        while front:
            nxt = []
            for u in front:
                for v in nb[u]:
                    if dist[a, v] == 99:
                        dist[a, v] = dist[a, u] + 1
                        nxt.append(v)
            front = nxt
    return dist


def _fep_sections(s: System, nb: list, heavy: list, seed: int) -> None:
    """The FEP sections beyond charges, types and bonded changes on the
    ligand (all of whose atoms are q-atoms, numbered as the topology atoms):
    electrostatic scaling of three q-q pairs, two soft (exponential
    repulsion) pairs, three excluded pairs (one per state, one in both), a
    Morse q-bond in state 1, and two extra SHAKE constraints with
    state-dependent distances (an X-H bond and an H...H distance across an
    angle).  Pairs are nonbonded (more than three bonds apart) and 3-7 A
    apart, so every term is non-trivial."""
    rng = np.random.default_rng(seed)
    x = np.array(s.x)
    dist = _graph_distances(nb)
    n = len(nb)
    cand = [(a, b) for a in range(n) for b in range(a + 1, n)
            if dist[a, b] > 3 and 3.0 < float(np.linalg.norm(x[a] - x[b])) < 7.0]
    pick = [cand[k] for k in rng.choice(len(cand), size=8, replace=False)]
    s.fep_extra["el_scale"] = [(a + 1, b + 1, f1, f2) for (a, b), (f1, f2) in
                               zip(pick[:3], [(0.5, 0.8), (0.2, 1.0), (1.0, 0.3)])]
    s.fep_extra["soft_pairs"] = [(a + 1, b + 1) for (a, b) in pick[3:5]]
    s.fep_extra["exp_repulsion"] = (30.0, 1.2)
    s.fep_extra["excluded_pairs"] = [(a + 1, b + 1, f1, f2) for (a, b), (f1, f2) in
                                     zip(pick[5:8], [(1, 0), (0, 1), (1, 1)])]
    # a Morse bond for state 1 of the first q-bond (qbond, md.f90:14060-14114)
    i, j, c1, c2 = s.qbonds[0]
    r0 = s.qbond_lib[c1 - 1][-1]
    s.qbond_lib.append((70.0, 2.0, r0))
    s.qbonds[0] = (i, j, len(s.qbond_lib), c2)
    # SHAKE: one X-H bond and the H...H distance of an H-X-H angle
    xh = [(a, b) for a in range(n) for b in nb[a] if heavy[a] and not heavy[b]]
    hh = [(a, b) for c in range(n) for a in nb[c] for b in nb[c] if a < b and not heavy[a] and not heavy[b]]
    sh = []
    for (a, b) in [xh[0]] + hh[:1]:
        d0 = round(float(np.linalg.norm(x[a] - x[b])), 4)
        sh.append((a + 1, b + 1, d0, round(d0 + 0.02, 4)))
    s.fep_extra["shake"] = sh


def _library_charges(s: System, n_lig: int) -> None:
    """qq_use_library_charges on (qatom.f90:297): q-q pairs take the topology
    charges (nonbond_qq_lib_charges, md.f90:10872-10975), which are made to
    differ from both states' FEP charges (0.9 of state 1, kept neutral)."""
    q = [round(0.9 * s.crg[k], 4) for k in range(n_lig)]
    q[-1] = round(q[-1] - sum(q), 4)
    s.crg[:n_lig] = q
    s.fep_extra["qq_use_library_charges"] = True


def _restraint_sections(s: System, heavy: list, seed: int) -> list:
    """.inp restraint sections (md.f90:3118-3259) over the ligand and a few
    waters: sequence restraints to each atom's topology position and to the
    mass centre, position restraints for all states and for each state,
    flat-bottomed distance restraints (one below d1, one above d2, one
    state-dependent), angle restraints, and a wall whose radius leaves some
    ligand atoms outside (harmonic) and some inside (Morse)."""
    rng = np.random.default_rng(seed)
    x = np.array(s.x)
    n = s.nat_solute
    wo = n + 1 + 3 * 7                      # an oxygen of a water (1-based)
    out = ["", "[sequence_restraints]", f"1 {n // 2} 2.0 0 0", f"{n // 2 + 1} {n} 1.5 1 2",
           "", "[atom_restraints]"]
    for k, ipsi in ((3, 0), (11, 1), (27, 2), (wo, 0)):
        p = x[k - 1] + rng.normal(0.0, 0.3, 3)
        out.append(f"{k} {p[0]:.4f} {p[1]:.4f} {p[2]:.4f} 2.0 3.0 1.5 {ipsi}")
    out += ["", "[distance_restraints]"]
    hv = [k + 1 for k in range(n) if heavy[k]]
    pairs = [(hv[0], hv[-1]), (hv[2], hv[12]), (hv[5], wo), (hv[8], hv[20])]
    for m, (a, b) in enumerate(pairs):
        r = float(np.linalg.norm(x[a - 1] - x[b - 1]))
        d1, d2 = ((r + 0.5, r + 1.0), (r - 1.0, r - 0.5), (r - 0.2, r + 0.2), (r + 0.3, r + 0.6))[m]
        out.append(f"{a} {b} {d1:.3f} {d2:.3f} {4.0 + m:.1f} {(0, 0, 1, 2)[m]}")
    out += ["", "[angle_restraints]"]
    for m, (a, b, c) in enumerate([(hv[1], hv[4], hv[9]), (hv[3], hv[7], hv[15]), (hv[6], wo, hv[11])]):
        u, w = x[a - 1] - x[b - 1], x[c - 1] - x[b - 1]
        th = math.degrees(math.acos(float(np.dot(u, w) / np.linalg.norm(u) / np.linalg.norm(w))))
        out.append(f"{a} {b} {c} {th + 7.0 * (m + 1):.2f} {5.0 + m:.1f} {m}")
    out += ["", "[wall_restraints]", f"1 {n} 3.0 10.0 2.0 0.5 1"]
    return out


def ligand_in_water(radius: float = 20.0, seed: int = 0, n_heavy: int = 30, n_h: int = 20,
                    n_pert: int = 6, alpha: float = 10.0, vdw_rule: int = 1,
                    sc_max_potential: bool = False, offdiag: tuple = (), water: str = "spc",
                    extras: tuple = ()) -> System:
    """Seeded ligand in an SPC (or 3-atom) water sphere.  ``extras`` adds, on
    the same system: "fep" the further FEP sections, "libq" library charges
    for q-q pairs, "restr" the restraint sections of the .inp."""
    rng = np.random.default_rng(seed)
    s = System(name="ligwat", radius=radius, vdw_rule=vdw_rule, sc_max_potential=sc_max_potential,
               offdiag=list(offdiag), water=water)
    grid = _Grid()
    xyz, types, nb, heavy = _grow_ligand(rng, n_heavy, n_h, np.zeros(3), grid, 1)
    # re-centre the ligand on the sphere centre
    c = np.mean(xyz, axis=0)
    xyz = [p - c for p in xyz]
    grid = _Grid()
    for k, p in enumerate(xyz):
        grid.add(p, ("lig", k))
    n = len(xyz)
    s.x = list(xyz)
    s.iac = list(types)
    s.crg = _neutral_charges(rng, n)
    b, a, t, i = _topology_from_graph(nb, 1, heavy)
    s.bonds, s.angles, s.torsions, s.impropers = b, a, t, i
    for k in range(n):
        if heavy[k]:
            members = [k + 1] + [h + 1 for h in nb[k] if not heavy[h]]
            s.cgps.append((k + 1, sorted(members)))
    s.residues.append((1, "LIG"))
    s.molecules.append(1)
    s.nat_solute = n
    s.nbonds_solute = len(s.bonds)
    s.nangles_solute = len(s.angles)
    s.ncgp_solute = len(s.cgps)
    _strain_free(s)
    _ligand_fep(rng, s, 1, n, types, n_pert, alpha, nb, heavy)
    _add_water_lattice(rng, s, radius, grid)
    if "fep" in extras:
        _fep_sections(s, nb, heavy, seed + 1000)
    if "libq" in extras:
        _library_charges(s, n)
    if "restr" in extras:
        s.inp_extra = _restraint_sections(s, heavy, seed + 2000)
    return s


def ligand_dual(radius: float = 20.0, seed: int = 0, n_heavy: int = 30, n_h: int = 20, vmax: float = 20.0,
                vdw_rule: int = 1) -> System:
    """A QligFEP-style dual-topology perturbation (qligfep.py:311-362) in an
    SPC sphere: ligand A and ligand B (A's graph with some heavy-atom types
    changed, new charges, coordinates displaced by up to 0.4 A, so the two
    overlap as aligned ligands do) are both in the topology and all their
    atoms are q-atoms.  State 1 has A with its charges and types and B as
    dummies (charge 0, the zero-LJ type DU); state 2 the reverse.  Soft-core
    is given as maximum potentials: 0 / vmax for A's atoms, vmax / 0 for B's
    (softcore_use_max_potential on)."""
    rng = np.random.default_rng(seed)
    s = System(name="ligdual", radius=radius, vdw_rule=vdw_rule, sc_max_potential=True)
    grid = _Grid()
    xyz, types, nb, heavy = _grow_ligand(rng, n_heavy, n_h, np.zeros(3), grid, 1)
    c = np.mean(xyz, axis=0)
    xa = [p - c for p in xyz]
    n = len(xa)
    xb = [p + rng.uniform(-0.4, 0.4, 3) for p in xa]
    types_b = [int(rng.choice([u for u in HEAVY_TYPES if u != t])) if heavy[k] and rng.random() < 0.3 else t
               for k, t in enumerate(types)]
    qa = _neutral_charges(rng, n)
    qb = list(np.round(np.array(qa) + rng.normal(0.0, 0.1, n), 4))
    qb[-1] = round(qb[-1] - sum(qb), 4)
    qb = [float(v) for v in qb]
    grid = _Grid()
    for p in xa + xb:
        grid.add(p, "lig")
    s.x = list(xa) + list(xb)
    s.iac = list(types) + list(types_b)
    s.crg = list(qa) + list(qb)
    for off in (0, n):
        b, a, t, i = _topology_from_graph(nb, 1 + off, heavy)
        s.bonds += b
        s.angles += a
        s.torsions += t
        s.impropers += i
        for k in range(n):
            if heavy[k]:
                members = [k + 1 + off] + [h + 1 + off for h in nb[k] if not heavy[h]]
                s.cgps.append((k + 1 + off, sorted(members)))
        s.residues.append((1 + off, "LIG" if off == 0 else "LID"))
        s.molecules.append(1 + off)
    s.nat_solute = 2 * n
    s.nbonds_solute = len(s.bonds)
    s.nangles_solute = len(s.angles)
    s.ncgp_solute = len(s.cgps)
    _strain_free(s)
    names = {t: LJ_TYPES[t - 1][0] for t in range(1, len(LJ_TYPES) + 1)}
    du = names[T_DU]
    s.qatoms = list(range(1, 2 * n + 1))
    s.qcharges = [(qa[k], 0.0) for k in range(n)] + [(0.0, qb[k]) for k in range(n)]
    s.qtypes = [(names[types[k]], du) for k in range(n)] + [(du, names[types_b[k]]) for k in range(n)]
    s.qalpha = [(0.0, vmax)] * n + [(vmax, 0.0)] * n
    _add_water_lattice(rng, s, radius, grid)
    return s


# residue-like unit: 7 heavy atoms + 9 hydrogens = 16 atoms
# heavy (0-based within unit): N0-C1, C1-C2, C2-C3(=O4), C1-C5-C6 side chain;
# C3 of one unit bonds N0 of the next.  Hydrogens: N0 1, C1 1, C2 2, C5 2, C6 3.
_UNIT_HEAVY = [T_NT, T_CT, T_CT, T_CA, T_OS, T_CT, T_CT]
_UNIT_NB = [[1], [0, 2, 5], [1, 3], [2, 4], [3], [1, 6], [5]]
_UNIT_H = [(0, T_HN), (1, T_HC), (2, T_HC), (2, T_HC), (5, T_HC), (5, T_HC), (6, T_HC), (6, T_HC),
           (6, T_HC)]
_UNIT_CRG_HEAVY = [-0.50, 0.14, -0.12, 0.50, -0.50, -0.12, -0.18]
_UNIT_CRG_H = [0.30, 0.06, 0.06, 0.06, 0.06, 0.06, 0.06, 0.06, 0.06]


def _unit_template():
    crg = _UNIT_CRG_HEAVY + _UNIT_CRG_H
    tot = round(sum(crg), 4)
    crg[-1] = round(crg[-1] - tot, 4)  # neutral per unit
    return crg


def protein_leg(radius: float = 25.0, seed: int = 2, n_units: int = 160, n_lig: int = 40,
                n_pert: int = 6, alpha: float = 10.0, core: float = 21.5, vdw_rule: int = 1,
                fill: bool = True) -> System:
    """Synthetic chain of ``n_units`` 16-atom units wrapped around a central
    two-state ligand, solvated by SPC water up to ``radius`` (a lattice, then,
    with ``fill``, the voids it leaves next to the chain)."""
    rng = np.random.default_rng(seed)
    s = System(name="protlig", radius=radius, vdw_rule=vdw_rule)
    grid = _Grid()
    # ligand first in the topology (it is the FEP part), chain afterwards
    n_heavy = (n_lig * 3) // 5
    xyz, ltypes, lnb, lheavy = _grow_ligand(rng, n_heavy, n_lig - n_heavy, np.zeros(3), grid, 1)
    c = np.mean(xyz, axis=0)
    xyz = [p - c for p in xyz]
    grid = _Grid()
    hgrid = _Grid()
    for k, p in enumerate(xyz):
        grid.add(p, ("lig", k))
        if lheavy[k]:
            hgrid.add(p, ("lig", k))
    s.x = list(xyz)
    s.iac = list(ltypes)
    s.crg = _neutral_charges(rng, len(xyz))
    b, a, t, i = _topology_from_graph(lnb, 1, lheavy)
    s.bonds, s.angles, s.torsions, s.impropers = b, a, t, i
    for k in range(len(xyz)):
        if lheavy[k]:
            members = [k + 1] + [h + 1 for h in lnb[k] if not lheavy[h]]
            s.cgps.append((k + 1, sorted(members)))
    s.residues.append((1, "LIG"))
    s.molecules.append(1)
    lig_n = len(xyz)

    # grow the chain heavy atom by heavy atom as a confined self-avoiding walk
    crg_t = _unit_template()
    chain_nb: list = []
    chain_heavy: list = []
    chain_xyz: list = []
    chain_types: list = []
    chain_crg: list = []
    prev_c = None  # index (in chain arrays) of the previous unit's carbonyl C
    prev_tor_o = 0.0

    def ok_pos(p, near=()):
        if np.linalg.norm(p) > core:
            return False
        if np.linalg.norm(p) < 5.5:
            return False
        return hgrid.min_dist(p, skip=near) >= 2.45 and grid.min_dist(p, skip=near) >= 2.2

    start = None
    for _t in range(1000):
        d = rng.normal(size=3)
        p = d / np.linalg.norm(d) * rng.uniform(8.0, core - 2.0)
        if ok_pos(p):
            start = p
            break
    unit = 0
    stuck = 0
    while unit < n_units:
        base = len(chain_xyz)
        placed_unit = False
        for _try in range(200):
            hp = []  # heavy positions for this unit
            # N0
            if prev_c is None:
                n0 = start
            else:
                pc = chain_xyz[prev_c]
                a0 = chain_xyz[chain_nb[prev_c][0]]
                b0 = chain_xyz[chain_nb[chain_nb[prev_c][0]][0]] if chain_nb[chain_nb[prev_c][0]] else a0 + 1
                n0 = _place(b0, a0, pc, 1.46, 116.0, prev_tor_o + 180.0)
            hp.append(n0)
            ref_a = chain_xyz[prev_c] if prev_c is not None else n0 + rng.normal(size=3)
            ref_b = (chain_xyz[chain_nb[prev_c][0]] if prev_c is not None else ref_a + rng.normal(size=3))
            c1 = _place(ref_b, ref_a, n0, 1.47, 111.0, rng.uniform(0, 360))
            c2 = _place(ref_a, n0, c1, 1.53, 111.0, rng.choice([60.0, 180.0, 300.0]) + rng.normal(0, 20))
            c3 = _place(n0, c1, c2, 1.52, 112.0, rng.choice([60.0, 180.0, 300.0]) + rng.normal(0, 20))
            tor_o = rng.uniform(0, 360)
            o4 = _place(c1, c2, c3, 1.46, 121.0, tor_o)
            c5 = _place(c2, n0, c1, 1.53, 110.0, 120.0 + rng.normal(0, 10))
            c6 = _place(n0, c1, c5, 1.53, 112.0, 180.0 + rng.normal(0, 20))
            hp += [c1, c2, c3, o4, c5, c6]
            skip = {("ch", prev_c)} if prev_c is not None else set()
            if not all(ok_pos(p, skip) for p in hp):
                continue
            close = {(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (0, 2), (1, 3), (2, 4), (0, 5),
                     (2, 5), (1, 6), (0, 6), (2, 6), (1, 4), (0, 3), (3, 5)}
            if any(np.linalg.norm(hp[i] - hp[j]) < 2.2 for i in range(7) for j in range(i + 1, 7)
                   if (i, j) not in close):
                continue
            # hydrogens on the tetrahedral directions left free by the heavy bonds
            hlist = []
            for par in range(7):
                nh = sum(1 for (pp, _t) in _UNIT_H if pp == par)
                if nh == 0:
                    continue
                ht = [t for (pp, t) in _UNIT_H if pp == par][0]
                nbp = [hp[v] for v in _UNIT_NB[par]]
                if par == 0 and prev_c is not None:
                    nbp.append(chain_xyz[prev_c])
                bl = 1.09
                for d in _tetra_dirs(hp[par], nbp, nh, rng):
                    hlist.append((par, ht, hp[par] + bl * d))
            hxyz = [h[2] for h in hlist]
            if any(grid.min_dist(h, skip=skip) < 2.0 for h in hxyz):
                continue
            if any(np.linalg.norm(hxyz[i] - hxyz[j]) < 1.5 for i in range(len(hxyz))
                   for j in range(i + 1, len(hxyz))):
                continue
            if any(np.linalg.norm(h - hp[k]) < 1.9 for h, (par, _t, _p) in zip(hxyz, hlist)
                   for k in range(7) if k != par and k not in _UNIT_NB[par]):
                continue
            placed_unit = True
            break
        if not placed_unit:
            stuck += 1
            if stuck > 400:
                raise RuntimeError("protein chain growth got stuck")
            # restart the walk from a new random point (chain break kept bonded
            # to nothing: a new chain segment)
            prev_c = None
            for _t in range(1000):
                d = rng.normal(size=3)
                p = d / np.linalg.norm(d) * rng.uniform(8.0, core - 2.0)
                if ok_pos(p):
                    start = p
                    break
            continue
        stuck = 0
        # heavy atoms
        unb = _UNIT_NB
        for k in range(7):
            chain_xyz.append(hp[k])
            chain_types.append(_UNIT_HEAVY[k])
            chain_crg.append(crg_t[k])
            chain_heavy.append(True)
            chain_nb.append([base + v for v in unb[k]])
            grid.add(hp[k], ("ch", base + k))
            hgrid.add(hp[k], ("ch", base + k))
        if prev_c is not None:
            chain_nb[base].append(prev_c)
            chain_nb[prev_c].append(base)
        for hcount, (par, ht, newh) in enumerate(hlist):
            pi = base + par
            idx = len(chain_xyz)
            chain_xyz.append(newh)
            chain_types.append(ht)
            chain_crg.append(crg_t[7 + hcount])
            chain_heavy.append(False)
            chain_nb.append([pi])
            chain_nb[pi].append(idx)
            grid.add(newh, ("ch", idx))
        prev_c = base + 3
        prev_tor_o = tor_o
        unit += 1

    off = lig_n + 1
    s.x += chain_xyz
    s.iac += chain_types
    s.crg += chain_crg
    b, a, t, i = _topology_from_graph(chain_nb, off, chain_heavy)
    s.bonds += b
    s.angles += a
    s.torsions += t
    s.impropers += i
    # charge groups: each heavy atom with its hydrogens; residues of 16 atoms
    for k in range(len(chain_xyz)):
        if chain_heavy[k]:
            members = [k + off] + [h + off for h in chain_nb[k] if not chain_heavy[h]]
            s.cgps.append((k + off, sorted(members)))
    for u in range(n_units):
        s.residues.append((off + 16 * u, "UNK"))
    s.molecules.append(off)
    s.nat_solute = s.natom
    s.nbonds_solute = len(s.bonds)
    s.nangles_solute = len(s.angles)
    s.ncgp_solute = len(s.cgps)
    _strain_free(s)
    _ligand_fep(rng, s, 1, lig_n, ltypes, n_pert, alpha, lnb, lheavy)
    first_water = s.natom + 1
    if fill:
        # a 3.0 A lattice, then the voids next to the chain: ~1700 waters in
        # the 25 A sphere around the 2600 solute atoms
        _add_water_lattice(rng, s, radius, grid, spacing=3.0, heavy_grid=hgrid)
        _fill_water_gaps(rng, s, radius, grid, hgrid, first_water)
    else:
        _add_water_lattice(rng, s, radius, grid, heavy_grid=hgrid)
    return s


# ---------------------------------------------------------------------------
# writers
# ---------------------------------------------------------------------------

def _neighbour_sets(s: System):
    n = s.nat_solute
    adj = [set() for _ in range(n + 1)]
    for (i, j, _c) in s.bonds[: s.nbonds_solute]:
        adj[i].add(j)
        adj[j].add(i)
    excl, one4 = set(), set()
    for i in range(1, n + 1):
        d1 = adj[i]
        d2 = set()
        for j in d1:
            d2 |= adj[j]
        d2 -= {i}
        d3 = set()
        for j in d2:
            d3 |= adj[j]
        d3 -= d1 | d2 | {i}
        for j in d1 | d2:
            excl.add((min(i, j), max(i, j)))
        for j in d3:
            one4.add((min(i, j), max(i, j)))
    one4 -= excl
    return excl, one4


def _bitmap_lines(pairs: set, n: int, rng_max: int = 25):
    """80i1 bitmap of (j-i) in 1..25 per solute atom plus the long-range pair list."""
    bits = np.zeros((n, rng_max), dtype=np.int8)
    long = []
    for (i, j) in sorted(pairs):
        if j - i <= rng_max:
            bits[i - 1, j - i - 1] = 1
        else:
            long.append((i, j))
    flat = "".join(str(int(v)) for v in bits.reshape(-1))
    lines = [flat[k:k + 80] for k in range(0, len(flat), 80)]
    return lines, long


def write_top(s: System, path: str) -> None:
    n = s.natom
    nsol = s.nat_solute
    out = []
    w = out.append
    w("Q topology file")
    w(f"TITLE      {s.name}")
    w("DATE       20260101")
    w("VERSION     5.06")
    w("FORCEFIELD synthetic-OPLS-like")
    w("FF_TYPE    2")
    w("PRM_FILE   synthetic")
    w("END        of header")
    w(f"{n:8d}{nsol:8d} = Total no. of atoms, no. of solute atoms. Coordinates: (2*3 per line)")
    xs = np.asarray(s.x, float).reshape(-1)
    for k in range(0, len(xs), 6):
        w("".join(f"{v:12.5f}" for v in xs[k:k + 6]))
    w(f"{n:8d} = No. of integer atom codes. iac's:")
    for k in range(0, n, 16):
        w("".join(f"{v:5d}" for v in s.iac[k:k + 16]))
    w(f"{len(s.bonds):8d}{s.nbonds_solute:8d} = No. of bonds, no. of solute bonds. i - j - icode: (5 per line)")
    for k in range(0, len(s.bonds), 5):
        w("".join(f" {b[0]:6d} {b[1]:6d} {b[2]:5d}" for b in s.bonds[k:k + 5]))
    w(f"{len(s.bond_lib):8d} = No. of bond codes. Parameters:")
    for c, (fk, r0) in enumerate(s.bond_lib, 1):
        w(f"{c:6d} {fk:10.3f} {r0:10.5f}")
    w(f"{len(s.angles):8d}{s.nangles_solute:8d} = No. of angles, no. of solute angles. i - j - k - icode: (3 per line)")
    for k in range(0, len(s.angles), 3):
        w("".join(f" {a[0]:6d} {a[1]:6d} {a[2]:6d} {a[3]:5d}" for a in s.angles[k:k + 3]))
    w(f"{len(s.angle_lib):8d} = No. of angle codes. Parameters:")
    for c, (fk, t0) in enumerate(s.angle_lib, 1):
        w(f"{c:6d} {fk:10.3f} {t0:10.3f}   0.000   0.00000")
    ntor = len(s.torsions)
    w(f"{ntor:8d}{ntor:8d} = No. of torsions, solute torsions. i - j - k - l - icode: (2 per line)")
    for k in range(0, ntor, 2):
        w("".join(f" {t[0]:6d} {t[1]:6d} {t[2]:6d} {t[3]:6d} {t[4]:5d}" for t in s.torsions[k:k + 2]))
    w(f"{len(TORSION_LIB):8d} = No. of torsion codes. Parameters:")
    for c, (fk, m, dl, p) in enumerate(TORSION_LIB, 1):
        w(f"{c:5d}{fk:10.4f}{m:8.3f}{dl:10.3f}{p:8.3f}")
    nimp = len(s.impropers)
    w(f"{nimp:8d}{nimp:8d} = No. of impropers, solute impr. i - j - k - l - icode: (2 per line)")
    for k in range(0, nimp, 2):
        w("".join(f" {t[0]:6d} {t[1]:6d} {t[2]:6d} {t[3]:6d} {t[4]:5d}" for t in s.impropers[k:k + 2]))
    w(f"{len(IMPROPER_LIB):8d}       2 = No. of improper codes, type (1=harmonic,2=periodic). Parameters:")
    for c, (fk, i0) in enumerate(IMPROPER_LIB, 1):
        w(f"{c:5d}{fk:10.4f}{i0:10.3f}")
    w(f"{n:8d} = No. of atomic charges")
    for k in range(0, n, 10):
        w("".join(f"{v:8.4f}" for v in s.crg[k:k + 10]))
    w(f"{len(s.cgps):8d}{s.ncgp_solute:8d}       1 = No. of charge groups, no of solvent cgps, switch atoms flag."
      " nat_cgp, iswitch / atom list:")
    for (sw, members) in s.cgps:
        w(f"{len(members):6d}{sw:6d}")
        w("".join(f"{m:6d}" for m in members))
    types = LJ_TYPES + (TIP3_TYPES if s.water == "tip3" else [])
    nty = len(types)
    w(f"{nty:8d} = No. of atom types")
    w(f"{s.vdw_rule:8d} = vdW combination rule (1 = Geom. / 2 = Arit.)")
    w(" 0.50000 332.0716 = Electrostatic 1-4 scaling factor and  Coulomb constant")
    w("Masses:")
    w("".join(f"{t[3]:9.4f}" for t in types))
    cols = [lj_columns(t[1], t[2], s.vdw_rule) for t in types]
    la, lb = ("sqrt (Aii)", "sqrt (Bii)") if s.vdw_rule == 1 else ("R*ii", "epsilon ii")
    for k, lab in enumerate(("normal", "polar", "1-4")):
        w(f"{la} {lab}:")
        w("".join(f"{c[k][0]:12.4f}" for c in cols))
        w(f"{lb} {lab}:")
        w("".join(f"{c[k][1]:12.5f}" for c in cols))
    w("       0 = No. of type-2 vdW interactions. pairs:")
    excl, one4 = _neighbour_sets(s)
    l14, long14 = _bitmap_lines(one4, nsol)
    lex, longex = _bitmap_lines(excl, nsol)
    w(f"{len(one4):8d} = No. of 1-4 neighbours. nborlist (range=max_nbr_range):")
    out += l14
    w(f"{len(long14):8d} = No. of long 1-4 nbrs (>max_nbr_range). pairlist:")
    for (i, j) in long14:
        w(f"{i:8d}{j:8d}")
    w(f"{len(excl):8d} = No. of exclusions. exclusion list (range=max_nbr_range):")
    out += lex
    w(f"{len(longex):8d} = No. of long exclusions (>max_nbr_range). pairlist:")
    for (i, j) in longex:
        w(f"{i:8d}{j:8d}")
    nres_sol = sum(1 for r in s.residues if r[0] <= nsol)
    w(f"{len(s.residues):8d}{nres_sol:8d} = No. of residues, No of solute residues. start atoms:")
    starts = [r[0] for r in s.residues]
    for k in range(0, len(starts), 13):
        w("".join(f"{v:6d}" for v in starts[k:k + 13]))
    w("Sequence:")
    names = [r[1] for r in s.residues]
    for k in range(0, len(names), 16):
        w("".join(f"{v:<4s} " for v in names[k:k + 16]))
    w(f"{len(s.molecules):8d} = No. of separate molecules. start atoms:")
    for k in range(0, len(s.molecules), 13):
        w("".join(f"{v:6d}" for v in s.molecules[k:k + 13]))
    w(f"{nty:8d} = No. of atom types:")
    tn = [t[0] for t in types]
    for k in range(0, nty, 8):
        w("".join(f"{v:<8s} " for v in tn[k:k + 8]))
    w(f"{nty:8d} = No. of SYBYL atom types:")
    for k in range(0, nty, 13):
        w("".join(f"{'C.3':<5s} " for _v in tn[k:k + 13]))
    w(f"{1 if s.water == 'tip3' else 0:8d} = solvent type (0=SPC,1=3-atom,2=general)")
    w(f"{s.radius:8.3f}{s.radius:8.3f} = Exclusion, solvent radii")
    w("   0.000   0.000   0.000 = Solute centre")
    w("   0.000   0.000   0.000 = Solvent centre")
    w("       0       0 = No. of excluded atoms (incl. water), no. of excluded waters")
    flags = "F" * n
    for k in range(0, n, 80):
        w(flags[k:k + 80])
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_fep(s: System, path: str) -> None:
    if not s.qatoms:
        return
    fx = s.fep_extra
    out = ["!synthetic two-state perturbation", "[FEP]", "states 2",
           f"softcore_use_max_potential {'on' if s.sc_max_potential else 'off'}"]
    if fx.get("qq_use_library_charges"):
        out.append("qq_use_library_charges on")
    out += ["", "[atoms]"]
    for k, a in enumerate(s.qatoms, 1):
        out.append(f"{k:5d}{a:6d}")
    out += ["", "[change_charges]"]
    for k, (q1, q2) in enumerate(s.qcharges, 1):
        out.append(f"{k:5d}{q1:10.4f}{q2:10.4f}")
    out += ["", "[atom_types]"]
    for (nm, sig, eps, mass) in LJ_TYPES:
        (a, b), (ap, bp), (a14, b14) = lj_columns(sig, eps, s.vdw_rule)
        if "exp_repulsion" in fx:
            # the second column of q types gives soft pairs their repulsion
            # A exp(-B r) (nonbond_qq, md.f90:10838-10842)
            ap, bp = fx["exp_repulsion"]
        out.append(f"{nm:<6s}{a:12.4f}{b:10.5f}{ap:12.4f}{bp:10.5f}{a14:12.4f}{b14:10.5f}{mass:9.4f}")
    out += ["", "[change_atoms]"]
    for k, (t1, t2) in enumerate(s.qtypes, 1):
        out.append(f"{k:5d}  {t1:<6s}{t2:<6s}")
    out += ["", "[softcore]"]
    for k, (a1, a2) in enumerate(s.qalpha, 1):
        out.append(f"{k:5d}{a1:8.2f}{a2:8.2f}")
    for sec, key, fmt in (("soft_pairs", "soft_pairs", "{:5d}{:5d}"),
                          ("el_scale", "el_scale", "{:5d}{:5d}{:8.3f}{:8.3f}"),
                          ("excluded_pairs", "excluded_pairs", "{:6d}{:6d}{:4d}{:4d}")):
        if fx.get(key):
            out += ["", f"[{sec}]"] + [fmt.format(*r) for r in fx[key]]
    if s.qbonds:
        out += ["", "[bond_types]"]
        for c, par in enumerate(s.qbond_lib, 1):
            if len(par) == 3:   # Morse: E_diss, alpha, b0
                out.append(f"{c:6d} {par[0]:10.2f} {par[1]:10.4f} {par[2]:10.5f}")
            else:
                out.append(f"{c:6d} {par[0]:10.2f} {par[1]:10.5f}")
        out += ["", "[change_bonds]"]
        for (i, j, c1, c2) in s.qbonds:
            out.append(f"{i:6d} {j:6d} {c1:5d} {c2:5d}")
    if s.qangles:
        out += ["", "[angle_types]"]
        for c, (fk, t0) in enumerate(s.qangle_lib, 1):
            out.append(f"{c:6d} {fk:10.2f} {t0:10.3f}")
        out += ["", "[change_angles]"]
        for (i, j, k, c1, c2) in s.qangles:
            out.append(f"{i:6d} {j:6d} {k:6d} {c1:5d} {c2:5d}")
    if s.qtorsions:
        out += ["", "[torsion_types]"]
        out.append(f"{1:4d}{0.60:10.3f}{1.0:8.2f}{0.0:10.3f}")
        out.append(f"{2:4d}{0.30:10.3f}{3.0:8.2f}{180.0:10.3f}")
        out += ["", "[change_torsions]"]
        for (i, j, k, l, c1, c2) in s.qtorsions:
            out.append(f"{i:6d}{j:6d}{k:6d}{l:6d}{c1:4d}{c2:4d}")
    if fx.get("shake"):
        out += ["", "[shake_constraints]"] + [f"{i:6d}{j:6d}{d1:10.4f}{d2:10.4f}" for (i, j, d1, d2) in fx["shake"]]
    if s.offdiag:
        out += ["", "[off_diagonals]"]
        for (i, j, k, l, a, mu) in s.offdiag:
            out.append(f"{i:3d}{j:3d}{k:5d}{l:5d}{a:10.3f}{mu:8.3f}")
    if s.qimpropers:
        out += ["", "[improper_types]"]
        out.append(f"{1:4d}{10.5:10.3f}{0.0:10.3f}")
        out += ["", "[change_impropers]"]
        for (i, j, k, l, c1, c2) in s.qimpropers:
            out.append(f"{i:6d}{j:6d}{k:6d}{l:6d}{c1:4d}{c2:4d}")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_inp(path: str, top: str, fep: str | None, final: str, *, steps: int, stepsize: float = 1.0,
              temperature: float = 300.0, seed: int = 1, lambdas: tuple | None = None, lrf: bool = True,
              rc_solute: float = 11.0, rc_solvent: float = 10.0, rc_q: float = 99.0, rc_lrf: float = 99.0,
              shake_solvent: bool = True, shake_hydrogens: bool = False, shake_solute: bool = False,
              nb_cycle: int = 25,
              output: int = 5, energy: int = 10, energy_file: str | None = None,
              restart: str | None = None, bath_coupling: float = 10.0,
              separate_scaling: bool = True, polarization: bool = True, trajectory: int = 0,
              trajectory_file: str | None = None, trajectory_atoms: list | None = None,
              extra_lines: list | None = None) -> None:
    onoff = lambda v: "on" if v else "off"  # noqa: E731
    out = ["[MD]", f"steps {steps}", f"stepsize {stepsize}", f"temperature {temperature}",
           f"bath_coupling {bath_coupling}"]
    if restart is None:
        out += [f"random_seed {seed}", f"initial_temperature {temperature}"]
    out += [f"shake_solvent {onoff(shake_solvent)}", f"shake_hydrogens {onoff(shake_hydrogens)}",
            f"shake_solute {onoff(shake_solute)}", f"lrf {onoff(lrf)}", f"separate_scaling {onoff(separate_scaling)}", "",
            "[cut-offs]", f"solute_solute {rc_solute}", f"solute_solvent {rc_solute}",
            f"solvent_solvent {rc_solvent}", f"q_atom {rc_q}"]
    if lrf:
        out.append(f"lrf {rc_lrf}")
    out += ["", "[sphere]", "shell_force 10.0", "shell_radius 0.85", "",
            "[solvent]", "radial_force 60.0", f"polarization {onoff(polarization)}",
            "polarization_force 20.0", "",
            "[intervals]", f"output {output}", f"non_bond {nb_cycle}"]
    if energy_file:
        out.append(f"energy {energy}")
    if trajectory_file:
        out.append(f"trajectory {trajectory}")
    out += ["", "[files]", f"topology {top}", f"final {final}"]
    if energy_file:
        out.append(f"energy {energy_file}")
    if trajectory_file:
        out.append(f"trajectory {trajectory_file}")
    if restart:
        out.append(f"restart {restart}")
    if fep:
        out.append(f"fep {fep}")
    if lambdas:
        out += ["", "[lambdas]", " ".join(f"{v:.6f}" for v in lambdas)]
    if trajectory_file and trajectory_atoms:
        out += ["", "[trajectory_atoms]"] + list(trajectory_atoms)
    if extra_lines:
        out += list(extra_lines)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


CONFIGS = {
    # name: (builder, builder kwargs, inp kwargs)
    "water": (water_droplet, {"radius": 15.0}, {"lrf": False, "rc_solute": 10.0}),
    "ligwat": (ligand_in_water, {"radius": 20.0}, {"lrf": True}),

    "protein": (protein_leg, {"radius": 25.0}, {"lrf": True, "shake_hydrogens": True}),
    # soft-core given as maximum potentials (qatom.f90:1600-1664), under
    # either LJ combination rule
    "ligmaxpot": (ligand_in_water, {"radius": 20.0, "sc_max_potential": True}, {"lrf": True}),
    "ligarithmax": (ligand_in_water, {"radius": 20.0, "vdw_rule": 2, "sc_max_potential": True}, {"lrf": True}),
    # 3-atom solvent with LJ on the hydrogens (nonbond_3atomsolvent,
    # nonbond_qw_3atom, md.f90:12317-12591, 11545-11670)
    "ligtip3": (ligand_in_water, {"radius": 20.0, "water": "tip3"}, {"lrf": True}),
    # EVB off-diagonal functions Hij = A exp(-mu r_kl) (md.f90:12911-12937)
    "ligevb": (ligand_in_water, {"radius": 20.0, "offdiag": ((1, 2, 1, 5, 25.0, 0.5), (2, 1, 3, 7, 10.0, 0.2))},
               {"lrf": True}),
    # the protein leg with the arithmetic LJ combination rule
    "protarith": (protein_leg, {"radius": 25.0, "vdw_rule": 2}, {"lrf": True, "shake_hydrogens": True}),
    # the ligand in water with the further FEP sections: el_scale,
    # soft_pairs, excluded_pairs, a Morse q-bond, shake_constraints
    "ligfepx": (ligand_in_water, {"radius": 20.0, "extras": ("fep",)}, {"lrf": True}),
    # library charges for q-q pairs under either LJ rule
    # (nonbond_qq_lib_charges / nonbon2_qq_lib_charges)
    "liglibq": (ligand_in_water, {"radius": 20.0, "extras": ("libq",)}, {"lrf": True}),
    "ligarithlibq": (ligand_in_water, {"radius": 20.0, "vdw_rule": 2, "extras": ("libq",)}, {"lrf": True}),
    # sequence, position, distance, angle and wall restraints (p_restrain)
    "ligrestr": (ligand_in_water, {"radius": 20.0, "extras": ("restr",)}, {"lrf": True}),
    # a QligFEP dual-topology perturbation: both ligands, all atoms q-atoms,
    # dummies of the zero-LJ type, soft-core maximum potentials 0 / 20
    "ligdual": (ligand_dual, {"radius": 20.0}, {"lrf": True}),
    # Q-atom lists rebuilt at every list update: a q_atom cut-off within the
    # sphere's exclusion radius (nbqplist / nbqwlist, md.f90:7903, 8086)
    "ligqcut": (ligand_in_water, {"radius": 20.0}, {"lrf": True, "rc_q": 9.0}),
    "protqcut": (protein_leg, {"radius": 25.0}, {"lrf": True, "shake_hydrogens": True, "rc_q": 12.0}),
    # every solute bond constrained (shake_solute): the ligand is one SHAKE
    # component of some fifty constraints
    "ligshake": (ligand_in_water, {"radius": 20.0}, {"lrf": True, "shake_solute": True}),
}


def make_window(config: str, outdir: str, *, seed: int = 0, md_seed: int = 1, lam: float = 0.5,
                steps: int = 100, energy: int = 10, output: int = 5, nb_cycle: int = 25,
                restart: str | None = None, tag: str | None = None, system: System | None = None,
                **inp_overrides) -> str:
    """Write topology/fep/inp of one window of ``config`` into ``outdir``;
    return the path of the .inp file.  ``seed`` seeds the system builder,
    ``md_seed`` the Maxwell velocities of qdyn."""
    builder, bkw, ikw = CONFIGS[config]
    os.makedirs(outdir, exist_ok=True)
    tag = tag or f"{config}_s{seed}"
    top = f"{config}_s{seed}.top"
    fep = f"{config}_s{seed}.fep"
    if system is None and not os.path.exists(os.path.join(outdir, top)):
        system = builder(seed=seed, **bkw)
    # restraint sections of the system's .inp files, kept beside the topology
    rst = os.path.join(outdir, f"{config}_s{seed}.rst")
    if system is not None:
        write_top(system, os.path.join(outdir, top))
        write_fep(system, os.path.join(outdir, fep))
        if system.inp_extra:
            with open(rst, "w") as f:
                f.write("\n".join(system.inp_extra) + "\n")
    extra = open(rst).read().splitlines() if os.path.exists(rst) else None
    has_fep = os.path.exists(os.path.join(outdir, fep))
    kw = dict(ikw)
    kw.update(inp_overrides)
    inp = os.path.join(outdir, f"{tag}.inp")
    write_inp(inp, top, fep if has_fep else None, f"{tag}.re", steps=steps, seed=md_seed,
              lambdas=(1.0 - lam, lam) if has_fep else None, nb_cycle=nb_cycle, output=output,
              energy=energy, energy_file=f"{tag}.en" if has_fep else None, restart=restart,
              extra_lines=extra, **kw)
    return inp
