"""Free energies from per-state energy files: the qfep analysis.

Reads the Fortran-unformatted energy files that qdyn (and this engine) write
-- per energy step one record (state number, 15 doubles of type q_energies,
nrgy.f90:51-58) for each state and one off-diagonal record (nrgy.f90:94-119)
-- and computes what qfep prints (src/q6/qfep.f90):

  Part 0  averages per file and state                   (qfep.f90:205-380)
  Part 1  Zwanzig forward/reverse sums and <dG>         (382-444, 594-600)
  Part 2/3 reaction free energy on the energy gap       (606-722)
  Part 4  thermodynamic integration                     (447-474)
  Part 5  overlap sampling (Lu, Kofke, Woolf 2004)      (477-520)
  Part 6  Bennett acceptance ratio, iterated to 1e-3    (523-583)

Sums run in the reference's order in double precision, so the printed values
agree with qfep's.  EVB: off-diagonal functions given to qfep (i, j, A, mu,
eta, r0: Hij = A exp(-mu (r - r0)) exp(-eta (r - r0)^2), r the first r_kl of
the energy file's off-diagonal record) set the ground-state energy and the
c1**2, c2**2 and <r_xy> columns (qfep.f90:155-170, 275-322, 606-722).
Off-diagonal elements taken from the files (a non-zero second number on the
states line) are refused: the reference reads none of them in that case.

Command line, reading qfep's own stdin answers:
    python -m q_amd.qfep < qfep.inp
"""
from __future__ import annotations

import math
import sys
from dataclasses import dataclass, field

import numpy as np

# field order of one q_energies record (nrgy.f90:51-58)
Q_FIELDS = ["lambda", "total", "q.bond", "q.angle", "q.torsion", "q.improper", "qx.el", "qx.vdw",
            "qq.el", "qq.vdw", "qp.el", "qp.vdw", "qw.el", "qw.vdw", "restraint"]


class QfepError(RuntimeError):
    pass


def read_en(path: str, nstates: int, noffd: int = 0):
    """All records of an energy file as an array (npts, nstates, 15); with
    ``noffd`` > 0 also the r_kl of the first ``noffd`` entries of each
    off-diagonal record, (npts, noffd) (get_ene, nrgy.f90:123-160)."""
    raw = open(path, "rb").read()
    out, rkl, k = [], [], 0
    n_raw = len(raw)
    while k < n_raw:
        block = []
        for s in range(nstates):
            if k + 4 > n_raw:
                raise QfepError(f"{path}: unexpected end of file in state {s + 1}")
            n = int(np.frombuffer(raw, np.int32, 1, k)[0])
            if n < 4 + 15 * 8 or k + 8 + n > n_raw:
                raise QfepError(f"{path}: bad record for state {s + 1} (maybe there are only {s} states?)")
            block.append(np.frombuffer(raw, np.float64, 15, k + 8))
            k += 8 + n
        if k + 4 > n_raw:
            raise QfepError(f"{path}: missing off-diagonal record")
        n = int(np.frombuffer(raw, np.int32, 1, k)[0])
        if noffd:
            if n < 24 * noffd:
                raise QfepError(f"{path}: maybe there are less than {noffd} off-diagonal elements?")
            rkl.append([float(np.frombuffer(raw, np.float64, 1, k + 4 + 24 * m + 16)[0]) for m in range(noffd)])
        k += 8 + n
        out.append(np.array(block))
    E = np.array(out).reshape(-1, nstates, 15)
    return (E, np.array(rkl).reshape(-1, noffd)) if noffd else E


@dataclass
class FepFile:
    name: str
    lam: np.ndarray          # (nstates,)
    v: np.ndarray            # (npts, nstates): EQ total + alpha
    vg: np.ndarray           # (npts,): ground state energy
    gap: np.ndarray          # (npts,)
    avg: np.ndarray          # (nstates, 15) averages over the points kept
    c1: np.ndarray = None    # (npts,): ground-state weights (EVB, two states)
    c2: np.ndarray = None
    r: np.ndarray = None     # (npts,): first r_kl of the file (EVB)


@dataclass
class QfepResult:
    files: list
    dgf: np.ndarray = field(default=None)
    dgfsum: np.ndarray = field(default=None)
    dgr: np.ndarray = field(default=None)
    dgrsum: np.ndarray = field(default=None)
    dG: np.ndarray = field(default=None)
    dgti: np.ndarray = field(default=None)
    dgtisum: np.ndarray = field(default=None)
    dglu: np.ndarray = field(default=None)
    dglusum: np.ndarray = field(default=None)
    dgbar: np.ndarray = field(default=None)
    dgbarsum: np.ndarray = field(default=None)
    gapmin: float = 999.0
    gapmax: float = -999.0
    bins: list = field(default_factory=list)      # Part 2 rows
    binavg: list = field(default_factory=list)    # Part 3 rows

    @property
    def dG_bar(self) -> float:
        """Total BAR free energy from the first to the last file."""
        return float(self.dgbarsum[len(self.files)])

    @property
    def dG_fep(self) -> float:
        """Total Zwanzig <dG> (average of forward and reverse)."""
        return float(self.dG[len(self.files)])


def _veff(lam, v):
    """sum over states of lambda_s * v_s, accumulated in state order"""
    acc = np.zeros(v.shape[0])
    for s in range(v.shape[1]):
        acc = acc + lam[s] * v[:, s]
    return acc


def _hij(funcs, nstates, r):
    """Off-diagonal matrix elements from the functions at distances r
    (qfep.f90:284-300): (i, j) takes A(i,j), or A(j,i) where A(i,j) is 0."""
    A = np.zeros((nstates, nstates))
    mu, eta, r0 = np.zeros_like(A), np.zeros_like(A), np.zeros_like(A)
    for (i, j, a, m, et, x0) in funcs:
        A[i - 1, j - 1], mu[i - 1, j - 1], eta[i - 1, j - 1], r0[i - 1, j - 1] = a, m, et, x0
    H = np.zeros((len(r), nstates, nstates))
    for i in range(nstates):
        for j in range(nstates):
            if i == j:
                continue
            a, b = (j, i) if A[i, j] == 0.0 else (i, j)
            H[:, i, j] = A[a, b] * np.exp(-mu[a, b] * (r - r0[a, b])) * np.exp(-eta[a, b] * (r - r0[a, b]) ** 2)
    return H


def qfep(paths, nstates: int, kT: float, nskip: int = 0, nbins: int = 50, nmin: int = 10,
         alpha=None, coeff=None, gas: int = 0, offdiag=None) -> QfepResult:
    """Run the qfep analysis on energy files `paths` (ordered along lambda).
    ``offdiag``: EVB off-diagonal functions (i, j, A, mu, eta, r0)."""
    nfiles = len(paths)
    funcs = list(offdiag or [])
    alfa = np.zeros(nstates)
    if alpha is not None:
        alfa[1:] = np.asarray(alpha, float)[: nstates - 1]
    coeff = np.asarray(coeff if coeff is not None else ([1.0, -1.0] + [0.0] * (nstates - 2))[:nstates], float)
    files = []
    gapmin, gapmax = 999.0, -999.0
    for p in paths:
        if funcs:
            rec, rkl = read_en(p, nstates, len(funcs))
        else:
            rec, rkl = read_en(p, nstates), None
        if len(rec) == 0:
            raise QfepError(f"{p}: no energies")
        lam = rec[0, :, 0].copy()
        E = rec.copy()
        if gas > 0:
            # total = q bonded + qq + restraint (qfep.f90:262-270)
            E[:, :, 1] = E[:, :, 2] + E[:, :, 3] + E[:, :, 4] + E[:, :, 5] + E[:, :, 8] + E[:, :, 9] + E[:, :, 14]
        npts = len(E)
        if npts <= nskip:
            raise QfepError(f"number of data sets in {p} is less than number of points to skip")
        avg = E[nskip:].sum(axis=0) * (1.0 / (npts - nskip))
        v = E[:, :, 1] + alfa[None, :]
        # ground state of the secular determinant (qfep.f90:284-322): the
        # diagonal is v, the off-diagonal the functions of the first r_kl
        H = _hij(funcs, nstates, rkl[:, 0] if funcs else np.zeros(len(v)))
        c1 = np.zeros(len(v))
        c2 = np.zeros(len(v))
        r1 = rkl[:, 0].copy() if funcs else None
        if nstates == 2:
            h12 = H[:, 0, 1]
            vg = 0.5 * (v[:, 0] + v[:, 1]) - 0.5 * np.sqrt((v[:, 0] - v[:, 1]) ** 2 + 4. * h12 ** 2)
            if funcs:
                with np.errstate(divide="ignore", invalid="ignore"):
                    c1 = 1. / (1. + ((vg - v[:, 0]) / h12) ** 2)
                c2 = 1 - c1
        elif funcs:
            for s in range(nstates):
                H[:, s, s] = v[:, s]
            vg = np.linalg.eigvalsh(H, UPLO="L").min(axis=1)
        else:
            vg = v.min(axis=1)
        gap = np.zeros(npts)
        for s in range(nstates):
            gap = gap + v[:, s] * coeff[s]
        if npts > nskip:
            gapmin = min(gapmin, float(gap[nskip:].min()))
            gapmax = max(gapmax, float(gap[nskip:].max()))
        files.append(FepFile(p, lam, v, vg, gap, avg, c1, c2, r1))
    R = QfepResult(files=files, gapmin=gapmin, gapmax=gapmax)
    if nfiles < 2:
        return R
    rt = kT
    n = nfiles
    # Zwanzig forward / reverse (qfep.f90:382-444)
    dgf = np.zeros(n + 2)
    dgfsum = np.zeros(n + 2)
    for i in range(1, n):
        F, G = files[i - 1], files[i]
        v = F.v[nskip:]
        dv = _veff(G.lam, v) - _veff(F.lam, v)
        ssum = 0.0
        for x in dv:
            ssum += math.exp(-x / rt)
        ssum /= float(len(v))
        dgf[i] = -rt * math.log(ssum)
        dgfsum[i + 1] = dgfsum[i] + dgf[i]
    dgr = np.zeros(n + 2)
    dgrsum = np.zeros(n + 2)
    for i in range(n, 1, -1):
        F, G = files[i - 1], files[i - 2]
        v = F.v[nskip:]
        dv = _veff(G.lam, v) - _veff(F.lam, v)
        ssum = 0.0
        for x in dv:
            ssum += math.exp(-x / rt)
        ssum /= float(len(v))
        dgr[i] = -rt * math.log(ssum)
        dgrsum[i - 1] = dgrsum[i] + dgr[i]
    # thermodynamic integration along lambda of state 1 (qfep.f90:447-474)
    dgti = np.zeros(n + 2)
    dgtisum = np.zeros(n + 2)
    for i in range(1, n + 1):
        v = files[i - 1].v[nskip:]
        ssum = 0.0
        for x in v[:, 1] - v[:, 0]:
            ssum += x
        dgti[i] = ssum / float(len(v))
    for i in range(2, n + 1):
        dlam = files[i - 2].lam[0] - files[i - 1].lam[0]
        dgtisum[i] = dgtisum[i - 1] + (0.5 * (dgti[i - 1] + dgti[i]) * dlam)
    # overlap sampling (qfep.f90:477-520)
    dglu = np.zeros(n + 2)
    dglusum = np.zeros(n + 2)
    for i in range(1, n):
        F, G = files[i - 1], files[i]
        vf = F.v[nskip:]
        dvf = (_veff(G.lam, vf) - _veff(F.lam, vf)) / 2
        sf = 0.0
        for x in dvf:
            sf += math.exp(-x / rt)
        sumf = sf / float(len(vf))
        vb = G.v[nskip:]
        dvb = (_veff(G.lam, vb) - _veff(F.lam, vb)) / 2
        sb = 0.0
        for x in dvb:
            sb += math.exp(x / rt)
        sumb = sb / float(len(vb))
        dglu[i] = -rt * math.log(sumf / sumb)
        dglusum[i + 1] = dglusum[i] + dglu[i]
    # Bennett acceptance ratio, self-consistent from the overlap estimate (523-583)
    dgbar = np.zeros(n + 2)
    dgbarsum = np.zeros(n + 2)
    for i in range(1, n):
        F, G = files[i - 1], files[i]
        vf, vb = F.v[nskip:], G.v[nskip:]
        dvf = _veff(G.lam, vf) - _veff(F.lam, vf)
        dvb = _veff(G.lam, vb) - _veff(F.lam, vb)
        nfnr = float(len(vf)) / float(len(vb))
        nrnf = float(len(vb)) / float(len(vf))
        konst = dglu[i]
        fel = 1.0
        it = 0
        while fel > 0.001:
            sf = 0.0
            for x in dvf:
                sf += 1 / ((1 + (nfnr * math.exp((x - konst) / rt))))
            sumf = sf / float(len(vf))
            sb = 0.0
            for x in dvb:
                sb += 1 / ((1 + (nrnf * math.exp((-x + konst) / rt))))
            sumb = sb / float(len(vb))
            dgbar[i] = -rt * math.log((sumf / sumb) * math.exp(-konst / rt) * nfnr)
            fel = abs(konst - dgbar[i])
            konst = dgbar[i]
            it += 1
            if it > 100000:
                raise QfepError("BAR iteration did not converge")
        dgbarsum[i + 1] = dgbarsum[i] + dgbar[i]
    dG = np.zeros(n + 2)
    for i in range(2, n + 1):
        dG[i] = dG[i - 1] + 0.5 * (dgf[i - 1] - dgr[i])
    R.dgf, R.dgfsum, R.dgr, R.dgrsum, R.dG = dgf, dgfsum, dgr, dgrsum, dG
    R.dgti, R.dgtisum, R.dglu, R.dglusum, R.dgbar, R.dgbarsum = dgti, dgtisum, dglu, dglusum, dgbar, dgbarsum
    _reaction_free_energy(R, nstates, rt, nskip, nbins, nmin)
    return R


def _reaction_free_energy(R: QfepResult, nstates, rt, nskip, nbins, nmin):
    """Parts 2 and 3 (qfep.f90:606-722), no off-diagonal couplings."""
    gaprange = R.gapmax - R.gapmin
    xint = gaprange / float(nbins)
    binsum = np.zeros((nbins + 2, 4))
    evb = R.files[0].r is not None
    ptsum = np.zeros(nbins + 2, dtype=np.int64)
    for fi, F in enumerate(R.files):
        avdvv = np.zeros((nstates, nbins + 2))
        avdvg = np.zeros(nbins + 2)
        nbinpts = np.zeros(nbins + 2, dtype=np.int64)
        sumv = np.zeros((nstates, nbins + 2))
        sumg = np.zeros(nbins + 2)
        avc1 = np.zeros(nbins + 2)
        avc2 = np.zeros(nbins + 2)
        avr = np.zeros(nbins + 2)
        kept = range(nskip, len(F.v))
        ibins = []
        for ip in kept:
            ib = int((F.gap[ip] - R.gapmin) / xint) + 1 if xint > 0 else 1
            ibins.append(ib)
            veff = 0.0
            for s in range(nstates):
                veff += F.lam[s] * F.v[ip, s]
            avdvv[:, ib] += F.v[ip, :] - veff
            avdvg[ib] += F.vg[ip] - veff
            avc1[ib] += F.c1[ip]
            avc2[ib] += F.c2[ip]
            if evb:
                avr[ib] += F.r[ip]
            nbinpts[ib] += 1
        for ib in range(1, nbins + 1):
            if nbinpts[ib] != 0:
                avc1[ib] /= float(nbinpts[ib])
                avc2[ib] /= float(nbinpts[ib])
                avr[ib] /= float(nbinpts[ib])
                avdvv[:, ib] /= nbinpts[ib]
                avdvg[ib] /= nbinpts[ib]
        for ip, ib in zip(kept, ibins):
            veff = 0.0
            for s in range(nstates):
                veff += F.lam[s] * F.v[ip, s]
            dvv = F.v[ip, :] - veff - avdvv[:, ib]
            dvg = F.vg[ip] - veff - avdvg[ib]
            sumv[:, ib] += np.exp(-dvv / rt)
            sumg[ib] += math.exp(-dvg / rt)
        for ib in range(1, nbins + 1):
            if nbinpts[ib] >= nmin:
                binsum[ib, 1] += avc1[ib] * nbinpts[ib]
                binsum[ib, 2] += avc2[ib] * nbinpts[ib]
                binsum[ib, 3] += avr[ib] * nbinpts[ib]
                sumv[:, ib] /= float(nbinpts[ib])
                sumg[ib] /= float(nbinpts[ib])
                ptsum[ib] += nbinpts[ib]
                sumv2 = -rt * np.log(sumv[:, ib]) + avdvv[:, ib]
                sumg2 = -rt * math.log(sumg[ib]) + avdvg[ib]
                dGv = R.dG[fi + 1] + sumv2
                dGg = R.dG[fi + 1] + sumg2
                binsum[ib, 0] += dGg * int(nbinpts[ib])
                R.bins.append((F.lam[0], ib, R.gapmin + ib * xint - xint / 2., dGv[0], dGv[1] if nstates > 1 else 0.0,
                               dGg, int(nbinpts[ib]), avc1[ib], avc2[ib]))
    valid = [ib for ib in range(1, nbins + 1) if ptsum[ib] >= nmin]
    for ib in valid:
        binsum[ib, :] /= float(ptsum[ib])
    mn = binsum[1:nbins + 1, 0].min()
    for ib in valid:
        R.binavg.append((ib, R.gapmin + ib * xint - xint / 2., binsum[ib, 0], binsum[ib, 0] - mn, int(ptsum[ib]),
                         binsum[ib, 1], binsum[ib, 2], binsum[ib, 3]))


def report(R: QfepResult, nstates: int) -> str:
    """The Part 0-6 tables in qfep's layout."""
    L = []
    L.append("# Part 0: Average energies for all states in all files")
    for F in R.files:
        for s in range(nstates):
            a = F.avg[s]
            L.append(f"{F.name:<22s}{s + 1:2d} {len(F.v):6d} {F.lam[s]:8.6f}" +
                     "".join(f"{a[k]:8.2f}" for k in range(1, 15)))
    n = len(R.files)
    if n > 1:
        L.append("# Part 1: Free energy perturbation summary:")
        L.append("# lambda(1)      dGf sum(dGf)      dGr sum(dGr)     <dG>")
        for i in range(1, n + 1):
            L.append(f"   {R.files[i - 1].lam[0]:8.6f}{R.dgf[i - 1]:9.3f}{R.dgfsum[i]:9.3f}{R.dgr[i + 1]:9.3f}"
                     f"{R.dgrsum[i]:9.3f}{R.dG[i]:9.3f}")
        L.append(f"# Min energy-gap is: {R.gapmin:9.2f}")
        L.append(f"# Max energy-gap is: {R.gapmax:9.2f}")
        L.append("# Part 2: Reaction free energy summary:")
        for row in R.bins:
            L.append(f"   {row[0]:8.6f}{row[1]:5d}  {row[2]:9.2f}{row[3]:9.2f}{row[4]:9.2f}{row[5]:9.2f}  {row[6]:5d}"
                     f"{row[7]:9.3f}{row[8]:9.3f}")
        L.append("# Part 3: Bin-averaged summary:")
        for row in R.binavg:
            L.append(f"{row[0]:4d} {row[1]:9.2f}{row[2]:9.2f}{row[3]:9.2f}  {row[4]:5d}{row[5]:8.3f}{row[6]:8.3f}"
                     f"{row[7]:8.3f}")
        for title, part, psum in (("# Part 4: Termodynamic integration:", R.dgti, R.dgtisum),
                                  ("# Part 5: Overlap sampling Lu et al:", R.dglu, R.dglusum),
                                  ("# Part 6: BAR Bennet:", R.dgbar, R.dgbarsum)):
            L.append(title)
            for i in range(1, n + 1):
                L.append(f"   {R.files[i - 1].lam[0]:8.6f}{part[i - 1]:9.3f}{psum[i]:9.3f}")
    return "\n".join(L) + "\n"


def _tokens(stream):
    for line in stream:
        t = line.split("!")[0].split("#")[0].split()
        if t:
            yield t


def main(argv=None):
    """Read qfep's stdin answers and print its report."""
    toks = _tokens(sys.stdin)
    nfiles = int(next(toks)[0])
    t = next(toks)
    nstates, noffd = int(t[0]), int(t[1])
    if noffd != 0:
        raise QfepError("off-diagonal elements from the energy files are not read by qfep (qfep.f90:251, 279): "
                        "give 0 and the off-diagonal functions instead")
    t = next(toks)
    kT, nskip = float(t[0]), int(t[1])
    gas = int(t[2]) if len(t) > 2 else 0
    nbins = int(next(toks)[0])
    nmin = int(next(toks)[0])
    alpha = [float(next(toks)[0]) for _ in range(2, nstates + 1)]
    nnoffd = int(next(toks)[0])
    offdiag = []
    for _ in range(nnoffd):
        t = next(toks)
        offdiag.append((int(t[0]), int(t[1])) + tuple(float(v) for v in t[2:6]))
    coeff = [float(v) for v in next(toks)[:nstates]]
    paths = [next(toks)[0] for _ in range(nfiles)]
    R = qfep(paths, nstates, kT, nskip, nbins, nmin, alpha, coeff, gas, offdiag)
    sys.stdout.write(report(R, nstates))


if __name__ == "__main__":
    main()
