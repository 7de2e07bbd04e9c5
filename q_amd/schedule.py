"""Learned lambda schedule for FEP windows.

The reference spaces lambda windows by hand (QligFEP writes a fixed uniform
ladder).  Here short pilot windows, run by the engine on an 11-point lambda
grid, give the energy gap dE = E(state 2) - E(state 1).  A small MLP
(q_amd.mlp, trained and run on the GPU by the library's own kernels) maps the
histogram of the reduced work u = 0.1 * (dE - <dE>) / kT seen in one pilot
window, with lambda and the soft-core alpha, to the BAR overlap of that
window with its grid neighbour 0.1 away.  Each predicted overlap is turned
into the local spread of the work (the width of Gaussian work with that
overlap), which integrates to a thermodynamic length along lambda; the
proposed ladder puts windows at equal steps of that length, with as many
windows (16 to the time budget) as the predicted neighbour overlap needs.
When the predicted interval overlaps disagree with the ones the pilots
measure by more than FALLBACK_DISAGREEMENT on average, the uniform ladder
with as many windows runs instead (the fallback).
The free energy of the ladder (BAR, by q_amd.qfep) is judged against a
uniform 51-window ladder; the uniform ladder with as many windows is the
method without the network.

    python -m q_amd.schedule build --out DIR     pilot data, training, evaluation
"""
from __future__ import annotations

import argparse
import json
import math
import os
import shutil
import sys
import time
from dataclasses import dataclass

import numpy as np

KT = 0.001986 * 300.0          # kcal/mol at 300 K
NBINS = 32
U_LO, U_HI = -8.0, 8.0         # histogram range of the centred reduced work
D0 = 0.1                       # spacing of the pilot grid
N_IN = NBINS + 2               # histogram, lambda, alpha / 10
PILOT_LAMBDAS = np.linspace(0.0, 1.0, 11)
TARGET_OVERLAP = 0.85          # neighbour overlap the proposed ladder aims for


# ---------------------------------------------------------------- data --
def energy_gap(en_path: str, nskip: int = 0) -> tuple:
    """(lambda of state 2, E2 - E1 per record) of a two-state energy file."""
    from q_amd.qfep import read_en
    rec = read_en(en_path, 2)[nskip:]
    return float(rec[0, 1, 0]), rec[:, 1, 1] - rec[:, 0, 1]


def histogram(u: np.ndarray) -> np.ndarray:
    """normalised 32-bin histogram of u on [U_LO, U_HI), outliers in the end bins"""
    idx = np.clip(np.floor((u - U_LO) / (U_HI - U_LO) * NBINS).astype(int), 0, NBINS - 1)
    h = np.bincount(idx, minlength=NBINS).astype(np.float32)
    return h / max(1.0, h.sum())


def features(dE: np.ndarray, lam: float, alpha: float) -> np.ndarray:
    """Inputs of the overlap network for a pilot window at lambda with energy
    gaps dE: the histogram of the reduced work of a step of D0 about its
    mean (a constant shift of the work changes the free energy, not the
    overlap), then lambda and alpha / 10."""
    u = D0 * (dE - np.mean(dE)) / KT
    return np.concatenate([histogram(u), [lam, alpha / 10.0]]).astype(np.float32)


def bar_df(wF: np.ndarray, wR: np.ndarray) -> float:
    """BAR free energy (kT units) from forward works wF (sampled in A) and
    reverse works wR (sampled in B), solved by bisection."""
    nF, nR = len(wF), len(wR)
    M = math.log(nF / nR)

    def g(df):
        a = np.sum(1.0 / (1.0 + np.exp(np.clip(M + wF - df, -700, 700))))
        b = np.sum(1.0 / (1.0 + np.exp(np.clip(-M + wR + df, -700, 700))))
        return a - b

    lo, hi = -50.0 + min(wF.min(), -wR.max()), 50.0 + max(wF.max(), -wR.min())
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if g(mid) > 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def bar_overlap(wF: np.ndarray, wR: np.ndarray) -> float:
    """Overlap of two windows in (0, 1]: the mean BAR acceptance of the
    forward and of the reverse samples (1 for identical states)."""
    df = bar_df(wF, wR)
    a = np.mean(1.0 / (1.0 + np.exp(np.clip(wF - df, -700, 700))))
    b = np.mean(1.0 / (1.0 + np.exp(np.clip(wR + df, -700, 700))))
    return float(a + b)


def pilot_samples(lams, dEs, alpha):
    """Training rows from one pilot ladder: for neighbouring windows k and
    k+1, the features of each, both labelled with the BAR overlap of the pair."""
    X, Y = [], []
    for k in range(len(lams) - 1):
        delta = lams[k + 1] - lams[k]
        o = bar_overlap(delta * dEs[k] / KT, -delta * dEs[k + 1] / KT)
        X.append(features(dEs[k], lams[k], alpha))
        Y.append(o)
        X.append(features(dEs[k + 1], lams[k + 1], alpha))
        Y.append(o)
    return np.array(X, np.float32), np.array(Y, np.float32)


# ------------------------------------------------------------ schedule --
_GH_X, _GH_W = np.polynomial.hermite_e.hermegauss(80)


def gaussian_overlap(sw) -> np.ndarray:
    """BAR overlap (as bar_overlap) of two windows whose works are Gaussian
    with standard deviation sw (kT): 2 <1 / (1 + e^x)> over x ~ N(sw^2/2, sw^2)."""
    sw = np.atleast_1d(np.asarray(sw, float))
    x = sw[:, None] ** 2 / 2 + sw[:, None] * _GH_X[None, :]
    return 2.0 * (_GH_W[None, :] / np.sqrt(2 * np.pi) / (1.0 + np.exp(np.clip(x, -700, 700)))).sum(1)


def work_spread(o) -> np.ndarray:
    """Inverse of gaussian_overlap: the work spread (kT) that gives overlap o."""
    o = np.clip(np.atleast_1d(np.asarray(o, float)), 1e-6, 1.0 - 1e-9)
    lo, hi = np.zeros_like(o), np.full_like(o, 60.0)
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        big = gaussian_overlap(mid) < o        # overlap falls with the spread
        hi = np.where(big, mid, hi)
        lo = np.where(big, lo, mid)
    return 0.5 * (lo + hi)


def ladder_from_overlaps(lams, ovl, nmin=16, nmax=22, target=TARGET_OVERLAP, n=None):
    """Windows at equal steps of the thermodynamic length implied by the
    neighbour overlaps `ovl` of the grid `lams` (work spread per unit lambda
    constant on each grid interval); `n` windows, or as many as the target
    overlap needs within nmin..nmax.  Returns the ladder and the overlap it
    implies between neighbours."""
    lams = np.asarray(lams, float)
    rate = work_spread(ovl) / np.diff(lams)                # spread per unit lambda
    L = np.concatenate([[0.0], np.cumsum(rate * np.diff(lams))])
    if n is None:
        n = int(np.clip(np.ceil(L[-1] / work_spread(target)[0]) + 1, nmin, nmax))
    sched = np.interp(np.linspace(0.0, L[-1], n), L, lams) if L[-1] > 0 else np.linspace(0, 1, n)
    sched[0], sched[-1] = 0.0, 1.0
    return sched, float(gaussian_overlap(L[-1] / (n - 1))[0])


def predicted_overlaps(predict, lams, dEs, alpha) -> np.ndarray:
    """Network estimate of the overlap of each grid interval: the mean of the
    predictions from the windows at its two ends."""
    p = predict(np.stack([features(d, l, alpha) for l, d in zip(lams, dEs)]))
    return 0.5 * (p[:-1] + p[1:])


def measured_overlaps(lams, dEs) -> np.ndarray:
    """BAR overlap of each pair of neighbouring pilot windows, measured"""
    return np.array([bar_overlap((lams[k + 1] - lams[k]) * dEs[k] / KT, -(lams[k + 1] - lams[k]) * dEs[k + 1] / KT)
                     for k in range(len(lams) - 1)])


# mean |predicted - measured| interval overlap of the pilot grid beyond which
# the network's ladder is not used and the uniform ladder with as many windows
# runs instead (the held-out pilot grids of the shipped network disagree by
# 0.03-0.20, 0.10 on average)
FALLBACK_DISAGREEMENT = 0.15


def learned_schedule(predict, lams, dEs, alpha, nmin=16, nmax=22, bound=FALLBACK_DISAGREEMENT):
    """The schedule of one perturbation from its pilot windows: the ladder
    at equal thermodynamic length from the network's interval overlaps, or,
    when those disagree with the overlaps the pilots measure by more than
    `bound` on average, the uniform ladder with as many windows (the
    fallback).  Returns the ladder and a record of the decision."""
    pred = predicted_overlaps(predict, lams, dEs, alpha)
    meas = measured_overlaps(lams, dEs)
    sched, implied = ladder_from_overlaps(lams, pred, nmin, nmax)
    dis = float(np.mean(np.abs(pred - meas)))
    fallback = dis > bound
    if fallback:
        sched = np.linspace(0.0, 1.0, len(sched))
    return sched, dict(disagreement=dis, fallback=fallback, implied_neighbour_overlap=implied,
                       predicted_overlap=pred, pilot_overlap=meas)


def propose_n(predict, lams, dEs, alpha, nmin=16, nmax=22):
    """The learned ladder of one perturbation from its pilot windows."""
    return learned_schedule(predict, lams, dEs, alpha, nmin, nmax)[0]


# ------------------------------------------------------------ running --
@dataclass
class Perturbation:
    seed: int
    leg: str                 # "ligwat" or "protein"
    n_pert: int
    alpha: float = 10.0
    scaffold: int = -1       # protein leg: index of the environment it is drawn in


def build_system(p: Perturbation):
    """the synthetic system of a ligand-in-water perturbation; the ligand
    generator retries with derived seeds when a growth attempt fails"""
    from q_amd import synth
    for k in range(50):
        s = p.seed + 100003 * k
        try:
            if p.leg == "ligwat":
                return synth.ligand_in_water(seed=s, radius=20.0, n_pert=p.n_pert, alpha=p.alpha)
            return synth.protein_leg(seed=s, radius=25.0, n_pert=p.n_pert, alpha=p.alpha)
        except RuntimeError:
            continue
    raise RuntimeError(f"no system for {p}")


def build_scaffold(seed: int):
    """a protein-leg environment (chain, ligand and water); perturbations
    are redrawn in it by synth.reperturb"""
    return build_system(Perturbation(seed, "protein", 6))


def write_ladder(p: Perturbation, system, lams, outdir, steps, energy=5, md_seed=1, tag="w"):
    from q_amd import synth
    inps = []
    for k, lam in enumerate(lams):
        inps.append(synth.make_window(p.leg, outdir, seed=p.seed, md_seed=md_seed, lam=float(lam), steps=steps,
                                      energy=energy, output=max(steps, 1), nb_cycle=25, tag=f"{tag}{k:03d}",
                                      system=system if k == 0 else None))
    return inps


def run_batched(inps, steps, batch=64, log=None):
    """Run every window for `steps` steps, up to `batch` windows per engine.
    Returns the energy file of each window, or None for a window that the
    engine stopped (SHAKE failure, list overflow); the others run on."""
    from q_amd.engine import Engine, QmiError
    out = []
    for b0 in range(0, len(inps), batch):
        chunk = inps[b0:b0 + batch]
        e = Engine(0)
        for p in chunk:
            e.add_window(p)
            e.set_log(len(e.windows) - 1, None)
        e.prepare()
        bad = set()
        try:
            e.run(steps, finalize=True)
        except QmiError as err:
            bad = {k for k in range(len(chunk)) if e.info(k)["status"] != 0}
            if not bad:
                raise
            if log:
                log(f"stopped windows {sorted(b0 + k for k in bad)}: {err}")
        e.close()
        out += [None if k in bad else os.path.splitext(p)[0] + ".en" for k, p in enumerate(chunk)]
        if log:
            log(f"ran windows {b0}..{b0 + len(chunk) - 1} for {steps} steps")
    return out


def bar_var(wF: np.ndarray, wR: np.ndarray) -> float:
    """Bennett's large-sample variance of the BAR free energy (kT^2 units),
    for uncorrelated samples."""
    nF, nR = len(wF), len(wR)
    M = math.log(nF / nR)
    df = bar_df(wF, wR)
    fF = 1.0 / (1.0 + np.exp(np.clip(M + wF - df, -700, 700)))
    fR = 1.0 / (1.0 + np.exp(np.clip(-M + wR + df, -700, 700)))
    return float((np.mean(fF ** 2) / np.mean(fF) ** 2 - 1.0) / nF + (np.mean(fR ** 2) / np.mean(fR) ** 2 - 1.0) / nR)


def ladder_stats(en_paths, nskip=0) -> dict:
    """BAR free energy of a ladder (qfep's estimator), the smallest
    neighbour overlap and Bennett's standard error (kcal/mol)."""
    from q_amd.qfep import qfep
    gaps = [energy_gap(p, nskip) for p in en_paths]
    ovl, var = [], 0.0
    for (la, da), (lb, db) in zip(gaps[:-1], gaps[1:]):
        d = lb - la
        wF, wR = d * da / KT, -d * db / KT
        ovl.append(bar_overlap(wF, wR))
        var += bar_var(wF, wR)
    return dict(dG=float(qfep(en_paths, 2, KT, nskip, 20, 5).dG_bar), min_overlap=float(min(ovl)),
                sigma=float(KT * math.sqrt(var)))


def ladder_dG(en_paths, nskip=0) -> float:
    from q_amd.qfep import qfep
    return qfep(en_paths, 2, KT, nskip, 20, 5).dG_bar


def simulated_fraction(n, prod_steps, pilot_steps, n_ref=51):
    """simulated time of pilots + proposed ladder over that of the reference ladder"""
    return (n * prod_steps + len(PILOT_LAMBDAS) * pilot_steps) / float(n_ref * prod_steps)


def load_predictor(path=None):
    """the trained overlap network shipped with the package, on the GPU"""
    from q_amd.mlp import MLP
    path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "overlap_mlp.npz")
    z = np.load(path)
    if int(z["n_in"]) != N_IN or int(z["nbins"]) != NBINS:
        raise ValueError(f"{path}: network built for other features")
    net = MLP(N_IN, int(z["hidden"]), seed=0)
    net.params = z["params"].astype(np.float32)
    return net


def evaluate(p: Perturbation, system, predict, workdir, pilot_steps, prod_steps, pilot_skip, prod_skip, nmin=16,
             nmax=None, pilots=None, log=None) -> dict:
    """Pilots (unless given), then four ladders of one perturbation: the
    learned one, the uniform one with as many windows (no network), the
    equal-length ladder from the measured pilot overlaps (the same spacing
    rule without the network), and the uniform 51-window reference."""
    if nmax is None:
        # the largest ladder that keeps pilots + ladder within half of the 51-window time
        nmax = max(nmin, min(24, int((0.5 * 51 * prod_steps - len(PILOT_LAMBDAS) * pilot_steps) // prod_steps)))
    if pilots is None:
        inps = write_ladder(p, system, PILOT_LAMBDAS, os.path.join(workdir, "pilot"), pilot_steps, tag="p")
        ens = run_batched(inps, pilot_steps)
        if any(en is None for en in ens):
            raise RuntimeError("a pilot window was stopped")
        pilots = [energy_gap(en, pilot_skip) for en in ens]
    lams = [a for a, _ in pilots]
    dEs = [b for _, b in pilots]
    sched, dec = learned_schedule(predict, lams, dEs, p.alpha, nmin, nmax)
    pred, meas, implied = dec["predicted_overlap"], dec["pilot_overlap"], dec["implied_neighbour_overlap"]
    n = len(sched)
    ladders = dict(ref51=np.linspace(0.0, 1.0, 51), learned=sched, uniform=np.linspace(0.0, 1.0, n),
                   direct=ladder_from_overlaps(lams, meas, n=n)[0])
    inps, owner = [], []
    for name, lamv in ladders.items():
        w = write_ladder(p, system, lamv, os.path.join(workdir, name), prod_steps, tag="e")
        inps += w
        owner += [name] * len(w)
    ens = run_batched(inps, prod_steps, log=log)
    if any(en is None for en in ens):
        raise RuntimeError("a window of the evaluation ladders was stopped")
    st = {name: ladder_stats([e for e, o in zip(ens, owner) if o == name], prod_skip) for name in ladders}
    ref = st["ref51"]["dG"]
    return dict(perturbation=p.seed, leg=p.leg, n_pert=p.n_pert, windows=n,
                schedule=[round(float(v), 4) for v in sched], predicted_overlap=[round(float(v), 3) for v in pred],
                pilot_overlap=[round(float(v), 3) for v in meas], implied_neighbour_overlap=implied,
                overlap_disagreement=dec["disagreement"], fallback=dec["fallback"], stats=st,
                err_learned=abs(st["learned"]["dG"] - ref), err_uniform=abs(st["uniform"]["dG"] - ref),
                err_direct=abs(st["direct"]["dG"] - ref),
                simulated_fraction=simulated_fraction(n, prod_steps, pilot_steps))


# ------------------------------------------------------------- driver --
def perturbation_sets(seed, n_train, n_eval, nsc):
    """Training and held-out perturbations; every perturbation runs in both
    legs: its own ligand in water, and a redrawn perturbation of one of nsc
    protein environments (held-out ones in two environments of their own)."""
    rng = np.random.default_rng(seed)
    train, held = [], []
    for k in range(n_train):
        n_pert = int(rng.integers(2, 11))
        train += [Perturbation(1000 + k, "ligwat", n_pert), Perturbation(1000 + k, "protein", n_pert, scaffold=k % nsc)]
    for k in range(n_eval):
        n_pert = int(rng.integers(2, 11))
        held += [Perturbation(9000 + k, "ligwat", n_pert), Perturbation(9000 + k, "protein", n_pert, scaffold=nsc + k % 2)]
    return train, held


def systems_for(perts, nsc, procs):
    """system builder for the given perturbations (ligands and protein
    environments built in parallel); the perturbation itself is drawn by
    synth.reperturb in both legs"""
    from concurrent.futures import ProcessPoolExecutor
    from q_amd import synth
    lig = [p for p in perts if p.leg == "ligwat"]
    scs = sorted({p.scaffold for p in perts if p.leg == "protein"})
    with ProcessPoolExecutor(max_workers=procs) as ex:
        lig_sys = list(ex.map(build_system, lig))
        sc_sys = list(ex.map(build_scaffold, [7000 + j for j in scs]))
    lig_of = {(p.seed, p.n_pert): s for p, s in zip(lig, lig_sys)}
    sc_of = dict(zip(scs, sc_sys))

    def system_of(p):
        base = lig_of[(p.seed, p.n_pert)] if p.leg == "ligwat" else sc_of[p.scaffold]
        return synth.reperturb(base, p.seed, p.n_pert, p.alpha)

    return system_of


def evaluate_replicas(args):
    """Held-out evaluation with replicas: for each held-out perturbation and
    leg, pilots, the learned ladder, and the four ladders of `evaluate` run
    with `replicas` Maxwell seeds; reports the replica mean and spread of
    each ladder's free energy."""
    t0 = time.time()

    def log(msg):
        print(f"[schedule {time.time() - t0:7.1f}s] {msg}", flush=True)

    net = load_predictor(args.weights)
    _, held = perturbation_sets(args.seed, args.n_train, args.n_eval, args.n_scaffolds)
    legs = set(getattr(args, "legs", "ligwat,protein").split(","))
    picks = None if getattr(args, "cases", None) is None else {int(c) for c in args.cases.split(",")}
    todo = [(j, p) for j, p in enumerate(held) if p.leg in legs and (picks is None or j in picks)]
    system_of = systems_for([p for _, p in todo], args.n_scaffolds, args.procs)
    log(f"built {len(todo)} held-out systems")
    os.makedirs(args.out, exist_ok=True)
    energy, batch = getattr(args, "energy", 5), getattr(args, "batch", 64)
    rows = []
    for j, p in todo:
        system = system_of(p)
        base = os.path.join(args.out, "work", f"case_{j:03d}")
        inps = write_ladder(p, system, PILOT_LAMBDAS, os.path.join(base, "pilot"), args.pilot_steps, tag="p")
        ens = run_batched(inps, args.pilot_steps)
        if any(en is None for en in ens):
            log(f"{p}: a pilot window was stopped, left out")
            continue
        pilots = [energy_gap(en, args.pilot_skip) for en in ens]
        lams = [a for a, _ in pilots]
        dEs = [b for _, b in pilots]
        nmax = max(16, min(24, int((0.5 * 51 * args.prod_steps - len(PILOT_LAMBDAS) * args.pilot_steps)
                                   // args.prod_steps)))
        sched, dec = learned_schedule(net.forward, lams, dEs, p.alpha, 16, nmax)
        implied = dec["implied_neighbour_overlap"]
        n = len(sched)
        ladders = dict(ref51=np.linspace(0.0, 1.0, 51), learned=sched, uniform=np.linspace(0.0, 1.0, n),
                       direct=ladder_from_overlaps(lams, measured_overlaps(lams, dEs), n=n)[0])
        inps, owner = [], []
        for r in range(args.replicas):
            for name, lamv in ladders.items():
                w = write_ladder(p, system, lamv, os.path.join(base, f"{name}_r{r}"), args.prod_steps, md_seed=1 + r,
                                 tag="e", energy=energy)
                inps += w
                owner += [(name, r)] * len(w)
        ens = run_batched(inps, args.prod_steps, batch=batch, log=log)
        if any(en is None for en in ens):
            log(f"{p}: a ladder window was stopped, left out")
            continue
        dG = {name: [] for name in ladders}
        stats = {name: [] for name in ladders}
        for name in ladders:
            for r in range(args.replicas):
                st = ladder_stats([e for e, o in zip(ens, owner) if o == (name, r)], args.prod_skip)
                dG[name].append(st["dG"])
                stats[name].append(st)
        mean = {k: float(np.mean(v)) for k, v in dG.items()}
        row = dict(perturbation=p.seed, leg=p.leg, n_pert=p.n_pert, windows=n,
                   schedule=[round(float(v), 4) for v in sched], implied_neighbour_overlap=implied,
                   overlap_disagreement=dec["disagreement"], fallback=dec["fallback"], dG=dG,
                   dG_mean=mean, replica_spread={k: float(np.std(v)) for k, v in dG.items()},
                   min_overlap={k: float(np.mean([s["min_overlap"] for s in v])) for k, v in stats.items()},
                   sigma={k: float(np.mean([s["sigma"] for s in v])) for k, v in stats.items()},
                   err={k: abs(mean[k] - mean["ref51"]) for k in ("learned", "uniform", "direct")},
                   simulated_fraction=simulated_fraction(n, args.prod_steps, args.pilot_steps))
        row["case"] = j
        rows.append(row)
        json.dump(dict(prod_steps=args.prod_steps, prod_skip=args.prod_skip, pilot_steps=args.pilot_steps,
                       pilot_skip=args.pilot_skip, replicas=args.replicas, energy=energy, row=row),
                  open(os.path.join(args.out, f"case_{j:03d}.json"), "w"), indent=1)
        if not getattr(args, "keep", True):
            shutil.rmtree(base, ignore_errors=True)
        log(json.dumps({k: v for k, v in row.items() if k not in ("schedule", "dG")}))
    rep = report(rows, dict(prod_steps=args.prod_steps, prod_skip=args.prod_skip, pilot_steps=args.pilot_steps,
                            pilot_skip=args.pilot_skip, replicas=args.replicas))
    json.dump(rep, open(os.path.join(args.out, "schedule_eval.json"), "w"), indent=1)
    log({k: v for k, v in rep.items() if k != "cases"})


def report(rows, head: dict) -> dict:
    """The evaluation summary over the case rows: mean and largest error of
    each ladder against the 51-window reference, overlaps, spreads, fallbacks."""
    rep = dict(head, cases=rows)
    for k in ("learned", "uniform", "direct"):
        rep["mae_" + k] = float(np.mean([r["err"][k] for r in rows]))
        rep["max_err_" + k] = float(np.max([r["err"][k] for r in rows]))
        rep["mean_min_overlap_" + k] = float(np.mean([r["min_overlap"][k] for r in rows]))
        rep["mean_sigma_" + k] = float(np.mean([r["sigma"][k] for r in rows]))
    rep["mean_ref51_replica_spread"] = float(np.mean([r["replica_spread"]["ref51"] for r in rows]))
    rep["fallbacks"] = int(sum(1 for r in rows if r["fallback"]))
    rep["mean_simulated_fraction"] = float(np.mean([r["simulated_fraction"] for r in rows]))
    return rep


def merge_reports(args):
    """one report from the per-case files of several evaluation runs"""
    rows, head = [], None
    for d in args.dirs:
        for f in sorted(os.listdir(d)):
            if f.startswith("case_") and f.endswith(".json"):
                c = json.load(open(os.path.join(d, f)))
                rows.append(c["row"])
                head = head or {k: c[k] for k in ("prod_steps", "prod_skip", "pilot_steps", "pilot_skip", "replicas",
                                                  "energy")}
    rows.sort(key=lambda r: (r.get("case", 0), r["leg"]))
    rep = report(rows, head)
    json.dump(rep, open(args.out, "w"), indent=1)
    print({k: v for k, v in rep.items() if k != "cases"})


def build(args):
    from q_amd.mlp import MLP
    os.makedirs(args.out, exist_ok=True)
    t0 = time.time()

    def log(msg):
        print(f"[schedule {time.time() - t0:7.1f}s] {msg}", flush=True)

    train, held = perturbation_sets(args.seed, args.n_train, args.n_eval, args.n_scaffolds)
    work = os.path.join(args.out, "work")
    os.makedirs(work, exist_ok=True)
    system_of = systems_for(train + held, args.n_scaffolds, args.procs)
    log("built the ligand-in-water systems and protein environments")
    # pilot ladders, one batch per leg so that windows of one size share engines
    pil = {}
    for leg in ("ligwat", "protein"):
        inps, owners = [], []
        for k, p in enumerate(train + held):
            if p.leg != leg:
                continue
            w = write_ladder(p, system_of(p), PILOT_LAMBDAS, os.path.join(work, f"pilot_{leg}_{k:04d}"),
                             args.pilot_steps, tag="p")
            inps += w
            owners += [k] * len(w)
        log(f"{leg}: {len(inps)} pilot windows written")
        for k, en in zip(owners, run_batched(inps, args.pilot_steps, log=log)):
            pil.setdefault(k, []).append(energy_gap(en, args.pilot_skip) if en else None)
    # perturbations with a stopped pilot window are left out
    failed = {k for k, v in pil.items() if any(g is None for g in v)}
    if failed:
        log(f"left out {len(failed)} perturbations with stopped pilot windows: {sorted(failed)}")
    X, Y = [], []
    for k, p in enumerate(train):
        if k in failed:
            continue
        x, y = pilot_samples([a for a, _ in pil[k]], [b for _, b in pil[k]], p.alpha)
        X.append(x)
        Y.append(y)
    X = np.concatenate(X)
    Y = np.concatenate(Y)
    log(f"training set: {len(Y)} rows, overlap mean {Y.mean():.3f}")
    net = MLP(N_IN, 128, seed=args.seed)
    l0, _ = net.grad(X, Y)
    net.train(X, Y, epochs=args.epochs, batch=256, lr=1e-3, seed=args.seed)
    l1, _ = net.grad(X, Y)
    log(f"trained: loss {l0:.5f} -> {l1:.5f}")
    np.savez(os.path.join(args.out, "overlap_mlp.npz"), params=net.params, n_in=N_IN, hidden=128,
             u_lo=U_LO, u_hi=U_HI, nbins=NBINS, kt=KT)
    np.savez_compressed(os.path.join(args.out, "pilot_data.npz"), X=X, Y=Y)
    report = dict(n_perturbations=args.n_train, rows=int(len(Y)), loss_before=l0, loss_after=l1,
                  pilot_steps=args.pilot_steps, prod_steps=args.prod_steps, pilot_skip=args.pilot_skip,
                  prod_skip=args.prod_skip, evals=[])
    for j, p in enumerate(held):
        kk = len(train) + j
        if kk in failed:
            continue
        try:
            row = evaluate(p, system_of(p), net.forward, os.path.join(work, f"eval_{j:03d}"), args.pilot_steps,
                           args.prod_steps, args.pilot_skip, args.prod_skip, pilots=pil[kk])
        except RuntimeError as err:
            log(f"evaluation of {p} left out: {err}")
            continue
        report["evals"].append(row)
        log(json.dumps({k: v for k, v in row.items() if k != "schedule"}))
    ev = report["evals"]
    for key in ("err_learned", "err_uniform", "err_direct"):
        report["mae_" + key[4:]] = float(np.mean([r[key] for r in ev]))
        report["max_" + key] = float(np.max([r[key] for r in ev]))
    report["mean_simulated_fraction"] = float(np.mean([r["simulated_fraction"] for r in ev]))
    for name in ("learned", "uniform", "direct"):
        report["mean_sigma_" + name] = float(np.mean([r["stats"][name]["sigma"] for r in ev]))
        report["mean_min_overlap_" + name] = float(np.mean([r["stats"][name]["min_overlap"] for r in ev]))
    json.dump(report, open(os.path.join(args.out, "schedule_report.json"), "w"), indent=1)
    log(f"MAE vs 51 windows: learned {report['mae_learned']:.3f}, uniform with as many windows "
        f"{report['mae_uniform']:.3f}, measured-overlap spacing {report['mae_direct']:.3f} kcal/mol; "
        f"sigma {report['mean_sigma_learned']:.3f} / {report['mean_sigma_uniform']:.3f} / "
        f"{report['mean_sigma_direct']:.3f}; min overlap {report['mean_min_overlap_learned']:.3f} / "
        f"{report['mean_min_overlap_uniform']:.3f} / {report['mean_min_overlap_direct']:.3f}")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m q_amd.schedule")
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build", help="pilot data, GPU training of the overlap MLP, held-out evaluation")
    b.add_argument("--out", required=True)
    b.add_argument("--n-train", type=int, default=200)
    b.add_argument("--n-scaffolds", type=int, default=16)
    b.add_argument("--n-eval", type=int, default=8)
    b.add_argument("--pilot-steps", type=int, default=1000)
    b.add_argument("--prod-steps", type=int, default=4000)
    b.add_argument("--pilot-skip", type=int, default=40, help="energy records dropped from each pilot window")
    b.add_argument("--prod-skip", type=int, default=200, help="energy records dropped from each ladder window")
    b.add_argument("--epochs", type=int, default=200)
    b.add_argument("--procs", type=int, default=16)
    b.add_argument("--seed", type=int, default=0)
    v = sub.add_parser("evaluate", help="held-out evaluation of a trained network with replicas")
    v.add_argument("--out", required=True)
    v.add_argument("--weights", default=None, help="network file (default: the one shipped in q_amd/data)")
    v.add_argument("--n-train", type=int, default=200)
    v.add_argument("--n-scaffolds", type=int, default=16)
    v.add_argument("--n-eval", type=int, default=4)
    v.add_argument("--replicas", type=int, default=2)
    v.add_argument("--pilot-steps", type=int, default=1000)
    v.add_argument("--prod-steps", type=int, default=15000)
    v.add_argument("--pilot-skip", type=int, default=40)
    v.add_argument("--prod-skip", type=int, default=1000,
                   help="energy records (not steps) dropped from the start of each ladder window")
    v.add_argument("--procs", type=int, default=16)
    v.add_argument("--seed", type=int, default=0)
    v.add_argument("--legs", default="ligwat,protein", help="comma-separated legs to evaluate")
    v.add_argument("--cases", default=None, help="comma-separated held-out case indices (default: all)")
    v.add_argument("--energy", type=int, default=5, help="energy interval of the production windows")
    v.add_argument("--batch", type=int, default=64, help="windows per engine")
    v.add_argument("--keep", action="store_true", help="keep the windows' files (default: removed per case)")
    m = sub.add_parser("merge", help="one report from the per-case files of evaluation runs")
    m.add_argument("--out", required=True)
    m.add_argument("dirs", nargs="+")
    args = ap.parse_args(argv)
    if args.cmd == "build":
        build(args)
    elif args.cmd == "evaluate":
        evaluate_replicas(args)
    elif args.cmd == "merge":
        merge_reports(args)


if __name__ == "__main__":
    main()
