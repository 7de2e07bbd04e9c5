"""Learned lambda schedule for FEP windows.

The reference spaces lambda windows by hand (QligFEP writes a fixed uniform
ladder).  Here short pilot windows, run by the engine, give the energy gap
dE = E(state 2) - E(state 1) along a coarse lambda grid.  For a candidate
step lambda -> lambda + delta the reduced work of the forward move is
u = delta * dE / kT; a 32-bin histogram of u, lambda, delta and the soft-core
alpha are the inputs of a small MLP (q_amd.mlp, trained and run on the GPU by
the library's kernels) that predicts the BAR overlap of the two windows.  A
schedule is then built greedily: from lambda = 0, take the largest step
whose predicted overlap reaches a target, with the target set so that the
ladder has 16-24 windows.  The free energy of the proposed ladder (BAR, by
q_amd.qfep) is judged against a uniform 51-window ladder; the uniform ladder
with the same number of windows is the method without the network.

    python -m q_amd.schedule build --out DIR     pilot data, training, evaluation
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from dataclasses import dataclass

import numpy as np

KT = 0.001986 * 300.0          # kcal/mol at 300 K
NBINS = 32
U_LO, U_HI = -8.0, 8.0         # histogram range of the centred reduced work
N_IN = NBINS + 3               # histogram, lambda, delta, alpha / 10
PILOT_LAMBDAS = np.linspace(0.0, 1.0, 11)


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


def features(dE: np.ndarray, lam: float, delta: float, alpha: float) -> np.ndarray:
    """Inputs of the overlap network for a step lambda -> lambda + delta from
    a window with energy gaps dE: the histogram of the reduced work
    u = delta * dE / kT about its mean (a constant shift of u changes the
    free energy, not the overlap), then lambda, delta and alpha / 10."""
    u = delta * (dE - np.mean(dE)) / KT
    return np.concatenate([histogram(u), [lam, delta, alpha / 10.0]]).astype(np.float32)


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


def pilot_samples(lams, dEs, alpha, max_step=4):
    """Training rows from one pilot ladder: for windows k and k+m (m = 1..4),
    features of the forward move (from window k) and of the reverse move
    (from window k+m), both labelled with the BAR overlap of the pair."""
    X, Y = [], []
    for k in range(len(lams)):
        for m in range(1, max_step + 1):
            j = k + m
            if j >= len(lams):
                break
            delta = lams[j] - lams[k]
            wF = delta * dEs[k] / KT
            wR = -delta * dEs[j] / KT
            o = bar_overlap(wF, wR)
            X.append(features(dEs[k], lams[k], delta, alpha))
            Y.append(o)
            X.append(features(dEs[j], lams[j], -delta, alpha))
            Y.append(o)
    return np.array(X, np.float32), np.array(Y, np.float32)


# ------------------------------------------------------------ schedule --
def propose(predict, lams, dEs, alpha, target, dmin=0.005, dmax=0.5):
    """Greedy ladder from 0 to 1: at each lambda the largest step whose
    predicted overlap (features from the nearest pilot window) is >= target."""
    steps = np.unique(np.concatenate([np.arange(dmin, dmax + 1e-9, 0.005)]))
    sched = [0.0]
    lam = 0.0
    while lam < 1.0 - 1e-9:
        k = int(np.argmin(np.abs(np.asarray(lams) - lam)))
        cand = steps[steps <= 1.0 - lam + 1e-9]
        if len(cand) == 0:
            break
        F = np.stack([features(dEs[k], lam, float(d), alpha) for d in cand])
        ok = cand[predict(F) >= target]
        d = float(ok.max()) if len(ok) else float(cand.min())
        lam = min(1.0, lam + d)
        if 1.0 - lam < 0.5 * dmin:
            lam = 1.0
        sched.append(lam)
    return np.array(sched)


def propose_n(predict, lams, dEs, alpha, nmin=16, nmax=24):
    """Schedule with nmin..nmax windows: bisection on the overlap target."""
    lo, hi = 0.0, 1.0
    best = None
    for _ in range(40):
        tgt = 0.5 * (lo + hi)
        s = propose(predict, lams, dEs, alpha, tgt)
        if len(s) < nmin:
            lo = tgt
        elif len(s) > nmax:
            hi = tgt
        else:
            best = s
            if len(s) <= (nmin + nmax) // 2:
                break
            hi = tgt
    if best is None:
        best = np.linspace(0.0, 1.0, nmin)
    return best


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
    """run every window for `steps` steps, up to `batch` windows per engine"""
    from q_amd.engine import Engine
    for b0 in range(0, len(inps), batch):
        e = Engine(0)
        for p in inps[b0:b0 + batch]:
            e.add_window(p)
            e.set_log(len(e.windows) - 1, None)
        e.prepare()
        e.run(steps, finalize=True)
        e.close()
        if log:
            log(f"ran windows {b0}..{min(len(inps), b0 + batch) - 1} for {steps} steps")
    return [os.path.splitext(p)[0] + ".en" for p in inps]


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


def evaluate(p: Perturbation, system, predict, workdir, pilot_steps, prod_steps, nskip, nmin=16, nmax=None,
             pilots=None, log=None) -> dict:
    """pilots (unless given), the proposed ladder, the uniform ladder with as
    many windows and the uniform 51-window ladder of one perturbation"""
    if nmax is None:
        # the largest ladder that keeps pilots + ladder within half of the 51-window time
        nmax = max(nmin, min(24, int((0.5 * 51 * prod_steps - len(PILOT_LAMBDAS) * pilot_steps) // prod_steps)))
    if pilots is None:
        inps = write_ladder(p, system, PILOT_LAMBDAS, os.path.join(workdir, "pilot"), pilot_steps, tag="p")
        pilots = [energy_gap(en, nskip) for en in run_batched(inps, pilot_steps)]
    lams = [a for a, _ in pilots]
    dEs = [b for _, b in pilots]
    sched = propose_n(predict, lams, dEs, p.alpha, nmin, nmax)
    n = len(sched)
    ladders = dict(ref51=np.linspace(0.0, 1.0, 51), learned=sched, uniform=np.linspace(0.0, 1.0, n))
    inps, owner = [], []
    for name, lamv in ladders.items():
        w = write_ladder(p, system, lamv, os.path.join(workdir, name), prod_steps, tag="e")
        inps += w
        owner += [name] * len(w)
    ens = run_batched(inps, prod_steps, log=log)
    st = {name: ladder_stats([e for e, o in zip(ens, owner) if o == name], nskip) for name in ladders}
    return dict(perturbation=p.seed, leg=p.leg, n_pert=p.n_pert, windows=n,
                schedule=[round(float(v), 4) for v in sched], stats=st,
                err_learned=abs(st["learned"]["dG"] - st["ref51"]["dG"]),
                err_uniform=abs(st["uniform"]["dG"] - st["ref51"]["dG"]),
                simulated_fraction=simulated_fraction(n, prod_steps, pilot_steps))


# ------------------------------------------------------------- driver --
def build(args):
    from concurrent.futures import ProcessPoolExecutor
    from q_amd import synth
    from q_amd.mlp import MLP
    os.makedirs(args.out, exist_ok=True)
    t0 = time.time()

    def log(msg):
        print(f"[schedule {time.time() - t0:7.1f}s] {msg}", flush=True)

    rng = np.random.default_rng(args.seed)
    nsc = args.n_scaffolds
    # every perturbation runs in both legs: its own ligand in water, and a
    # redrawn perturbation of one of nsc protein environments
    train, held = [], []
    for k in range(args.n_train):
        n_pert = int(rng.integers(2, 11))
        train += [Perturbation(1000 + k, "ligwat", n_pert), Perturbation(1000 + k, "protein", n_pert, scaffold=k % nsc)]
    for k in range(args.n_eval):
        n_pert = int(rng.integers(2, 11))
        held += [Perturbation(9000 + k, "ligwat", n_pert), Perturbation(9000 + k, "protein", n_pert, scaffold=nsc + k % 2)]
    work = os.path.join(args.out, "work")
    os.makedirs(work, exist_ok=True)
    lig = [p for p in train + held if p.leg == "ligwat"]
    with ProcessPoolExecutor(max_workers=args.procs) as ex:
        lig_sys = list(ex.map(build_system, lig))
        scaffolds = list(ex.map(build_scaffold, [7000 + j for j in range(nsc + 2)]))
    lig_of = {id(p): s for p, s in zip(lig, lig_sys)}

    def system_of(p):
        # the perturbation itself is drawn by synth.reperturb in both legs
        base = lig_of[id(p)] if p.leg == "ligwat" else scaffolds[p.scaffold]
        return synth.reperturb(base, p.seed, p.n_pert, p.alpha)

    log(f"built {len(lig_sys)} ligand-in-water systems and {len(scaffolds)} protein environments")
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
            pil.setdefault(k, []).append(energy_gap(en, args.nskip))
    X, Y = [], []
    for k, p in enumerate(train):
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
                  pilot_steps=args.pilot_steps, prod_steps=args.prod_steps, nskip=args.nskip, evals=[])
    for j, p in enumerate(held):
        kk = len(train) + j
        row = evaluate(p, system_of(p), net.forward, os.path.join(work, f"eval_{j:03d}"), args.pilot_steps,
                       args.prod_steps, args.nskip, pilots=pil[kk])
        report["evals"].append(row)
        log(json.dumps({k: v for k, v in row.items() if k != "schedule"}))
    ev = report["evals"]
    for key in ("err_learned", "err_uniform"):
        report["mae_" + key[4:]] = float(np.mean([r[key] for r in ev]))
        report["max_" + key] = float(np.max([r[key] for r in ev]))
    report["mean_simulated_fraction"] = float(np.mean([r["simulated_fraction"] for r in ev]))
    for name in ("learned", "uniform"):
        report["mean_sigma_" + name] = float(np.mean([r["stats"][name]["sigma"] for r in ev]))
        report["mean_min_overlap_" + name] = float(np.mean([r["stats"][name]["min_overlap"] for r in ev]))
    json.dump(report, open(os.path.join(args.out, "schedule_report.json"), "w"), indent=1)
    log(f"MAE vs 51 windows: learned {report['mae_learned']:.3f}, uniform with as many windows "
        f"{report['mae_uniform']:.3f} kcal/mol; sigma {report['mean_sigma_learned']:.3f} vs "
        f"{report['mean_sigma_uniform']:.3f}")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m q_amd.schedule")
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build", help="pilot data, GPU training of the overlap MLP, held-out evaluation")
    b.add_argument("--out", required=True)
    b.add_argument("--n-train", type=int, default=200)
    b.add_argument("--n-scaffolds", type=int, default=16)
    b.add_argument("--n-eval", type=int, default=8)
    b.add_argument("--pilot-steps", type=int, default=500)
    b.add_argument("--prod-steps", type=int, default=2000)
    b.add_argument("--nskip", type=int, default=10)
    b.add_argument("--epochs", type=int, default=200)
    b.add_argument("--procs", type=int, default=16)
    b.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    if args.cmd == "build":
        build(args)


if __name__ == "__main__":
    main()
