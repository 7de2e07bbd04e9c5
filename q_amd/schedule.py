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
U_LO, U_HI = -10.0, 22.0       # histogram range of u = delta * dE / kT
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
    u = delta * dE / KT
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
    leg: str            # "ligwat" or "protein"
    n_pert: int
    alpha: float = 10.0


def build_system(p: Perturbation):
    """the synthetic system of a perturbation; the ligand generator retries
    with derived seeds when a growth attempt fails"""
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


def ladder_dG(en_paths, nskip=0) -> float:
    from q_amd.qfep import qfep
    return qfep(en_paths, 2, KT, nskip, 20, 5).dG_bar


# ------------------------------------------------------------- driver --
def build(args):
    from q_amd.mlp import MLP
    os.makedirs(args.out, exist_ok=True)
    t0 = time.time()

    def log(msg):
        print(f"[schedule {time.time() - t0:7.1f}s] {msg}", flush=True)

    rng = np.random.default_rng(args.seed)
    train = [Perturbation(1000 + k, "ligwat", int(rng.integers(2, 11))) for k in range(args.n_train)]
    train += [Perturbation(5000 + k, "protein", int(rng.integers(2, 11))) for k in range(args.n_train_protein)]
    held = [Perturbation(9000 + k, "ligwat", int(rng.integers(2, 11))) for k in range(args.n_eval)]
    # pilot ladders
    from concurrent.futures import ProcessPoolExecutor
    work = os.path.join(args.out, "work")
    os.makedirs(work, exist_ok=True)
    with ProcessPoolExecutor(max_workers=args.procs) as ex:
        systems = list(ex.map(build_system, train + held))
    log(f"built {len(systems)} systems")
    pilot_inps, owners = [], []
    for k, (p, s) in enumerate(zip(train + held, systems)):
        d = os.path.join(work, f"pilot_{k:03d}")
        inps = write_ladder(p, s, PILOT_LAMBDAS, d, args.pilot_steps, tag="p")
        pilot_inps += inps
        owners += [k] * len(inps)
    log(f"wrote {len(pilot_inps)} pilot windows")
    ens = run_batched(pilot_inps, args.pilot_steps, log=log)
    pil = {}
    for k, en in zip(owners, ens):
        lam, dE = energy_gap(en, args.nskip)
        pil.setdefault(k, []).append((lam, dE))
    X, Y = [], []
    for k, p in enumerate(train):
        lams = [a for a, _ in pil[k]]
        dEs = [b for _, b in pil[k]]
        x, y = pilot_samples(lams, dEs, p.alpha)
        X.append(x)
        Y.append(y)
    X = np.concatenate(X)
    Y = np.concatenate(Y)
    log(f"training set: {len(Y)} rows, overlap mean {Y.mean():.3f}")
    net = MLP(N_IN, 128, seed=args.seed)
    l0, _ = net.grad(X, Y)
    loss = net.train(X, Y, epochs=args.epochs, batch=256, lr=1e-3, seed=args.seed)
    l1, _ = net.grad(X, Y)
    log(f"trained: loss {l0:.4f} -> {l1:.4f}")
    np.savez(os.path.join(args.out, "overlap_mlp.npz"), params=net.params, n_in=N_IN, hidden=128,
             u_lo=U_LO, u_hi=U_HI, nbins=NBINS, kt=KT)
    # evaluation on held-out perturbations
    report = dict(n_train=len(train), rows=int(len(Y)), loss_before=l0, loss_after=l1, evals=[])
    for k, p in enumerate(held):
        kk = len(train) + k
        lams = [a for a, _ in pil[kk]]
        dEs = [b for _, b in pil[kk]]
        sched = propose_n(net.forward, lams, dEs, p.alpha)
        n = len(sched)
        uni = np.linspace(0.0, 1.0, n)
        ref = np.linspace(0.0, 1.0, 51)
        s = systems[kk]
        base = os.path.join(work, f"eval_{k:03d}")
        runs = {}
        for name, lamv in (("ref51", ref), ("learned", sched), ("uniform", uni)):
            inps = write_ladder(p, s, lamv, os.path.join(base, name), args.prod_steps, tag="e")
            ens = run_batched(inps, args.prod_steps)
            runs[name] = ladder_dG(ens, args.nskip)
        sim_frac = (n * args.prod_steps + len(PILOT_LAMBDAS) * args.pilot_steps) / (51.0 * args.prod_steps)
        row = dict(perturbation=p.seed, n_pert=p.n_pert, windows=n, schedule=[round(float(v), 4) for v in sched],
                   dG_ref51=runs["ref51"], dG_learned=runs["learned"], dG_uniform_same_n=runs["uniform"],
                   err_learned=abs(runs["learned"] - runs["ref51"]), err_uniform=abs(runs["uniform"] - runs["ref51"]),
                   simulated_fraction=sim_frac)
        report["evals"].append(row)
        log(json.dumps(row))
    ev = report["evals"]
    report["mae_learned"] = float(np.mean([r["err_learned"] for r in ev]))
    report["mae_uniform_same_n"] = float(np.mean([r["err_uniform"] for r in ev]))
    report["max_err_learned"] = float(np.max([r["err_learned"] for r in ev]))
    report["mean_simulated_fraction"] = float(np.mean([r["simulated_fraction"] for r in ev]))
    json.dump(report, open(os.path.join(args.out, "schedule_report.json"), "w"), indent=1)
    log(f"MAE learned {report['mae_learned']:.3f}  uniform(same n) {report['mae_uniform_same_n']:.3f} kcal/mol")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m q_amd.schedule")
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build", help="pilot data, GPU training of the overlap MLP, held-out evaluation")
    b.add_argument("--out", required=True)
    b.add_argument("--n-train", type=int, default=200)
    b.add_argument("--n-train-protein", type=int, default=0)
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
