"""Benchmark: aggregate ns/day of 64 batched FEP windows on one MI355X.

Workload (one "step" = one MD step of every window): a lambda ladder of 64
windows -- 32 windows of the ligand-in-water leg (50-atom two-state ligand in
a 20 A SPC sphere, ~3.4k atoms) and 32 windows of the protein leg (160
16-atom units + 40-atom two-state ligand in a 25 A sphere, ~5.7k atoms),
lambda_k = k/31 per leg, replica seeds 1-2 (water leg) and 3-4 (protein
leg), 1 fs steps, LRF, SHAKE, pair lists every 25 steps, energies every 10
and summaries every 5 steps -- all windows advanced together by one engine.

Printed: one JSON line with the aggregate ns/day, the one-window-per-engine
rate of the same windows run one after the other (the plain GPU form), the
live roofline of the dominant kernel family and the reference CPU baseline
(Fortran qdyn built from the reference sources, timed on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def make_ladder(work: str, n_water: int, n_protein: int, steps_in_inp: int):
    from q_amd import synth
    inps = []
    legs = []
    if n_water:
        legs.append(("ligwat", 0, n_water, 1))
    if n_protein:
        legs.append(("protein", 2, n_protein, 3))
    for cfg, sys_seed, n, md_seed0 in legs:
        d = os.path.join(work, cfg)
        os.makedirs(d, exist_ok=True)
        system = None
        for k in range(n):
            lam = k / (n - 1) if n > 1 else 0.5
            inp = synth.make_window(cfg, d, seed=sys_seed, md_seed=md_seed0 + (k % 2), lam=lam,
                                    steps=steps_in_inp, energy=10, output=5, nb_cycle=25,
                                    tag=f"{cfg}_w{k:02d}", system=system if k == 0 else None)
            inps.append(inp)
    return inps


def run_engine(inps, warmup, steps, timing=False):
    from q_amd.engine import Engine
    e = Engine(0)
    for p in inps:
        e.add_window(p)
    e.prepare()
    if warmup:
        e.run(warmup, finalize=False)
    e.set_timing(timing)
    t0 = time.perf_counter()
    e.run(steps, finalize=False)
    t1 = time.perf_counter()
    return e, t1 - t0


def cpu_reference(inp_water, inp_protein, n_water_steps, n_protein_steps):
    """Serial Fortran qdyn (oracle/_ref/qdyn) on one window of each leg."""
    from oracle import qprobe_dump as Q
    if not os.path.exists(Q.QDYN):
        return None
    out = {}
    for name, inp, n in (("ligwat", inp_water, n_water_steps), ("protein", inp_protein, n_protein_steps)):
        d = os.path.dirname(inp)
        src = open(inp).read()
        src = re.sub(r"steps \d+", f"steps {n}", src, count=1)
        p = os.path.join(d, f"cpu_{name}.inp")
        open(p, "w").write(src.replace(f"{os.path.splitext(os.path.basename(inp))[0]}.", f"cpu_{name}."))
        t0 = time.perf_counter()
        r = subprocess.run([Q.QDYN, os.path.basename(p)], cwd=d, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        m = re.search(r"Total time of main loop:\s+([0-9.]+)", r.stdout)
        loop = float(m.group(1)) if m else wall
        out[name] = dict(steps=n, seconds=loop if loop > 0 else wall, wall=wall)
    return out


def traffic_per_launch(kernel: str):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor), as recorded
    in profiles/traffic.txt ("kernel bytes" lines); None when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        t = line.split()
        if len(t) >= 2 and t[0] == kernel and not line.startswith("#"):
            return float(t[1])
    return None


def copy_probe(nbytes: int):
    """A plain device copy of nbytes (torch clone), the calibration kernel
    for rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import torch
    x = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    x.fill_(1)
    torch.cuda.synchronize()
    y = x.clone()
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "ok": bool(y[0].item() == 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--windows", type=int, default=64, help="total windows (half per leg)")
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=400)
    ap.add_argument("--copy-probe", type=int, default=0, help="only run a plain device copy of this many bytes")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs shard windows and are not "
                                   "part of this build", "n_gpus": args.gpus}))
        sys.exit(2)

    work = tempfile.mkdtemp(prefix="qmi_bench_")
    nw = args.windows // 2
    n_timing = max(20, min(args.steps, 100))
    # every window must still be running through the timing pass
    steps_inp = args.warmup + args.steps + n_timing + 10
    t = time.perf_counter()
    inps = make_ladder(work, nw, args.windows - nw, steps_inp)
    t_gen = time.perf_counter() - t
    print(f"[bench] generated {len(inps)} windows in {t_gen:.1f}s", file=sys.stderr, flush=True)

    e, el = run_engine(inps, args.warmup, args.steps, timing=False)
    infos = [e.info(w) for w in range(len(inps))]
    tuning = e.tuning()
    dt_fs = 1.0
    ns_sim = len(inps) * args.steps * dt_fs * 1e-6
    value = ns_sim / el * 86400.0
    ms_per_step = el / args.steps * 1e3
    print(f"[bench] batched: {ms_per_step:.3f} ms/step, {value:.1f} ns/day", file=sys.stderr, flush=True)

    # live per-family device timing over a second timed region of the same engine
    e.set_timing(True)
    e.run(n_timing, finalize=False)
    fam = e.timings()
    e.close()

    # work counts for the roofline: water-water molecule pairs of every window
    from q_amd.engine import LIST_WW, Engine
    e2 = Engine(0)
    for p in inps[:1] + inps[nw:nw + 1]:
        e2.add_window(p)
    e2.prepare()
    e2.evaluate(rebuild_lists=True)
    n_ww = [len(e2.pairlist(0, LIST_WW)[0]), len(e2.pairlist(1, LIST_WW)[0])]
    e2.close()
    n_ww_total = nw * n_ww[0] + (len(inps) - nw) * n_ww[1]
    # water-water tile kernel: 200 FP32 FLOP per listed molecule pair
    # (1 LJ + Coulomb and 8 Coulomb-only site pairs) over the kernel's own
    # device time, measured with HIP events on the engine's stream
    t_ww = fam[4] * 1e-6 if len(fam) > 4 and fam[4] > 0 else None
    achieved = (200.0 * n_ww_total / t_ww / 1e12) if t_ww else None

    result = {
        "metric": "aggregate ns/day, 64 batched FEP windows on 1 GPU",
        "value": round(value, 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair math, int64 fixed-point sums, f64 integrator/SHAKE/LRF",
        "data": "synthetic",
        "config": {
            "workload": "batched lambda ladder: 32 ligand-in-water windows + 32 protein-leg windows, 1 fs, "
                        "one engine, all windows advanced per step",
            "windows": len(inps),
            "atoms_per_window": sorted({i["natom"] for i in infos}),
            "total_atoms": sum(i["natom"] for i in infos),
            "parallelism": "batched windows on 1 GPU",
        },
        "family_us_per_step": {
            "lists_lrf_build": round(fam[0], 2) if fam else None,
            "nonbonded_tiles": round(fam[1], 2) if fam else None,
            "lrf_bonded_restraints": round(fam[2], 2) if fam else None,
            "integrate_shake_thermostat": round(fam[3], 2) if fam else None,
            "ww_kernel": round(fam[4], 2) if fam else None,
            "wall_per_step_timed_pass": round(fam[5], 2) if fam else None,
        },
        "autotune": tuning,
        "roofline": {
            "bound": "mfma", "kernel": "k_ww<false> (SPC water-water rows, FP32 vector ALU; compute roof = FP32 peak)",
            "achieved": round(achieved, 4) if achieved else None, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 5) if achieved else None,
            "traffic": traffic_per_launch("k_ww<false>"),
            "work": {"ww_molecule_pairs_per_step": n_ww_total, "flop_per_pair": 200,
                     "kernel_us_per_launch": round(fam[4], 2) if fam else None},
        },
    }

    if not args.no_baselines:
        # plain GPU form: one window per engine, windows one after the other
        per = []
        for p in (inps[0], inps[nw]):
            e1, t1 = run_engine([p], 10, max(50, args.steps // 5))
            per.append(t1 / max(50, args.steps // 5))
            e1.close()
        seq = nw * per[0] + (len(inps) - nw) * per[1]
        result["one_window_per_launch"] = {
            "ms_per_step_ligwat": round(per[0] * 1e3, 4), "ms_per_step_protein": round(per[1] * 1e3, 4),
            "aggregate_ns_per_day": round(len(inps) * dt_fs * 1e-6 / seq * 86400.0, 3),
            "batched_speedup": round(seq / (el / args.steps), 2),
        }
        cpu = cpu_reference(inps[0], inps[nw], args.cpu_steps, args.cpu_steps // 2)
        if cpu:
            s_l = cpu["ligwat"]["seconds"] / cpu["ligwat"]["steps"]
            s_p = cpu["protein"]["seconds"] / cpu["protein"]["steps"]
            seq_cpu = nw * s_l + (len(inps) - nw) * s_p
            result["cpu_baseline"] = {
                "value": round(len(inps) * dt_fs * 1e-6 / seq_cpu * 86400.0, 4),
                "unit": "ns/day", "cores": 1, "kind": "reference",
                "sample": f"serial Fortran qdyn (reference sources, amdflang -O2): {cpu['ligwat']['steps']} steps "
                          f"of one ligand-in-water window ({s_l * 1e3:.1f} ms/step) and "
                          f"{cpu['protein']['steps']} steps of one protein-leg window ({s_p * 1e3:.1f} ms/step); "
                          "aggregate for the 64-window set run one window after another on 1 core",
            }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
