"""Benchmark: aggregate ns/day of 64 batched FEP windows on one MI355X.

Workload (one "step" = one MD step of every window): a lambda ladder of 64
windows -- 32 windows of the ligand-in-water leg (50-atom two-state ligand in
a 20 A SPC sphere, ~3.4k atoms) and 32 windows of the protein leg (160
16-atom units + 40-atom two-state ligand in a 25 A sphere with ~1700 SPC
waters, ~7.7k atoms),
lambda_k = k/31 per leg, replica seeds 1-2 (water leg) and 3-4 (protein
leg), 1 fs steps, LRF, SHAKE, pair lists every 25 steps, energies every 10
and summaries every 5 steps -- all windows advanced together by one engine.

The timed region starts at a pair-list update step (the warm-up is followed
by the few untimed steps that reach the next multiple of 25), so list
updates and the LRF build are timed at their 1-in-25 rate (or more often, in
short runs) for any --steps / --warmup.

Printed: one JSON line with the aggregate ns/day, the per-system rates, the
one-window-per-engine rate of the same windows run one after the other over
the same step pattern (the plain GPU form), the live roofline of the
water-water kernel and the reference CPU baseline (Fortran qdyn built from
the reference sources, one serial process per host core, timed on this
host).
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

FP32_PEAK_TFLOPS = 157.3           # dense FP32 vector rate with packed FP32 instructions
FP32_NONPACKED_TFLOPS = 78.6       # one FP32 FMA per lane per cycle (packed FP32 is disabled in this build)
HBM_PEAK_GBS = 8000.0
NB_CYCLE = 25                      # pair-list update interval of the benchmark windows


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


def list_aligned(warmup: int) -> int:
    """Untimed steps after `warmup` that make the timed region start at a
    pair-list update step (a multiple of NB_CYCLE)."""
    return (-warmup) % NB_CYCLE


def list_steps(first: int, n: int) -> int:
    return sum(1 for k in range(first, first + n) if k % NB_CYCLE == 0)


def run_engine(inps, warmup, steps, timing=False):
    """Warm-up, then `steps` timed steps from a list-update step."""
    from q_amd.engine import Engine
    e = Engine(0)
    for p in inps:
        e.add_window(p)
    e.prepare()
    pre = warmup + list_aligned(warmup)
    if pre:
        e.run(pre, finalize=False)
    e.set_timing(timing)
    t0 = time.perf_counter()
    e.run(steps, finalize=False)
    t1 = time.perf_counter()
    return e, t1 - t0


def cpu_reference(inps_water, inps_protein, water_steps, protein_steps, cores):
    """Serial Fortran qdyn (oracle/_ref/qdyn) as the reference runs a ladder
    on a CPU node: one process per core, here `cores` windows at once (half
    of each leg), each for a bounded number of steps.  Returns per-window
    main-loop seconds per step."""
    from oracle import qprobe_dump as Q
    if not os.path.exists(Q.QDYN):
        return None
    procs = []
    nw = (cores + 1) // 2
    picks = [("ligwat", p, water_steps) for p in inps_water[:nw]] + \
            [("protein", p, protein_steps) for p in inps_protein[:cores - nw]]
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(cores))
    t0 = time.perf_counter()
    for k, (name, inp, n) in enumerate(picks):
        d = os.path.dirname(inp)
        stem = os.path.splitext(os.path.basename(inp))[0]
        src = open(inp).read()
        src = re.sub(r"steps \d+", f"steps {n}", src, count=1)
        p = os.path.join(d, f"cpu_{stem}.inp")
        open(p, "w").write(src.replace(f"{stem}.", f"cpu_{stem}."))
        cmd = [Q.QDYN, os.path.basename(p)]
        if os.path.exists("/usr/bin/taskset") and k < len(cpus):
            cmd = ["taskset", "-c", str(cpus[k])] + cmd
        procs.append((name, n, subprocess.Popen(cmd, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                                text=True)))
    out = {"ligwat": [], "protein": []}
    for name, n, pr in procs:
        txt, _ = pr.communicate(timeout=600)
        m = re.search(r"Total time of main loop:\s+([0-9.]+)", txt)
        if pr.returncode != 0 or not m or "terminated normally" not in txt:
            return None
        out[name].append(float(m.group(1)) / n)
    out["wall"] = time.perf_counter() - t0
    return out


def traffic_per_launch(kernel: str):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor), as recorded
    in profiles/traffic.txt (tab-separated "kernel, bytes, ..." lines); None
    when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        t = line.rstrip("\n").split("\t")
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
    ap.add_argument("--cpu-steps", type=int, default=1000, help="steps of each CPU ligand-in-water window "
                    "(protein-leg windows run a third of it)")
    ap.add_argument("--cpu-cores", type=int, default=16, help="CPU baseline: concurrent serial qdyn processes")
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
    n_one = 4 * NB_CYCLE                       # one-window runs: 100 steps from a list step
    pre = args.warmup + list_aligned(args.warmup)
    # every window must still be running through the timing pass
    steps_inp = pre + args.steps + n_timing + 2 * NB_CYCLE
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

    # live per-family device timing over a second timed region of the same
    # engine (the kernels one after the other on one stream, HIP events)
    e.set_timing(True)
    e.run(n_timing, finalize=False)
    fam = e.timings()
    # the water-water force kernel alone: 50 launches in a row on the
    # engine's stream between two HIP events, at the timed run's lists
    us_ww = e.time_kernel(0, 50)
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
    # water-water kernel: 200 FP32 FLOP per listed molecule pair (1 LJ +
    # Coulomb and 8 Coulomb-only site pairs) over the kernel's own device
    # time per launch, measured with HIP events on the engine's stream
    t_ww = us_ww * 1e-6 if us_ww > 0 else None
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
            "waters_per_window": sorted({i["nwat"] for i in infos}),
            "total_atoms": sum(i["natom"] for i in infos),
            "parallelism": "batched windows on 1 GPU",
            "timed_steps": [pre, pre + args.steps - 1],
            "list_updates_timed": list_steps(pre, args.steps),
            "untimed_alignment_steps": list_aligned(args.warmup),
        },
        "per_system_ns_per_day": {"batched_each_window": round(value / len(inps), 3)},
        "family_us_per_step": {
            "lists_lrf_build": round(fam[0], 2) if fam else None,
            "nonbonded_tiles": round(fam[1], 2) if fam else None,
            "bonded_restraints_polarisation": round(fam[2], 2) if fam else None,
            "lrf_eval_integrate_shake_thermostat": round(fam[3], 2) if fam else None,
            "ww_kernel": round(fam[4], 2) if fam else None,
            "wall_per_step_timed_pass": round(fam[5], 2) if fam else None,
        },
        "autotune": tuning,
        "roofline": {
            "bound": "valu_fp32",
            "kernel": "k_ww<false, false> (SPC water-water rows, force-only instance; FP32 vector ALU; compute roof = FP32 vector peak)",
            "achieved": round(achieved, 4) if achieved else None, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 5) if achieved else None,
            "frac_nonpacked": round(achieved / FP32_NONPACKED_TFLOPS, 5) if achieved else None,
            "traffic": traffic_per_launch("k_ww<false, false>"),
            "work": {"ww_molecule_pairs_per_step": n_ww_total, "flop_per_pair": 200,
                     "kernel_us_per_launch": round(us_ww, 2),
                     "timing": "50 back-to-back launches at the timed run's lists, HIP events"},
        },
    }

    if not args.no_baselines:
        # plain GPU form: one window per engine, windows one after the other,
        # each timed over 100 steps from a list-update step
        per = []
        for p in (inps[0], inps[nw]):
            e1, t1 = run_engine([p], NB_CYCLE, n_one)
            per.append(t1 / n_one)
            e1.close()
        seq = nw * per[0] + (len(inps) - nw) * per[1]
        result["one_window_per_launch"] = {
            "ms_per_step_ligwat": round(per[0] * 1e3, 4), "ms_per_step_protein": round(per[1] * 1e3, 4),
            "aggregate_ns_per_day": round(len(inps) * dt_fs * 1e-6 / seq * 86400.0, 3),
            "batched_speedup": round(seq / (el / args.steps), 2),
            "timed_steps": [NB_CYCLE, NB_CYCLE + n_one - 1],
        }
        result["per_system_ns_per_day"].update({
            "one_window_ligwat": round(dt_fs * 1e-6 / per[0] * 86400.0, 3),
            "one_window_protein": round(dt_fs * 1e-6 / per[1] * 86400.0, 3),
        })
        cores = max(1, min(args.cpu_cores, len(inps),
                           len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1))
        cpu = cpu_reference(inps[:nw], inps[nw:], args.cpu_steps, args.cpu_steps // 3, cores)
        if cpu:
            s_l = sum(cpu["ligwat"]) / len(cpu["ligwat"]) if cpu["ligwat"] else None
            s_p = sum(cpu["protein"]) / len(cpu["protein"]) if cpu["protein"] else None
            rate = sum(dt_fs * 1e-6 / t * 86400.0 for t in cpu["ligwat"] + cpu["protein"])
            model = ""
            try:
                for line in open("/proc/cpuinfo"):
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
            except OSError:
                pass
            result["cpu_baseline"] = {
                "value": round(rate, 4),
                "unit": "ns/day", "cores": cores, "kind": "reference",
                "sample": f"Fortran qdyn (reference sources, amdflang -O2), one serial process per core on "
                          f"{cores} cores ({model}): {len(cpu['ligwat'])} ligand-in-water windows x "
                          f"{args.cpu_steps} steps ({s_l * 1e3:.1f} ms/step) and {len(cpu['protein'])} protein-leg "
                          f"windows x {args.cpu_steps // 3} steps ({s_p * 1e3:.1f} ms/step) at once, "
                          f"{cpu['wall']:.1f} s; value = the summed ns/day of the concurrent windows",
            }
            result["per_system_ns_per_day"].update({
                "cpu_ligwat": round(dt_fs * 1e-6 / s_l * 86400.0, 3) if s_l else None,
                "cpu_protein": round(dt_fs * 1e-6 / s_p * 86400.0, 3) if s_p else None,
            })
    print(json.dumps(result))


if __name__ == "__main__":
    main()
