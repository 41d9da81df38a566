#!/usr/bin/env python3
"""Benchmark of the chiML time-stepping path on one MI355X.

Workload: 3D Yee 256^3, dx = 2 nm, 16-cell CPML, diagonal anisotropic
background eps = (2.0, 2.25, 2.5), Drude + 2-pole Lorentz sphere of radius 40
cells, soft plane-wave sheet, 40-frequency DFT flux box of 120^3 points.
A step is one pass of the reference's step() over the whole domain; the metric
is Mcell-updates/s = cells * steps / seconds / 1e6 with the fields resident in
HBM.  The FP32 path is the headline (`value`); the FP64 path is measured on
the same deck and reported beside it.

roofline: the whole Yee + CPML + ADE step (the field kernels) against the
bytes a step must move when every state word is read once and written once
(fields, one material byte per cell and E component, CPML psi, D and two P
words per pole), over the field kernels' mean time per step.  That time is
measured with HIP events on the simulation's stream around each field-kernel
launch in the timed region.  traffic is the field kernels' HBM bytes per step
from the committed rocprofv3 --pmc summary (profiles/traffic_config2_<precision>.json).
kernel_roofline: the dominant kernel alone, against its own per-launch bytes.
accuracy: FP32 against FP64 on fixed runs, whatever --steps is: the 256^3
deck with six point detectors for 1000 steps (detector series and flux
spectrum), and a 2D Maxwell-Liouville deck for 4000 steps (populations).
cpu_baseline: the FP64 CPU restatement of the reference step loop (the
oracle) on this host for a few steps of the same 256^3 deck, with its update
lists split over y slabs across P threads (P = the host's CPU share), and on
one thread beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--n", type=int, default=256, help="points per axis")
    p.add_argument("--no-fp64", action="store_true", help="skip the FP64 leg")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-steps", type=int, default=16)  # about 8 s of one core on 256^3
    p.add_argument("--cpu-steps-par", type=int, default=48, help="timed oracle steps on the threaded CPU leg")
    p.add_argument("--no-accuracy", action="store_true", help="skip the fixed FP32-vs-FP64 accuracy runs")
    p.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    p.add_argument("--sweep-requests", type=int, default=32,
                   help="requests in the serving-sweep sample (0 skips the sweep leg)")
    p.add_argument("--sweep-steps", type=int, nargs=2, default=[100, 1000],
                   help="step range of the sweep sample (the full sweep uses 1000 10000)")
    p.add_argument("--no-sweep-baseline", action="store_true", help="skip the one-at-a-time replay")
    p.add_argument("--sweep-active", type=int, default=16, help="simulations the server holds at once")
    p.add_argument("--sweep-progress", type=int, default=0, help="print a line every this many finished requests")
    p.add_argument("--only-sweep", action="store_true", help="run the serving sweep alone and print its line")
    p.add_argument("--no-configs", action="store_true",
                   help="skip the other workloads (2D TMz, 2D and 3D Maxwell-Liouville)")
    p.add_argument("--only", choices=["tmz", "ml2d", "ml3d", "tfsf3d"],
                   help="run one of the other workloads alone and print its line (for profiling)")
    p.add_argument("--only-steps", type=int, default=0, help="step count for --only (0: the workload's own)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="first run a plain 256 MiB device copy three times (the known-size kernel that scales "
                        "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    return p.parse_args()


KERNELS = ("update_fused", "update_ade", "update_h", "update_h_pml", "update_e", "update_e_pml", "dft_flush", "ml_density")
FIELD_KERNELS = ("update_fused", "update_ade", "update_h", "update_h_pml", "update_e", "update_e_pml")


def run_gpu(n, steps, warmup, precision):
    import torch
    import chiml_amd
    from chiml_amd import decks

    prof = min(steps, 200)
    deck = decks.config2_3d(n=n, steps=warmup + prof + steps, precision=precision)
    sim = chiml_amd.FDTDField(deck)
    cells = sim.cost()[0]
    sim.step(warmup)
    sim.sync()
    # untimed pass: every launch timed with HIP events -> per-kernel breakdown
    sim.profile(True)
    sim.step(prof)
    sim.sync()
    kern = {}
    for k in KERNELS:
        ms, nl = sim.profile_read(k)
        if nl > 0:
            kern[k] = {"ms_per_launch": ms / nl, "launches": nl, "ms_per_step": ms / max(prof, 1),
                       "bytes_per_launch": sim.kernel_bytes(k)}
    cand = [k for k in kern if kern[k]["bytes_per_launch"] > 0]
    dominant = max(cand, key=lambda k: kern[k]["ms_per_step"]) if cand else "update_e"
    # timed region: one event pair per step around the field launches
    field = [k for k in FIELD_KERNELS if k in kern]
    sim.profile_only("field_step")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    step_ms, step_n = sim.profile_read("field_step")
    dom_ms, dom_n = kern[dominant]["ms_per_launch"] * kern[dominant]["launches"], kern[dominant]["launches"]
    sim.profile(False)
    flux = sim.fluxArr()[0].getFlux() if sim.fluxArr() else None
    wall = t1 - t0
    res = {
        "cells": cells,
        "wall_s": wall,
        "mcells": cells * steps / wall / 1e6,
        "ms_per_step": wall / steps * 1e3,
        "kernels": kern,
        "dominant": dominant,
        "dominant_ms": dom_ms / max(dom_n, 1),
        "dominant_launches": dom_n,
        "dominant_bytes": sim.kernel_bytes(dominant),
        "field_ms_per_step": step_ms / max(step_n, 1),
        "field_steps_timed": step_n,
        "field_launches": {k: kern[k]["launches"] for k in field},
        "yee_step_min": sim.kernel_bytes("yee_step_min"),
        "footprint": sim.footprint(),
        "flux0": None if flux is None else [float(flux[0].real), float(flux[0].imag)],
    }
    sim.close()
    return res


def load_traffic(precision):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)."""
    path = os.path.join(ROOT, "profiles", f"traffic_config2_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def kernel_roofline(r, precision):
    """The dominant kernel alone: its algorithmic bytes per launch over its
    mean launch time in the profiled pass just before the timed region
    (HIP events around every launch)."""
    name = r["dominant"]
    ms = r["dominant_ms"]
    b = r["dominant_bytes"]
    achieved = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traffic = None
    t = load_traffic(precision)
    if t and name in t.get("kernels", {}):
        traffic = t["kernels"][name]["hbm_bytes_per_launch"]
    return {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": name,
        "ms_per_launch": round(ms, 5),
        "launches": r["dominant_launches"],
        "algorithmic_bytes_per_launch": b,
    }


def step_roofline(r, precision):
    """The Yee + CPML + ADE step against the compulsory bytes of a step in
    which every state word is read once and written once, over the field
    launches' mean time per step in the timed region (one HIP event pair per
    step around them); traffic sums the field kernels' measured HBM bytes per
    step."""
    b = r["yee_step_min"]
    ms = r["field_ms_per_step"]
    if b <= 0 or ms <= 0:
        return None
    achieved = b / (ms * 1e-3) / 1e9
    traffic = None
    t = load_traffic(precision)
    if t:
        ks = t.get("kernels", {})
        if all(k in ks for k in r["field_launches"]):
            traffic = int(sum(ks[k]["hbm_bytes_per_launch"] for k in r["field_launches"]))
    return {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernels": sorted(r["field_launches"]),
        "algorithmic_bytes_per_step": b,
        "field_kernels_ms_per_step": round(ms, 5),
        "frac_of_wall_step": round(b / (r["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }


def kernel_table(r):
    out = {}
    for k, v in r["kernels"].items():
        e = {"ms_per_launch": round(v["ms_per_launch"], 5), "launches": v["launches"], "ms_per_step": round(v["ms_per_step"], 5)}
        if v["bytes_per_launch"] > 0:
            e["GBps"] = round(v["bytes_per_launch"] / (v["ms_per_launch"] * 1e-3) / 1e9, 1)
        out[k] = e
    return out


def run_workload(deck, graph_steps, profile_kernel=None):
    """Mcell-updates/s of one deck run to its end on the GPU (after a short warmup)."""
    import torch
    import chiml_amd

    sim = chiml_amd.FDTDField(deck)
    cells = sim.cost()[0]
    warm = min(64, sim.nsteps // 10)
    if graph_steps:
        sim.set_graph(graph_steps)
    sim.step(warm)
    sim.sync()
    if profile_kernel and not graph_steps:
        sim.profile_only(profile_kernel)
    steps = sim.nsteps - warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    t1 = time.perf_counter()
    out = {"mcells_per_s": round(cells * steps / (t1 - t0) / 1e6, 1), "ms_per_step": round((t1 - t0) / steps * 1e3, 4),
           "steps": steps, "cells": int(cells), "graph_steps": graph_steps}
    if profile_kernel and not graph_steps:
        ms, n = sim.profile_read(profile_kernel)
        if n:
            b = sim.kernel_bytes(profile_kernel)
            out[profile_kernel] = {"ms_per_launch": round(ms / n, 5), "GBps": round(b / (ms / n * 1e-3) / 1e9, 1)}
    sim.close()
    return out


def other_workload(name, steps=0):
    """(deck, graph steps, profiled kernel) of one of the other workloads at its
    full size; steps > 0 overrides its step count."""
    from chiml_amd import decks

    if name == "tmz":
        return decks.config1_2d(n=256, steps=steps or 2000, profile="gaussian"), 64, None
    if name == "ml2d":
        return decks.config3_ml2d(n=1024, steps=steps or 20000), 64, None
    if name == "tfsf3d":
        return decks.tfsf_deck((256, 256, 256), steps or 1000, (201, 201, 201), m=(0, 0, 1), pml=16, precision="fp32"), 0, None
    return decks.config4_ml3d(n=192, steps=steps or 5000), 0, "ml_density"


def run_configs():
    """The other workloads of the path, each at its full size: 2D TMz 256^2
    (2000 steps, FP64), 2D Maxwell-Liouville 1024^2 with 32,768 2-level
    emitters (20,000 steps, FP32 fields / FP64 rho) and 3D Maxwell-Liouville
    192^3 with 368,640 (192 x 192 x 10) 6-level emitters over a 20-plane
    Drude film (5000 steps), and 3D 256^3 FP32 vacuum with a 201^3 TFSF box
    (normal incidence, 1000 steps)."""
    res = {}
    res["tmz_256_cylinder_fp64"] = run_workload(*other_workload("tmz"))
    res["ml2d_1024_2level_fp32"] = run_workload(*other_workload("ml2d"))
    res["ml3d_192_6level_fp32"] = run_workload(*other_workload("ml3d"))
    res["tfsf_256_box201_fp32"] = run_workload(*other_workload("tfsf3d"))
    return res


def run_sweep(args):
    """Serving leg: seeded requests (Poisson 20/s; 2D 128^2-1024^2 and 3D
    64^3-192^3; dielectric or Drude object; a quarter with 2- or 4-level
    emitters; FP32 fields) through the sweep server, then the same requests
    one at a time on the plain GPU path."""
    import torch
    from chiml_amd import decks, sweep

    reqs = decks.sweep_requests(args.sweep_requests, steps_range=tuple(args.sweep_steps))
    total = torch.cuda.get_device_properties(0).total_memory
    res = {"sample": f"{args.sweep_requests} requests, steps {args.sweep_steps[0]}-{args.sweep_steps[1]}, "
                     f"seed 5, Poisson 20/s",
           "server": sweep.run_server(reqs, hbm_budget=int(0.6 * total), max_active=args.sweep_active,
                                      chunk_steps=64, graph_steps=0, batch=True, progress=args.sweep_progress),
           "server_streams": sweep.run_server(reqs, hbm_budget=int(0.6 * total), max_active=args.sweep_active,
                                              chunk_steps=64, graph_steps=32, progress=args.sweep_progress)}
    res["server"]["mode"] = "shared launches (chiml_dev_step_batch), 64-step chunks"
    res["server_streams"]["mode"] = "one stream per simulation, 32-step captured graphs"
    if not args.no_sweep_baseline:
        res["one_at_a_time"] = sweep.run_one_at_a_time(reqs, progress=args.sweep_progress)
        res["throughput_ratio"] = round(res["server"]["mcells_per_s"] / res["one_at_a_time"]["mcells_per_s"], 3)
    return res


def rel_linf(a, b):
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    den = float(np.max(np.abs(b))) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / den) if den > 0 else None


def accuracy_leg(n=256):
    """FP32 fields against FP64 fields on fixed runs (independent of --steps):
    config 2 at n^3 with six point detectors (Ex..Hz) for 1000 steps, and a
    2D Maxwell-Liouville deck (256^2, 128 x 16 two-level emitters, resonant
    pulse) for 4000 steps.  Errors are normalised L-inf per series."""
    import chiml_amd
    from chiml_amd import decks

    out = {}
    res = {}
    for prec in ("fp32", "fp64"):
        sim = chiml_amd.FDTDField(decks.config2_3d(n=n, steps=1000, precision=prec, ndet=6))
        sim.step(sim.nsteps)
        sim.sync()
        res[prec] = ([d.series()[1] for d in sim.dtcArr()], sim.fluxArr()[0].getFlux())
        sim.close()
    import numpy as np
    det = [rel_linf(a, b) for a, b in zip(res["fp32"][0], res["fp64"][0])]
    # FP32 rounding follows the magnitude of the strong components: each
    # series' error is also stated against the peak of its field family (E or
    # H) at the same point, the scale its rounding is relative to
    fam = []
    for k, (a, b) in enumerate(zip(res["fp32"][0], res["fp64"][0])):
        grp = range(0, 3) if k < 3 else range(3, 6)
        scale = max(float(np.max(np.abs(res["fp64"][0][j]))) for j in grp)
        fam.append(float(np.max(np.abs(a - b))) / scale if scale > 0 else None)
    names = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]
    out["config2_1000_steps"] = {
        "detectors_err_over_family_peak": {c: None if e is None else float(f"{e:.3e}") for c, e in zip(names, fam)},
        "detectors_rel_linf": {c: None if e is None else float(f"{e:.3e}") for c, e in zip(names, det)},
        "flux_spectrum_rel_linf": float(f"{rel_linf(res['fp32'][1], res['fp64'][1]):.3e}"),
    }
    pops = {}
    for prec in ("fp32", "fp64"):
        sim = chiml_amd.FDTDField(decks.config3_ml2d(n=256, steps=4000, pml=20, slab=(128, 16), precision=prec))
        sim.step(sim.nsteps)
        sim.sync()
        q = sim.qeArr()[0]
        pops[prec] = [q.population(0), q.population(1)]
        sim.close()
    p32 = np.concatenate(pops["fp32"])
    p64 = np.concatenate(pops["fp64"])
    # the excited population is the small one: its error is also stated
    # against its own peak
    out["ml2d_256_4000_steps"] = {
        "populations_abs_linf": float(f"{float(np.max(np.abs(p32 - p64))):.3e}"),
        "excited_population_rel_linf": float(f"{rel_linf(pops['fp32'][1], pops['fp64'][1]):.3e}"),
        "excited_population_peak": float(f"{float(np.max(np.abs(pops['fp64'][1]))):.3e}"),
    }
    out["bound"] = 1e-5
    return out


def cpu_share():
    """Threads for the CPU baseline: the CPUs this process may run on, capped
    by OMP_NUM_THREADS when it is set (the host's CPU share for one GPU)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_time(n, steps, threads):
    from chiml_amd import decks
    from oracle import pyoracle

    pyoracle.set_threads(threads)
    deck = decks.config2_3d(n=n, steps=steps + 1, precision="fp64")
    t0 = time.perf_counter()
    o = pyoracle.Oracle(deck)
    setup = time.perf_counter() - t0
    o.step(1)  # first touch
    t1 = time.perf_counter()
    o.step(steps)
    t2 = time.perf_counter()
    del o
    pyoracle.set_threads(1)
    return n ** 3 * steps / (t2 - t1) / 1e6, t2 - t1, setup


def cpu_baseline(n, steps, steps_par):
    p = cpu_share()
    v1, s1, setup = cpu_time(n, steps, 1)
    vp, sp, _ = cpu_time(n, steps_par, p)
    return {
        "value": round(vp, 3),
        "unit": "Mcell-updates/s",
        "cores": p,
        "kind": "port",
        "sample": f"FP64 oracle (reference step loop restated, BLAS-1 per x-run), same {n}^3 deck; update lists "
                  f"split over y slabs across {p} threads: {steps_par} timed steps ({sp:.1f} s) after 1 untimed step; "
                  f"setup {setup:.1f} s excluded",
        "one_thread": {"value": round(v1, 3), "steps": steps, "seconds": round(s1, 1)},
        "nproc": os.cpu_count(),
        "cpu_share": p,
        "cpu_model": cpu_model(),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU slabs are a design note only)"}))
        sys.exit(2)
    if args.copy_calibration:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.only:
        print(json.dumps({"workload": args.only, **run_workload(*other_workload(args.only, args.only_steps))}), flush=True)
        return
    if args.only_sweep:
        print(json.dumps({"workload": "serving sweep", **run_sweep(args)}), flush=True)
        return
    r32 = run_gpu(args.n, args.steps, args.warmup, args.precision)
    out = {
        "metric": "Mcell-updates/s + % HBM roofline (FP32 vs FP64); sweep p50/p99 latency",
        "value": round(r32["mcells"], 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D Yee {args.n}^3 dx=2nm, 16-cell CPML, aniso eps (2.0,2.25,2.5), Drude+2 Lorentz sphere r=40, "
                        f"soft plane-wave sheet, 40-freq DFT flux box 120^3",
            "cells": int(r32["cells"]),
            "precision": args.precision,
            "parallelism": "single GPU, one device-resident domain",
        },
        "roofline": step_roofline(r32, args.precision),
        "kernel_roofline": kernel_roofline(r32, args.precision),
        "kernels": kernel_table(r32),
    }
    if not args.no_fp64 and args.precision == "fp32":
        r64 = run_gpu(args.n, args.steps, args.warmup, "fp64")
        out["fp64"] = {
            "value": round(r64["mcells"], 1),
            "ms_per_step": round(r64["ms_per_step"], 4),
            "roofline": step_roofline(r64, "fp64"),
            "kernel_roofline": kernel_roofline(r64, "fp64"),
            "kernels": kernel_table(r64),
        }
        out["fp32_over_fp64"] = round(r32["mcells"] / r64["mcells"], 3)
    if not args.no_accuracy and args.n == 256:
        out["accuracy_fp32_vs_fp64"] = accuracy_leg(args.n)
    if not args.no_configs:
        out["other_workloads"] = run_configs()
    if args.sweep_requests > 0:
        out["sweep"] = run_sweep(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.n, args.cpu_steps, args.cpu_steps_par)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
