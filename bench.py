"""Benchmark: MHD cell-updates/s of the Godunov stage on one MI355X.

Workload: seeded 3-D MHD turbulence (PLM + HLLD + CT + Heun RK2), 256^3
periodic cells as 512 blocks of 32^3, gamma = 5/3, beta = 1, rms Mach 0.5.
A "step" is one Heun cycle (two RK substeps) over every cell; cell-updates/s
= cells x cycles / wall seconds, the reference's Speed definition
(mesh/cycle/cycle.cpp:184-190).  The state is resident in HBM before the
timed region; the timed region is K cycles replayed from the captured
hipGraph, bracketed by device synchronisation.

Both precision forms run: the plain FP64 form and the mixed form (FP32
reconstruction / Riemann / EMF, FP64 conserved state, CT, energy and dt).
``value`` is the mixed form's rate; the FP64 rate, the speed-up and the
mixed-vs-FP64 L1 errors are reported beside it.

roofline: the dominant kernel (largest share of the cycle) with its
algorithmic bytes per launch divided by its mean launch time (HIP events on
the library's stream).  cpu_baseline: the CPU oracle (FP64 C++ restatement of
the reference algorithm, "port") on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per launch unit (cell or face) of each staged kernel,
# per FP64/FP32 form: reads + writes of the arrays the kernel must touch once.
#   c2p    : U (5 x 8 B) + 6 face-B reads (3 distinct x 8 B) ; W write (8 x fs)
#   flux   : W read (8 x fs, once per cell) ; F write (8 x fs per face)
#   emf    : F reads (3 axes x 3 slots x fs) + W (6 x fs) ; EE write (3 x fs)
#   update : U src (+tgt at substep 1) + F reads + EE + BF r/w
def alg_bytes(kind, fs, step_avg=True):
    if kind == "c2p":
        return (5 + 3) * 8 + 8 * fs
    if kind.startswith("flux"):
        return 8 * fs + 8 * fs
    if kind == "emf":
        return 9 * fs + 6 * fs + 3 * fs
    if kind == "update":
        u = (5 + 5) * 8 if not step_avg else 7.5 * 8 + 5 * 8
        return u + 3 * 5 * fs + 3 * fs + (3 + 3) * 8
    return 0.0


def make_problem(n, seed=0, beta=1.0):
    from kratos_amd import problems
    return problems.turbulence((n, n, n), seed=seed, beta=beta)


def rel_l1(a, b):
    s = np.abs(b).sum()
    return float(np.abs(a - b).sum() / s) if s > 0 else float(np.abs(a - b).sum())


def run_form(spec, phys, u, b, precision, steps, warmup, profile_cycles, path="auto"):
    import kratos_amd
    ctx = kratos_amd.Context(spec, phys, precision, path=path)
    ctx.upload(u, b)
    if warmup:
        ctx.advance(warmup)
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass
    ctx.lib.kr_sync(ctx.h)
    t0 = time.perf_counter()
    done = ctx.advance(steps)
    ctx.lib.kr_sync(ctx.h)
    wall = time.perf_counter() - t0
    st = ctx.stats()
    prof = ctx.profile(profile_cycles) if profile_cycles else None
    return ctx, {"wall": wall, "cycles": done, "gpu_ms": st["gpu_ms"], "prof": prof,
                 "n_cells": st["n_cells"]}


def cpu_baseline(seconds_target=15.0, threads=None):
    """Oracle on a bounded sample: 64^3 cells of the same recipe (8 blocks of 32^3)."""
    import kratos_amd
    import oracle
    threads = threads or min(16, os.cpu_count() or 1)
    n = 64
    spec = kratos_amd.MeshSpec((n, n, n), (32, 32, 32))
    phys = kratos_amd.Physics(mhd=True, gamma=5.0 / 3.0, cfl=0.3)
    u, b, _ = make_problem(n, seed=0)
    o = oracle.OracleMesh(spec, phys, n_threads=threads)
    o.upload(u, b)
    o.advance(1)
    cycles, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds_target and cycles < 200:
        o.advance(1)
        cycles += 1
    wall = time.perf_counter() - t0
    return {"value": n ** 3 * cycles / wall, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"64^3 MHD turbulence (8 blocks of 32^3), {cycles} cycles, "
                      f"{wall:.1f} s, FP64 C++ oracle on {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--block", type=int, default=32)
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--path", default="auto", choices=["auto", "staged", "fused"])
    ap.add_argument("--profile-cycles", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is designed but not built", file=sys.stderr)
        sys.exit(2)

    import kratos_amd
    n, nb = args.n, args.block
    spec = kratos_amd.MeshSpec((n, n, n), (nb, nb, nb))
    phys = kratos_amd.Physics(mhd=True, gamma=5.0 / 3.0, cfl=0.3)
    t_ic = time.perf_counter()
    u, b, _ = make_problem(n)
    t_ic = time.perf_counter() - t_ic
    print(f"[bench] IC {n}^3 generated in {t_ic:.1f} s", file=sys.stderr, flush=True)

    forms = [f for f in args.forms.split(",") if f]
    res, states = {}, {}
    for form in forms:
        ctx, r = run_form(spec, phys, u, b, form, args.steps, args.warmup, args.profile_cycles,
                          args.path)
        r["rate"] = r["n_cells"] * r["cycles"] / r["wall"]
        states[form] = ctx.download()
        r["divb"] = ctx.divb_max()
        r["t"] = ctx.time()[0]
        ctx.close()
        res[form] = r
        print(f"[bench] {form}: {r['rate']:.4e} cell-updates/s, {r['wall'] / max(r['cycles'], 1) * 1e3:.3f}"
              f" ms/cycle, divB {r['divb']:.2e}", file=sys.stderr, flush=True)

    head = "mixed" if "mixed" in res else forms[0]
    r = res[head]
    fs = 4 if head == "mixed" else 8
    kms = r["prof"]["kernel_ms"] if r["prof"] else {}
    calls = r["prof"]["kernel_calls"] if r["prof"] else {}
    # dominant kernel: largest total time per cycle
    dom, dom_t = None, -1.0
    for k, ms in kms.items():
        tot = ms * calls.get(k, 0)
        if tot > dom_t:
            dom, dom_t = k, tot
    roof = None
    if dom:
        nc = r["n_cells"]
        units = nc
        if dom.startswith("flux"):
            units = nc * (nb + 1) * (nb + 2) ** 2 / nb ** 3   # faces incl. transverse ghost lines
        bytes_launch = alg_bytes(dom, fs) * units
        achieved = bytes_launch / (kms[dom] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dom, "mean_launch_ms": round(kms[dom], 4),
                "alg_bytes_per_launch": bytes_launch}
    acc = {}
    if "mixed" in states and "fp64" in states:
        (um, bm), (ud, bd) = states["mixed"], states["fp64"]
        acc = {"l1_rho": rel_l1(um[0], ud[0]), "l1_E": rel_l1(um[1], ud[1]),
               "l1_B": rel_l1(np.concatenate([x.ravel() for x in bm]),
                              np.concatenate([x.ravel() for x in bd])),
               "t": res["mixed"]["t"]}
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    line = {
        "metric": "MHD cell-updates/s (mixed vs FP64 at stated L1 error), % of HBM roofline",
        "value": r["rate"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] / max(r["cycles"], 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic (seeded turbulence IC)",
        "config": {"workload": f"3-D MHD turbulence {n}^3, {spec.n_blocks} blocks of {nb}^3, "
                               "PLM+HLLD+CT+RK2, gamma 5/3, beta 1, Mach 0.5",
                   "form": head, "cells": r["n_cells"]},
        "forms": {f: {"cell_updates_per_s": v["rate"], "ms_per_cycle": v["wall"] / max(v["cycles"], 1) * 1e3,
                      "divb_max": v["divb"],
                      "kernel_ms": v["prof"]["kernel_ms"] if v["prof"] else None}
                  for f, v in res.items()},
        "speedup_mixed_over_fp64": (res["mixed"]["rate"] / res["fp64"]["rate"]
                                    if "mixed" in res and "fp64" in res else None),
        "accuracy_mixed_vs_fp64": acc,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
