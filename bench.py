"""Benchmark: MHD cell-updates/s of the Godunov stage on one MI355X.

Workload: seeded 3-D MHD turbulence (PLM + HLLD + CT + Heun RK2), 256^3
periodic cells as 512 blocks of 32^3, gamma = 5/3, beta = 1, rms Mach 0.5.
A "step" is one Heun cycle (two RK substeps) over every cell; cell-updates/s
= cells x cycles / wall seconds, the reference's Speed definition
(mesh/cycle/cycle.cpp:184-190).  The state is resident in HBM before the
timed region; the timed region is K cycles replayed from the captured
hipGraph, bracketed by device synchronisation.  Before it, after the W
warmup cycles, further untimed cycles run until the tile autotuner has
timed every candidate and the graph is instantiated (warmup_cycles_run
reports the total).

Both precision forms run: the plain FP64 form and the mixed form (FP32
reconstruction / Riemann / EMF, FP64 conserved state, CT, energy and dt).
``value`` is the mixed form's rate; the FP64 rate and the speed-up are
reported beside it.

accuracy: the same recipe at beta = 1 and beta = 0.01, both forms run to
t = 0.1 (t_lim top-up), and beta = 0.01 once more to t = 0.003; relative L1
of rho, E and B of the mixed form against the FP64 form (which equals the
CPU oracle bit for bit on every case the parity tests run), and max
|div B| (absolute, and times dx over max |B|).  At beta = 0.01 also the L1
distance that a 1e-12 relative perturbation of the initial density reaches
in the FP64 form alone: the problem's own sensitivity floor, which by
t = 0.1 has grown past the mixed form's distance and at t = 0.003 is still
at round-off, so the early run measures the mixed form's own error.

hydro: the hydro module (PLM + HLLC + RK2) on the same 256^3 recipe without
the face field, 512 blocks of 32^3, both forms (fused hydro kernel).

smr_blast: the MHD blast on a three-level statically refined mesh (128^3
root cells in 16^3 blocks, ~2,000 blocks), both forms on the staged kernels
with the cross-level corrections; cell-updates/s counts every block's cells.

ensemble: 64 independent 64^3 periodic boxes (seeds 0..63, beta drawn from
U[0.1, 10]) in one context, one time group per box (per-box dt), all boxes
advanced by the same launches to t = 0.05; aggregate cell-updates/s of the
mixed form and the per-box relative density L1 against the FP64 form.

roofline: the dominant kernel (largest share of the cycle) with its
algorithmic bytes per launch divided by its mean launch time (HIP events on
the library's stream).  traffic: HBM bytes per launch of that kernel from
the rocprofv3 FETCH_SIZE / WRITE_SIZE summary committed under profiles/
for the same kernel and workload, when present; valu_frac: the time that
kernel's VALU instruction count (SQ_INSTS_VALU of the same summary) takes to
issue at one wave64 instruction per 4 cycles per SIMD, over the launch time
measured here; stream_copy_gbs: a plain 1 GiB device copy timed on the
same GPU, the rate a streaming kernel reaches there (frac_of_stream_copy
puts the dominant kernel against it).  cpu_baseline: the CPU
oracle (FP64 C++ restatement of the reference algorithm, "port") on a
bounded sample of the same recipe: 128^3 as 512 blocks of 16^3, on the
cores the process may use (at most 16), with nproc and the CPU model.
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
# state per cell in the algorithmic byte count: the reference's conserved u
# (rho, E, m, B_cc: 8 doubles) and one face B per axis (3); 5 passes of it
# per Heun cycle = 440 B per MHD cell-update.  This build keeps no u[B_cc]
# (c2p recomputes it from the face field), so it needs only 8 of the 11.
STATE_DOUBLES = 8 + 3


def alg_bytes(kind, fs, fused):
    """Algorithmic HBM bytes per launch unit (cell, or face for the staged
    flux kernels) of one kernel kind, averaged over the two RK substeps.

    fused stage kernel ("update" in the fused path): the state it must read
    and write once -- substep 0 reads slot 0 and writes slot 1 (2 passes of
    11 doubles), substep 1 also reads the target slot (3 passes), i.e. the
    440 B per cell-update of the cycle split over its two launches.
    Staged kernels: their own inputs and outputs read / written once.
    """
    if kind == "update" and fused:
        return (2 * STATE_DOUBLES + 3 * STATE_DOUBLES) / 2 * 8
    if kind == "c2p":
        return (5 + 3) * 8 + 8 * fs
    if kind.startswith("flux"):
        return 8 * fs + 8 * fs
    if kind == "emf":
        return 9 * fs + 6 * fs + 3 * fs
    if kind == "update":
        u = 7.5 * 8 + 5 * 8
        return u + 3 * 5 * fs + 3 * fs + (3 + 3) * 8
    return 0.0


def make_problem(n, seed=0, beta=1.0, beta_range=None):
    from kratos_amd import problems
    return problems.turbulence((n, n, n), seed=seed, beta=beta, beta_range=beta_range)


def rel_l1(a, b):
    s = np.abs(b).sum()
    return float(np.abs(a - b).sum() / s) if s > 0 else float(np.abs(a - b).sum())


def divb_rel(divb, bfs, dx):
    """max |div B| made dimensionless: times the cell width, over max |B|
    of the face field (round-off of the face sums sits near 1e-16 times the
    number of cycles' worth of accumulated updates)."""
    bmax = max(float(np.abs(x).max()) for x in bfs)
    return divb * dx / bmax if bmax > 0 else divb


def sync_all(ctx):
    ctx.lib.kr_sync(ctx.h)


def settle(ctx, max_cycles=64):
    """Untimed cycles until the tile autotuner has timed every candidate
    and the cycle's hipGraph is instantiated, so that a timed kr_advance
    afterwards is graph replays only.  Returns the cycles it ran."""
    extra = 0
    while extra < max_cycles:
        st = ctx.stats()
        if st["tune_pending"] == 0 and st["graph_ready"]:
            break
        if ctx.advance(1) == 0:
            break                               # every time group at t_lim
        extra += 1
    st = ctx.stats()
    if st["tune_pending"] or not st["graph_ready"]:
        raise RuntimeError(f"autotuner / graph not settled after {extra} cycles: {st}")
    return extra


def run_form(spec, phys, u, b, precision, steps, warmup, profile_cycles, path="auto",
             tile=(0, 0, 0)):
    import kratos_amd
    ctx = kratos_amd.Context(spec, phys, precision, path=path, tile=tile)
    ctx.upload(u, b)
    if warmup:
        ctx.advance(warmup)
    extra = settle(ctx)
    sync_all(ctx)
    t0 = time.perf_counter()
    done = ctx.advance(steps)
    sync_all(ctx)
    wall = time.perf_counter() - t0
    st = ctx.stats()
    prof = ctx.profile(profile_cycles) if profile_cycles else None
    return ctx, {"wall": wall, "cycles": done, "gpu_ms": st["gpu_ms"], "prof": prof,
                 "warmup_cycles_run": warmup + extra,
                 "n_cells": st["n_cells"], "sweep_tile": st["sweep_tile"],
                 "sweep_kernel": st["sweep_kernel"],
                 "tune_ms": [round(t, 4) for t in st["tune_ms"]]}


def accuracy_leg(n, nb, beta, t_lim, path, sensitivity=False):
    """Both forms to t_lim from one IC; mixed-vs-FP64 L1 and div B.  With
    `sensitivity`, also the FP64 form from the same IC with the density
    multiplied by 1 + 1e-12 * N(0, 1) (numpy seed 1): its L1 distance to
    the unperturbed FP64 run is the floor that round-off-level differences
    reach on this problem by t_lim, whatever the precision."""
    import kratos_amd
    spec = kratos_amd.MeshSpec((n, n, n), (nb, nb, nb))
    phys = kratos_amd.Physics(mhd=True, gamma=5.0 / 3.0, cfl=0.3, t_lim=t_lim, top_up=True)
    u, b, _ = make_problem(n, beta=beta)
    runs = [("fp64", "fp64", u), ("mixed", "mixed", u)]
    if sensitivity:
        up = u.copy()
        up[0] *= 1 + 1e-12 * np.random.default_rng(1).standard_normal(up[0].shape)
        runs.append(("fp64_perturbed", "fp64", up))
    out = {}
    for name, form, u0 in runs:
        ctx = kratos_amd.Context(spec, phys, form, path=path)
        ctx.upload(u0, b)
        done = ctx.advance(1000000)
        out[name] = (ctx.download(), ctx.divb_max(), ctx.time()[0], done)
        ctx.close()
    (um, bm), divb_m, tm, cm = out["mixed"]
    (ud, bd), divb_d, td, cd = out["fp64"]
    cat = lambda bs: np.concatenate([x.ravel() for x in bs])
    res = {"beta": beta, "t": tm, "cycles_mixed": cm, "cycles_fp64": cd,
           "l1_rho": rel_l1(um[0], ud[0]), "l1_E": rel_l1(um[1], ud[1]),
           "l1_B": rel_l1(cat(bm), cat(bd)),
           "divb_max_mixed": divb_m, "divb_max_fp64": divb_d,
           "divb_rel_mixed": divb_rel(divb_m, bm, 1.0 / n),
           "divb_rel_fp64": divb_rel(divb_d, bd, 1.0 / n)}
    if sensitivity:
        (up_, _), _, _, _ = out["fp64_perturbed"]
        res["l1_rho_fp64_perturbed_1e-12"] = rel_l1(up_[0], ud[0])
    return res


def ensemble_leg(n_box, n, t_lim, path, nb=None):
    """n_box independent n^3 boxes, one time group each, both forms; each
    box is one n^3 block, or (nb) blocks of nb^3."""
    import kratos_amd
    nb = nb or n
    spec = kratos_amd.MeshSpec((n, n, n), (nb, nb, nb))
    phys = kratos_amd.Physics(mhd=True, gamma=5.0 / 3.0, cfl=0.3, t_lim=t_lim, top_up=True)
    ics = [make_problem(n, seed=s, beta_range=(0.1, 10.0))[:2] for s in range(n_box)]
    res = {}
    for form in ("fp64", "mixed"):
        ctx = kratos_amd.Context(spec, phys, form, path=path, members=n_box)
        for m, (u, b) in enumerate(ics):
            ctx.upload(u, b, member=m)
        settle(ctx)                         # untimed: autotuner cycles and graph capture
        sync_all(ctx)
        c0 = [ctx.time(member=m)[2] for m in range(n_box)]
        t0 = time.perf_counter()
        ctx.advance(1000000)
        sync_all(ctx)
        wall = time.perf_counter() - t0
        c1 = [ctx.time(member=m)[2] for m in range(n_box)]
        upd = sum((c1[m] - c0[m]) * n ** 3 for m in range(n_box))
        states = [ctx.download(member=m) for m in range(n_box)]
        res[form] = {"rate": upd / wall, "wall": wall, "cycles": c1,
                     "t": [ctx.time(member=m)[0] for m in range(n_box)],
                     "divb": ctx.divb_max(), "states": states}
        ctx.close()
    l1 = [rel_l1(res["mixed"]["states"][m][0][0], res["fp64"]["states"][m][0][0])
          for m in range(n_box)]
    return {"boxes": n_box, "box_cells": n ** 3, "box_blocks": spec.n_blocks, "t": t_lim,
            "cell_updates_per_s_mixed": res["mixed"]["rate"],
            "cell_updates_per_s_fp64": res["fp64"]["rate"],
            "wall_s_mixed": res["mixed"]["wall"],
            "cycles_min_max": [min(res["mixed"]["cycles"]), max(res["mixed"]["cycles"])],
            "all_boxes_at_t": all(abs(t - t_lim) == 0 for t in res["mixed"]["t"]),
            "l1_rho_per_box_max": max(l1), "l1_rho_per_box_median": float(np.median(l1)),
            "divb_max_mixed": res["mixed"]["divb"]}


def hydro_leg(n, nb, steps, warmup, path):
    """The hydro module on the same recipe (the turbulence state without
    the face field, gamma 5/3; the fused hydro kernel under path auto):
    cell-updates/s of both forms, per-cycle kernel times, and the mixed
    form's density L1 against FP64 after warmup + timed cycles.  cycle_hbm_frac prices a
    cycle at 200 B per cell (stage 1 reads U and writes U', stage 2 reads
    U', U and writes U: five passes of 5 conserved doubles), the least
    any two-stage scheme moves."""
    import kratos_amd
    spec = kratos_amd.MeshSpec((n, n, n), (nb, nb, nb))
    phys = kratos_amd.Physics(mhd=False, gamma=5.0 / 3.0, cfl=0.3)
    u, _, _ = make_problem(n)
    out = {"workload": f"hydro turbulence {n}^3, {spec.n_blocks} blocks of {nb}^3"}
    states = {}
    for form in ("fp64", "mixed"):
        ctx = kratos_amd.Context(spec, phys, form, path=path)
        ctx.upload(u, None)
        ctx.advance(warmup)
        settle(ctx)
        sync_all(ctx)
        t0 = time.perf_counter()
        done = ctx.advance(steps)
        sync_all(ctx)
        wall = time.perf_counter() - t0
        rate = n ** 3 * done / wall
        out[f"cell_updates_per_s_{form}"] = rate
        out[f"ms_per_cycle_{form}"] = wall / max(done, 1) * 1e3
        out[f"cycle_hbm_frac_{form}"] = round(200 * rate / 1e9 / HBM_PEAK_GBS, 4)
        states[form] = ctx.download()[0]
        prof = ctx.profile(2)
        out[f"kernel_ms_per_cycle_{form}"] = {
            k: round(v * prof["kernel_calls"][k] / 2, 4) for k, v in prof["kernel_ms"].items()}
        ctx.close()
    out["l1_rho_mixed_vs_fp64"] = rel_l1(states["mixed"][0], states["fp64"][0])
    out["speedup_mixed_over_fp64"] = out["cell_updates_per_s_mixed"] / out["cell_updates_per_s_fp64"]
    return out


def smr_leg(n_root, nb, steps, warmup, cycles_acc):
    """MHD blast on a statically refined mesh: [-0.5, 0.5]^3 periodic,
    n_root^3 root cells in nb^3 blocks, level 1 on |x_a| <= 0.25 and level 2
    on |x_a| <= 0.125 (tree-balanced), all blocks one size; both forms on the
    staged kernels with the cross-level halo, flux and EMF corrections.
    Reports each form's cell-updates/s over `steps` replayed cycles, and the
    mixed form's density L1 against FP64 after `cycles_acc` cycles; then the
    same mesh given with blocks of two sizes (groups of eight leaves as
    blocks of 2 nb), mixed form: its rate and whether its state after
    `cycles_acc` cycles equals the one-size mesh's bit for bit."""
    import kratos_amd
    from kratos_amd import problems
    spec = kratos_amd.RefinedMeshSpec((n_root,) * 3, (nb,) * 3,
                                      refine=[(1, (-0.25,) * 3, (0.25,) * 3),
                                              (2, (-0.125,) * 3, (0.125,) * 3)],
                                      x_min=(-0.5,) * 3, x_max=(0.5,) * 3)
    phys = kratos_amd.Physics(mhd=True, gamma=5.0 / 3.0, cfl=0.3)
    # the same mesh with blocks of two sizes: every group of eight same-level
    # leaves at an even index as one block of 2 nb (its state is kept with
    # one value per shared face, as such a block holds it)
    msp = kratos_amd.MergedMeshSpec(spec)
    blocks = msp.split_blocks(msp.merge_blocks(problems.blast_blocks(spec)))
    out = {"workload": f"MHD blast, {n_root}^3 root cells, {spec.n_blocks} blocks of {nb}^3 on "
                       f"{spec.n_levels} levels", "blocks": spec.n_blocks,
           "blocks_per_level": [sum(1 for lv in spec.leaves if lv[0] == L)
                                for L in range(spec.n_levels)],
           "cells": spec.n_cells}
    states = {}
    for form in ("fp64", "mixed"):
        ctx = kratos_amd.Context(spec, phys, form)
        ctx.upload_blocks(blocks)
        ctx.advance(cycles_acc)
        sync_all(ctx)
        states[form] = ctx.download_blocks()
        out[f"divb_max_{form}"] = ctx.divb_max()
        ctx.advance(warmup)
        settle(ctx)
        sync_all(ctx)
        t0 = time.perf_counter()
        done = ctx.advance(steps)
        sync_all(ctx)
        wall = time.perf_counter() - t0
        out[f"cell_updates_per_s_{form}"] = spec.n_cells * done / wall
        out[f"ms_per_cycle_{form}"] = wall / max(done, 1) * 1e3
        prof = ctx.profile(2)
        # per cycle: a kind groups several launches here (the staged kernel,
        # the SMR corrections of that step, the fused kernel of the pure
        # blocks), so the per-launch mean of the stats would mislead
        out[f"kernel_ms_per_cycle_{form}"] = {
            k: round(v * prof["kernel_calls"][k] / 2, 4) for k, v in prof["kernel_ms"].items()}
        ctx.close()
    # the two-size mesh, mixed form: same cycles, same state bit for bit
    ctx = kratos_amd.Context(msp, phys, "mixed")
    ctx.upload_blocks(msp.merge_blocks(blocks))
    ctx.advance(cycles_acc)
    sync_all(ctx)
    same = all(np.array_equal(np.asarray(x[0]), np.asarray(y[0]))
               for x, y in zip(msp.split_blocks(ctx.download_blocks()), states["mixed"]))
    ctx.advance(warmup)
    settle(ctx)
    sync_all(ctx)
    t0 = time.perf_counter()
    done = ctx.advance(steps)
    sync_all(ctx)
    wall = time.perf_counter() - t0
    ctx.close()
    n2 = sum(msp.merged(b) for b in range(msp.n_blocks))
    out["two_block_sizes"] = {
        "blocks": msp.n_blocks, f"blocks_of_{2 * nb}^3": n2, f"blocks_of_{nb}^3": msp.n_blocks - n2,
        "cell_updates_per_s_mixed": spec.n_cells * done / wall,
        "state_after_cycles_acc_equals_one_size_mesh": bool(same)}
    num = den = 0.0
    for b in range(spec.n_blocks):
        _, dx = spec.geometry(b)
        v = dx[0] * dx[1] * dx[2]
        m = np.asarray(states["mixed"][b][0]).reshape(-1, 8)[:, 0]
        d = np.asarray(states["fp64"][b][0]).reshape(-1, 8)[:, 0]
        num += np.abs(m - d).sum() * v
        den += np.abs(d).sum() * v
    out["cycles_acc"] = cycles_acc
    out["l1_rho_mixed_vs_fp64"] = num / den
    out["speedup_mixed_over_fp64"] = out["cell_updates_per_s_mixed"] / out["cell_updates_per_s_fp64"]
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seconds_target=15.0, threads=None):
    """Oracle on a bounded sample: 128^3 cells of the same recipe as 512
    blocks of 16^3 (the oracle spreads blocks over its threads, so this
    keeps every thread busy), on every core the process may use, capped at
    16 (the host share of one GPU on the measurement machine)."""
    import kratos_amd
    import oracle
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = threads or min(16, avail or 1)
    n = 128
    spec = kratos_amd.MeshSpec((n, n, n), (16, 16, 16))
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
    assert spec.n_blocks >= 4 * threads
    return {"value": n ** 3 * cycles / wall, "unit": "cell-updates/s", "cores": threads,
            "kind": "port", "nproc": os.cpu_count(), "cpus_available": avail,
            "cpu_model": cpu_model(),
            "sample": f"128^3 MHD turbulence (512 blocks of 16^3), {cycles} cycles, "
                      f"{wall:.1f} s, FP64 C++ oracle on {threads} threads"}


def copy_probe():
    import torch
    n = 1 << 28
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(5):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel": "torch.mul(a, 1.0, out=b)", "bytes_read_per_launch": 4 * n,
                      "bytes_written_per_launch": 4 * n, "launches": 5}))


def stream_copy_gbs(n_bytes=1 << 30, launches=20):
    """Measured HBM rate of a plain device copy on this GPU (torch.mul(a,
    1.0, out=b) over n_bytes in and n_bytes out, timed with events on
    torch's stream): what a streaming kernel reaches here, beside the spec
    sheet's 8 TB/s."""
    import torch
    a = torch.ones(n_bytes // 4, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(launches):
        torch.mul(a, 1.0, out=b)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / launches
    del a, b
    return 2 * n_bytes / (ms * 1e-3) / 1e9


def committed_profile(kernel, workload):
    """The committed rocprofv3 PMC summary row of a kernel and workload:
    HBM bytes and VALU wave-instructions per launch."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        rows = json.load(f)
    for r in rows:
        if r.get("kernel") == kernel and r.get("workload") == workload:
            return r
    return {}


# VALU issue rate of one MI355X: 256 CUs x 4 SIMDs, one wave64 instruction
# per 4 cycles per SIMD (16 lanes; packed FP32 is not used), 2.4 GHz
VALU_WAVE_INSTR_PER_S = 256 * 4 / 4 * 2.4e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=8,
                    help="untimed cycles; the first ones also run the tile autotuner")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--block", type=int, default=32)
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--path", default="auto", choices=["auto", "staged", "fused"])
    ap.add_argument("--profile-cycles", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--t-acc", type=float, default=0.1)
    ap.add_argument("--t-acc-low", type=float, default=0.003,
                    help="second, early time of the beta = 0.01 accuracy run, before the "
                         "flow amplifies round-off (at 256^3 the perturbed-FP64 floor is "
                         "3e-13 at t = 0.003 and 8e-6 at t = 0.01; 0: skip)")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ens-boxes", type=int, default=64)
    ap.add_argument("--ens-n", type=int, default=64)
    ap.add_argument("--t-ens", type=float, default=0.05)
    ap.add_argument("--no-smr", action="store_true")
    ap.add_argument("--no-hydro", action="store_true")
    ap.add_argument("--hydro-only", action="store_true",
                    help="run only the hydro leg and print its JSON object")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the ensemble leg (one JSON line), e.g. under a profiler")
    ap.add_argument("--smr-only", action="store_true",
                    help="run only the refined-mesh blast leg and print its JSON")
    ap.add_argument("--smr-root", type=int, default=128)
    ap.add_argument("--smr-block", type=int, default=16)
    ap.add_argument("--tile", default="",
                    help="force the fused-kernel variant as TX,TY,WAVES (profiling runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB fp32 copy kernel (calibrates the HBM "
                         "byte counters of rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is designed but not built", file=sys.stderr)
        sys.exit(2)

    if args.hydro_only:
        hyd = hydro_leg(args.n, args.block, args.steps, args.warmup, args.path)
        print(json.dumps({k: (float(v) if isinstance(v, np.floating) else v) for k, v in hyd.items()}))
        return

    if args.ensemble_only:
        ens = ensemble_leg(args.ens_boxes, args.ens_n, args.t_ens, args.path)
        print(json.dumps({k: (float(v) if isinstance(v, np.floating) else v) for k, v in ens.items()}))
        return

    if args.smr_only:
        smr = smr_leg(args.smr_root, args.smr_block, max(args.steps // 2, 4), args.warmup, 20)
        print(json.dumps({k: (float(v) if isinstance(v, np.floating) else v) for k, v in smr.items()}))
        return

    import kratos_amd
    n, nb = args.n, args.block
    spec = kratos_amd.MeshSpec((n, n, n), (nb, nb, nb))
    phys = kratos_amd.Physics(mhd=True, gamma=5.0 / 3.0, cfl=0.3)
    t_ic = time.perf_counter()
    u, b, _ = make_problem(n)
    t_ic = time.perf_counter() - t_ic
    print(f"[bench] IC {n}^3 generated in {t_ic:.1f} s", file=sys.stderr, flush=True)

    forms = [f for f in args.forms.split(",") if f]
    res = {}
    fused = None
    for form in forms:
        tile = tuple(int(x) for x in args.tile.split(",")) if args.tile else (0, 0, 0)
        ctx, r = run_form(spec, phys, u, b, form, args.steps, args.warmup, args.profile_cycles,
                          args.path, tile if form == "mixed" or tile[2] in (0, 1) else (0, 0, 0))
        r["rate"] = r["n_cells"] * r["cycles"] / r["wall"]
        r["divb"] = ctx.divb_max()
        r["divb_rel"] = divb_rel(r["divb"], ctx.download()[1], 1.0 / n)
        r["t"] = ctx.time()[0]
        if fused is None:
            # the fused stage kernel reports under "update" with no flux launches
            k = r["prof"]["kernel_calls"] if r["prof"] else {}
            fused = k.get("update", 0) > 0 and k.get("flux_x", 0) == 0
        ctx.close()
        res[form] = r
        print(f"[bench] {form}: {r['rate']:.4e} cell-updates/s, "
              f"{r['wall'] / max(r['cycles'], 1) * 1e3:.3f} ms/cycle, divB {r['divb']:.2e}",
              file=sys.stderr, flush=True)
    del u, b

    head = "mixed" if "mixed" in res else forms[0]
    r = res[head]
    fs = 4 if head == "mixed" else 8
    kms = r["prof"]["kernel_ms"] if r["prof"] else {}
    calls = r["prof"]["kernel_calls"] if r["prof"] else {}
    dom, dom_t = None, -1.0
    for k, ms in kms.items():
        tot = ms * calls.get(k, 0)
        if tot > dom_t:
            dom, dom_t = k, tot
    workload = f"mhd_turbulence_{n}^3_{nb}^3_{head}"
    roof = None
    if dom:
        nc = r["n_cells"]
        units = nc
        if dom.startswith("flux"):
            units = nc * (nb + 1) * (nb + 2) ** 2 / nb ** 3   # faces incl. transverse ghost lines
        bytes_launch = alg_bytes(dom, fs, fused) * units
        achieved = bytes_launch / (kms[dom] * 1e-3) / 1e9
        tl = r["sweep_tile"]
        if fused and dom == "update" and r.get("sweep_kernel"):
            kname = f"k_sweep_mx<{tl[0]},{tl[1]},{tl[2]}>"
        elif fused and dom == "update":
            kname = f"k_sweep<{'float' if head == 'mixed' else 'double'},{tl[0]},{tl[1]},{tl[2]},{tl[3]}>"
        else:
            kname = f"k_{dom}"
        cp = committed_profile(kname, workload)
        traffic = cp.get("bytes_per_launch")
        valu = cp.get("valu_wave_instr_per_launch")
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "kernel": kname, "mean_launch_ms": round(kms[dom], 4),
                "alg_bytes_per_launch": bytes_launch,
                "cycle_frac": round(STATE_DOUBLES * 5 * 8 * r["rate"] / 1e9 / HBM_PEAK_GBS, 4),
                # the VALU issue time of the launch's instruction count (from the
                # committed counters) over the launch time measured here
                "valu_frac": (round(valu / VALU_WAVE_INSTR_PER_S / (kms[dom] * 1e-3), 4)
                              if valu else None)}
        copy = stream_copy_gbs()
        roof["stream_copy_gbs"] = round(copy, 1)
        roof["frac_of_stream_copy"] = round(achieved / copy, 4)

    acc = []
    if not args.no_accuracy:
        runs = [(1.0, args.t_acc), (0.01, args.t_acc)]
        if args.t_acc_low > 0:
            # low beta at an early time, where the problem's own sensitivity
            # floor (the perturbed FP64 run) is still far below the 1e-5 bar
            runs.append((0.01, args.t_acc_low))
        for beta, t_end in runs:
            a = accuracy_leg(n, nb, beta, t_end, args.path, sensitivity=beta < 0.1)
            acc.append(a)
            print(f"[bench] accuracy beta={beta} t={t_end}: {a}", file=sys.stderr, flush=True)
    ens = None
    if not args.no_ensemble:
        ens = ensemble_leg(args.ens_boxes, args.ens_n, args.t_ens, args.path)
        print(f"[bench] ensemble: {ens}", file=sys.stderr, flush=True)
        if args.ens_n >= 64:
            # the same boxes as 8 blocks of (n/2)^3: a block's z march is half
            # as long, which shortens the cycles at the end, when only the
            # boxes with the smallest dt are still running
            ens["as_8_blocks_per_box"] = ensemble_leg(args.ens_boxes, args.ens_n, args.t_ens,
                                                      args.path, args.ens_n // 2)
            print(f"[bench] ensemble, 8 blocks per box: {ens['as_8_blocks_per_box']}",
                  file=sys.stderr, flush=True)
    hyd = None
    if not args.no_hydro:
        hyd = hydro_leg(n, nb, args.steps, args.warmup, args.path)
        print(f"[bench] hydro: {hyd}", file=sys.stderr, flush=True)
    smr = None
    if not args.no_smr:
        smr = smr_leg(args.smr_root, args.smr_block, max(args.steps // 2, 4), 2, 20)
        print(f"[bench] smr: {smr}", file=sys.stderr, flush=True)
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
                   "form": head, "cells": r["n_cells"], "kernel_path": "fused" if fused else "staged"},
        "forms": {f: {"cell_updates_per_s": v["rate"],
                      "ms_per_cycle": v["wall"] / max(v["cycles"], 1) * 1e3,
                      "divb_max": v["divb"], "divb_rel": v["divb_rel"],
                      "sweep_tile": v["sweep_tile"], "autotune_ms": v["tune_ms"],
                      "warmup_cycles_run": v["warmup_cycles_run"],
                      "kernel_ms": v["prof"]["kernel_ms"] if v["prof"] else None}
                  for f, v in res.items()},
        "speedup_mixed_over_fp64": (res["mixed"]["rate"] / res["fp64"]["rate"]
                                    if "mixed" in res and "fp64" in res else None),
        "accuracy_mixed_vs_fp64": acc,
        "ensemble": ens,
        "smr_blast": smr,
        "hydro": hyd,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
