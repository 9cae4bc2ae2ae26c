"""Benchmark of the three-body FS2+FS3 energy+forces call on one MI355X.

Workload: 512 rigid waters, L = 24.84 A (0.997 g/cm^3), 8x8x8 jittered lattice
with random orientations (seeded), FS2+FS3 variant with the shipped fs2fs3
parameters.  One step = one full call (list build, FS2, FS3, reductions) with
inputs resident in HBM, replayed from a HIP graph.  The metric is FS3 triplet
evaluations per second at a relative force RMS error below 1e-5 against the
FP64 oracle: the FP64 path.  The mixed-precision path is timed beside it and
its errors are reported; it meets 1e-5 on the site forces but not on the COM
and molecular forces (the short-range energies cancel ~1e4-fold, so FP32
rounding of the coefficients or of the pair vectors alone costs 6e-4 / 6e-5 of
a triplet's energy; tests/test_mixed_limits.py).  The CPU oracle is timed on
the host cores.

Roofline: the dominant kernel is the FS3 evaluation (k_fs3_eval_f64), bound by
the vector ALU.  Its algorithmic work is 221,479 ops per region evaluation (a
triplet in the 4.5-5.5 A switch band is two); its duration comes from HIP
events recorded on the evaluation stream around that kernel.  HBM traffic per
launch is read from the committed rocprofv3 counter summary when present
(FETCH_SIZE + WRITE_SIZE, calibrated on a copy of known size).

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "3-body triplet evals/s (E+F), 512 H2O: mixed vs FP64 at force RMS err<1e-5"

# Algorithmic work per unit (fixed figures): one FS3 region evaluation of one
# triplet is 217,139 flops + 4,340 transcendentals = 221,479 ops; a triplet
# with 4.5 < R_avg < 5.5 A is two region evaluations.
W3_OPS = 221_479
PEAK_FP32_VECTOR_TFLOPS = 157.3   # MI355X datasheet, packed FP32 vector
PEAK_FP64_VECTOR_TFLOPS = 78.6    # MI355X datasheet, FP64 vector
PEAK_FP32_NONPACKED_TFLOPS = 78.6  # 256 CU x 4 SIMD x 16 lanes x 2 flops x 2.4 GHz, one FMA per lane per clock
TRAFFIC_FILE = ROOT / "profiles" / "k_fs3_eval_512w_traffic.txt"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-boxes and gradient-path measurements")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the timed region, run one device copy of this many MiB (counter calibration)")
    return ap.parse_args()


def timed_steps(tb, com, site, box, steps, warmup):
    import torch
    fs = torch.zeros_like(site)
    c2 = torch.zeros_like(com)
    c3 = torch.zeros_like(com)
    en = torch.zeros(3, dtype=torch.float64, device=com.device)
    for _ in range(warmup):
        tb.eval_device(com, site, box, fs, c2, c3, en)
    torch.cuda.synchronize()
    ms_eval, ms_fs3 = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        tb.eval_device(com, site, box, fs, c2, c3, en)
        st = tb.stats()  # waits for the call's events
        ms_eval.append(st["ms_fs3_eval"])
        ms_fs3.append(st["ms_fs3"])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3 / steps, float(np.mean(ms_eval)), float(np.mean(ms_fs3)), tb.stats()


def pack_boxes(boxes, dev):
    import torch
    com = np.concatenate([b.com for b in boxes])
    site = np.concatenate([b.site for b in boxes])
    ms = np.cumsum([0] + [b.nm for b in boxes]).astype(np.int32)
    bl = np.array([b.box for b in boxes], float)
    return (ms, bl, torch.tensor(com.T.copy(), device=dev),
            torch.tensor(site.transpose(2, 0, 1).copy(), device=dev))


def wall_ms(fn, reps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def bench_batched(p, seed, reps=10):
    """96 independent periodic boxes, N log-uniform in [64, 1024], mixed FS2+FS3:
    per-box sequential calls, one batched call, and autotuned graph replays."""
    import torch
    from ccpol23_plus_2025_mbmd_amd.geometry import uneven_boxes
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    dev = torch.device("cuda:0")
    boxes = uneven_boxes(96, seed=seed)
    ms, bl, com, site = pack_boxes(boxes, dev)
    per_box = [(torch.tensor(b.com.T.copy(), device=dev),
                torch.tensor(b.site.transpose(2, 0, 1).copy(), device=dev), b.box) for b in boxes]
    out = {"n_boxes": len(boxes), "n_molecules": int(ms[-1]),
           "box_sizes": [int(b.nm) for b in boxes]}
    res = {}
    for prec in ("mixed", "fp64"):
        tb = ThreeBody(p, prec)
        res[prec] = [t.cpu().numpy() for t in tb.eval_batch_device(ms, bl, com, site)]
        n3 = tb.stats()["n_fs3"]
        seq = wall_ms(lambda: [tb.eval_device(c, s, L) for c, s, L in per_box], max(2, reps // 4))
        batch = wall_ms(lambda: tb.eval_batch_device(ms, bl, com, site), reps)
        tune, tune_ms = tb.autotune(ms, bl, com, site, reps=2)
        g = tb.graph(ms, bl, com, site)
        graph = wall_ms(g.launch, reps)
        assert g.status() == 0
        g.close()
        out[prec] = {"n_fs3_triplets": n3, "ms_per_box_sequential": seq, "ms_batched_call": batch,
                     "ms_graph_replay": graph, "tuning": tune,
                     "evals_per_s_graph": n3 / (graph * 1e-3), "evals_per_s_sequential": n3 / (seq * 1e-3),
                     "graph_over_sequential": seq / graph}
        log(f"batched {prec}: sequential {seq:.3f} ms, batched {batch:.3f} ms, graph {graph:.3f} ms, tuning {tune}")
    # mixed against FP64 over the whole batch: site, COM and molecular forces,
    # and the largest relative error of a box's FS3 and FS2 energies
    (fm, c2m, c3m, em), (f6, c26, c36, e6) = res["mixed"], res["fp64"]
    rel = lambda a, b: float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    mol_m = fm[:, :, :17].sum(2) + c2m + c3m
    mol_6 = f6[:, :, :17].sum(2) + c26 + c36
    out["mixed_error_vs_fp64"] = {
        "site_force_rel_rms": rel(fm[:, :, :17], f6[:, :, :17]),
        "com_force_rel_rms": rel(np.concatenate([c2m, c3m]), np.concatenate([c26, c36])),
        "molecular_force_rel_rms": rel(mol_m, mol_6),
        "fs3_energy_rel_max": float(np.max(np.abs(em[:, 1] - e6[:, 1]) / np.abs(e6[:, 1]))),
        "fs2_energy_rel_max": float(np.max(np.abs(em[:, 0] - e6[:, 0]) / np.abs(e6[:, 0])))}
    log(f"batched mixed vs fp64: {out['mixed_error_vs_fp64']}")
    return out


def bench_neighbours(p, seed, sizes=(1024, 2048, 4096, 8192), reps=7):
    """The list build (conversion, neighbour search, offsets: ms_lists) at larger
    boxes with the cell grid (tuning neighbour_cells = 2: from 4 cells a side)
    and with the sweep over the whole box (0); the lists are the same.  The
    default (1) takes the grid from 6 cells a side."""
    import torch
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    dev = torch.device("cuda:0")
    out = {}
    for n in sizes:
        box = lattice_box(n, seed=seed)
        com = torch.tensor(box.com.T.copy(), device=dev).contiguous()
        site = torch.tensor(box.site.transpose(2, 0, 1).copy(), device=dev).contiguous()
        row = {"box_A": box.box}
        for name, cells in (("cells", 2), ("sweep", 0), ("default", 1)):
            tb = ThreeBody(p, "fp64")
            tb.set_tuning(neighbour_cells=cells)
            lists, total = [], []
            for r in range(reps + 2):
                tb.eval_device(com, site, box.box)
                torch.cuda.synchronize()
                st = tb.stats()
                if r >= 2:
                    lists.append(st["ms_lists"])
                    total.append(st["ms_total"])
            row[name] = {"ms_lists": float(np.median(lists)), "ms_call": float(np.median(total)),
                         "n_cells": st["n_cells"]}
        out[str(n)] = row
        log(f"neighbours {n}: cells {row['cells']}, sweep {row['sweep']}, default {row['default']}")
    return out


def bench_gradient(p, seed, reps=10):
    """4096 isolated trimers: energies, forces and dE/dc (728 per trimer), and the
    728 x 728 Gram matrix of the coefficient gradients on the matrix cores."""
    import torch
    from ccpol23_plus_2025_mbmd_amd.geometry import trimer_set
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    dev = torch.device("cuda:0")
    tri = trimer_set(4096, seed=seed)
    ms, bl, com, site = pack_boxes(tri, dev)
    out = {"n_trimers": len(tri)}
    for prec in ("mixed", "fp64"):
        tb = ThreeBody(p, prec)
        phi = tb.eval_batch_grad_device(ms, bl, com, site)[4]
        t_grad = wall_ms(lambda: tb.eval_batch_grad_device(ms, bl, com, site, phi=phi), reps)
        t_gram = wall_ms(lambda: tb.gram(phi), reps)
        out[prec] = {"n_fs3_triplets": tb.stats()["n_fs3"], "ms_energy_forces_coef_grad": t_grad,
                     "ms_gram_4096x728": t_gram,
                     "gram_tflops_equivalent_fp64": 2.0 * 4096 * 728 * 728 / 2 / (t_gram * 1e-3) / 1e12}
        log(f"gradient {prec}: E+F+dE/dc {t_grad:.3f} ms, Gram {t_gram:.3f} ms")
    return out


def bench_md(p, n, seed, steps=40, reps=5, dt_fs=0.5):
    """Throughput of the device NVE driver (velocity Verlet + NO_SQUISH with the
    three-body forces, sites from quaternions on the device) at n waters, in
    each precision: ms per step and ns/day of the three-body part alone (the
    reference's two-body, Ewald and induction terms stay on its host).  With
    the three-body term alone the liquid collapses after ~20 fs, so the timed
    work is `reps` runs of `steps` steps (20 fs) from the same 298 K state."""
    import torch
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    from ccpol23_plus_2025_mbmd_amd.md import RigidMD, maxwell_state, water_constants
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    box = lattice_box(n, seed=seed)
    pc, mass, pmi = water_constants()
    v, j = maxwell_state(n, 298.0, mass, pmi, seed=seed + 1)
    out = {"n_molecules": n, "dt_fs": dt_fs, "steps": steps * reps}
    for prec in ("mixed", "fp64"):
        md = RigidMD(ThreeBody(p, prec), box.com, box.quats, box.box, v, j)
        md.run(steps, dt_fs * 1e-3)  # sizes the lists
        sec = 0.0
        for _ in range(reps):
            md.set_state(box.com, box.quats, v, j)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            md.run(steps, dt_fs * 1e-3)  # returns after the device has finished
            sec += time.perf_counter() - t0
        md.close()
        nst = steps * reps
        out[prec] = {"ms_per_step": sec * 1e3 / nst, "ns_per_day_threebody_only": nst * dt_fs * 1e-6 / sec * 86400}
        log(f"md {prec}: {sec * 1e3 / nst:.3f} ms/step")
    return out


def force_errors(box, r64, rmx):
    """Relative RMS errors of mixed against FP64 on one box: site forces, COM
    forces (FS2 and FS3), molecular forces (sites + COM, what an integrator
    uses), torques about the COM, and the relative energy error."""
    def rel(a, b):
        return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    f64 = np.stack([r64[k][:, :17] for k in ("fxs", "fys", "fzs")], -1)
    fmx = np.stack([rmx[k][:, :17] for k in ("fxs", "fys", "fzs")], -1)
    c64 = np.stack([r64[k] for k in ("fxfs2", "fyfs2", "fzfs2", "fxfs3", "fyfs3", "fzfs3")])
    cmx = np.stack([rmx[k] for k in ("fxfs2", "fyfs2", "fzfs2", "fxfs3", "fyfs3", "fzfs3")])
    m64 = f64.sum(1) + c64[:3].T + c64[3:].T
    mmx = fmx.sum(1) + cmx[:3].T + cmx[3:].T
    sites = box.site[:, :17]
    e64 = r64["qfs2"] + r64["qfs3"]
    return {"site_force_rel_rms": rel(fmx, f64), "com_force_rel_rms": rel(cmx, c64),
            "molecular_force_rel_rms": rel(mmx, m64),
            "torque_rel_rms": rel(np.cross(sites, fmx).sum(1), np.cross(sites, f64).sum(1)),
            "energy_rel": float(abs(rmx["qfs2"] + rmx["qfs3"] - e64) / abs(e64))}


def snapshot_errors(p, tb_mix, tb_64, n, count=50):
    """Mixed vs FP64 over independent seeded snapshots of the benchmark box
    (a three-body-only NVE trajectory collapses within ~20 fs, since the
    short-range repulsion is the host's two-body term)."""
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    rows = [force_errors(box, tb_64.evaluate(box), tb_mix.evaluate(box))
            for box in (lattice_box(n, seed=10_000 + s) for s in range(count))]
    out = {"snapshots": count}
    for k in rows[0]:
        v = [r[k] for r in rows]
        out[k + "_max"] = max(v)
        out[k + "_mean"] = float(np.mean(v))
    log("snapshots: " + ", ".join(f"{k} max {out[k + '_max']:.2e}" for k in rows[0]))
    return out


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary."""
    if not TRAFFIC_FILE.exists():
        return None
    m = re.search(rf"^traffic_bytes_per_launch\[{kernel}\]\s*=\s*([0-9.eE+]+)", TRAFFIC_FILE.read_text(), re.M)
    return float(m.group(1)) if m else None


def host_cpu():
    """CPU model and the cores this process may use (affinity, cgroup quota)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        quota = None if q == "max" else float(q) / float(per)
    except (OSError, ValueError):
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "cgroup_cpu_quota": quota}


def cpu_baseline(p, box, threads):
    from oracle import oracle as O
    log(f"cpu baseline: oracle (faithful per-triplet table rebuild), {threads} threads, full {box.nm}-water call")
    t0 = time.perf_counter()
    r = O.threebody_term(p, box, nthreads=threads, rebuild=True)
    wall = time.perf_counter() - t0
    rate = r["n_fs3"] / wall
    n3, n2 = r["n_fs3"], r["n_fs2"]
    # one core: the lists in full, the evaluations on the first 1/16 of each list
    # (list order), scaled to the whole call
    w3, w2 = max(1, n3 // 16), max(1, n2 // 16)
    log(f"cpu baseline: one thread, first {w3} FS3 triplets and {w2} FS2 items")
    s = O.threebody_term(p, box, nthreads=1, rebuild=True, fs3_window=(0, w3), fs2_window=(0, w2))
    t1 = s["t_fs3_list"] + s["t_fs2_list"] + s["t_fs3_eval"] * n3 / w3 + s["t_fs2_eval"] * n2 / w2
    # the orbit table built once per thread instead of per triplet
    log(f"cpu baseline: hoisted orbit table, {threads} threads, full call")
    t0 = time.perf_counter()
    O.threebody_term(p, box, nthreads=threads, rebuild=False)
    wall_h = time.perf_counter() - t0
    # the optimised CPU algorithm: lists from neighbour lists, FS3 by the
    # factorised contraction (best of 3 calls), on all threads and on one
    log(f"cpu baseline: optimised algorithm, {threads} threads and 1 thread, full call")
    opt = []
    for nt in (threads, 1):
        best = 1e300
        for _ in range(3 if nt > 1 else 1):
            t0 = time.perf_counter()
            O.threebody_term(p, box, nthreads=nt, optimised=True)
            best = min(best, time.perf_counter() - t0)
        opt.append(best)
    return {"value": rate, "unit": "triplet evals/s", "cores": threads, "kind": "port",
            "sample": (f"one full FS2+FS3 E+F call at {box.nm} waters ({n3} FS3 triplets, {n2} "
                       f"FS2 items) of the C FP64 restatement in the reference's loop order with the "
                       f"per-triplet orbit-table rebuild, {threads} OpenMP threads, {wall:.2f} s wall"),
            "ms_per_call": wall * 1e3,
            "one_core": {"value": n3 / t1, "cores": 1, "ms_per_call_estimate": t1 * 1e3,
                         "sample": (f"the same restatement on one thread: both lists in full, the first {w3} "
                                    f"FS3 triplets and {w2} FS2 items evaluated, eval time scaled by the "
                                    f"list lengths")},
            "hoisted_table": {"value": n3 / wall_h, "cores": threads, "ms_per_call": wall_h * 1e3,
                              "sample": ("the same full call with the orbit table built once per thread "
                                         "instead of per triplet (the reference's loop order otherwise)")},
            "optimised": {"value": n3 / opt[0], "cores": threads, "ms_per_call": opt[0] * 1e3,
                          "one_core_value": n3 / opt[1], "one_core_ms_per_call": opt[1] * 1e3,
                          "sample": ("the same full call with the optimised CPU algorithm: FS3 and FS2 lists "
                                     "from neighbour lists, each FS3 region by the factorised contraction "
                                     "(site-pair functions once per site pair), FS2 as the reference; FP64, "
                                     "OpenMP over triplets; best of 3 calls")}}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the three-body path is "
              "measured on a single MI355X; across GPUs it would shard by cost-balanced chunks of the triplet "
              "and FS2 lists, with one all-reduce of the per-molecule sums.", file=sys.stderr)
        sys.exit(2)
    import torch
    from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box
    from ccpol23_plus_2025_mbmd_amd.params import load_variant
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody

    assert torch.cuda.is_available(), "needs a GPU"
    dev = torch.device("cuda:0")
    p = load_variant("fs2fs3")
    box = lattice_box(args.n, seed=args.seed)
    com = torch.tensor(box.com.T.copy(), dtype=torch.float64, device=dev).contiguous()
    site = torch.tensor(np.stack([box.xs.reshape(box.nm, box.ld), box.ys.reshape(box.nm, box.ld),
                                  box.zs.reshape(box.nm, box.ld)]), dtype=torch.float64, device=dev).contiguous()

    tb_mix = ThreeBody(p, "mixed")
    tb_64 = ThreeBody(p, "fp64")
    log(f"box: {box.nm} waters, L={box.box:.4f} A")
    ms_mix, ms_eval_mix, ms_fs3_mix, st_mix = timed_steps(tb_mix, com, site, box.box, args.steps, args.warmup)
    log(f"mixed: {ms_mix:.3f} ms/step, FS3 eval kernel {ms_eval_mix:.3f} ms, stats {st_mix}")
    ms_64, ms_eval_64, ms_fs3_64, st_64 = timed_steps(tb_64, com, site, box.box, args.steps, args.warmup)
    log(f"fp64:  {ms_64:.3f} ms/step, FS3 eval kernel {ms_eval_64:.3f} ms")

    # the same evaluation recorded into a HIP graph and replayed (fixed tensors,
    # no host synchronisation per step): the path an on-device MD loop uses
    def graph_ms(tb):
        g = tb.graph(np.array([0, box.nm], dtype=np.int32), np.array([box.box]), com, site)
        for _ in range(args.warmup):
            g.launch()
        t = wall_ms(g.launch, args.steps)
        torch.cuda.synchronize()
        assert g.status() == 0
        g.close()
        return t
    gms_mix, gms_64 = graph_ms(tb_mix), graph_ms(tb_64)
    log(f"graph replay: mixed {gms_mix:.3f} ms/step, fp64 {gms_64:.3f} ms/step")

    if args.copy_mb > 0:
        a = torch.empty(args.copy_mb << 20, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_mb} MiB")

    # mixed vs FP64 on this snapshot (host entry point): site forces, COM forces
    # (FS2 + FS3), molecular forces (sites + COM) and energy
    r64 = tb_64.evaluate(box)
    rmx = tb_mix.evaluate(box)
    snap = force_errors(box, r64, rmx)

    n3 = st_64["n_fs3"]
    nreg = st_64["n_fs3_sr"] + st_64["n_fs3_lr"] + 2 * st_64["n_fs3_switch"]
    value_64 = n3 / (gms_64 * 1e-3)
    value_mix = n3 / (gms_mix * 1e-3)
    ops = nreg * W3_OPS
    achieved_64 = ops / (ms_eval_64 * 1e-3) / 1e12
    achieved_mix = ops / (ms_eval_mix * 1e-3) / 1e12

    out = {
        "metric": METRIC, "value": value_64, "unit": "triplet evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": gms_64, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{box.nm} rigid H2O, L={box.box:.2f} A, jittered 8x8x8 lattice, random "
                               f"quaternions (seed {args.seed}), FS2+FS3 fs2fs3 parameters, FP64 (the path "
                               f"that meets the 1e-5 force bar; mixed precision reported beside it)",
                   "n_molecules": box.nm, "n_fs3_triplets": n3, "n_fs2_items": st_64["n_fs2_items"],
                   "n_fs2_pairs": st_64["n_fs2_pairs"],
                   "fs2_items_per_s": st_64["n_fs2_items"] / (gms_64 * 1e-3),
                   "region_hist_sr_switch_lr": [st_64["n_fs3_sr"], st_64["n_fs3_switch"], st_64["n_fs3_lr"]]},
        "roofline": {"bound": "valu", "kernel": "k_fs3_eval_f64", "achieved": achieved_64,
                     "peak": PEAK_FP64_VECTOR_TFLOPS, "unit": "TFLOP/s", "frac": achieved_64 / PEAK_FP64_VECTOR_TFLOPS,
                     "traffic": read_traffic("k_fs3_eval_f64"), "ops_per_launch": ops, "ms_per_launch": ms_eval_64},
        "launch": "hipGraph replay of the full call (lists, FS2, FS3, reductions); per-call entry point below",
        "eager_call": {"fp64_value": n3 / (ms_64 * 1e-3), "fp64_ms_per_step": ms_64,
                       "mixed_value": n3 / (ms_mix * 1e-3), "mixed_ms_per_step": ms_mix},
        "mixed": {"value": value_mix, "ms_per_step": gms_mix, "fs3_eval_kernel": "k_fs3_eval_f32",
                  "fs3_eval_kernel_ms": ms_eval_mix, "achieved_tflops": achieved_mix,
                  "frac_of_fp32_vector_peak": achieved_mix / PEAK_FP32_VECTOR_TFLOPS,
                  "frac_of_fp32_nonpacked_peak": achieved_mix / PEAK_FP32_NONPACKED_TFLOPS,
                  "traffic": read_traffic("k_fs3_eval_f32"),
                  "over_fp64": value_mix / value_64,
                  "error_vs_fp64_this_snapshot": snap,
                  "meets_1e-5_force_bar": {"site_forces": snap["site_force_rel_rms"] <= 1e-5,
                                           "com_forces": snap["com_force_rel_rms"] <= 1e-5,
                                           "molecular_forces": snap["molecular_force_rel_rms"] <= 1e-5}},
        "stats_fp64_ms": {k: st_64[k] for k in ("ms_lists", "ms_fs2", "ms_fs3", "ms_fs3_eval", "ms_reduce",
                                                "ms_total")},
        "stats_mixed_ms": {k: st_mix[k] for k in ("ms_lists", "ms_fs2", "ms_fs3", "ms_fs3_eval", "ms_reduce",
                                                  "ms_total")},
    }
    if not args.no_extras:
        out["mixed"]["error_vs_fp64_50_snapshots"] = snapshot_errors(p, tb_mix, tb_64, args.n)
        out["batched_96_boxes"] = bench_batched(p, args.seed)
        out["gradient_4096_trimers"] = bench_gradient(p, args.seed)
        out["md_nve_512"] = bench_md(p, args.n, args.seed)
        out["neighbour_search"] = bench_neighbours(p, args.seed)
    if not args.no_cpu_baseline:
        host = host_cpu()
        threads = args.cpu_threads or host["affinity"]
        cb = cpu_baseline(p, box, threads)
        cb["host"] = host
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value_64 / cb["value"]
        out["gpu_over_cpu_optimised"] = value_64 / cb["optimised"]["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
