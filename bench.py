"""Benchmark of the lookahead decomposition on one MI355X; prints ONE JSON line.

Workload: 64 watertight SDF meshes (smooth union of 3-8 random capsules/boxes
minus 0-3 spheres, marching tetrahedra on a 48^3 grid, 2k-50k triangles,
extent 2, seed 4), decomposed with lookahead_decompose at tau = 0.05,
width/width2 30/5, depth 2, quick depth 0, 16 concave edges, 16 parts cut in
parallel, at most 100 iterations, hull merge on. A step is one batched search
over all 64 meshes (the meshes are independent; every launch spans all of
them). Metric: meshes per second. After the timed plain steps, one step of
the learned-proposal form (the shipped network keeps the top 16 level-0
candidates of each cutting part) is timed and reported under "learned".

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--meshes N]

`--workload score` measures the batched exact cut scorer instead (candidate
cuts scored per second): 256 SDF meshes (marching tetrahedra on a 64^3 grid,
500-60k triangles, seed 3), 30 candidate planes each (up to 16 through
sampled concave edges, the rest axis cuts), tau 0.05; a step scores all 7680
cuts (cut, hulls of both halves, Hausdorff where the volume cost is under
tau). `--autotune` tunes the hull launch shapes on the workload first
(untimed).
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
PARAMS = dict(max_iters=100, width=30, width2=5, threshold=0.05, depth=2, max_n_cutting=16,
              n_concave_edges=16, merge_hulls=True)


def kernel_name(cls, wide_min, half_cap):
    """Kernel behind a hull class reported by the library."""
    if cls == "kdop_prefilter":
        return "la_kdop_kernel"
    if cls.endswith("_retry"):
        base = cls[:-6]
        n = int(base[6:])
        return "la_hull_block<4>" if base.startswith("hull_b") else f"la_hull_lds4<{n // 256}>"
    if cls.startswith("hull_b"):
        n = int(cls[6:])
        ppl = n // 512
        if n == 2048 and half_cap:
            return "la_hull_ldsv<4,8>"
        return f"la_hull_block<{ppl}>"
    n = int(cls[6:])
    ppl = n // 64
    if wide_min > 0 and ppl >= 4 and n > wide_min:
        if ppl >= 8 and half_cap:
            return f"la_hull_ldsv<{ppl // 4},4>"
        return f"la_hull_lds4<{ppl // 4}>"
    return f"la_hull_wave<{ppl}>"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(n):
    from cuacd_amd import sdf
    return sdf.sdf_meshes(n, seed=4, res=48, min_tris=2000, max_tris=50000)


def cpu_baseline(ctx, meshes, gpu_iters, budget_s=20.0):
    """The serial lookahead oracle (Python control over the C oracle's cut,
    hull, scoring and Hausdorff pieces, one core) on a bounded sample: the
    meshes of the set from the smallest up, each run until it converges,
    until about budget_s of oracle time is spent; a mesh the budget cuts
    short counts by the fraction of its search iterations done. The input
    hulls (an input of the search) are computed outside the timed calls."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lookahead_oracle as LO
    sizes = [len(t) for _, t in meshes]
    kw = dict(PARAMS)
    kw.pop("merge_hulls")
    spent, worth, used = 0.0, 0.0, []
    for m in np.argsort(sizes, kind="stable"):
        m = int(m)
        if spent >= budget_s:
            break
        v, t = meshes[m]
        hv, ht = ctx._input_hull(v)
        t0 = time.perf_counter()
        _, done = LO.lookahead(v, t, hv, ht, deadline=t0 + (budget_s - spent), **kw)
        spent += time.perf_counter() - t0
        frac = min(max(done, 1) / max(int(gpu_iters[m]), 1), 1.0)
        worth += frac
        used.append(m)
        if frac < 1.0:
            break
    lo, hi = min(sizes[m] for m in used), max(sizes[m] for m in used)
    which = (f"the smallest mesh of the set ({lo} tris)" if len(used) == 1 else
             f"the {len(used)} smallest meshes of the set ({lo}-{hi} tris)")
    return {"value": worth / spent, "unit": "meshes/s", "cores": 1, "kind": "port",
            "sample": f"{which}: {worth:.2f} meshes' worth of search iterations in {spent:.1f} s "
                      f"without the merge pass"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--meshes", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["lookahead", "score"], default="lookahead")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only run the known-size device copy (HBM counter calibration) and exit")
    ap.add_argument("--top-k", type=int, default=16,
                    help="learned-proposal step after the plain steps (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: one GPU only (asked for {args.gpus}); the meshes of the batch are independent and "
              "shard across GPUs without communication (DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import cuacd_amd

    if args.copy_probe:
        ctx = cuacd_amd.Context(device=0)
        nb = args.copy_probe << 20
        ms = cuacd_amd._gpu.copy_probe(nb, 3)
        print(json.dumps({"copy_probe_bytes": nb, "launches": 3, "ms_per_copy": ms,
                          "gbs": 2 * nb / (ms * 1e-3) / 1e9}), flush=True)
        ctx.close()
        return
    if args.workload == "score":
        return main_score(args)
    t0 = time.perf_counter()
    meshes = workload(args.meshes)
    tris = np.array([len(t) for _, t in meshes])
    log(f"workload: {len(meshes)} meshes, {tris.min()}-{tris.max()} tris (mean {tris.mean():.0f}), "
        f"generated in {time.perf_counter() - t0:.1f} s")

    ctx = cuacd_amd.Context(device=0)
    name, n_cu, hbm = ctx.device_info()
    log(f"device {name}, {n_cu} CUs, {hbm / 2**30:.0f} GiB")
    # first touch of every kernel (not a step): one small mesh
    ctx.batch_lookahead_decompose(meshes[:1], **dict(PARAMS, max_iters=2))

    for _ in range(args.warmup):
        ctx.batch_lookahead_decompose(meshes, **PARAMS)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        parts, info = ctx.batch_lookahead_decompose(meshes, return_info=True, **PARAMS)
        times.append(time.perf_counter() - t0)
    stats = ctx.last_lookahead_stats
    ms_per_step = 1e3 * float(np.mean(times))
    n_parts = np.array([len(p) for p in parts])
    iters = np.array([i["iters"] for i in info])
    max_conc = 0.0
    for i in info:
        rv = 0.3 * np.cbrt(3 / (4 * np.pi) * np.maximum(i["hull_vol"] - i["mesh_vol"], 0))
        c = np.maximum(rv, i["hausdorff"])
        max_conc = max(max_conc, float(c.max()) if len(c) else 0.0)
    log(f"step {ms_per_step / 1e3:.2f} s; parts {n_parts.sum()} (per mesh {n_parts.min()}-{n_parts.max()}), "
        f"iterations {iters.min()}-{iters.max()}, meshes at max_iters {(iters >= PARAMS['max_iters']).sum()}")
    gpu_ms = sum(s["ms"] for s in stats.values())
    for k, s in stats.items():
        log(f"  {k:14s} {s['ms']:9.1f} ms  {s['launches']:5d} launches  {s['alg_bytes'] / 1e9:8.3f} GB "
            f"({s['alg_bytes'] / max(s['ms'], 1e-9) / 1e6:7.1f} GB/s)")
    log(f"  GPU time in phases {gpu_ms / 1e3:.2f} s of {ms_per_step / 1e3:.2f} s")
    sec = ctx.last_lookahead_sections
    log("  host sections: " + ", ".join(f"{k} {v / 1e3:.2f} s" for k, v in sec.items()) +
        f" (sum {sum(sec.values()) / 1e3:.2f} s)")
    # single-kernel timings: HIP events around each launch of that kernel alone
    kern = {}
    for ph, name in (("leaf_scoring", "la_score_kernel"), ("level0_cuts", "la_cut_kernel"),
                     ("hausdorff", "hd_task_kernel")):
        kern[name] = stats[ph]
    wide_min = ctx.get_tuning("hull_wide_min")
    half_cap = ctx.get_tuning("hull_half_cap") != 0
    for cls, k in ctx.last_lookahead_kernels.items():
        kern[kernel_name(cls, wide_min, half_cap)] = k
    for name, k in sorted(kern.items(), key=lambda x: -x[1]["ms"])[:6]:
        log(f"  kernel {name:22s} {k['ms']:9.1f} ms  {k['launches']:5d} launches  "
            f"{k['alg_bytes'] / max(k['ms'], 1e-9) / 1e6:7.1f} GB/s algorithmic")
    dom = max(kern, key=lambda k: kern[k]["ms"])
    d = kern[dom]
    achieved = d["alg_bytes"] / (d["ms"] * 1e-3) / 1e9 if d["ms"] > 0 else 0.0

    learned = None
    if args.top_k > 0:
        learned = learned_run(ctx, meshes, ms_per_step, int(n_parts.sum()), args.top_k)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, meshes, iters)
        log(f"cpu baseline {cpu['value']:.4f} meshes/s ({cpu['sample']})")

    value = len(meshes) / (ms_per_step * 1e-3)
    line = {
        "metric": "lookahead meshes/sec at tau=0.05 (plain)",
        "value": value,
        "unit": "meshes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 geometry / fp64 volumes / int32 exact hull grid",
        "data": "synthetic",
        "config": {"workload": "plain lookahead_decompose: 64 SDF meshes (marching tetrahedra 48^3, 2k-50k tris, "
                               "extent 2, seed 4), tau 0.05, width/width2 30/5, depth 2, quick 0, max_iters 100, "
                               "merge_hulls on, n_concave_edges 16, max_n_cutting 16",
                   "meshes": len(meshes), "tris_total": int(tris.sum()), "parts_total": int(n_parts.sum()),
                   "max_part_concavity": max_conc, "iterations_max": int(iters.max())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": dom, "kernel_ms_total": d["ms"], "launches": d["launches"],
                     "avg_launch_ms": d["ms"] / max(d["launches"], 1),
                     "alg_bytes_per_launch": d["alg_bytes"] / max(d["launches"], 1)},
        "cpu_baseline": cpu,
        "learned": learned,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def max_concavity(info):
    m = 0.0
    for i in info:
        rv = 0.3 * np.cbrt(3 / (4 * np.pi) * np.maximum(i["hull_vol"] - i["mesh_vol"], 0))
        c = np.maximum(rv, i["hausdorff"])
        m = max(m, float(c.max()) if len(c) else 0.0)
    return m


def learned_run(ctx, meshes, plain_ms, plain_parts, top_k):
    """One step of the learned-proposal form (the shipped network ranks each
    cutting part's level-0 candidates; only the top_k are evaluated exactly)."""
    from cuacd_amd import proposal as PR
    net = PR.ProposalNet.load_default(ctx)
    try:
        t0 = time.perf_counter()
        parts, info = ctx.batch_lookahead_decompose(meshes, return_info=True, proposal=net, top_k=top_k, **PARAMS)
        ms = 1e3 * (time.perf_counter() - t0)
    finally:
        net.close()
    pr = ctx.last_lookahead_proposal
    n_parts = sum(len(p) for p in parts)
    iters = max(i["iters"] for i in info)
    out = {"metric": "lookahead meshes/sec at tau=0.05 (learned proposal)", "value": len(meshes) / (ms * 1e-3),
           "unit": "meshes/s", "ms_per_step": ms, "top_k": top_k, "speedup_vs_plain": plain_ms / ms,
           "parts_total": n_parts, "parts_vs_plain": n_parts / plain_parts - 1.0,
           "max_part_concavity": max_concavity(info), "iterations_max": int(iters),
           "candidates_ranked": pr["ranked"], "candidates_kept": pr["kept"], "fallbacks": pr["fallbacks"],
           "proposal_ms": ctx.last_lookahead_stats["proposal"]["ms"]}
    log(f"learned top{top_k}: {ms / 1e3:.2f} s ({out['value']:.1f} meshes/s, x{out['speedup_vs_plain']:.2f}), "
        f"parts {n_parts} ({100 * out['parts_vs_plain']:+.1f}%), max concavity {out['max_part_concavity']:.4f}, "
        f"kept {pr['kept']} of {pr['ranked']} candidates, network {out['proposal_ms']:.1f} ms")
    st = ctx.last_lookahead_stats
    log("  learned phases: " + ", ".join(f"{k} {v['ms']:.0f} ms" for k, v in st.items()) +
        f" (GPU {sum(v['ms'] for v in st.values()) / 1e3:.2f} s)")
    sec = ctx.last_lookahead_sections
    log("  learned host sections: " + ", ".join(f"{k} {v / 1e3:.2f} s" for k, v in sec.items()))
    return out


def score_workload(n):
    from cuacd_amd import datasets, sdf
    meshes = sdf.sdf_meshes(n, seed=3, res=64, min_tris=500, max_tris=60000)
    planes = [datasets.cut_candidates(v, t, n_edges=16, n_total=30, seed=m) for m, (v, t) in enumerate(meshes)]
    return meshes, planes


def cpu_score_baseline(meshes, planes, tau, budget_s=20.0):
    """The serial oracle scorer (one core) on the cuts of the smallest mesh,
    as many as fit in the time budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lookahead_oracle as LO
    m = int(np.argmin([len(t) for _, t in meshes]))
    v, t = meshes[m]
    done = 0
    t0 = time.perf_counter()
    for pl in planes[m]:
        LO.cut_cost(v, t, pl, tau)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "cuts/s", "cores": 1, "kind": "port",
            "sample": f"{done} cuts of the smallest mesh ({len(t)} tris) by the serial oracle scorer "
                      f"(C cut, exact hull, reference-sampler Hausdorff) in {dt:.1f} s"}


def main_score(args):
    import cuacd_amd
    from cuacd_amd import autotune
    tau = 0.05
    n = 256 if args.meshes == 64 else args.meshes
    t0 = time.perf_counter()
    meshes, planes = score_workload(n)
    tris = np.array([len(t) for _, t in meshes])
    mesh_of = np.concatenate([np.full(len(p), m, np.int32) for m, p in enumerate(planes)])
    pl = np.concatenate(planes)
    log(f"score workload: {len(meshes)} meshes, {tris.min()}-{tris.max()} tris (mean {tris.mean():.0f}), "
        f"{len(pl)} cuts, generated in {time.perf_counter() - t0:.1f} s")
    ctx = cuacd_amd.Context(device=0)
    name, n_cu, hbm = ctx.device_info()
    log(f"device {name}, {n_cu} CUs, {hbm / 2**30:.0f} GiB")
    ctx.batch_score_cuts(meshes[:1], mesh_of[:30], pl[:30], threshold=tau)  # first touch
    if args.autotune:
        best, trials = autotune.tune(ctx, lambda: ctx.batch_score_cuts(meshes, mesh_of, pl, threshold=tau), log=log)
        log(f"autotune: {best} ({len(trials)} trials)")
    for _ in range(args.warmup):
        ctx.batch_score_cuts(meshes, mesh_of, pl, threshold=tau)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        score, detail, st = ctx.batch_score_cuts(meshes, mesh_of, pl, threshold=tau, return_detail=True)
        times.append(time.perf_counter() - t0)
    ms_per_step = 1e3 * float(np.mean(times))
    scored = int(detail[:, 9].sum())
    n_h = int((detail[:, 7:9] >= 0).sum())
    log(f"step {ms_per_step:.1f} ms; {scored} of {len(pl)} cuts scored, {n_h} halves with Hausdorff; "
        f"stages: cuts {st[0]:.1f} ms, hulls {st[1]:.1f} ms, hausdorff {st[2]:.1f} ms")
    kern = {"la_cut_kernel": {"ms": st[0], "launches": 1, "alg_bytes": st[3]}}
    wide_min = ctx.get_tuning("hull_wide_min")
    half_cap = ctx.get_tuning("hull_half_cap") != 0
    for i, cls in enumerate(cuacd_amd.LOOKAHEAD_KERNELS):
        ms, la, by = st[6 + 3 * i: 9 + 3 * i]
        kern[kernel_name(cls, wide_min, half_cap)] = {"ms": ms, "launches": int(la), "alg_bytes": by}
    for name, k in sorted(kern.items(), key=lambda x: -x[1]["ms"])[:6]:
        log(f"  kernel {name:22s} {k['ms']:9.1f} ms  {k['launches']:5d} launches  "
            f"{k['alg_bytes'] / max(k['ms'], 1e-9) / 1e6:7.1f} GB/s algorithmic")
    dom = max(kern, key=lambda k: kern[k]["ms"])
    d = kern[dom]
    achieved = d["alg_bytes"] / (d["ms"] * 1e-3) / 1e9 if d["ms"] > 0 else 0.0
    cpu = None
    if not args.no_cpu:
        cpu = cpu_score_baseline(meshes, planes, tau)
        log(f"cpu baseline {cpu['value']:.3f} cuts/s ({cpu['sample']})")
    line = {
        "metric": "candidate cuts scored/sec (plain, batched exact scorer)",
        "value": len(pl) / (ms_per_step * 1e-3),
        "unit": "cuts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 geometry / fp64 volumes / int32 exact hull grid",
        "data": "synthetic",
        "config": {"workload": f"batched cut scoring: {len(meshes)} SDF meshes (marching tetrahedra 64^3, "
                               f"500-60k tris, extent 2, seed 3), 30 candidate planes each (16 concave-edge "
                               f"planes + axis cuts), tau 0.05", "cuts": int(len(pl)),
                   "tris_total": int(tris.sum()), "autotuned": bool(args.autotune),
                   "tuning": ctx.get_tuning()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": dom, "kernel_ms_total": d["ms"], "launches": d["launches"],
                     "avg_launch_ms": d["ms"] / max(d["launches"], 1),
                     "alg_bytes_per_launch": d["alg_bytes"] / max(d["launches"], 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
