"""Benchmark on one MI355X; prints ONE JSON line.

The line's metric is the plain lookahead decomposition (meshes per second):
64 watertight SDF meshes (smooth union of 3-8 random capsules/boxes minus 0-3
spheres, marching tetrahedra on a 48^3 grid, 2k-50k triangles, extent 2,
seed 4), lookahead_decompose at tau 0.05, width/width2 30/5, depth 2, quick
depth 0, 16 concave edges, 16 parts cut in parallel, at most 100 iterations,
hull merge on. A step is one batched search over all 64 meshes. --steps and
--warmup apply to these steps.

The same line carries the other workloads as objects, each with its own
roofline and CPU baseline:
  learned    the learned proposal on the same 64 meshes at tau 0.05 and 0.03
             (the shipped network keeps the top 16 and top 24 level-0
             candidates of each cutting part: the fewest that keep the part
             count within 10 % of the plain search's), against the plain
             search at the same tau;
  hull_batch batch exact hulls + FP64 volumes of 1024 clouds x 256 points
             uniform in the unit ball (seed 0), inputs resident in HBM;
  hausdorff  directed Hausdorff distances both ways between 512 noisy level-3
             icospheres (1280 triangles, radial scale 1 + 0.15 N(0,1), seed 0)
             and their exact hulls, 4096 stratified area-uniform samples per
             mesh, as a replayed HIP-graph plan over resident meshes;
  score      the batched exact cut scorer: 256 SDF meshes (64^3 grid,
             500-60k triangles, seed 3), 30 candidate planes each, tau 0.05;
             eager (host-sized stages), and as a score plan replayed as one
             captured HIP graph (object "graph").

CPU baselines run first, before the GPU is touched: the oracle on all host
cores (worker processes or OpenMP threads, at most 16) and on one core, each
on a bounded sample of its workload. They are baselines, not targets.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--workload all|lookahead|learned|hull|hausdorff|score]
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
PARAMS = dict(max_iters=100, width=30, width2=5, threshold=0.05, depth=2, max_n_cutting=16,
              n_concave_edges=16, merge_hulls=True)
# VALU busy fraction and HBM bytes per launch (FETCH_SIZE x 2.00 +
# WRITE_SIZE x 1.00, calibrated on a copy of known size) of the hull tiers,
# from their counters (profiles/lookahead_64sdf_hull_tier_counters.md):
# counters cannot be read inside the timed run, so the line carries the
# profiled figures beside the live HBM fraction
HULL_VALU_BUSY = {"la_hull_ldsv<2,8>": 0.042, "la_hull_ldsv<4,8>": 0.024, "la_hull_ldsv<2,4>": 0.034,
                  "la_hull_ldsv<8,8>": 0.024, "la_hull_block<8>": 0.004}
HULL_TRAFFIC = {"la_hull_ldsv<2,8>": 74.90e6, "la_hull_ldsv<4,8>": 49.78e6, "la_hull_ldsv<2,4>": 59.41e6,
                "la_hull_ldsv<8,8>": 78.93e6}
LA_LIMITER = ("latency: each hull is a chain of dependent insertion steps (LDS reads and writes, block barriers); "
              "HBM traffic is ~1x the algorithmic bytes, so the HBM fraction only says how far from a bandwidth "
              "bound this work sits")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cores():
    """Cores the CPU baselines may use: OMP_NUM_THREADS when set (the GPU box
    sets the one-GPU share), else this process's CPU affinity; at most 16."""
    env = os.environ.get("OMP_NUM_THREADS")
    n = int(env) if env and env.isdigit() else len(os.sched_getaffinity(0))
    return max(1, min(16, n))


def _oracle():
    p = os.path.join(ROOT, "oracle")
    if p not in sys.path:
        sys.path.insert(0, p)
    import oracle
    import lookahead_oracle
    return oracle, lookahead_oracle


def _worker_init():
    O, _ = _oracle()
    O.set_threads(1)


# ---------------------------------------------------------------------------
# workloads (host)

def lookahead_workload(n=64):
    from cuacd_amd import sdf
    return sdf.sdf_meshes(n, seed=4, res=48, min_tris=2000, max_tris=50000)


def score_workload(n=256):
    from cuacd_amd import datasets, sdf
    meshes = sdf.sdf_meshes(n, seed=3, res=64, min_tris=500, max_tris=60000)
    planes = [datasets.cut_candidates(v, t, n_edges=16, n_total=30, seed=m) for m, (v, t) in enumerate(meshes)]
    return meshes, planes


def hull_workload():
    from cuacd_amd import datasets
    return datasets.ball_clouds(1024, 256, seed=0)


def hausdorff_workload():
    from cuacd_amd import datasets
    return datasets.noisy_icospheres(512, 3, 0.15, seed=0)


# ---------------------------------------------------------------------------
# CPU baselines (the oracle; run before the GPU is initialised)

def _la_task(args):
    m, v, t, deadline = args
    O, LO = _oracle()
    if time.perf_counter() >= deadline:
        return m, 0, 0.0, False
    hv, ht, _ = LO.hull_mesh(v)  # the canonical exact hull, as the GPU builds it
    kw = dict(PARAMS)
    kw.pop("merge_hulls")
    t0 = time.perf_counter()
    _, done = LO.lookahead(v, t, hv, ht, deadline=deadline, **kw)
    dt = time.perf_counter() - t0
    return m, done, dt, time.perf_counter() < deadline


def cpu_lookahead(meshes, cores, budget_s):
    """The serial lookahead oracle (Python control over the C oracle's cut,
    hull, scoring and Hausdorff pieces) on the meshes from the smallest up,
    each until it converges, for about budget_s: on one core in this process,
    and on `cores` worker processes (one mesh each). A mesh cut short counts
    by the fraction of its search iterations done (the GPU run gives each
    mesh's iteration count). Without the merge post-pass."""
    O, _ = _oracle()
    O.set_threads(1)
    order = [int(m) for m in np.argsort([len(t) for _, t in meshes], kind="stable")]
    single = []
    t_end = time.perf_counter() + budget_s
    for m in order:
        r = _la_task((m, meshes[m][0], meshes[m][1], t_end))
        single.append(r)
        if not r[3]:
            break
    ctx = mp.get_context("spawn")
    with ctx.Pool(cores, initializer=_worker_init) as pool:
        pool.map(int, range(cores))  # workers up before the clock starts
        t0 = time.perf_counter()
        deadline = t0 + budget_s
        multi = pool.map(_la_task, [(m, meshes[m][0], meshes[m][1], deadline) for m in order], chunksize=1)
        wall = time.perf_counter() - t0
    return {"single": single, "multi": [r for r in multi if r[1] > 0], "wall": wall, "cores": cores}


def finish_cpu_lookahead(raw, meshes, gpu_iters):
    sizes = [len(t) for _, t in meshes]

    def worth(rs):
        return sum(min(max(done, 1) / max(int(gpu_iters[m]), 1), 1.0) for m, done, _, _ in rs if done > 0)

    def which(rs):
        ms = [m for m, done, _, _ in rs if done > 0]
        return f"{len(ms)} meshes of {min(sizes[m] for m in ms)}-{max(sizes[m] for m in ms)} tris" if ms else "none"

    s_time = sum(dt for _, _, dt, _ in raw["single"])
    ws, wm = worth(raw["single"]), worth(raw["multi"])
    return {"value": wm / raw["wall"], "unit": "meshes/s", "cores": raw["cores"], "kind": "port",
            "sample": f"the set's smallest meshes, each to convergence, on {raw['cores']} worker processes for "
                      f"{raw['wall']:.1f} s: {which(raw['multi'])}, {wm:.2f} meshes' worth of search iterations "
                      f"(without the merge pass)",
            "single_core": {"value": ws / max(s_time, 1e-9), "cores": 1,
                            "sample": f"{which(raw['single'])}, {ws:.2f} meshes' worth in {s_time:.1f} s"}}


def cpu_hull(clouds, cores, budget_s):
    """The oracle's exact hull + FP64 volume (C, OpenMP over clouds) on the
    full 1024-cloud batch: passes on one thread, then on `cores` threads."""
    O, _ = _oracle()
    packed = np.ascontiguousarray(np.concatenate(clouds, 0), np.float32)
    off = np.zeros(len(clouds) + 1, np.int32)
    off[1:] = np.cumsum([len(c) for c in clouds])
    out = {}
    for thr in (1, cores):
        O.set_threads(thr)
        O.hull_batch(packed, off, check_general=False)
        reps, t0 = 0, time.perf_counter()
        while reps == 0 or time.perf_counter() - t0 < budget_s / 2:
            O.hull_batch(packed, off, check_general=False)
            reps += 1
        dt = time.perf_counter() - t0
        out[thr] = (reps * len(clouds) / dt, f"{reps} passes over the {len(clouds)}-cloud batch in {dt:.2f} s")
    O.set_threads(cores)
    return {"value": out[cores][0], "unit": "clouds/s", "cores": cores, "kind": "port",
            "sample": out[cores][1] + f" on {cores} threads",
            "single_core": {"value": out[1][0], "cores": 1, "sample": out[1][1]}}


def _surface_samples(v, t, n, rng):
    a, b, c = v[t[:, 0]].astype(np.float64), v[t[:, 1]].astype(np.float64), v[t[:, 2]].astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    tri = rng.choice(len(t), size=n, p=area / area.sum())
    r1, r2 = np.sqrt(rng.uniform(size=(n, 1))), rng.uniform(size=(n, 1))
    p = (1 - r1) * a[tri] + r1 * (1 - r2) * b[tri] + r1 * r2 * c[tri]
    return p.astype(np.float32)


def cpu_hausdorff(meshes, cores, budget_s):
    """FP64 brute force (the oracle: exact point-triangle distances to every
    triangle, C, OpenMP over pairs) of both directed distances between each
    icosphere and its exact hull, 4096 area-uniform samples per surface (the
    workload's sample count; numpy's sampler - the distance work does not
    depend on where the samples fall): batches of meshes on one thread, then
    on `cores` threads, for about budget_s / 2 each; only the distance calls
    are timed."""
    O, LO = _oracle()
    rng = np.random.default_rng(0)
    hulls = [LO.hull_mesh(v)[:2] for v, _ in meshes]
    res = {}
    for thr in (1, cores):
        O.set_threads(thr)
        done, spent, i = 0, 0.0, 0
        while spent < budget_s / 2 and done < len(meshes):
            idx = [(i + k) % len(meshes) for k in range(thr)]
            i += thr
            sm = np.stack([_surface_samples(*meshes[k], 4096, rng) for k in idx])
            sh = np.stack([_surface_samples(*hulls[k], 4096, rng) for k in idx])
            t0 = time.perf_counter()
            O.directed_batch(sm, [hulls[k] for k in idx])
            O.directed_batch(sh, [meshes[k] for k in idx])
            spent += time.perf_counter() - t0
            done += thr
        res[thr] = (done, spent)
    O.set_threads(cores)
    return {"value": res[cores][0] / res[cores][1], "unit": "meshes/s", "cores": cores, "kind": "port",
            "sample": f"{res[cores][0]} meshes (both directions) in {res[cores][1]:.2f} s on {cores} threads",
            "single_core": {"value": res[1][0] / res[1][1], "cores": 1,
                            "sample": f"{res[1][0]} meshes (both directions) in {res[1][1]:.2f} s on one thread"}}


def _score_task(args):
    m, v, t, planes, tau, deadline = args
    _, LO = _oracle()
    done, t0 = 0, time.perf_counter()
    for pl in planes:
        if time.perf_counter() >= deadline:
            break
        LO.cut_cost(v, t, pl, tau)
        done += 1
    return m, done, time.perf_counter() - t0


def cpu_score(meshes, planes, tau, cores, budget_s):
    """The serial oracle scorer (C cut, exact hull, reference-sampler
    Hausdorff) on cuts of the set's meshes from the smallest up: one core in
    this process, then `cores` worker processes (one mesh each)."""
    O, _ = _oracle()
    O.set_threads(1)
    order = [int(m) for m in np.argsort([len(t) for _, t in meshes], kind="stable")]
    m0 = order[0]
    _, d1, s1 = _score_task((m0, meshes[m0][0], meshes[m0][1], planes[m0], tau, time.perf_counter() + budget_s / 2))
    ctx = mp.get_context("spawn")
    with ctx.Pool(cores, initializer=_worker_init) as pool:
        pool.map(int, range(cores))
        t0 = time.perf_counter()
        deadline = t0 + budget_s / 2
        res = pool.map(_score_task, [(m, meshes[m][0], meshes[m][1], planes[m], tau, deadline)
                                     for m in order[:4 * cores]], chunksize=1)
        wall = time.perf_counter() - t0
    cuts = sum(d for _, d, _ in res)
    sizes = [len(meshes[m][1]) for m, d, _ in res if d > 0]
    return {"value": cuts / wall, "unit": "cuts/s", "cores": cores, "kind": "port",
            "sample": f"{cuts} cuts of {len(sizes)} of the smallest meshes ({min(sizes)}-{max(sizes)} tris) on "
                      f"{cores} worker processes in {wall:.1f} s",
            "single_core": {"value": d1 / s1, "cores": 1,
                            "sample": f"{d1} cuts of the smallest mesh ({len(meshes[m0][1])} tris) in {s1:.1f} s"}}


# ---------------------------------------------------------------------------
# GPU legs

def kernel_name(cls, wide_min, half_cap, ldsb8=0, w8=0):
    """Kernel behind a hull class reported by the library (w8: the hull_w8
    knob, eight-wave multi-apex kernels for the 513-1024-point (bit 0) and
    257-512-point (bit 1) wave tiers, bits 2 and 3 for their retries)."""
    if cls == "kdop_prefilter":
        return "la_kdop (extremes + rough hull + filter)"
    if cls.endswith("_retry"):
        base = cls[:-6]
        n = int(base[6:])
        if base.startswith("hull_b"):
            return f"la_hull_block<{n // 512}>"
        if (w8 >> (2 if n == 1024 else 3)) & 1:
            return f"la_hull_lds8<{n // 512}>"
        return f"la_hull_lds4<{n // 256}>"
    if cls.startswith("hull_b"):
        n = int(cls[6:])
        ppl = n // 512
        if n == 2048 and half_cap:
            return "la_hull_ldsv<4,8>"
        if n == 4096 and half_cap and ldsb8:
            return "la_hull_ldsv<8,8>"
        return f"la_hull_block<{ppl}>"
    n = int(cls[6:])
    ppl = n // 64
    if wide_min > 0 and ppl >= 4 and n > wide_min:
        if ppl >= 8 and half_cap:
            if (w8 >> (0 if ppl == 16 else 1)) & 1:
                return f"la_hull_ldsv<{ppl // 8},8>"
            return f"la_hull_ldsv<{ppl // 4},4>"
        return f"la_hull_lds4<{ppl // 4}>"
    return f"la_hull_wave<{ppl}>"


def max_concavity(info):
    m = 0.0
    for i in info:
        rv = 0.3 * np.cbrt(3 / (4 * np.pi) * np.maximum(i["hull_vol"] - i["mesh_vol"], 0))
        c = np.maximum(rv, i["hausdorff"])
        m = max(m, float(c.max()) if len(c) else 0.0)
    return m


def _roofline(kern, bound="hbm"):
    dom = max(kern, key=lambda k: kern[k]["ms"])
    d = kern[dom]
    launches = max(d["launches"], 1)
    avg_ms = d["ms"] / launches
    per_launch = d["alg_bytes"] / launches
    achieved = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return {"bound": bound, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom, "kernel_ms_total": d["ms"],
            "launches": d["launches"], "avg_launch_ms": avg_ms, "alg_bytes_per_launch": per_launch}


def gpu_lookahead(ctx, meshes, steps, warmup, tau=0.05, verbose=True):
    kw = dict(PARAMS, threshold=tau)
    for _ in range(warmup):
        ctx.batch_lookahead_decompose(meshes, **kw)
    times = []
    for _ in range(max(steps, 1)):
        t0 = time.perf_counter()
        parts, info = ctx.batch_lookahead_decompose(meshes, return_info=True, **kw)
        times.append(time.perf_counter() - t0)
    ms = 1e3 * float(np.mean(times))
    n_parts = np.array([len(p) for p in parts])
    iters = np.array([i["iters"] for i in info])
    out = {"ms": ms, "parts": int(n_parts.sum()), "iters": iters, "max_conc": max_concavity(info),
           "stats": ctx.last_lookahead_stats, "sections": ctx.last_lookahead_sections,
           "sections_gpu": ctx.last_lookahead_sections_gpu, "sections_syncs": ctx.last_lookahead_sections_syncs,
           "kernels": ctx.last_lookahead_kernels}
    if verbose:
        log(f"plain tau {tau}: step {ms / 1e3:.2f} s; parts {n_parts.sum()} (per mesh {n_parts.min()}-"
            f"{n_parts.max()}), iterations {iters.min()}-{iters.max()}, meshes at max_iters "
            f"{(iters >= PARAMS['max_iters']).sum()}, max concavity {out['max_conc']:.4f}")
    return out


def _add_kernel(kern, name, k):
    """Several timer classes can run the same kernel (a size class and the
    retries of a smaller tier): their times, launches and bytes add up."""
    if name in kern:
        o = kern[name]
        kern[name] = {"ms": o["ms"] + k["ms"], "launches": o["launches"] + k["launches"],
                      "alg_bytes": o["alg_bytes"] + k["alg_bytes"]}
    else:
        kern[name] = dict(k)


def lookahead_kernels(ctx, r):
    stats = r["stats"]
    gpu_ms = sum(s["ms"] for s in stats.values())
    for k, s in stats.items():
        log(f"  {k:14s} {s['ms']:9.1f} ms  {s['launches']:5d} launches  {s['alg_bytes'] / 1e9:8.3f} GB "
            f"({s['alg_bytes'] / max(s['ms'], 1e-9) / 1e6:7.1f} GB/s)")
    log(f"  GPU time in phases {gpu_ms / 1e3:.2f} s of {r['ms'] / 1e3:.2f} s")
    sec = r["sections"]
    log("  host sections: " + ", ".join(f"{k} {v / 1e3:.2f} s" for k, v in sec.items()) +
        f" (sum {sum(sec.values()) / 1e3:.2f} s)")
    sg, sy = r["sections_gpu"], r["sections_syncs"]
    log("  outside GPU phases, by section: " +
        ", ".join(f"{k} {(sec[k] - sg[k]) / 1e3:.3f} s / {sy[k]} waits" for k in sec) +
        f" (sum {(sum(sec.values()) - sum(sg.values())) / 1e3:.2f} s, {sum(sy.values())} host waits)")
    kern = {}
    for ph, name in (("leaf_scoring", "la_score_kernel"), ("level0_cuts", "la_cut_kernel")):
        kern[name] = stats[ph]
    wide_min = ctx.get_tuning("hull_wide_min")
    half_cap = ctx.get_tuning("hull_half_cap") != 0
    ldsb8 = ctx.get_tuning("hull_ldsb8_vcap")
    w8 = ctx.get_tuning("hull_w8")
    for cls, k in r["kernels"].items():
        _add_kernel(kern, kernel_name(cls, wide_min, half_cap, ldsb8, w8), k)
    for name, k in sorted(kern.items(), key=lambda x: -x[1]["ms"]):
        if k["launches"] == 0:
            continue
        log(f"  kernel {name:22s} {k['ms']:9.1f} ms  {k['launches']:5d} launches  "
            f"{k['alg_bytes'] / max(k['ms'], 1e-9) / 1e6:7.1f} GB/s algorithmic")
    return kern


def gpu_learned(ctx, meshes, plain, top_ks, reps):
    """The learned proposal at each tau against the plain search at the same
    tau; top_ks maps tau to the number of level-0 candidates kept per part."""
    from cuacd_amd import proposal as PR
    net = PR.ProposalNet.load_default(ctx)
    out = {}
    try:
        for tau, pl in plain.items():
            top_k = top_ks[tau]
            kw = dict(PARAMS, threshold=tau)
            ctx.batch_lookahead_decompose(meshes, proposal=net, top_k=top_k, **kw)  # warm
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                parts, info = ctx.batch_lookahead_decompose(meshes, return_info=True, proposal=net, top_k=top_k,
                                                            **kw)
                times.append(time.perf_counter() - t0)
            ms = 1e3 * float(np.mean(times))
            pr = ctx.last_lookahead_proposal
            n_parts = sum(len(p) for p in parts)
            iters = np.array([i["iters"] for i in info])
            o = {"metric": f"lookahead meshes/sec at tau={tau} (learned proposal)",
                 "value": len(meshes) / (ms * 1e-3), "unit": "meshes/s", "ms_per_step": ms,
                 "ms_per_step_runs": [1e3 * x for x in times], "top_k": top_k,
                 "plain_ms_per_step": pl["ms"], "speedup_vs_plain": pl["ms"] / ms, "parts_total": n_parts,
                 "plain_parts_total": pl["parts"], "parts_vs_plain": n_parts / pl["parts"] - 1.0,
                 "max_part_concavity": max_concavity(info), "plain_max_part_concavity": pl["max_conc"],
                 "every_part_within_tau": bool(max_concavity(info) <= tau + 1e-6),
                 "meshes_at_max_iters": int((iters >= PARAMS["max_iters"]).sum()),
                 "candidates_ranked": pr["ranked"], "candidates_kept": pr["kept"], "fallbacks": pr["fallbacks"],
                 "proposal_ms": ctx.last_lookahead_stats["proposal"]["ms"]}
            out[f"tau_{tau}"] = o
            log(f"learned top{top_k} tau {tau}: {ms / 1e3:.3f} s ({o['value']:.1f} meshes/s, "
                f"x{o['speedup_vs_plain']:.2f} vs plain {pl['ms'] / 1e3:.2f} s), parts {n_parts} "
                f"({100 * o['parts_vs_plain']:+.1f}%), max concavity {o['max_part_concavity']:.4f}, kept "
                f"{pr['kept']} of {pr['ranked']} candidates, network {o['proposal_ms']:.1f} ms; runs "
                + ", ".join(f"{x:.3f}" for x in times))
            st = ctx.last_lookahead_stats
            log("  learned phases: " + ", ".join(f"{k} {v['ms']:.0f} ms" for k, v in st.items()) +
                f" (GPU {sum(v['ms'] for v in st.values()) / 1e3:.2f} s)")
            sec = ctx.last_lookahead_sections
            log("  learned host sections: " + ", ".join(f"{k} {v / 1e3:.2f} s" for k, v in sec.items()))
    finally:
        net.close()
    return out


def gpu_hull(ctx, clouds, launches):
    """Exact hulls + FP64 volumes of the resident batch: gpu_hull_dandc_dev,
    one hull_wave_kernel<4> launch per step; HIP events on the context's
    stream around `launches` back-to-back launches."""
    from cuacd_amd import _gpu
    n, npts = len(clouds), len(clouds[0])
    packed = np.ascontiguousarray(np.concatenate(clouds, 0), np.float32)
    off = np.arange(n + 1, dtype=np.int32) * npts
    mhv, mht = npts, 2 * npts - 4
    bufs = []

    def dalloc(nbytes):
        bufs.append(_gpu.dev_alloc(int(nbytes)))
        return bufs[-1]

    d_pts, d_off = dalloc(packed.nbytes), dalloc(off.nbytes)
    d_v, d_t = dalloc(n * mhv * 12), dalloc(n * mht * 12)
    d_nv, d_nt, d_e, d_vol = dalloc(n * 4), dalloc(n * 4), dalloc(n * 4), dalloc(n * 8)
    try:
        _gpu.memcpy_h2d(d_pts, packed.ctypes.data, packed.nbytes)
        _gpu.memcpy_h2d(d_off, off.ctypes.data, off.nbytes)

        def step():
            _gpu.hull_dandc_dev(d_pts, d_off, n, npts, mhv, mht, d_v, d_t, d_nv, d_nt, d_e, d_vol)

        for _ in range(5):
            step()
        _gpu.sync()
        t0 = time.perf_counter()
        _gpu.timer_start()
        for _ in range(launches):
            step()
        gpu_ms = _gpu.timer_stop()
        _gpu.sync()
        wall_ms = (time.perf_counter() - t0) * 1e3 / launches
        nv, nt, err = (np.empty(n, np.int32) for _ in range(3))
        vol = np.empty(n, np.float64)
        for a, d in ((nv, d_nv), (nt, d_nt), (err, d_e), (vol, d_vol)):
            _gpu.memcpy_d2h(a.ctypes.data, d, a.nbytes)
    finally:
        for b in bufs:
            _gpu.dev_free(b)
    if (err != 0).any():
        raise RuntimeError(f"hull batch reported errors {np.unique(err)}")
    kms = gpu_ms / launches
    # algorithmic bytes per launch: the clouds and offsets read, each hull's
    # vertices and triangles, counts, error word and FP64 volume written
    alg = packed.nbytes + off.nbytes + int(nv.sum()) * 12 + int(nt.sum()) * 12 + n * (12 + 8)
    achieved = alg / (kms * 1e-3) / 1e9
    log(f"hull batch: {kms:.4f} ms per launch ({n / (kms * 1e-3):.0f} clouds/s), wall {wall_ms:.4f} ms, "
        f"mean hull nv {nv.mean():.1f} nt {nt.mean():.1f}, mean volume {vol.mean():.4f}")
    return {"metric": "batch_hull_volume clouds/sec (exact hull + FP64 volume)", "value": n / (wall_ms * 1e-3),
            "unit": "clouds/s", "ms_per_step": wall_ms, "launches": launches,
            "config": {"workload": "batch_hull_volume: 1024 clouds x 256 fp32 points uniform in the unit ball "
                                   "(rejection-sampled from [-1,1]^3, seed 0), resident in HBM",
                       "hull_nv_mean": float(nv.mean()), "hull_nt_mean": float(nt.mean()),
                       "volume_mean": float(vol.mean())},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "hull_wave_kernel<4>",
                         "avg_launch_ms": kms, "alg_bytes_per_launch": alg, "limiter": LA_LIMITER}}


def gpu_hausdorff(ctx, meshes, replays, n_samples=4096):
    """Both directed distances between each icosphere and its exact hull on
    resident meshes: a Hausdorff plan (preparation + query launches captured
    as a HIP graph) replayed `replays` times; HIP events on the context's
    stream around the replays. The hulls are built first, outside the timed
    region."""
    from cuacd_amd import _gpu
    hulls = [(hv, ht) for hv, ht, _ in ctx.batch_hull_dandc_mesh([v for v, _ in meshes])]
    n = len(meshes)

    def pack(ms):
        vs = [np.ascontiguousarray(v, np.float32) for v, _ in ms]
        ts = [np.ascontiguousarray(t, np.int32) for _, t in ms]
        vo = np.zeros(n + 1, np.int32)
        to = np.zeros(n + 1, np.int32)
        vo[1:] = np.cumsum([len(v) for v in vs])
        to[1:] = np.cumsum([len(t) for t in ts])
        return np.concatenate(vs), vo, np.concatenate(ts), to

    av, avo, at, ato = pack(meshes)
    bv, bvo, bt, bto = pack(hulls)
    bufs = []

    def up(a):
        bufs.append(_gpu.dev_alloc(max(int(a.nbytes), 16)))
        _gpu.memcpy_h2d(bufs[-1], a.ctypes.data, a.nbytes)
        return bufs[-1]

    plan = 0
    try:
        d_av, d_at, d_bv, d_bt = up(av), up(at), up(bv), up(bt)
        d_out = up(np.zeros(2 * n, np.float32))
        plan = _gpu.hausdorff_plan_create(d_av, d_at, d_bv, d_bt, avo.ctypes.data, ato.ctypes.data,
                                          bvo.ctypes.data, bto.ctypes.data, n, 1, n_samples, d_out, d_out + 4 * n,
                                          0, 0)
        for _ in range(3):
            _gpu.hausdorff_plan_run(plan)
        _gpu.sync()
        t0 = time.perf_counter()
        _gpu.timer_start()
        for _ in range(replays):
            _gpu.hausdorff_plan_run(plan)
        gpu_ms = _gpu.timer_stop()
        _gpu.sync()
        wall_ms = (time.perf_counter() - t0) * 1e3 / replays
        d2 = np.empty(2 * n, np.float32)
        _gpu.memcpy_d2h(d2.ctypes.data, d_out, d2.nbytes)
    finally:
        if plan:
            _gpu.hausdorff_plan_destroy(plan)
        for b in bufs:
            _gpu.dev_free(b)
    d = np.sqrt(d2)
    rms = gpu_ms / replays
    # compulsory bytes: both meshes of each pair read once, two distances written
    alg = av.nbytes + at.nbytes + bv.nbytes + bt.nbytes + 8 * n
    achieved = alg / (rms * 1e-3) / 1e9
    # brute-force-equivalent work: every sample against every triangle of the
    # other surface, ~70 flops per point-triangle distance
    flops = float(n_samples) * (len(at) + len(bt)) * 70.0
    tfl = flops / (rms * 1e-3) / 1e12
    log(f"hausdorff: {rms:.3f} ms per replay ({n / (rms * 1e-3):.0f} meshes/s), wall {wall_ms:.3f} ms; "
        f"H(mesh->hull) mean {d[:n].mean():.4f}, H(hull->mesh) mean {d[n:].mean():.4f}; brute-force-equivalent "
        f"{tfl:.1f} TFLOP/s")
    return {"metric": "concavity metric meshes/sec (directed Hausdorff both ways, 4096 samples)",
            "value": n / (wall_ms * 1e-3), "unit": "meshes/s", "ms_per_step": wall_ms, "replays": replays,
            "config": {"workload": "512 level-3 icospheres (642 v / 1280 t, radial scale 1 + 0.15 N(0,1), seed 0) "
                                   "against their exact hulls, 4096 stratified area-uniform samples per surface, "
                                   "resident in HBM, one HIP-graph plan replay per step",
                       "hull_nv_mean": float(np.mean([len(h[0]) for h in hulls])),
                       "h_mesh_to_hull_mean": float(d[:n].mean()), "h_hull_to_mesh_mean": float(d[n:].mean())},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "hd_prep_kernel + hd_query_kernel (one graph replay)", "avg_launch_ms": rms,
                         "alg_bytes_per_launch": alg,
                         "valu_equivalent": {"gflop_per_launch": flops / 1e9, "achieved_tflops": tfl,
                                             "peak_tflops": FP32_PEAK_TFLOPS, "frac": tfl / FP32_PEAK_TFLOPS,
                                             "note": "brute-force-equivalent work (every sample against every "
                                                     "triangle, 70 flops each); the BVH does less"}}}


def gpu_score_graph(ctx, meshes, mesh_of, pl, tau, steps, want_score, want_detail):
    """The same batch as a score plan: every stage sized on the device and the
    whole level replayed as one captured HIP graph; a step is one replay plus
    the readback and the host's cost formula, checked against the eager run."""
    try:
        plan = ctx.score_plan(meshes, mesh_of, pl, threshold=tau)
    except RuntimeError as e:
        log(f"score graph: not planned ({e})")
        return {"planned": False, "reason": str(e)}
    try:
        plan.run()  # warm
        times = []
        same = True
        for _ in range(steps):
            t0 = time.perf_counter()
            sc, dt = plan.run(return_detail=True)
            times.append(time.perf_counter() - t0)
            same = same and sc.tobytes() == want_score.tobytes() and dt.tobytes() == want_detail.tobytes()
    finally:
        plan.close()
    ms = 1e3 * float(np.mean(times))
    log(f"score graph: {ms:.1f} ms per replay ({len(pl) / (ms * 1e-3):.0f} cuts/s), identical to eager: {same}")
    return {"planned": True, "value": len(pl) / (ms * 1e-3), "unit": "cuts/s", "ms_per_step": ms,
            "ms_per_step_runs": [1e3 * x for x in times], "identical_to_eager": bool(same),
            "host_syncs_per_step": 1}


def gpu_score(ctx, meshes, planes, steps, tau=0.05):
    import cuacd_amd
    mesh_of = np.concatenate([np.full(len(p), m, np.int32) for m, p in enumerate(planes)])
    pl = np.concatenate(planes)
    ctx.batch_score_cuts(meshes[:1], mesh_of[:30], pl[:30], threshold=tau)  # first touch
    ctx.batch_score_cuts(meshes, mesh_of, pl, threshold=tau)  # warm
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        score, detail, st = ctx.batch_score_cuts(meshes, mesh_of, pl, threshold=tau, return_detail=True)
        times.append(time.perf_counter() - t0)
    ms = 1e3 * float(np.mean(times))
    scored = int(detail[:, 9].sum())
    n_h = int((detail[:, 7:9] >= 0).sum())
    log(f"score: {ms:.1f} ms per step ({len(pl) / (ms * 1e-3):.0f} cuts/s); {scored} of {len(pl)} cuts scored, "
        f"{n_h} halves with Hausdorff; stages: cuts {st[0]:.1f} ms, hulls {st[1]:.1f} ms, hausdorff {st[2]:.1f} ms")
    kern = {"la_cut_kernel": {"ms": st[0], "launches": 1, "alg_bytes": st[3]}}
    wide_min = ctx.get_tuning("hull_wide_min")
    half_cap = ctx.get_tuning("hull_half_cap") != 0
    ldsb8 = ctx.get_tuning("hull_ldsb8_vcap")
    w8 = ctx.get_tuning("hull_w8")
    for i, cls in enumerate(cuacd_amd.LOOKAHEAD_KERNELS):
        k_ms, la, by = st[6 + 3 * i: 9 + 3 * i]
        _add_kernel(kern, kernel_name(cls, wide_min, half_cap, ldsb8, w8), {"ms": k_ms, "launches": int(la), "alg_bytes": by})
    for name, k in sorted(kern.items(), key=lambda x: -x[1]["ms"]):
        if k["launches"] == 0:
            continue
        log(f"  kernel {name:22s} {k['ms']:9.1f} ms  {k['launches']:5d} launches  "
            f"{k['alg_bytes'] / max(k['ms'], 1e-9) / 1e6:7.1f} GB/s algorithmic")
    tris = np.array([len(t) for _, t in meshes])
    roof = _roofline(kern)
    roof["limiter"] = LA_LIMITER
    graph = gpu_score_graph(ctx, meshes, mesh_of, pl, tau, steps, score, detail)
    return {"metric": "candidate cuts scored/sec (plain, batched exact scorer)", "value": len(pl) / (ms * 1e-3),
            "unit": "cuts/s", "ms_per_step": ms, "steps": steps, "graph": graph,
            "config": {"workload": f"batched cut scoring: {len(meshes)} SDF meshes (marching tetrahedra 64^3, "
                                   f"500-60k tris, extent 2, seed 3), 30 candidate planes each (16 concave-edge "
                                   f"planes + axis cuts), tau 0.05", "cuts": int(len(pl)),
                       "tris_total": int(tris.sum()), "stage_ms": {"cuts": st[0], "hulls": st[1], "hausdorff": st[2]}},
            "roofline": roof}


# ---------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--meshes", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["all", "lookahead", "learned", "hull", "hausdorff", "score"],
                    default="all")
    ap.add_argument("--autotune", action="store_true", help="tune the hull launch shapes on the score workload")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only run the known-size device copy (HBM counter calibration) and exit")
    ap.add_argument("--top-k", default="16,24",
                    help="level-0 candidates kept per part by the learned proposal at tau 0.05 and 0.03 "
                         "(one value for both, or two; 0 skips the learned step)")
    ap.add_argument("--learned-reps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: one GPU only (asked for {args.gpus}); the meshes of the batch are independent and "
              "shard across GPUs without communication", file=sys.stderr)
        sys.exit(2)
    w = args.workload
    want = {k: w in ("all", k) for k in ("lookahead", "learned", "hull", "hausdorff", "score")}
    want["lookahead"] = want["lookahead"] or want["learned"]
    cores = host_cores()

    # ---- workloads and CPU baselines, before the GPU is touched
    t0 = time.perf_counter()
    meshes = lookahead_workload(args.meshes) if want["lookahead"] else None
    clouds = hull_workload() if want["hull"] else None
    spheres = hausdorff_workload() if want["hausdorff"] else None
    smeshes, splanes = score_workload() if want["score"] else (None, None)
    log(f"workloads generated in {time.perf_counter() - t0:.1f} s; CPU baselines on {cores} cores")
    if meshes is not None:
        tris = np.array([len(t) for _, t in meshes])
        log(f"lookahead workload: {len(meshes)} meshes, {tris.min()}-{tris.max()} tris (mean {tris.mean():.0f})")
    cpu = {}
    if not args.no_cpu:
        t0 = time.perf_counter()
        if want["score"]:
            cpu["score"] = cpu_score(smeshes, splanes, 0.05, cores, budget_s=12.0)
        if want["lookahead"] and w in ("all", "lookahead"):
            cpu["lookahead_raw"] = cpu_lookahead(meshes, cores, budget_s=12.0)
        if want["hull"]:
            cpu["hull"] = cpu_hull(clouds, cores, budget_s=6.0)
        if want["hausdorff"]:
            cpu["hausdorff"] = cpu_hausdorff(spheres, cores, budget_s=10.0)
        log(f"CPU baselines done in {time.perf_counter() - t0:.1f} s")

    import cuacd_amd
    ctx = cuacd_amd.Context(device=0)
    name, n_cu, hbm = ctx.device_info()
    log(f"device {name}, {n_cu} CUs, {hbm / 2**30:.0f} GiB")
    if args.copy_probe:
        nb = args.copy_probe << 20
        ms = cuacd_amd._gpu.copy_probe(nb, 3)
        print(json.dumps({"copy_probe_bytes": nb, "launches": 3, "ms_per_copy": ms,
                          "gbs": 2 * nb / (ms * 1e-3) / 1e9}), flush=True)
        ctx.close()
        return
    if args.autotune and want["score"]:
        from cuacd_amd import autotune
        mesh_of = np.concatenate([np.full(len(p), m, np.int32) for m, p in enumerate(splanes)])
        pl = np.concatenate(splanes)
        best, trials = autotune.tune(ctx, lambda: ctx.batch_score_cuts(smeshes, mesh_of, pl, threshold=0.05),
                                     log=log)
        log(f"autotune: {best} ({len(trials)} trials)")

    objs = {}
    line = None
    if want["lookahead"]:
        ctx.batch_lookahead_decompose(meshes[:1], **dict(PARAMS, max_iters=2))  # first touch
        plain = gpu_lookahead(ctx, meshes, args.steps, args.warmup)
        kern = lookahead_kernels(ctx, plain)
        roof = _roofline(kern)
        roof["limiter"] = LA_LIMITER
        if roof["kernel"] in HULL_VALU_BUSY:
            roof["valu_busy"] = {"frac": HULL_VALU_BUSY[roof["kernel"]], "measured": "rocprofv3 SQ counters",
                                 "profile": "profiles/lookahead_64sdf_hull_tier_counters.md"}
        if roof["kernel"] in HULL_TRAFFIC:
            roof["traffic"] = HULL_TRAFFIC[roof["kernel"]]
            roof["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this workload, "
                                      "profiles/lookahead_64sdf_hull_tier_counters.md")
        cb = None
        if "lookahead_raw" in cpu:
            cb = finish_cpu_lookahead(cpu["lookahead_raw"], meshes, plain["iters"])
            log(f"cpu baseline lookahead {cb['value']:.4f} meshes/s on {cb['cores']} cores "
                f"({cb['single_core']['value']:.4f} on one)")
        tris = np.array([len(t) for _, t in meshes])
        line = {
            "metric": "lookahead meshes/sec at tau=0.05 (plain)",
            "value": len(meshes) / (plain["ms"] * 1e-3),
            "unit": "meshes/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": plain["ms"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "fp32 geometry / fp64 volumes / int32 exact hull grid",
            "data": "synthetic",
            "config": {"workload": "plain lookahead_decompose: 64 SDF meshes (marching tetrahedra 48^3, 2k-50k "
                                   "tris, extent 2, seed 4), tau 0.05, width/width2 30/5, depth 2, quick 0, "
                                   "max_iters 100, merge_hulls on, n_concave_edges 16, max_n_cutting 16",
                       "meshes": len(meshes), "tris_total": int(tris.sum()), "parts_total": plain["parts"],
                       "max_part_concavity": plain["max_conc"], "iterations_max": int(plain["iters"].max())},
            "roofline": roof,
            "cpu_baseline": cb,
            # host wall time of the step by section of the search, the GPU-phase
            # time inside each section and the host waits on the stream
            "host_sections": {k: {"wall_ms": plain["sections"][k], "gpu_phase_ms": plain["sections_gpu"][k],
                                  "host_waits": plain["sections_syncs"][k]} for k in plain["sections"]},
        }
        ks = [int(x) for x in str(args.top_k).split(",")]
        top_ks = {0.05: ks[0], 0.03: ks[-1]}
        if want["learned"] and min(ks) > 0:
            plains = {0.05: plain, 0.03: gpu_lookahead(ctx, meshes, 1, 0, tau=0.03)}
            objs["learned"] = gpu_learned(ctx, meshes, plains, top_ks, args.learned_reps)
    if want["hull"]:
        o = gpu_hull(ctx, clouds, 50)
        o["cpu_baseline"] = cpu.get("hull")
        objs["hull_batch"] = o
    if want["hausdorff"]:
        o = gpu_hausdorff(ctx, spheres, 20)
        o["cpu_baseline"] = cpu.get("hausdorff")
        objs["hausdorff"] = o
    if want["score"]:
        o = gpu_score(ctx, smeshes, splanes, steps=3)
        o["cpu_baseline"] = cpu.get("score")
        objs["score"] = o
    for k, o in objs.items():
        c = o.get("cpu_baseline") if isinstance(o, dict) else None
        if c:
            log(f"cpu baseline {k}: {c['value']:.4g} {c['unit']} on {c['cores']} cores, "
                f"{c['single_core']['value']:.4g} on one")
    if line is None:
        # a single non-lookahead workload: its object is the line
        k, o = next(iter(objs.items()))
        line = {"metric": o["metric"], "value": o["value"], "unit": o["unit"], "n_gpus": 1,
                "steps": o.get("steps", o.get("launches", o.get("replays"))), "warmup": args.warmup,
                "ms_per_step": o["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "fp32 geometry / fp64 volumes / int32 exact hull grid", "data": "synthetic",
                "config": o["config"], "roofline": o["roofline"], "cpu_baseline": o.get("cpu_baseline")}
    else:
        line.update(objs)
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
