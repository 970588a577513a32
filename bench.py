"""Benchmark of the hot path on one MI355X; prints ONE JSON line.

Workload (current): batch exact hull + FP64 volume of 1024 clouds x 256 points
uniform in the unit ball (seed 0), inputs resident in HBM. A step is one pass
of the hull kernel over the whole batch.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline_hull(clouds, reps):
    """The CPU oracle (exact hull + FP64 volume, OpenMP over clouds) on the same
    clouds: the CPU baseline, not the product path."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    packed = np.concatenate(clouds, 0)
    off = np.zeros(len(clouds) + 1, np.int32)
    off[1:] = np.cumsum([len(c) for c in clouds])
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    oracle.hull_batch(packed, off, check_general=False)  # warm
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.hull_batch(packed, off, check_general=False)
    dt = time.perf_counter() - t0
    return {"value": reps * len(clouds) / dt, "unit": "clouds/s", "cores": threads, "kind": "port",
            "sample": f"{reps} passes over the full {len(clouds)}-cloud batch ({dt:.2f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this workload runs on one GPU only (asked for {args.gpus}); "
              "per-mesh work shards across GPUs without communication, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import cuacd_amd
    from cuacd_amd import _gpu, datasets

    n_clouds, n_pts = 1024, 256
    clouds = datasets.ball_clouds(n_clouds, n_pts, seed=0)
    packed = np.ascontiguousarray(np.concatenate(clouds, 0), np.float32)
    offsets = (np.arange(n_clouds + 1, dtype=np.int32) * n_pts)
    max_hv, max_ht = n_pts, 2 * n_pts - 4

    ctx = cuacd_amd.Context(device=0)
    name, n_cu, hbm = ctx.device_info()
    log(f"device {name}, {n_cu} CUs, {hbm / 2**30:.0f} GiB")

    def dalloc(nbytes):
        return _gpu.dev_alloc(int(nbytes))

    d_pts = dalloc(packed.nbytes)
    d_off = dalloc(offsets.nbytes)
    d_v = dalloc(n_clouds * max_hv * 12)
    d_t = dalloc(n_clouds * max_ht * 12)
    d_nv = dalloc(n_clouds * 4)
    d_nt = dalloc(n_clouds * 4)
    d_e = dalloc(n_clouds * 4)
    d_vol = dalloc(n_clouds * 8)
    _gpu.memcpy_h2d(d_pts, packed.ctypes.data, packed.nbytes)
    _gpu.memcpy_h2d(d_off, offsets.ctypes.data, offsets.nbytes)

    def step():
        _gpu.hull_dandc_dev(d_pts, d_off, n_clouds, n_pts, max_hv, max_ht, d_v, d_t, d_nv, d_nt, d_e, d_vol)

    for _ in range(args.warmup):
        step()
    _gpu.sync()
    t0 = time.perf_counter()
    _gpu.timer_start()
    for _ in range(args.steps):
        step()
    gpu_ms = _gpu.timer_stop()
    _gpu.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    kernel_ms = gpu_ms / args.steps

    nv = np.empty(n_clouds, np.int32)
    nt = np.empty(n_clouds, np.int32)
    err = np.empty(n_clouds, np.int32)
    vol = np.empty(n_clouds, np.float64)
    _gpu.memcpy_d2h(nv.ctypes.data, d_nv, nv.nbytes)
    _gpu.memcpy_d2h(nt.ctypes.data, d_nt, nt.nbytes)
    _gpu.memcpy_d2h(err.ctypes.data, d_e, err.nbytes)
    _gpu.memcpy_d2h(vol.ctypes.data, d_vol, vol.nbytes)
    if (err != 0).any():
        log("kernel reported errors:", np.unique(err))
        sys.exit(1)
    # algorithmic bytes per launch: read the clouds + offsets, write hull vertex
    # coordinates, triangles, counts/error and the FP64 volume
    alg_bytes = packed.nbytes + offsets.nbytes + int(nv.sum()) * 12 + int(nt.sum()) * 12 + n_clouds * (12 + 8)
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    log(f"kernel {kernel_ms:.4f} ms/launch, wall {ms_per_step:.4f} ms/step, mean hull nv {nv.mean():.1f} "
        f"nt {nt.mean():.1f}, volume {vol.mean():.4f}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_hull(clouds, reps=3)
        log(f"cpu baseline {cpu['value']:.0f} clouds/s on {cpu['cores']} threads")

    value = n_clouds / (ms_per_step * 1e-3)
    line = {
        "metric": "batch_hull_volume clouds/sec (exact hull + FP64 volume)",
        "value": value,
        "unit": "clouds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 grid / fp32 filter / fp64 exact",
        "data": "synthetic",
        "config": {"workload": "batch_hull_volume: 1024 clouds x 256 fp32 points uniform in the unit ball, seed 0",
                   "n_clouds": n_clouds, "n_pts": n_pts},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "hull_wave_kernel<4>", "kernel_ms": kernel_ms, "alg_bytes": alg_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
