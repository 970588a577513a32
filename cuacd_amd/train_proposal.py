"""Train the learned cut proposal on the GPU and save its weights.

    python -m cuacd_amd.train_proposal [--meshes 1024] [--epochs 70] [--source scorer] [--out PATH]

Training meshes: seeded SDF solids (the benchmark's generator, seed stream
1000+, marching tetrahedra on a 48^3 grid, 2k-50k triangles) - disjoint from
the benchmark's seed-4 set. Triples (--source scorer, the default and the
source of the shipped weights): the search's level-0 candidates of parts
(each mesh and the pieces after 1-3 iterations) scored by the batched exact
scorer. --source search instead records, in plain lookahead runs, every
cutting part of every iteration and each level-0 candidate's path cost (the
value the search minimises; a recording run scores every leaf). 10% of the
parts are held out to report the ranking quality (is the exact best
candidate among the predicted top 8, and the regret).
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "proposal_weights.npz")
SEARCH = dict(width=30, width2=5, threshold=0.05, depth=2, max_n_cutting=16, n_concave_edges=16)


def log(*a):
    print("[train]", *a, file=sys.stderr, flush=True)


def build_dataset(ctx, n_meshes, seed=1000, iters=(0, 1, 2, 3), source="scorer"):
    from . import proposal as PR
    from . import sdf
    t0 = time.perf_counter()
    meshes = sdf.sdf_meshes(n_meshes, seed=seed, res=48, min_tris=2000, max_tris=50000)
    log(f"{len(meshes)} training meshes in {time.perf_counter() - t0:.1f} s")
    if source == "search":
        t0 = time.perf_counter()
        data = PR.record_dataset(ctx, meshes, max_iters=100, merge_hulls=False, **SEARCH)
        log(f"{len(data[0])} cutting parts, {len(data[3])} triples recorded in {time.perf_counter() - t0:.1f} s")
        return data
    t0 = time.perf_counter()
    parts = PR.training_parts(ctx, meshes, iters=iters, **SEARCH)
    log(f"{len(parts)} parts in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    data = PR.make_dataset(ctx, parts, threshold=SEARCH["threshold"], width=SEARCH["width"],
                           n_concave_edges=SEARCH["n_concave_edges"], log=log)
    log(f"{len(data[3])} triples in {time.perf_counter() - t0:.1f} s")
    return data


def split(data, frac=0.1, seed=0):
    pts, pl, po, tg = data
    P = len(pts)
    rng = np.random.default_rng(seed)
    held = np.zeros(P, bool)
    held[rng.choice(P, max(1, int(frac * P)), replace=False)] = True

    def take(mask):
        sel = np.nonzero(mask)[0]
        remap = -np.ones(P, np.int64)
        remap[sel] = np.arange(len(sel))
        keep = mask[po]
        return pts[sel], pl[keep], remap[po[keep]].astype(np.int32), tg[keep]

    return take(~held), take(held)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=70)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch-parts", type=int, default=64)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--source", choices=["search", "scorer"], default="scorer")
    args = ap.parse_args()
    import cuacd_amd
    from . import proposal as PR
    ctx = cuacd_amd.Context(device=0)
    data = build_dataset(ctx, args.meshes, source=args.source)
    train, held = split(data)
    t0 = time.perf_counter()
    net, hist = PR.train(ctx, train, epochs=args.epochs, batch_parts=args.batch_parts, lr=args.lr, log=log)
    log(f"trained in {time.perf_counter() - t0:.1f} s")
    for name, d in (("train", train), ("held-out", held)):
        hit, regret, n = PR.ranking_quality(net, d, top_k=8)
        log(f"{name}: best exact candidate in the predicted top 8 for {100 * hit:.1f}% of {n} parts, "
            f"mean regret {regret:.4f}")
    rng = np.random.default_rng(0)

    class Rand:
        def predict(self, pts, pl, po):
            return rng.uniform(size=len(pl)).astype(np.float32)

    hit, regret, n = PR.ranking_quality(Rand(), held, top_k=8)
    log(f"held-out, random ranking: {100 * hit:.1f}% of {n} parts, mean regret {regret:.4f}")
    net.save(args.out)
    log(f"saved {args.out}")
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
