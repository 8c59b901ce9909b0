"""Autotuner of the kernel shapes (writes the library's tuning table).

The brick pipeline has shape choices that do not change a bitstream but do change the time of a
batch: the largest tables designed one per thread instead of one per wave, the resident workgroups
of the four LDS table-design classes, whether levels whose rows fit them run the fused x+y
passes, how many z-slices one fused workgroup serves at most, and -- the one choice that does change the
words -- the coding block size.  The reference fixes all of them
(32x4-thread DWT blocks, src/cudaCompress/util/DWTFloat.cu:92-100; coding block 128,
src/cudaCompress/InstanceImpl.cpp:65-69).  This module times candidates on the device with hipGraph
plans of a brick batch and keeps the fastest, by coordinate descent from the current table:

  1. tables designed one per thread: up to 0, 32, 64, 128 symbols;
  2. each design class's grid: x1/2 and x2 of the current value;
  3. fused x+y passes on / off;
  4. z-slices per fused workgroup: all whose x pairs fit its threads, at most 1, 2, 4;
  5. coding block size 32, 64, 128, 256 (a new instance each).

Every candidate must give the words of the starting configuration at the same coding block size
and the same reconstructions (which no coding block size changes); a candidate that does not is
an error, not a slower point.  Step times drift by a few per cent over a run, so a candidate is
timed in alternation with the current best (best, candidate, best, candidate, ...) and replaces it
only when its mean is at least 1 % lower.  Writes tuning/gfx950.txt (read by the library at instance
creation) and tuning/gfx950_sweep.json (every point measured).

    GPU_MAX_HW_QUEUES=16 python -m cudacompress_amd.autotune [--bricks 2048] [--steps 8] [--inflight 3]

(16 hardware queues as bench.py sets them: the plans in flight need their own queues.)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "gfx950.txt")
SWEEP = os.path.join(HERE, "tuning", "gfx950_sweep.json")


class Bench:
    """A brick batch resident on the device and timed through hipGraph plans."""

    def __init__(self, cc, torch, bricks, inflight, steps):
        self.cc, self.torch = cc, torch
        dev = torch.device("cuda:0")
        self.shapes = [b.shape for b in bricks]
        self.steps_q = [0.005 * float(b.std()) for b in bricks]
        self.srcs = [torch.from_numpy(b).to(dev) for b in bricks]
        self.raw = int(sum(b.nbytes for b in bricks))
        nvox = sum(b.size for b in bricks)
        self.outs = [torch.empty(nvox + 4096 * len(bricks), dtype=torch.int32, device=dev) for _ in range(inflight)]
        self.dsts = [[torch.empty_like(s) for s in self.srcs] for _ in range(inflight)]
        self.inflight, self.steps = inflight, steps
        self.ref = {}  # coding block size -> words of the first run at that size
        self.rec = None  # reconstructions of the first run (no coding block size changes them)

    def run(self, tuning):
        cc, torch = self.cc, self.torch
        inst = cc.createInstance(0, 7, 128 ** 3 // 8, tuning["codingBlockSize"], 14)
        inst.setTuning(tuning)
        plans = [cc.BricksPlan(inst, self.srcs, self.dsts[k], self.shapes, self.steps_q, 2, self.outs[k])
                 for k in range(self.inflight)]
        pending = [False] * self.inflight

        def step(i):
            k = i % self.inflight
            if pending[k]:
                plans[k].wait()
            plans[k].launch()
            pending[k] = True

        for i in range(self.inflight):
            step(i)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(self.steps):
            step(i)
        results = [plans[k].wait() if pending[k] else None for k in range(self.inflight)]
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / self.steps
        offs = results[0][0]
        words = self.outs[0][: int(offs[-1])].cpu().numpy().view(np.uint32).copy()
        for p in plans:
            p.close()
        inst.close()
        cbs = tuning["codingBlockSize"]
        if cbs not in self.ref:
            self.ref[cbs] = words
        if self.rec is None:
            self.rec = [d.clone() for d in self.dsts[0]]
        same_words = np.array_equal(words, self.ref[cbs])
        same_rec = all(torch.equal(a, b) for a, b in zip(self.rec, self.dsts[0]))
        if not (same_words and same_rec):
            raise RuntimeError(f"tuning {tuning} changed the bitstream or the reconstruction")
        return ms, {"GBps": round(2 * self.raw / (ms * 1e-3) / 1e9, 2), "compressed_bytes": int(words.size * 4)}


def tune(bench, start, log=print, rounds=2, margin=0.01):
    points = []

    def measure(t):
        ms, extra = bench.run(t)
        points.append({"tuning": json.loads(json.dumps(t)), "ms_per_step": round(ms, 4), **extra})
        log(f"  {t} -> {ms:.3f} ms ({extra['GBps']} GB/s)")
        return ms

    def better(best, cand):
        """Alternating measurements of the current best and a candidate: (candidate wins, means)."""
        a, b = [], []
        for _ in range(rounds):
            a.append(measure(best))
            b.append(measure(cand))
        ma, mb = sum(a) / len(a), sum(b) / len(b)
        return mb < ma * (1.0 - margin), ma, mb

    best = json.loads(json.dumps(start))
    best_ms = None

    def consider(cand):
        nonlocal best, best_ms
        won, ma, mb = better(best, cand)
        if won:
            best, best_ms = cand, mb
        elif best_ms is None or ma < best_ms:
            best_ms = ma

    for k in (0, 32, 64, 128):
        if k != best.get("laneDesignK", 0):
            consider(dict(best, laneDesignK=k))
    for c in range(4):
        for f in (0.5, 2.0):
            t = json.loads(json.dumps(best))
            t["designGrid"][c] = max(64, int(best["designGrid"][c] * f))
            consider(t)
    for fused in (1, 0):
        if fused != best["fusedXY"]:
            consider(dict(best, fusedXY=fused))
    if best["fusedXY"]:
        for k in (1, 2, 4, 0):
            if k != best["xySlicesMax"]:
                consider(dict(best, xySlicesMax=k))
    for cbs in (32, 64, 128, 256):
        if cbs != best["codingBlockSize"]:
            consider(dict(best, codingBlockSize=cbs))
    # the starting point against the choice once more
    if best != start:
        _, best_ms, start_ms = better(best, json.loads(json.dumps(start)))
    else:
        start_ms = best_ms
    return best, best_ms, start_ms, points


def write_table(best, best_ms, start, start_ms, points, workload, path=TABLE, sweep=SWEEP):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    first = points[0]["ms_per_step"]
    lines = [
        "# gfx950 tuning table, written by cudacompress_amd/autotune.py (python -m cudacompress_amd.autotune)",
        f"# workload: {workload}",
        f"# objective: ms per step (candidates timed in alternation with the current best); first point "
        f"{first:.3f} ms, starting point {start_ms:.3f} ms against the chosen {best_ms:.3f} ms at the end",
        "design_grid " + " ".join(str(g) for g in best["designGrid"]),
        f"fused_xy {best['fusedXY']}",
        f"xy_slices_max {best['xySlicesMax']}",
        f"coding_block_size {best['codingBlockSize']}",
        f"lane_design_k {best.get('laneDesignK', 0)}",
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(sweep, "w") as f:
        json.dump({"workload": workload, "start": start, "chosen": best, "chosen_ms_per_step": round(best_ms, 4),
                   "start_ms_again": round(start_ms, 4), "points": points}, f, indent=1)


def main():
    # the plans' streams (1 + 4 each) need their own hardware queues, as in bench.py; set before
    # the HIP runtime starts
    os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
    ap = argparse.ArgumentParser()
    ap.add_argument("--bricks", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--inflight", type=int, default=3)
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--sweep", default=SWEEP)
    args = ap.parse_args()
    import torch

    import cudacompress_amd as cc
    from cudacompress_amd import synthetic as S
    bricks = S.brick_set(args.bricks)
    bench = Bench(cc, torch, bricks, args.inflight, args.steps)
    start = cc.loadTuning()
    print(f"autotune: {len(bricks)} bricks, {args.inflight} plans in flight, {args.steps} steps per point; "
          f"start {start}", flush=True)
    best, best_ms, start_ms, points = tune(bench, start, log=lambda s: print(s, flush=True))
    workload = (f"{len(bricks)} float32 bricks (axes from {{16..128}}, seed 4), 2-level 9/7, step 0.005 x std, "
                f"log2 14; {args.inflight} plans in flight, {args.steps} steps per point")
    write_table(best, best_ms, start, start_ms, points, workload, args.out, args.sweep)
    print(json.dumps({"chosen": best, "ms_per_step": round(best_ms, 4), "start_again_ms": round(start_ms, 4)}))


if __name__ == "__main__":
    sys.exit(main())
