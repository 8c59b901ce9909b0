"""Benchmark of the wavelet compression path on one MI355X.

Default workload (`bricks`, the batched configuration the headline metric compares with plain
calls): 2048 float32 bricks, each axis drawn from {16,32,48,64,96,128} (seeded), cut from seeded
k^-5/3 Gaussian random fields; 2-level 3D CDF 9/7, deadzone quantization at 0.005 x the brick's
std, RLE+Huffman (log2 14, the coding block size of the library's tuning table).  One step =
batched compression of every brick into one device buffer + batched decompression, inputs and
bitstreams resident in HBM.  The same
bricks are also run through the plain per-brick API (compressVolumeFloat/decompressVolumeFloat,
host bitstreams) on a sample, and the batched bitstreams are checked against the plain ones.
Steps are replays of hipGraph plans; three plans of the same bricks (own outputs, bitstreams and
streams) take the steps in turn, so one step's serial table designs overlap the next step's
streaming passes (`--inflight`; every plan's words and reconstructions are checked equal).

Serving (on by default after the bricks run, alone with `--workload serving`): a 1024^3 k^-5/3
field (seed 5, generated on the device), 729 overlapping 128^3 level-0 bricks (stride 120) plus
729 64^3 bricks of its 2x box-downsampled level of detail, compressed in one batch at step
0.005 x std of each level and kept resident; decode requests of 1-64 bricks (uniform, each brick
level 0 or LOD with probability 1/2) arrive as a Poisson process at 1k-50k bricks/s and are
replayed in real time through the native serving loop (cc_server_run, 32 GB decode-workspace
budget).  Reports p50/p99 request latency (arrival to completion of the request's batch) and the
sustained decoded GB/s per arrival rate.

Other workloads: `lossless2d` (4096^2 uint16 fBm, 3-level 2D CDF 5/3, log2 16, plain API) and
`float3d` (128^3 k^-5/3 field, 2-level 9/7, step 0.01, plain API).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload bricks|lossless2d|float3d]

Prints one JSON line.  value = raw GB/s = (raw bytes encoded + raw bytes decoded) / step time.
roofline: the kernel with the largest total time over the timed steps -- its algorithmic bytes
per launch over its average launch time, from HIP events on its stream (cc_profile_*).
cpu_baseline: the oracle (C++ restatement of the reference path, one host thread) on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

# Every plan runs on its own stream with four side streams for the table-design classes; HIP maps
# streams onto hardware queues (4 by default), and streams sharing a queue serialise.  Sixteen
# queues let the in-flight plans' streams run concurrently.  Set before HIP initialises.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "raw GB/s encode+decode at fixed PSNR (batched vs plain) + p99 brick decode ms"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="bricks", choices=["bricks", "serving", "lossless2d", "float3d"])
    ap.add_argument("--no-serving", action="store_true", help="skip the serving run after the bricks run")
    ap.add_argument("--serving-n", type=int, default=1024, help="side of the serving field")
    ap.add_argument("--serving-slots", type=int, default=16)
    ap.add_argument("--serving-cbs", type=int, default=0,
                    help="coding block size of the served bricks (0: the tuning table's)")
    ap.add_argument("--serving-requests", type=int, default=1000, help="requests replayed per arrival rate")
    ap.add_argument("--serving-rates", default="1000,5000,10000,20000,50000", help="arrival rates, bricks/s")
    ap.add_argument("--bricks", type=int, default=2048)
    ap.add_argument("--groups", type=int, default=0,
                    help="brick groups of a batch: eager batches and plans (a plan runs its groups staggered, "
                         "each captured on its own); 0: library default (one group)")
    ap.add_argument("--inflight", type=int, default=3,
                    help="plan mode: plans (own buffers and streams) whose steps run concurrently; step i "
                         "launches plan i %% inflight after waiting for that plan's previous step")
    ap.add_argument("--tune", default="",
                    help="JSON object of kernel-shape tuning keys applied over the library's table for the brick "
                         "batches (e.g. '{\"laneDesignK\": 128}')")
    ap.add_argument("--single-plan-groups", type=int, default=2,
                    help="after the timed steps, also time one plan alone with its bricks in this many staggered "
                         "groups (0: skip)")
    ap.add_argument("--mode", default="plan", choices=["plan", "eager"],
                    help="plan: compression and decompression replayed from hipGraphs (BricksPlan.launch, "
                         "waited for before the plan's next step); "
                         "eager: compressBricksFloat + decompressBricksFloat calls")
    ap.add_argument("--plain-sample", type=int, default=256, help="bricks timed through the plain per-brick API")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="length of each timed CPU-baseline leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the C1/C2/C3 single-volume records")
    ap.add_argument("--config-reps", type=int, default=3, help="timed round trips per single-volume config")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic.json"))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device-to-device copy first (the reference kernel whose "
                         "FETCH_SIZE/WRITE_SIZE counters calibrate the traffic figures)")
    return ap.parse_args()


def timed(torch, lib, step, steps):
    """Runs `step` `steps` times between device synchronisations with the kernel profiler on.
    Returns (ms per step, per-kernel entries)."""
    from cudacompress_amd import _capi
    lib.cc_profile_reset()
    lib.cc_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lib.cc_profile_enable(0)
    kernels = _capi.profile_entries()
    lib.cc_profile_reset()
    return (t1 - t0) * 1e3 / steps, kernels


def roofline_of(kernels, steps, workload, traffic_file):
    """Roofline of the dominant streaming kernel: the largest total time among kernels that move
    algorithmic bytes.  Kernels with none (the serial table designs) are reported beside it."""
    kernels = sorted(kernels, key=lambda k: -k[2])
    streaming = [k for k in kernels if k[3] > 0]
    dom = streaming[0] if streaming else kernels[0]
    per_launch_ms = dom[2] / dom[1]
    per_launch_bytes = dom[3] / dom[1]
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    traffic = None
    try:
        with open(traffic_file) as f:
            t = json.load(f)
        key = f"{workload}:{dom[0]}"
        if key in t:
            traffic = t[key]["hbm_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        traffic = None
    roof = {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "algorithmic_bytes_per_launch": int(per_launch_bytes), "avg_launch_us": round(per_launch_ms * 1e3, 3),
            "launches_per_step": dom[1] / steps}
    if kernels[0] is not dom:
        roof["dominant_overall"] = {"kernel": kernels[0][0], "ms_per_step": round(kernels[0][2] / steps, 4),
                                    "bound": "latency: serial table construction, one wave per table, "
                                             "no HBM stream"}
    table = [{"kernel": k[0], "launches_per_step": round(k[1] / steps, 2), "ms_per_step": round(k[2] / steps, 4),
              "GBps": round(k[3] / max(k[2], 1e-9) / 1e6, 1)} for k in kernels[:14]]
    return roof, table, sum(k[2] for k in kernels) / steps


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pipeline_bytes(N, M, NH, C, levels, b=4, s=2, dims=3):
    """Algorithmic bytes of one compression + one decompression of the 9/7 pipeline: each stage
    reads its input once and writes its output once.  N coefficients (b bytes), symbols of s bytes,
    M zero-run pairs, NH symbols in Huffman-only streams, C compressed bytes, g = sum over levels
    of 2^(-dims*l) (the DWT touches the leading corner of every level)."""
    g = sum(2.0 ** (-dims * l) for l in range(levels))
    dwt = 2 * b * N * g
    quant = (b + s) * N
    rle = s * (N - NH) + 2 * s * M
    hist = s * (2 * M + NH)
    huff = s * (2 * M + NH) + C
    enc = dwt + quant + rle + hist + huff
    dec = huff + rle + quant + dwt  # Huffman decode, zero-run expansion, dequantize, inverse DWT
    return {"encode": int(enc), "decode": int(dec), "total": int(enc + dec)}


def band_layout(shape, levels, q0):
    """(slices, step) of every quantized band of compressVolumeFloat in bitstream order: the lowpass
    band at q0 / 2^(levels-1), then per level (coarsest first) the 7 highpass bands (band b at
    x = b % 2, y = b / 2 % 2, z = b / 4) at q0 / 2^level."""
    Z, Y, X = shape
    f = 1 << levels
    out = [((slice(0, Z // f), slice(0, Y // f), slice(0, X // f)), float(np.float32(q0) / np.float32(f // 2)))]
    for level in range(levels - 1, -1, -1):
        z, y, x = Z // (f // 2) // 2, Y // (f // 2) // 2, X // (f // 2) // 2
        q = float(np.float32(q0) / np.float32(f // 2))
        for bnd in range(1, 8):
            ox, oy, oz = bnd % 2, bnd // 2 % 2, bnd // 4
            out.append(((slice(oz * z, (oz + 1) * z), slice(oy * y, (oy + 1) * y), slice(ox * x, (ox + 1) * x)), q))
        f //= 2
    return out


def fp64_report(torch, cc, vol, levels, q0):
    """FP32-vs-FP64 error of the multi-level 9/7 analysis on the device: the FP32 levels
    (util.dwtFloat3DForward) against the FP64 twin (util.dwtDouble3DForward), and per band the
    coefficients whose FP64 value lies within 1e-5 (in steps) of a deadzone bin edge |c|/step = k
    >= 1, and the symbols an FP64 quantizer would give differently."""
    c32 = vol.clone()
    c64 = vol.double()
    Z, Y, X = vol.shape
    for lv in range(levels):
        sl = (slice(0, Z >> lv), slice(0, Y >> lv), slice(0, X >> lv))
        src = c32[sl].contiguous()
        dst, b1, b2 = torch.empty_like(src), torch.empty_like(src), torch.empty_like(src)
        cc.util.dwtFloat3DForward(dst, b2, b1, src, X >> lv, Y >> lv, Z >> lv)
        c32[sl] = dst
        src = c64[sl].contiguous()
        dst, buf = torch.empty_like(src), torch.empty_like(src)
        cc.util.dwtDouble3DForward(dst, buf, src, X >> lv, Y >> lv, Z >> lv)
        c64[sl] = dst
    d = c32.double() - c64
    near = differ = 0
    for sl, q in band_layout(vol.shape, levels, q0):
        a64 = c64[sl].abs() / q
        k = torch.round(a64)
        near += int(((k >= 1) & ((a64 - k).abs() < 1e-5)).sum())
        inv = np.float32(1.0) / np.float32(q)   # the quantizer's fp32 product with 1/step, truncated
        q32 = torch.trunc(c32[sl] * float(inv))
        q64 = torch.trunc(c64[sl] / q)
        differ += int((q32.double() != q64).sum())
    torch.cuda.synchronize()
    return {"dwt_max_abs": float(d.abs().max()), "dwt_rel_l2": float(d.norm() / c64.norm()),
            "near_boundary_1e-5": near, "symbols_differ_fp64": differ, "coefficients": int(vol.numel())}


def run_configs(args, torch, cc):
    """The single-volume configurations through the plain per-call API, inputs resident on the
    device: C1 (128^3, 2-level 9/7, step 0.01), C2 (4096^2 lossless 5/3, log2 16) and C3 (3 x 256^3
    velocity, 3-level 9/7, steps {0.1, 0.01, 0.001} x std, log2 16).  Raw GB/s = (raw bytes
    compressed + decompressed) / time of compress + decompress calls (host bitstreams)."""
    from cudacompress_amd import synthetic as S
    dev = torch.device("cuda:0")
    reps = max(1, args.config_reps)

    def timed_pair(compress, decompress):
        w = compress()
        decompress(w)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            w = compress()
            decompress(w)
        torch.cuda.synchronize()
        return w, (time.perf_counter() - t0) / reps

    out = {}
    # C1
    vol = S.gaussian_field((128, 128, 128), 1)
    src = torch.from_numpy(vol).to(dev)
    rec = torch.empty_like(src)
    inst = cc.createInstance(0, 7, 64 ** 3, 128, 14)
    w, t = timed_pair(lambda: cc.compressVolumeFloat(inst, src, 128, 128, 128, 2, 0.01),
                      lambda w: cc.decompressVolumeFloat(inst, rec, 128, 128, 128, 2, w, 0.01))
    out["C1"] = {"workload": "128^3 k^-5/3 field (seed 1), 2-level 9/7, step 0.01, log2 14, plain API",
                 "GBps": round(2 * vol.nbytes / t / 1e9, 3), "ms_per_roundtrip": round(t * 1e3, 3),
                 "compression_ratio": round(vol.nbytes / (4 * w.size), 3),
                 "psnr_db": round(S.psnr(vol, rec.cpu().numpy()), 3),
                 "fp32_vs_fp64": fp64_report(torch, cc, src, 2, 0.01), "input_sha256": S.sha256(vol)[:16]}
    inst.close()
    # C2
    img = S.fbm_heightfield(4096, 0.8, 2)
    src = torch.from_numpy(img).to(dev)
    rec = torch.empty_like(src)
    inst = cc.createInstance(0, 3, 2048 * 2048, 128, 16)
    w, t = timed_pair(lambda: cc.compressVolumeLossless(inst, src, 4096, 4096, 1, 3),
                      lambda w: cc.decompressVolumeLossless(inst, rec, 4096, 4096, 1, 3, w))
    out["C2"] = {"workload": "4096^2 uint16 fBm (H 0.8, seed 2) as int16, 3-level 2D 5/3, log2 16, plain API",
                 "GBps": round(2 * img.nbytes / t / 1e9, 3), "ms_per_roundtrip": round(t * 1e3, 3),
                 "compression_ratio": round(img.nbytes / (4 * w.size), 3),
                 "lossless_exact": bool(torch.equal(rec, src)), "input_sha256": S.sha256(img)[:16]}
    inst.close()
    del src, rec
    # C3
    u = S.velocity_field(256, 3)
    comps = [torch.from_numpy(np.ascontiguousarray(u[c])).to(dev) for c in range(3)]
    recs = [torch.empty_like(c) for c in comps]
    stds = [float(u[c].std()) for c in range(3)]
    lo, hi = float(u.min()), float(u.max())
    inst = cc.createInstance(0, 7, 128 ** 3, 128, 16)
    rows = []
    for rel in (0.1, 0.01, 0.001):
        qs = [float(np.float32(rel * s)) for s in stds]
        words = [None] * 3

        def compress():
            for c in range(3):
                words[c] = cc.compressVolumeFloat(inst, comps[c], 256, 256, 256, 3, qs[c])
            return words

        def decompress(ws):
            for c in range(3):
                cc.decompressVolumeFloat(inst, recs[c], 256, 256, 256, 3, ws[c], qs[c])
        ws, t = timed_pair(compress, decompress)
        se = sum(float(((r.double() - c.double()) ** 2).sum()) for r, c in zip(recs, comps))
        rows.append({"step_x_std": rel, "GBps": round(2 * u.nbytes / t / 1e9, 3),
                     "compression_ratio": round(u.nbytes / (4 * sum(x.size for x in ws)), 3),
                     "psnr_db": round(float(20 * np.log10((hi - lo) / np.sqrt(se / u.size))), 3),
                     "fp32_vs_fp64_u_x": fp64_report(torch, cc, comps[0], 3, qs[0])})
    inst.close()
    out["C3"] = {"workload": "3 x 256^3 divergence-free velocity (seed 3), 3-level 9/7 per component, "
                             "steps {0.1, 0.01, 0.001} x std, log2 16, plain API", "steps": rows}
    del comps, recs
    return out


def cpu_baseline_bricks(bricks, steps_q, seconds):
    """The oracle (C++ restatement of the reference path) on the first bricks of the batch: one
    host thread, then every host core (a thread pool over bricks; the oracle releases the GIL in
    its C calls); and the reference's own src/CPU entropy coder (oracle/_ref) against the oracle's
    coder on the same symbol streams, one thread."""
    import concurrent.futures as cf
    from oracle import pipeline as P
    from oracle import pyoracle as o

    def one(i):
        b = bricks[i]
        ow = P.compress_volume_float(b, 2, steps_q[i])
        P.decompress_volume_float(ow, b.shape, 2, steps_q[i])
        return b.nbytes

    t0 = time.perf_counter()
    runs, done = 0, 0
    while time.perf_counter() - t0 < seconds and runs < len(bricks):
        done += one(runs)
        runs += 1
    t1 = time.perf_counter() - t0
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(cores, 16))
    t0 = time.perf_counter()
    done_p, runs_p = 0, 0
    with cf.ThreadPoolExecutor(cores) as ex:
        futs = []
        i = 0
        while i < len(bricks) and (time.perf_counter() - t0 < seconds):
            batch = [ex.submit(one, j) for j in range(i, min(i + cores, len(bricks)))]
            i += len(batch)
            done_p += sum(f.result() for f in batch)
            runs_p += len(batch)
    tp = time.perf_counter() - t0
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    res = {"value": round(2.0 * done_p / tp / 1e9, 5), "unit": "GB/s", "cores": cores, "kind": "port",
           "cpu_model": model, "host_cpus": os.cpu_count(),
           "sample": f"bricks 0..{runs_p - 1} ({done_p} raw bytes) compressed and decompressed through the oracle "
                     f"(C++ restatement of the reference path), {cores} threads over bricks, {tp:.2f} s",
           "single_thread": {"value": round(2.0 * done / t1 / 1e9, 5), "cores": 1,
                             "sample": f"bricks 0..{runs - 1} ({done} raw bytes) one by one, {t1:.2f} s"}}
    if o.ref_available():
        # the entropy half: the reference's own CPU coder and the oracle's coder on the same streams
        k = max(1, min(runs, 64))
        calls = [P.volume_float_calls(bricks[i], 2, steps_q[i]) for i in range(k)]
        raw = sum(bricks[i].nbytes for i in range(k))
        t0 = time.perf_counter()
        for cl in calls:
            for rl, st in cl:
                w, nb = o.ref_encode(st, rl=rl, cbs=128)
                o.ref_decode(w, nb, len(st), st[0].size, rl=rl, cbs=128)
        tr = time.perf_counter() - t0
        t0 = time.perf_counter()
        for cl in calls:
            for rl, st in cl:
                w, nb = o.encode(st, rl=rl, log2=14, cbs=128)
                o.decode(w, nb, len(st), st[0].size, rl=rl, log2=14, cbs=128)
        to = time.perf_counter() - t0
        res["entropy_coder"] = {
            "reference_GBps": round(2.0 * raw / tr / 1e9, 5), "oracle_GBps": round(2.0 * raw / to / 1e9, 5),
            "kind": "reference", "cores": 1,
            "sample": f"the symbol streams of bricks 0..{k - 1} ({raw} raw bytes) coded and decoded by the "
                      f"reference's src/CPU coder (CPU dialect) and by the oracle's coder (GPU dialect), one thread"}
    return res


# ------------------------------------------------------------------------------------ bricks
def run_bricks(args, torch, cc, lib):
    from cudacompress_amd import synthetic as S
    dev = torch.device("cuda:0")
    bricks = S.brick_set(args.bricks)
    progress(f"bricks: {len(bricks)} generated")
    shapes = [b.shape for b in bricks]
    raw_bytes = int(sum(b.nbytes for b in bricks))
    steps_q = [0.005 * float(b.std()) for b in bricks]
    srcs = [torch.from_numpy(b).to(dev) for b in bricks]
    dsts = [torch.empty_like(s) for s in srcs]
    nvox = sum(b.size for b in bricks)
    out = torch.empty(nvox + 4096 * len(bricks), dtype=torch.int32, device=dev)
    # kernel shapes and coding block size from the library's tuning table (the autotuner's choice)
    tuning = dict(cc.loadTuning(), **(json.loads(args.tune) if args.tune else {}))
    cbs = int(tuning["codingBlockSize"])
    inst = cc.createInstance(0, 7, 128 * 128 * 128 // 8, cbs, 14)
    inst.setTuning(tuning)
    if args.groups:
        inst.setBrickGroups(args.groups)
    state = {}
    plan = cc.BricksPlan(inst, srcs, dsts, shapes, steps_q, 2, out) if args.mode == "plan" else None
    depth = max(1, args.inflight) if plan is not None else 1
    # further plans for concurrent steps: the same bricks into their own outputs and bitstreams
    extra = []
    for _ in range(depth - 1):
        d_k = [torch.empty_like(s) for s in srcs]
        o_k = torch.empty_like(out)
        extra.append((cc.BricksPlan(inst, srcs, d_k, shapes, steps_q, 2, o_k), d_k, o_k))
    plans = [plan] + [e[0] for e in extra] if plan is not None else []
    pending = [False] * depth
    results = [None] * depth
    state["i"] = 0

    def step():
        if plans:
            k = state["i"] % depth
            if pending[k]:
                results[k] = plans[k].wait()
            plans[k].launch()
            pending[k] = True
            state["i"] += 1
        else:
            offs, stat = cc.compressBricksFloat(inst, srcs, shapes, steps_q, 2, out)
            dstat = cc.decompressBricksFloat(inst, dsts, shapes, steps_q, 2, out, offs)
            results[0] = (offs, stat, dstat)

    def drain():
        for k in range(len(plans)):
            if pending[k]:
                results[k] = plans[k].wait()
                pending[k] = False

    if args.copy_calibration:
        a = torch.empty(256 << 20, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    progress("bricks: plans ready, timing")
    ms, kernels = timed(torch, lib, step, args.steps)
    drain()
    progress(f"bricks: {ms:.3f} ms per step")
    value = 2.0 * raw_bytes / (ms * 1e-3) / 1e9
    state["offs"], state["stat"], state["dstat"] = results[0]
    concurrent_equal = None
    if extra:
        # every plan compressed the same bricks: identical words and reconstructions
        n_w = int(results[0][0][-1])
        ran = [k for k in range(1, depth) if results[k] is not None]  # short runs may not reach every plan
        concurrent_equal = all(np.array_equal(results[k][0], results[0][0]) for k in ran) and all(
            torch.equal(extra[k - 1][2][:n_w], out[:n_w]) and all(torch.equal(a, b) for a, b in zip(extra[k - 1][1], dsts))
            for k in ran)
        assert concurrent_equal, "concurrent plans disagree"
    kernel_source = "HIP events around every kernel inside the timed steps"
    if not any(k[0] == "code_pack" for k in kernels):
        # graph replays carry no per-kernel events: time the same kernels in eager calls
        def eager():
            o, _ = cc.compressBricksFloat(inst, srcs, shapes, steps_q, 2, out)
            cc.decompressBricksFloat(inst, dsts, shapes, steps_q, 2, out, o)
        _, kernels = timed(torch, lib, eager, args.steps)
        kernel_source = "HIP events around every kernel of eager batch calls on the same bricks"
    offs, stat = state["offs"], state["stat"]
    assert np.all(stat == 0), "batched compression failed for some bricks"
    assert np.all(state["dstat"] == 0), "batched decompression failed for some bricks"
    compressed = int(offs[-1]) * 4

    # quality: aggregate PSNR (range over all bricks / RMSE over all voxels) and the worst brick
    se, vmin, vmax, worst = 0.0, np.inf, -np.inf, np.inf
    for s, d in zip(srcs, dsts):
        e = (s.double() - d.double())
        se += float((e * e).sum())
        vmin, vmax = min(vmin, float(s.min())), max(vmax, float(s.max()))
        mse_b = float((e * e).mean())
        rng_b = float(s.max() - s.min())
        worst = min(worst, 20 * np.log10(rng_b / np.sqrt(mse_b)) if mse_b > 0 else np.inf)
    psnr = 20 * np.log10((vmax - vmin) / np.sqrt(se / nvox))

    # plain per-brick API on a sample, and bitstream identity with the batch
    plain = None
    equal = None
    if not args.no_plain:
        progress("bricks: plain per-brick sample")
        words = out[: int(offs[-1])].cpu().numpy().view(np.uint32)
        k = min(args.plain_sample, len(bricks))
        pin = cc.createInstance(0, 7, 128 * 128 * 128 // 8, cbs, 14)
        ref = torch.empty_like(srcs[0])
        sample_bytes = sum(bricks[i].nbytes for i in range(k))
        for i in range(min(4, k)):  # warm the plain path
            z, y, x = shapes[i]
            w = cc.compressVolumeFloat(pin, srcs[i], x, y, z, 2, steps_q[i])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        equal = True
        plain_words = []
        for i in range(k):
            z, y, x = shapes[i]
            w = cc.compressVolumeFloat(pin, srcs[i], x, y, z, 2, steps_q[i])
            plain_words.append(w)
            d = torch.empty_like(srcs[i])
            cc.decompressVolumeFloat(pin, d, x, y, z, 2, w, steps_q[i])
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        for i in range(k):
            if not np.array_equal(words[offs[i]:offs[i + 1]], plain_words[i]):
                equal = False
        plain = {"GBps": round(2.0 * sample_bytes / t_plain / 1e9, 4), "bricks": k, "seconds": round(t_plain, 3),
                 "api": "compressVolumeFloat + decompressVolumeFloat per brick (host bitstreams)"}
        pin.close()
        del ref

    single = None
    if plans and (depth > 1 or args.groups <= 1) and args.single_plan_groups > 0:
        # one plan alone (one copy of every buffer), its bricks in staggered groups
        progress(f"bricks: one plan, {args.single_plan_groups} groups")
        for p in plans[1:]:
            p.close()
        inst1 = cc.createInstance(0, 7, 128 * 128 * 128 // 8, cbs, 14)
        inst1.setTuning(tuning)
        inst1.setBrickGroups(args.single_plan_groups)
        d1 = [torch.empty_like(s) for s in srcs]
        o1 = torch.empty_like(out)
        p1 = cc.BricksPlan(inst1, srcs, d1, shapes, steps_q, 2, o1)
        for _ in range(args.warmup):
            p1.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r1 = p1.run()
        torch.cuda.synchronize()
        ms1 = (time.perf_counter() - t0) * 1e3 / args.steps
        n_w = int(state["offs"][-1])
        same = bool(np.array_equal(r1[0], state["offs"]) and torch.equal(o1[:n_w], out[:n_w])
                    and all(torch.equal(a, b) for a, b in zip(d1, dsts)))
        assert same, "the single grouped plan disagrees with the batch"
        single = {"GBps": round(2.0 * raw_bytes / (ms1 * 1e-3) / 1e9, 3), "ms_per_step": round(ms1, 4),
                  "plans_in_flight": 1, "brick_groups": args.single_plan_groups,
                  "device_buffers": "one copy of the outputs, bitstreams and work buffers (+ group bitstream staging)",
                  "words_and_reconstructions_equal_batch": same}
        progress(f"bricks: one plan {ms1:.3f} ms per step")
        p1.close()
        inst1.close()
        del d1, o1
        plans = plans[:1]

    cpu = None
    if not args.no_cpu_baseline:
        progress("bricks: CPU baseline legs")
        cpu = cpu_baseline_bricks(bricks, steps_q, args.cpu_seconds)

    coding = None
    if plans:
        coding = plans[0].codingStats()
        coding["pipeline_bytes_per_step"] = pipeline_bytes(coding["N"], coding["M"], coding["N_H"], coding["C"], 2)
    for p in plans:
        p.close()
    del extra
    inst.close()
    cfg = {"workload": "C4 batched: %d float32 bricks, axes from {16,32,48,64,96,128} (seeded), cut from seeded "
                       "k^-5/3 fields; 2-level 3D CDF 9/7, deadzone step 0.005*std(brick), RLE+Huffman log2 14, "
                       "coding block %d (the tuning table's); one batched compress + one batched decompress per "
                       "step" % (len(bricks), cbs),
           "bricks": len(bricks), "raw_bytes_per_direction": raw_bytes, "compressed_bytes": compressed,
           "compression_ratio": round(raw_bytes / compressed, 3), "psnr_db": round(float(psnr), 3),
           "worst_brick_psnr_db": round(float(worst), 3),
           "mode": "batched, " + (("hipGraph plan replays, %d plan(s) in flight (each step one full compression + "
                                   "decompression of the batch)" % depth) if args.mode == "plan"
                                  else "eager batch calls"),
           "brick_groups": max(1, args.groups), "tuning": tuning,
           "device_buffers": "%d copies of the outputs, bitstreams and work buffers" % depth,
           "concurrent_plans_equal": concurrent_equal,
           "kernel_timing": kernel_source,
           "single_plan": single,
           "plain": plain, "batched_vs_plain": round(value / plain["GBps"], 2) if plain else None,
           "bitstreams_equal_plain_sample": equal, "coding": coding}
    return value, ms, raw_bytes, kernels, cfg, cpu, "f32"


# ------------------------------------------------------------------------------------ serving
def run_serving(args, torch, cc, lib):
    from cudacompress_amd import synthetic as S
    dev = torch.device("cuda:0")
    t_setup = time.perf_counter()
    field = S.gaussian_field_device(args.serving_n, 5, dev)
    l0, lod, lod_field = S.serving_bricks(field, 128, 8)
    std0, std1 = float(field.std()), float(lod_field.std())
    del field, lod_field
    bricks = l0 + lod
    n0 = len(l0)
    shapes = [tuple(b.shape) for b in bricks]
    steps = [0.005 * std0] * n0 + [0.005 * std1] * len(lod)
    nvox = sum(b.numel() for b in bricks)
    cbs = args.serving_cbs or int(cc.loadTuning()["codingBlockSize"])
    inst = cc.createInstance(0, 7, 128 * 128 * 128 // 8, cbs, 14)
    out = torch.empty(nvox, dtype=torch.int32, device=dev)
    # compressed in batches of 243 bricks (at most ~5.1e8 voxels per batch) into one word buffer
    offs = np.zeros(len(bricks) + 1, dtype=np.uint64)
    dec = [torch.empty_like(b) for b in bricks]
    for lo in range(0, len(bricks), 243):
        hi = min(lo + 243, len(bricks))
        w0 = int(offs[lo])
        o, stat = cc.compressBricksFloat(inst, bricks[lo:hi], shapes[lo:hi], steps[lo:hi], 2, out[w0:])
        assert np.all(stat == 0), "serving bricks failed to compress"
        # quality of every brick through the batched decode
        cc.decompressBricksFloat(inst, dec[lo:hi], shapes[lo:hi], steps[lo:hi], 2, out[w0:], o)
        offs[lo + 1:hi + 1] = w0 + o[1:]
    torch.cuda.synchronize()
    compressed = int(offs[-1]) * 4
    psnr = []
    for lo, hi in ((0, n0), (n0, len(bricks))):
        se, vmin, vmax, cnt = 0.0, np.inf, -np.inf, 0
        for b, d in zip(bricks[lo:hi], dec[lo:hi]):
            e = b.double() - d.double()
            se += float((e * e).sum())
            vmin, vmax, cnt = min(vmin, float(b.min())), max(vmax, float(b.max())), cnt + b.numel()
        psnr.append(round(float(20 * np.log10((vmax - vmin) / np.sqrt(se / cnt))), 3))
    del dec, bricks, l0, lod
    ranges = np.stack([offs[:-1], offs[1:]], axis=1)
    pool_n = 2048
    pool = torch.empty((pool_n, 128 ** 3), dtype=torch.float32, device=dev)
    base = pool.data_ptr()
    stride = 128 ** 3 * 4
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    srv = cc.Server(inst, out, ranges, shapes, steps, 2, slots=args.serving_slots, workspaceBytes=32 << 30)
    brick_bytes = np.array([4 * int(np.prod(sh)) for sh in shapes], dtype=np.int64)

    def trace(rng, lam, nreq):
        sizes = rng.integers(1, 65, nreq)
        gaps = rng.exponential(sizes.mean() / lam, nreq)  # request rate lam / E[size]
        arrivals = np.cumsum(gaps) - gaps[0]
        reqs, outs, k = [], [], 0
        for c in sizes:
            ids = rng.integers(0, n0, c) + n0 * rng.integers(0, 2, c)
            reqs.append(ids.astype(np.uint32))
            outs.append([base + ((k + j) % pool_n) * stride for j in range(c)])
            k += c
        return arrivals, reqs, outs

    # warm-up: the largest requests on every slot grow the slots' buffers once
    big = [np.arange(64, dtype=np.uint32) % n0 for _ in range(2 * args.serving_slots)]
    srv.run(np.zeros(len(big)), big, [[base + j * stride for j in range(64)] for _ in big])
    rows = []
    for lam in [int(x) for x in args.serving_rates.split(",") if x]:
        rng = np.random.default_rng(5000 + lam)
        nreq = args.serving_requests
        arrivals, reqs, outs = trace(rng, lam, nreq)
        lat, elapsed = srv.run(arrivals, reqs, outs)
        ok = lat >= 0
        dec_bytes = int(sum(brick_bytes[r].sum() for r in reqs))
        ms = np.sort(lat[ok] * 1e3)
        rows.append({"rate_bricks_per_s": lam, "requests": nreq, "bricks": int(sum(len(r) for r in reqs)),
                     "completed": int(ok.sum()), "failed": int((~ok).sum()),
                     "p50_ms": round(float(np.percentile(ms, 50)), 3),
                     "p99_ms": round(float(np.percentile(ms, 99)), 3),
                     "p99_9_ms": round(float(np.percentile(ms, 99.9)), 3), "max_ms": round(float(ms[-1]), 3),
                     "offered_GBps": round(dec_bytes / max(arrivals[-1], 1e-9) / 1e9, 2),
                     "decoded_GBps": round(dec_bytes / elapsed / 1e9, 2), "seconds": round(elapsed, 3)})
        print(f"serving rate {lam}: p50 {rows[-1]['p50_ms']} ms p99 {rows[-1]['p99_ms']} ms "
              f"{rows[-1]['decoded_GBps']} GB/s", file=sys.stderr, flush=True)
    srv.close()
    inst.close()
    del pool, out
    return {"workload": "C5 serving: %d^3 k^-5/3 field (seed 5), %d level-0 128^3 bricks (stride 120, last origin "
                        "clamped) + %d 64^3 LOD bricks of the 2x box-downsampled field, 2-level 9/7, step "
                        "0.005*std per level, resident in HBM; Poisson arrivals, requests of U{1..64} bricks, "
                        "each brick LOD with p=1/2; coding block %d; %d slots, 32 GB decode-workspace budget"
                        % (args.serving_n, n0, n0, cbs, args.serving_slots),
            "latency": "request arrival to completion of the batch holding its last brick, observed by the host loop",
            "admission": "brick by brick in arrival order, batches up to the slot's share of the workspace budget",
            "compressed_bytes": compressed, "raw_bytes": 4 * nvox, "compression_ratio": round(4 * nvox / compressed, 3),
            "psnr_db_level0": psnr[0], "psnr_db_lod": psnr[1], "setup_seconds": round(setup_s, 2), "rates": rows,
            "p99_ms_max_over_rates": max(r["p99_ms"] for r in rows)}


# ------------------------------------------------------------------------------------ plain volumes
def run_plain(args, torch, cc, lib):
    from cudacompress_amd import synthetic as S
    dev = torch.device("cuda:0")
    if args.workload == "lossless2d":
        n = 4096
        data = S.fbm_heightfield(n, 0.8, 2)
        src = torch.from_numpy(data).to(dev)
        out = torch.empty_like(src)
        inst = cc.createInstance(0, 3, (n // 2) * (n // 2), 128, 16)

        def compress():
            return cc.compressVolumeLossless(inst, src, n, n, 1, 3)

        def decompress(w):
            cc.decompressVolumeLossless(inst, out, n, n, 1, 3, w)
        desc = ("C2 lossless: 4096x4096 uint16 spectral fBm (H=0.8, seed 2) as int16, 3-level 2D CDF 5/3, "
                "3 RLHuff highpass streams per level + Huffman lowpass, log2 16, coding block 128, plain per-call API")
    else:
        n = 128
        data = S.gaussian_field((n, n, n), 1)
        src = torch.from_numpy(data).to(dev)
        out = torch.empty_like(src)
        inst = cc.createInstance(0, 7, (n // 2) ** 3, 128, 14)

        def compress():
            return cc.compressVolumeFloat(inst, src, n, n, n, 2, 0.01)

        def decompress(w):
            cc.decompressVolumeFloat(inst, out, n, n, n, 2, w, 0.01)
        desc = ("C1 float: 128^3 k^-5/3 Gaussian random field (seed 1), 2-level 3D CDF 9/7, deadzone step 0.01, "
                "7 RLHuff streams per level + Huffman lowpass, log2 14, coding block 128, plain per-call API")
    raw_bytes = data.nbytes
    state = {}

    def step():
        w = compress()
        decompress(w)
        state["w"] = w

    for _ in range(args.warmup):
        step()
    ms, kernels = timed(torch, lib, step, args.steps)
    value = 2.0 * raw_bytes / (ms * 1e-3) / 1e9
    words = state["w"]
    rec = out.cpu().numpy()
    cfg = {"workload": desc, "raw_bytes_per_direction": raw_bytes, "compressed_bytes": int(words.size * 4),
           "compression_ratio": round(raw_bytes / (words.size * 4), 3), "mode": "plain"}
    if args.workload == "lossless2d":
        cfg["lossless_exact"] = bool(np.array_equal(rec, data))
    else:
        cfg["psnr_db"] = round(S.psnr(data, rec), 3)
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pipeline as P
        runs, t_total = 0, 0.0
        while runs == 0 or (t_total < 10.0 and runs < 20):
            t0 = time.perf_counter()
            if args.workload == "lossless2d":
                ow = P.compress_lossless(data, 3, log2=16, cbs=128)
                P.decompress_lossless(ow, data.shape, 3, log2=16, cbs=128)
            else:
                ow = P.compress_volume_float(data, 2, 0.01)
                P.decompress_volume_float(ow, data.shape, 2, 0.01)
            t_total += time.perf_counter() - t0
            runs += 1
        cfg["bitstream_equals_oracle"] = bool(np.array_equal(ow, words))
        cpu = {"value": round(2.0 * raw_bytes * runs / t_total / 1e9, 5), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"{runs} full compress+decompress runs of the same input through the oracle "
                         f"(C++ restatement, one thread), {t_total / runs:.3f} s per run"}
    inst.close()
    return value, ms, raw_bytes, kernels, cfg, cpu, "int16" if args.workload == "lossless2d" else "f32"


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (requested --gpus {args.gpus}); "
              "bricks would shard by brick id with no exchange on the data path", file=sys.stderr)
        sys.exit(2)
    import torch

    import cudacompress_amd as cc
    from cudacompress_amd import _capi
    lib = _capi.load()
    if args.workload == "serving":
        srv = run_serving(args, torch, cc, lib)
        top = max(srv["rates"], key=lambda r: r["decoded_GBps"])
        res = {"metric": METRIC, "value": top["decoded_GBps"], "unit": "GB/s", "n_gpus": 1, "steps": len(srv["rates"]),
               "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f32", "data": "synthetic", "config": srv, "roofline": None, "cpu_baseline": None}
        print(json.dumps(res))
        return
    if args.workload == "bricks":
        value, ms, raw, kernels, cfg, cpu, dtype = run_bricks(args, torch, cc, lib)
        if not args.no_serving:
            cfg["serving"] = run_serving(args, torch, cc, lib)
    else:
        value, ms, raw, kernels, cfg, cpu, dtype = run_plain(args, torch, cc, lib)
    roof, table, gpu_ms = roofline_of(kernels, args.steps, args.workload, args.traffic_file)
    if roof["kernel"] == "code_pack" and "compressed_bytes" in cfg:
        # the launches also write the codewords (their share of the step's compressed bytes), whose
        # size is known only after they ran
        per = roof["algorithmic_bytes_per_launch"] + cfg["compressed_bytes"] / roof["launches_per_step"]
        roof["algorithmic_bytes_per_launch"] = int(per)
        roof["achieved"] = round(per / (roof["avg_launch_us"] * 1e-6) / 1e9, 2)
        roof["frac"] = round(roof["achieved"] / HBM_PEAK_GBS, 5)
    if roof["kernel"] == "dec_blocks" and cfg.get("coding"):
        # the block decode reads the codewords (and block offsets) and writes every decoded pair of
        # the RLE sequences (symbols and zero counts) and the plain sequences' symbols, 16-bit each:
        # C + 2 (2M + N_H) per step, over the time of both of its launches (long and short sequences)
        cd = cfg["coding"]
        per = (cd["C"] + 2 * (2 * cd["M"] + cd["N_H"])) / roof["launches_per_step"]
        short_ms = sum(k[2] for k in kernels if k[0] == "dec_blocks_short")
        t_us = roof["avg_launch_us"] + short_ms * 1e3 / (roof["launches_per_step"] * args.steps)
        roof["algorithmic_bytes_per_launch"] = int(per)
        roof["avg_launch_us_with_short"] = round(t_us, 3)
        roof["achieved"] = round(per / (t_us * 1e-6) / 1e9, 2)
        roof["frac"] = round(roof["achieved"] / HBM_PEAK_GBS, 5)
    cfg["gpu_kernel_ms_per_step"] = round(gpu_ms, 4)
    if cfg.get("coding"):
        pb = cfg["coding"]["pipeline_bytes_per_step"]["total"]
        ach = pb / (ms * 1e-3) / 1e9
        roof["pipeline"] = {"algorithmic_bytes_per_step": pb, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
                            "formula": "per stage one read of its input + one write of its output: DWT 2bNg, "
                                       "quantize (b+s)N, RLE s(N-N_H)+2sM, histogram s(2M+N_H), Huffman "
                                       "s(2M+N_H)+C; decode mirrors all but the histogram (b=4, s=2)"}
    if args.workload == "bricks" and not args.no_configs:
        progress("configs C1-C3")
        cfg["configs"] = run_configs(args, torch, cc)
    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": cfg, "roofline": roof,
        "cpu_baseline": cpu, "kernels": table,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
