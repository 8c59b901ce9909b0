"""Benchmark of the wavelet compression path on one MI355X.

Default workload (the lossless integer configuration): a 4096x4096 uint16 heightfield of spectral
fractional Brownian motion (Hurst 0.8, seed 2) stored as int16 = u - 32768, compressed with a
3-level 2D CDF 5/3 DWT, three RLE+Huffman highpass streams per level and a Huffman lowpass stream
(log2HuffmanDistinctSymbolCountMax 16, coding block 128), then decompressed.  One step = one
compress + one decompress with the input already resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lossless2d|float3d]

Prints one JSON line.  value = raw GB/s = (raw bytes encoded + raw bytes decoded) / step time.
roofline: the kernel with the largest total time over the timed steps, its algorithmic bytes per
launch divided by its average launch time (hipEvents on its stream, via cc_profile_*).
cpu_baseline: the oracle's restatement of the same pipeline (oracle/pipeline.py over the C++
oracle), one host thread, timed on the same input.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "raw GB/s encode+decode at fixed PSNR (batched vs plain) + p99 brick decode ms"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="lossless2d", choices=["lossless2d", "float3d"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic.json"))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device-to-device copy first (the reference kernel whose "
                         "FETCH_SIZE/WRITE_SIZE counters calibrate the traffic figures)")
    return ap.parse_args()


def lossless2d():
    from cudacompress_amd import synthetic as S
    n = 4096
    img = S.fbm_heightfield(n, 0.8, 2)
    return dict(name="lossless2d", data=img, shape=(n, n), levels=3, log2=16, cbs=128,
                desc=("C2 lossless: 4096x4096 uint16 spectral fBm (H=0.8, seed 2) as int16, 3-level 2D CDF 5/3, "
                      "3 RLHuff highpass streams per level + Huffman lowpass, log2 16, coding block 128, "
                      "plain per-call API"))


def float3d():
    from cudacompress_amd import synthetic as S
    n = 128
    vol = S.gaussian_field((n, n, n), 1)
    return dict(name="float3d", data=vol, shape=(n, n, n), levels=2, log2=14, cbs=128, q0=0.01,
                desc=("C1 float: 128^3 k^-5/3 Gaussian random field (seed 1), 2-level 3D CDF 9/7, deadzone step "
                      "0.01, 7 RLHuff streams per level + Huffman lowpass, log2 14, coding block 128"))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (requested --gpus {args.gpus}); "
              "see DESIGN.md for how bricks would shard", file=sys.stderr)
        sys.exit(2)

    import torch

    import cudacompress_amd as cc
    from cudacompress_amd import _capi
    from cudacompress_amd import synthetic as S

    wl = lossless2d() if args.workload == "lossless2d" else float3d()
    data = wl["data"]
    raw_bytes = data.nbytes
    dev = torch.device("cuda:0")
    src = torch.from_numpy(data).to(dev)
    out = torch.empty_like(src)
    if wl["name"] == "lossless2d":
        y, x = wl["shape"]
        inst = cc.createInstance(0, 3, (x // 2) * (y // 2), wl["cbs"], wl["log2"])

        def compress():
            return cc.compressVolumeLossless(inst, src, x, y, 1, wl["levels"])

        def decompress(words):
            cc.decompressVolumeLossless(inst, out, x, y, 1, wl["levels"], words)
    else:
        z, y, x = wl["shape"]
        inst = cc.createInstance(0, 7, (x // 2) * (y // 2) * (z // 2), wl["cbs"], wl["log2"])

        def compress():
            return cc.compressVolumeFloat(inst, src, x, y, z, wl["levels"], wl["q0"])

        def decompress(words):
            cc.decompressVolumeFloat(inst, out, x, y, z, wl["levels"], words, wl["q0"])

    def step():
        w = compress()
        decompress(w)
        return w

    if args.copy_calibration:
        a = torch.empty(256 << 20, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    for _ in range(args.warmup):
        words = step()
    torch.cuda.synchronize()

    lib = _capi.load()
    lib.cc_profile_reset()
    lib.cc_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        words = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lib.cc_profile_enable(0)
    kernels = _capi.profile_entries()
    lib.cc_profile_reset()

    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = 2.0 * raw_bytes / (ms_per_step * 1e-3) / 1e9

    # exactness / quality of the last step
    rec = out.cpu().numpy()
    if wl["name"] == "lossless2d":
        quality = {"lossless_exact": bool(np.array_equal(rec, data)), "psnr_db": "inf" if np.array_equal(rec, data) else S.psnr(data, rec)}
    else:
        quality = {"psnr_db": round(S.psnr(data, rec), 3)}
    compressed_bytes = int(words.size * 4)

    kernels.sort(key=lambda k: -k[2])
    dom = kernels[0]
    per_launch_ms = dom[2] / dom[1]
    per_launch_bytes = dom[3] / dom[1]
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    traffic = None
    try:
        with open(args.traffic_file) as f:
            t = json.load(f)
        key = f"{wl['name']}:{dom[0]}"
        if key in t:
            traffic = t[key]["hbm_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        traffic = None
    roofline = {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "algorithmic_bytes_per_launch": int(per_launch_bytes), "avg_launch_us": round(per_launch_ms * 1e3, 3),
                "launches_per_step": dom[1] / args.steps}
    kernel_table = [{"kernel": k[0], "launches_per_step": round(k[1] / args.steps, 2),
                     "ms_per_step": round(k[2] / args.steps, 4), "GBps": round(k[3] / max(k[2], 1e-9) / 1e6, 1)}
                    for k in kernels[:12]]
    gpu_kernel_ms = sum(k[2] for k in kernels) / args.steps

    cpu = None
    bitexact = None
    if not args.no_cpu_baseline:
        from oracle import pipeline as P
        runs, t_total = 0, 0.0
        while runs == 0 or (t_total < 10.0 and runs < 20):  # bounded sample: ~10 s of CPU work
            t0 = time.perf_counter()
            if wl["name"] == "lossless2d":
                ow = P.compress_lossless(data, wl["levels"], log2=wl["log2"], cbs=wl["cbs"])
                P.decompress_lossless(ow, data.shape, wl["levels"], log2=wl["log2"], cbs=wl["cbs"])
            else:
                ow = P.compress_volume_float(data, wl["levels"], wl["q0"], log2=wl["log2"], cbs=wl["cbs"])
                P.decompress_volume_float(ow, data.shape, wl["levels"], wl["q0"], log2=wl["log2"], cbs=wl["cbs"])
            t_total += time.perf_counter() - t0
            runs += 1
        tc = t_total / runs
        bitexact = bool(np.array_equal(ow, words))
        cpu = {"value": round(2.0 * raw_bytes / tc / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"{runs} full compress+decompress runs of the same {wl['name']} input ({raw_bytes} raw "
                         f"bytes each way) through the oracle (C++ restatement, one thread), {tc:.3f} s per run"}

    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int16" if wl["name"] == "lossless2d" else "f32", "data": "synthetic",
        "config": {"workload": wl["desc"], "raw_bytes_per_direction": raw_bytes, "compressed_bytes": compressed_bytes,
                   "compression_ratio": round(raw_bytes / compressed_bytes, 3), "mode": "plain",
                   "bitstream_equals_oracle": bitexact, **quality, "gpu_kernel_ms_per_step": round(gpu_kernel_ms, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernel_table,
    }
    print(json.dumps(res))
    inst.close()


if __name__ == "__main__":
    main()
