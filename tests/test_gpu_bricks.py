"""GPU parity of the brick-batched float codec: every brick's bitstream equals the single-volume
compressVolumeFloat bitstream of that brick (itself equal to the oracle's), and batched
decompression equals per-brick decompressVolumeFloat bit for bit."""
import numpy as np
import pytest

from oracle import pipeline as P
from tests.gpu_util import log

pytestmark = pytest.mark.gpu

AXES = (16, 32, 48, 64)


def make_bricks(rng, field, count, axes=AXES):
    bricks = []
    for _ in range(count):
        z, y, x = (int(rng.choice(axes)) for _ in range(3))
        oz, oy, ox = (int(rng.integers(0, field.shape[k] - s + 1)) for k, s in enumerate((z, y, x)))
        bricks.append(np.ascontiguousarray(field[oz:oz + z, oy:oy + y, ox:ox + x]))
    return bricks


def run_batch(cc, torch, inst, bricks, steps, levels, sym=2):
    srcs = [torch.from_numpy(b).cuda() for b in bricks]
    shapes = [b.shape for b in bricks]
    total = sum(b.size for b in bricks)
    out = torch.zeros(total + 4096 * len(bricks), dtype=torch.int32, device="cuda")
    offs, stat = cc.compressBricksFloat(inst, srcs, shapes, steps, levels, out, symbolSize=sym)
    torch.cuda.synchronize()
    return srcs, out, offs, stat


@pytest.fixture(scope="module")
def field():
    from cudacompress_amd import synthetic as S
    return S.gaussian_field((96, 96, 96), 4)


# (levels, brick axes): the last case has odd half-sizes on level 1 (x pairs 5, 7, 9, 11)
@pytest.mark.parametrize("levels,axes", [(1, AXES), (2, AXES), (3, (16, 32, 48)), (2, (20, 28, 36, 44))])
def test_bricks_equal_single_volume_calls(cc, torch_gpu, field, levels, axes):
    log("bricks levels", levels, "axes", axes)
    torch = torch_gpu
    rng = np.random.default_rng(levels + len(axes) * (axes[0] % 16))
    bricks = make_bricks(rng, field, 24, axes)
    std = float(field.std())
    steps = [float(rng.choice([0.005, 0.05, 0.5])) * std for _ in bricks]
    inst = cc.createInstance(0, 7, 64 * 64 * 64, 128, 14)
    srcs, out, offs, stat = run_batch(cc, torch, inst, bricks, steps, levels)
    assert np.all(stat == 0), stat
    words = out.cpu().numpy().view(np.uint32)
    dsts = [torch.zeros_like(s) for s in srcs]
    cc.decompressBricksFloat(inst, dsts, [b.shape for b in bricks], steps, levels, out, offs)
    torch.cuda.synchronize()
    for i, b in enumerate(bricks):
        z, y, x = b.shape
        want = cc.compressVolumeFloat(inst, srcs[i], x, y, z, levels, steps[i])
        got = words[offs[i]:offs[i + 1]]
        assert got.size == want.size and np.array_equal(got, want), f"brick {i} {b.shape}"
        ref = torch.zeros_like(srcs[i])
        cc.decompressVolumeFloat(inst, ref, x, y, z, levels, want, steps[i])
        torch.cuda.synchronize()
        assert np.array_equal(dsts[i].cpu().numpy(), ref.cpu().numpy()), f"brick {i} reconstruction"
    # two bricks straight against the oracle
    for i in (0, len(bricks) - 1):
        want = P.compress_volume_float(bricks[i], levels, steps[i])
        assert np.array_equal(words[offs[i]:offs[i + 1]], want)
    inst.close()


def test_bricks_wide_slices_take_separate_passes(cc, torch_gpu):
    """A level whose bricks include a z-slice too large for the fused x+y passes' LDS (96 x 192)
    runs the separate x and y passes; the next level (48 x 96) and a batch whose largest slice is
    exactly the fused budget (128 x 128) run fused.  Every brick equals its single-volume call."""
    from cudacompress_amd import synthetic as S
    log("bricks wide slices")
    torch = torch_gpu
    field = S.gaussian_field((16, 192, 192), 11)
    std = float(field.std())
    cases = [[(8, 96, 192), (16, 32, 48), (8, 16, 16)], [(8, 128, 128), (16, 48, 96), (8, 16, 32)]]
    inst = cc.createInstance(0, 7, 64 * 64 * 64, 128, 14)
    for shapes in cases:
        bricks = [np.ascontiguousarray(field[:z, :y, :x]) for z, y, x in shapes]
        steps = [0.01 * std, 0.05 * std, 0.2 * std]
        srcs, out, offs, stat = run_batch(cc, torch, inst, bricks, steps, 2)
        assert np.all(stat == 0), stat
        words = out.cpu().numpy().view(np.uint32)
        dsts = [torch.zeros_like(s) for s in srcs]
        cc.decompressBricksFloat(inst, dsts, [b.shape for b in bricks], steps, 2, out, offs)
        torch.cuda.synchronize()
        for i, b in enumerate(bricks):
            z, y, x = b.shape
            want = cc.compressVolumeFloat(inst, srcs[i], x, y, z, 2, steps[i])
            assert np.array_equal(words[offs[i]:offs[i + 1]], want), f"brick {b.shape}"
            ref = torch.zeros_like(srcs[i])
            cc.decompressVolumeFloat(inst, ref, x, y, z, 2, want, steps[i])
            torch.cuda.synchronize()
            assert np.array_equal(dsts[i].cpu().numpy(), ref.cpu().numpy()), f"brick {b.shape} reconstruction"
    inst.close()


def test_bricks_symbol32_and_overflow(cc, torch_gpu, field):
    """Symbol32 instance (log2 18) and, on a log2 14 instance, one brick whose symbols overflow:
    only that brick fails."""
    log("bricks symbol32 / overflow")
    torch = torch_gpu
    rng = np.random.default_rng(9)
    bricks = make_bricks(rng, field, 6)
    std = float(field.std())
    steps = [0.0005 * std] * len(bricks)
    inst = cc.createInstance(0, 7, 64 * 64 * 64, 128, 18)
    srcs, out, offs, stat = run_batch(cc, torch, inst, bricks, steps, 2, sym=4)
    assert np.all(stat == 0)
    words = out.cpu().numpy().view(np.uint32)
    for i, b in enumerate(bricks):
        z, y, x = b.shape
        want = cc.compressVolumeFloat(inst, srcs[i], x, y, z, 2, steps[i])
        assert np.array_equal(words[offs[i]:offs[i + 1]], want)
    inst.close()
    inst = cc.createInstance(0, 7, 64 * 64 * 64, 128, 14)
    steps = [0.01 * std] * len(bricks)
    steps[2] = 1e-6 * std
    srcs, out, offs, stat = run_batch(cc, torch, inst, bricks, steps, 2)
    assert stat[2] == cc.CC_ERR_SYMBOL_OVERFLOW
    assert np.all(np.delete(stat, 2) == 0)
    words = out.cpu().numpy().view(np.uint32)
    z, y, x = bricks[0].shape
    want = cc.compressVolumeFloat(inst, srcs[0], x, y, z, 2, steps[0])
    assert np.array_equal(words[offs[0]:offs[1]], want)
    inst.close()


def test_bricks_plan_graph_replays(cc, torch_gpu, field):
    """A plan (compression and decompression captured in hipGraphs) gives the eager batch's
    bitstreams and reconstructions, replay after replay, also after the source data changes."""
    log("bricks plan")
    torch = torch_gpu
    rng = np.random.default_rng(21)
    bricks = make_bricks(rng, field, 20)
    std = float(field.std())
    steps = [0.005 * std] * len(bricks)
    shapes = [b.shape for b in bricks]
    inst = cc.createInstance(0, 7, 64 * 64 * 64, 128, 14)
    srcs = [torch.from_numpy(b).cuda() for b in bricks]
    dsts = [torch.zeros_like(s) for s in srcs]
    total = sum(b.size for b in bricks)
    out = torch.zeros(total + 4096 * len(bricks), dtype=torch.int32, device="cuda")
    plan = cc.BricksPlan(inst, srcs, dsts, shapes, steps, 2, out)
    for rep in range(3):
        if rep == 2:  # new data of the same geometry
            for s in srcs:
                s.mul_(1.0 + 0.01 * rep).add_(0.001)
        offs, stat = plan.compress()
        assert np.all(stat == 0)
        pw = out[: int(offs[-1])].cpu().numpy().view(np.uint32).copy()
        assert np.all(plan.decompress() == 0)
        torch.cuda.synchronize()
        prec = [d.cpu().numpy().copy() for d in dsts]
        # eager batch on the same data
        out2 = torch.zeros_like(out)
        offs2, stat2 = cc.compressBricksFloat(inst, srcs, shapes, steps, 2, out2)
        assert np.array_equal(offs, offs2), rep
        assert np.array_equal(pw, out2[: int(offs2[-1])].cpu().numpy().view(np.uint32)), rep
        d2 = [torch.zeros_like(s) for s in srcs]
        cc.decompressBricksFloat(inst, d2, shapes, steps, 2, out2, offs2)
        torch.cuda.synchronize()
        for a, b in zip(prec, d2):
            assert np.array_equal(a, b.cpu().numpy()), rep
        # compression and decompression in one submission: the same words and reconstructions
        for d in dsts:
            d.zero_()
        out.zero_()
        offs3, cst, dst = plan.run()
        assert np.all(cst == 0) and np.all(dst == 0)
        assert np.array_equal(offs3, offs), rep
        assert np.array_equal(out[: int(offs3[-1])].cpu().numpy().view(np.uint32), pw), rep
        for a, b in zip(prec, dsts):
            assert np.array_equal(a, b.cpu().numpy()), rep
    # a second plan of the same bricks, both launched before either is waited for
    dsts_b = [torch.zeros_like(s) for s in srcs]
    out_b = torch.zeros_like(out)
    plan_b = cc.BricksPlan(inst, srcs, dsts_b, shapes, steps, 2, out_b)
    torch.cuda.synchronize()
    for _ in range(2):
        plan.launch()
        plan_b.launch()
        ra, rb = plan.wait(), plan_b.wait()
        for x, y in zip(ra, rb):
            assert np.array_equal(x, y)
        assert np.all(ra[1] == 0) and np.all(ra[2] == 0)
        nw = int(ra[0][-1])
        assert torch.equal(out[:nw], out_b[:nw])
        for a, b in zip(dsts, dsts_b):
            assert torch.equal(a, b)
    # coding statistics of the last compression against the oracle's streams
    st = plan.codingStats()
    from oracle import pyoracle as o
    n = m = nh = 0
    for s_, q in zip(srcs, steps):  # the sources as last compressed (changed in the third replay)
        b = s_.cpu().numpy()
        coef = P.dwt97_levels(b, 2)
        for k, (sl, qb) in enumerate(P.band_steps(b.shape, 2, q)):
            syms = o.quantize(np.ascontiguousarray(coef[sl]).ravel(), qb)
            n += syms.size
            if k == 0:
                nh += syms.size           # the lowpass band: Huffman only
            else:
                m += o.rle_encode(syms)[0].size
    log(f"   stats {st}; oracle N={n} M={m} N_H={nh}")
    assert (st["N"], st["M"], st["N_H"]) == (n, m, nh)
    assert st["C"] == 4 * int(ra[0][-1])
    plan_b.close()
    plan.close()
    inst.close()


@pytest.mark.parametrize("groups", [2, 3, 7])
def test_grouped_batches_equal_single_batch(cc, torch_gpu, field, groups):
    """A batch run as concurrent groups writes the single batch's words at the same offsets and
    reconstructs the same values; a plan made on the same instance (each group captured on its own,
    staggered, bitstreams concatenated on the device) replays the same, through separate
    compress / decompress replays and through one launch of both."""
    log("bricks groups", groups)
    torch = torch_gpu
    rng = np.random.default_rng(40 + groups)
    bricks = make_bricks(rng, field, 19)
    std = float(field.std())
    steps = [float(rng.choice([0.005, 0.05])) * std for _ in bricks]
    shapes = [b.shape for b in bricks]
    inst = cc.createInstance(0, 7, 64 * 64 * 64, 128, 14)
    srcs, out1, offs1, stat1 = run_batch(cc, torch, inst, bricks, steps, 2)
    assert np.all(stat1 == 0)
    d1 = [torch.zeros_like(s) for s in srcs]
    cc.decompressBricksFloat(inst, d1, shapes, steps, 2, out1, offs1)
    inst.setBrickGroups(groups)
    out2 = torch.full_like(out1, -1)
    offs2, stat2 = cc.compressBricksFloat(inst, srcs, shapes, steps, 2, out2)
    torch.cuda.synchronize()
    assert np.all(stat2 == 0)
    assert np.array_equal(offs1, offs2)
    n = int(offs1[-1])
    assert np.array_equal(out1[:n].cpu().numpy(), out2[:n].cpu().numpy())
    d2 = [torch.zeros_like(s) for s in srcs]
    cc.decompressBricksFloat(inst, d2, shapes, steps, 2, out2, offs2)
    torch.cuda.synchronize()
    for a, b in zip(d1, d2):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    out3 = torch.full_like(out1, -1)
    d3 = [torch.zeros_like(s) for s in srcs]
    plan = cc.BricksPlan(inst, srcs, d3, shapes, steps, 2, out3)
    for _ in range(2):
        offs3, stat3 = plan.compress()
        assert np.all(stat3 == 0) and np.array_equal(offs1, offs3)
        assert np.all(plan.decompress() == 0)
        torch.cuda.synchronize()
        assert np.array_equal(out1[:n].cpu().numpy(), out3[:n].cpu().numpy())
        for a, b in zip(d1, d3):
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    for _ in range(2):
        out3.fill_(-1)
        for d in d3:
            d.zero_()
        offs4, cst, dst = plan.run()
        assert np.all(cst == 0) and np.all(dst == 0) and np.array_equal(offs1, offs4)
        assert np.array_equal(out1[:n].cpu().numpy(), out3[:n].cpu().numpy())
        for a, b in zip(d1, d3):
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    st = plan.codingStats()
    assert st["C"] == 4 * n
    plan.close()
    inst.close()


def test_server_results_same_alone_and_in_the_mix(cc, torch_gpu, field):
    """Requests decoded through the serving loop -- batched with others, several slots in flight --
    equal the same requests decoded alone, bit for bit, and equal the eager batch decode."""
    import time
    log("server alone vs mix")
    torch = torch_gpu
    rng = np.random.default_rng(33)
    bricks = make_bricks(rng, field, 30)
    std = float(field.std())
    steps = [0.005 * std] * len(bricks)
    inst = cc.createInstance(0, 7, 64 * 64 * 64, 128, 14)
    srcs, out, offs, stat = run_batch(cc, torch, inst, bricks, steps, 2)
    assert np.all(stat == 0)
    shapes = [b.shape for b in bricks]
    ranges = np.stack([offs[:-1], offs[1:]], axis=1)
    reqs = [list(rng.choice(len(bricks), size=int(rng.integers(1, 9)), replace=True)) for _ in range(40)]

    def serve(requests, slots, budget):
        srv = cc.Server(inst, out, ranges, shapes, steps, 2, slots=slots, workspaceBytes=budget)
        outs = {i: [torch.full_like(srcs[b], -1.0) for b in r] for i, r in enumerate(requests)}
        torch.cuda.synchronize()  # the server's slot streams do not wait on torch's stream
        t0 = time.perf_counter()
        for i, r in enumerate(requests):
            srv.submit(i, r, outs[i], time.perf_counter() - t0)
            srv.poll(time.perf_counter() - t0)
        while srv.pending():
            srv.poll(time.perf_counter() - t0)
        done, _ = srv.completed()
        failed, _ = srv.failed()
        srv.close()
        torch.cuda.synchronize()
        assert failed.size == 0 and sorted(done.tolist()) == list(range(len(requests)))
        return {i: [o.cpu().numpy() for o in v] for i, v in outs.items()}

    mix = serve(reqs, slots=3, budget=64 << 20)
    alone = {}
    for i, r in enumerate(reqs[:10]):
        alone[i] = serve([r], slots=1, budget=1 << 30)[0]
    eager = [torch.zeros_like(s) for s in srcs]
    cc.decompressBricksFloat(inst, eager, shapes, steps, 2, out, offs)
    torch.cuda.synchronize()
    for i, r in enumerate(reqs):
        for j, b in enumerate(r):
            assert np.array_equal(mix[i][j], eager[b].cpu().numpy()), (i, j)
            if i in alone:
                assert np.array_equal(alone[i][j], mix[i][j]), (i, j)
    inst.close()


def test_server_index_reports_broken_bricks(cc, torch_gpu, field):
    """The serving index parses every brick once at creation: a brick whose stored range is empty
    or cut short fails the requests that hold it (status: corrupt stream); the other requests
    decode to the eager batch's values."""
    log("server index with broken bricks")
    torch = torch_gpu
    rng = np.random.default_rng(34)
    bricks = make_bricks(rng, field, 8)
    std = float(field.std())
    steps = [0.005 * std] * len(bricks)
    inst = cc.createInstance(0, 7, 64 * 64 * 64, 128, 14)
    srcs, out, offs, stat = run_batch(cc, torch, inst, bricks, steps, 2)
    assert np.all(stat == 0)
    shapes = [b.shape for b in bricks]
    ranges = np.stack([offs[:-1], offs[1:]], axis=1).copy()
    ranges[2, 1] = ranges[2, 0]                               # empty
    ranges[5, 1] = ranges[5, 0] + (ranges[5, 1] - ranges[5, 0]) // 2  # truncated
    reqs = [[0, 1], [2], [3, 5], [4, 6, 7], [1, 2, 3]]
    srv = cc.Server(inst, out, ranges, shapes, steps, 2, slots=2, workspaceBytes=1 << 30)
    outs = {i: [torch.full_like(srcs[b], -1.0) for b in r] for i, r in enumerate(reqs)}
    torch.cuda.synchronize()
    for i, r in enumerate(reqs):
        srv.submit(i, r, outs[i], 0.0)
        srv.poll(0.0)
    while srv.pending():
        srv.poll(0.0)
    done, _ = srv.completed()
    failed, status = srv.failed()
    srv.close()
    torch.cuda.synchronize()
    assert sorted(done.tolist()) == [0, 3]
    assert sorted(failed.tolist()) == [1, 2, 4]
    assert all(int(s) == cc.CC_ERR_CORRUPT_STREAM for s in status)
    eager = [torch.zeros_like(s) for s in srcs]
    cc.decompressBricksFloat(inst, eager, shapes, steps, 2, out, offs)
    torch.cuda.synchronize()
    for i in (0, 3):
        for j, b in enumerate(reqs[i]):
            assert np.array_equal(outs[i][j].cpu().numpy(), eager[b].cpu().numpy()), (i, j)
    inst.close()


def test_tuned_shapes_keep_words_and_values(cc, torch_gpu, field):
    """Kernel-shape tuning (design grids, separate or fused x+y passes, z-slices per fused
    workgroup, tables designed one per thread up to 32..256 symbols) changes no word and no
    reconstructed value; a coding block size from the table changes the words but not the values."""
    log("tuning: shapes keep the words")
    torch = torch_gpu
    rng = np.random.default_rng(29)
    bricks = make_bricks(rng, field, 16, (16, 32, 64))
    std = float(field.std())
    steps = [0.005 * std] * len(bricks)
    shapes = [b.shape for b in bricks]
    srcs = [torch.from_numpy(b).cuda() for b in bricks]
    total = sum(b.size for b in bricks)
    base = cc.defaultTuning()
    variants = [dict(base), dict(base, designGrid=[64, 64, 64, 64]), dict(base, fusedXY=0),
                dict(base, xySlicesMax=1), dict(base, xySlicesMax=2, designGrid=[8192, 512, 256, 128])]
    variants += [dict(base, laneDesignK=k) for k in (0, 32, 64, 128, 256) if k != base["laneDesignK"]]
    variants += [dict(base, laneDesignK=256, designGrid=[3, 64, 64, 64])]
    ref = None
    for t in variants:
        inst = cc.createInstance(0, 7, 64 ** 3, 128, 14)
        inst.setTuning(t)
        assert inst.getTuning() == t
        out = torch.zeros(total + 4096 * len(bricks), dtype=torch.int32, device="cuda")
        dsts = [torch.zeros_like(s) for s in srcs]
        plan = cc.BricksPlan(inst, srcs, dsts, shapes, steps, 2, out)
        offs, cst, dst = plan.run()
        assert np.all(cst == 0) and np.all(dst == 0)
        got = (out[: int(offs[-1])].cpu().numpy().copy(), [d.cpu().numpy() for d in dsts])
        if ref is None:
            ref = got
        assert np.array_equal(got[0], ref[0]), t
        for a, b in zip(got[1], ref[1]):
            assert np.array_equal(a, b), t
        plan.close()
        inst.setTuning(None)
        inst.close()
    inst = cc.createInstance(0, 7, 64 ** 3, 64, 14)
    out = torch.zeros(total + 4096 * len(bricks), dtype=torch.int32, device="cuda")
    dsts = [torch.zeros_like(s) for s in srcs]
    offs, stat = cc.compressBricksFloat(inst, srcs, shapes, steps, 2, out)
    assert np.all(stat == 0)
    cc.decompressBricksFloat(inst, dsts, shapes, steps, 2, out, offs)
    torch.cuda.synchronize()
    for i, b in enumerate(bricks):
        z, y, x = b.shape
        w = cc.compressVolumeFloat(inst, srcs[i], x, y, z, 2, steps[i])
        assert np.array_equal(out[int(offs[i]):int(offs[i + 1])].cpu().numpy().view(np.uint32), w)
    for a, b in zip(dsts, ref[1]):
        assert np.array_equal(a.cpu().numpy(), b)
    inst.close()
