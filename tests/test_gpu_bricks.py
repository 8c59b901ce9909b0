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


@pytest.mark.parametrize("levels", [1, 2, 3])
def test_bricks_equal_single_volume_calls(cc, torch_gpu, field, levels):
    log("bricks levels", levels)
    torch = torch_gpu
    rng = np.random.default_rng(levels)
    axes = AXES if levels < 3 else (16, 32, 48)
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
