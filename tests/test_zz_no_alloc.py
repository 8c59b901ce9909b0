"""cc_decode_device_noalloc (decode kernels on the library's static per-device scratch array)
against cc_decode_device (the same kernels on instance-owned buffers), through the C ABI: two
calls in a row on the same scratch, the second smaller, each bit-exact to the allocating decode
and to the encoded symbols; calls too large for the array are refused before any launch."""
import numpy as np
import pytest

from tests.gpu_util import laplace_symbols, log, to_dev, to_host

pytestmark = pytest.mark.gpu


def _encode(cc, torch, inst, streams, rl):
    dev = [to_dev(torch, s) for s in streams]
    cap = sum(s.size for s in streams) * 2 + 4096
    words = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    nw = cc.encodeDevice(inst, rl, dev, int(streams[0].size), words)
    assert nw is not None
    return words, nw


def _decode(cc, torch, inst, words, nw, ns, n, rl, noalloc):
    outs = [torch.full((n,), 0x5A5A, dtype=torch.int16, device="cuda") for _ in range(ns)]
    f = cc.decodeDeviceNoAlloc if noalloc else cc.decodeDevice
    used = f(inst, rl, words, outs, n, wordCount=nw)
    torch.cuda.synchronize()
    return used, [to_host(o, np.uint16) for o in outs]


@pytest.mark.parametrize("rl", [True, False], ids=["rlhuff", "huff"])
def test_noalloc_decode_equals_allocating_decode(cc, torch_gpu, rl):
    torch = torch_gpu
    inst = cc.createInstance(0, 16, 1 << 18, 128, 14)
    rng = np.random.default_rng(20261203 + int(rl))
    # first call: 7 streams of 64^3/2 symbols; second call: 3 streams of 5,003 symbols
    for ns, n in ((7, 131072), (3, 5003)):
        log(f"noalloc decode rl={rl} streams={ns} n={n}")
        streams = [laplace_symbols(rng, n, scale=6, density=0.4) for _ in range(ns)]
        words, nw = _encode(cc, torch, inst, streams, rl)
        used_a, got_a = _decode(cc, torch, inst, words, nw, ns, n, rl, noalloc=False)
        used_b, got_b = _decode(cc, torch, inst, words, nw, ns, n, rl, noalloc=True)
        assert used_a == used_b == nw
        for s in range(ns):
            assert np.array_equal(got_b[s], got_a[s]), f"stream {s} differs from cc_decode_device"
            assert np.array_equal(got_b[s], streams[s]), f"stream {s} differs from the encoded symbols"
    inst.close()


def test_noalloc_decode_refuses_calls_beyond_its_scratch(cc, torch_gpu):
    torch = torch_gpu
    inst = cc.createInstance(0, 8, 1 << 21, 128, 14)
    words = torch.zeros((1024,), dtype=torch.int32, device="cuda")
    # 8 RLE streams of 2^21 symbols need far more than the 32 MiB array: refused on the host
    n = 1 << 21
    outs = [torch.zeros((n,), dtype=torch.int16, device="cuda") for _ in range(8)]
    with pytest.raises(cc.CCError) as e:
        cc.decodeDeviceNoAlloc(inst, True, words, outs, n, wordCount=1024)
    assert e.value.status == cc.CC_ERR_INVALID_ARG
    # more than 128 streams in one call: refused on the host
    small = [torch.zeros((16,), dtype=torch.int16, device="cuda") for _ in range(129)]
    with pytest.raises(cc.CCError) as e:
        cc.decodeDeviceNoAlloc(inst, False, words, small, 16, wordCount=1024)
    assert e.value.status == cc.CC_ERR_INVALID_ARG
    inst.close()
