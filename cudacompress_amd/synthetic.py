"""Seeded synthetic inputs for the benchmark configurations (numpy, CPU).

* gaussian_field(shape, seed): isotropic Gaussian random field with energy spectrum E(k) ~ k^-5/3
  (per-mode amplitude ~ k^-(5/3 + d - 1)/2), random phases, zero mean, unit standard deviation.
* fbm_heightfield(n, hurst, seed): spectral fractional Brownian motion (2D amplitude ~ k^-(H+1)),
  min-max mapped to [0, 65535] and returned as int16 = u - 32768.
* velocity_field(n, seed): divergence-free u = curl A with |A_hat| ~ k^-17/6 (E_u ~ k^-5/3),
  unit RMS, shape (3, n, n, n).
* brick_set(count, axes, seed): bricks with each axis drawn uniformly from `axes`, cut at seeded
  offsets from seeded k^-5/3 fields.
* gaussian_field_device(n, seed, device): the k^-5/3 field of gaussian_field on a cube of side n,
  generated on a torch device (Philox white noise from torch.Generator(seed)) for sizes the host
  FFT makes slow (1024^3).
* serving_bricks(field, ...): the overlapping brick decomposition of a cube plus its 2x
  box-downsampled level of detail.
"""
import hashlib

import numpy as np


def _kmag(shape):
    ks = [np.fft.fftfreq(n) * n for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1]) * shape[-1]]
    grids = np.meshgrid(*ks, indexing="ij", sparse=True)
    k2 = sum(g.astype(np.float64) ** 2 for g in grids)
    return np.sqrt(k2), grids


def gaussian_field(shape, seed, slope=-5.0 / 3.0):
    rng = np.random.default_rng(seed)
    d = len(shape)
    white = rng.standard_normal(shape)
    spec = np.fft.rfftn(white)
    k, _ = _kmag(shape)
    with np.errstate(divide="ignore"):
        amp = np.where(k > 0, k ** ((slope - (d - 1)) / 2.0), 0.0)
    field = np.fft.irfftn(spec * amp, s=shape, axes=tuple(range(d)))
    field -= field.mean()
    field /= field.std()
    return field.astype(np.float32)


def fbm_heightfield(n, hurst, seed):
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n))
    spec = np.fft.rfftn(white)
    k, _ = _kmag((n, n))
    with np.errstate(divide="ignore"):
        amp = np.where(k > 0, k ** (-(hurst + 1.0)), 0.0)
    h = np.fft.irfftn(spec * amp, s=(n, n), axes=(0, 1))
    lo, hi = h.min(), h.max()
    u = np.rint((h - lo) / (hi - lo) * 65535.0).astype(np.int64)
    return (u - 32768).astype(np.int16)


def velocity_field(n, seed):
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    k, grids = _kmag(shape)
    with np.errstate(divide="ignore"):
        amp = np.where(k > 0, k ** (-17.0 / 6.0), 0.0)
    a_hat = [np.fft.rfftn(rng.standard_normal(shape)) * amp for _ in range(3)]
    kz, ky, kx = [2j * np.pi * g / n for g in grids]
    # u = curl A with components ordered (x, y, z) on a (z, y, x) array
    ax, ay, az = a_hat
    ux = ky * az - kz * ay
    uy = kz * ax - kx * az
    uz = kx * ay - ky * ax
    u = np.stack([np.fft.irfftn(c, s=shape, axes=(0, 1, 2)) for c in (ux, uy, uz)])
    u /= np.sqrt((u ** 2).sum(axis=0).mean())
    return u.astype(np.float32)


def brick_set(count, axes=(16, 32, 48, 64, 96, 128), seed=4, nfields=2, field_n=256):
    """`count` float32 bricks; brick i: sizes and offsets from default_rng(seed + i), cut from
    field i % nfields (gaussian_field((field_n,)*3, seed + 1000 + f)).  Returns a list of
    C-contiguous (z, y, x) arrays."""
    fields = [gaussian_field((field_n,) * 3, seed + 1000 + f) for f in range(nfields)]
    out = []
    for i in range(count):
        r = np.random.default_rng(seed + i)
        z, y, x = (int(r.choice(axes)) for _ in range(3))
        f = fields[i % nfields]
        oz, oy, ox = (int(r.integers(0, field_n - s + 1)) for s in (z, y, x))
        out.append(np.ascontiguousarray(f[oz:oz + z, oy:oy + y, ox:ox + x]))
    return out


def gaussian_field_device(n, seed, device, slope=-5.0 / 3.0):
    """float32 (n, n, n) tensor on `device`: white noise -> rfftn -> amplitude k^((slope - 2) / 2)
    (k = 0 removed) -> irfftn, then zero mean and unit std.  The spectrum is complex64."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    spec = torch.fft.rfftn(torch.randn((n, n, n), generator=g, device=device, dtype=torch.float32))
    f = torch.fft.fftfreq(n, d=1.0 / n, device=device, dtype=torch.float32)
    fr = torch.fft.rfftfreq(n, d=1.0 / n, device=device, dtype=torch.float32)
    e = (slope - 2.0) / 2.0
    for z in range(n):  # one z plane at a time keeps the amplitude temporary small
        k2 = f[z] ** 2 + f.view(n, 1) ** 2 + fr.view(1, -1) ** 2
        amp = torch.where(k2 > 0, k2.clamp_min(1e-30) ** (e / 2.0), torch.zeros_like(k2))
        spec[z].mul_(amp)
    field = torch.fft.irfftn(spec, s=(n, n, n))
    del spec
    field -= field.mean(dtype=torch.float64).float()
    field /= field.std().float()
    return field


def serving_offsets(n, brick, overlap):
    """Brick origins along one axis: stride brick - overlap, the last one clamped to n - brick."""
    stride = brick - overlap
    count = (n - brick + stride - 1) // stride + 1
    return [min(i * stride, n - brick) for i in range(count)]


def serving_bricks(field, brick=128, overlap=8):
    """Level-0 bricks of side `brick` overlapping by `overlap` voxels (origins serving_offsets, the
    last clamped to the edge; 1024 -> 9 per axis at stride 120) and the level-of-detail bricks: the
    2x box-downsampled field cut into bricks of side brick/2 at half the level-0 origins (aligned
    1:1 with the level-0 bricks).  Returns (l0, lod, lod_field): lists of contiguous tensors in
    z-major brick order."""
    n = field.shape[0]
    offs = serving_offsets(n, brick, overlap)
    h = n // 2
    lod_field = field.view(h, 2, h, 2, h, 2).mean(dim=(1, 3, 5))
    hb = brick // 2
    l0, lod = [], []
    for oz in offs:
        for oy in offs:
            for ox in offs:
                l0.append(field[oz:oz + brick, oy:oy + brick, ox:ox + brick].contiguous())
                lz, ly, lx = (min(o // 2, h - hb) for o in (oz, oy, ox))
                lod.append(lod_field[lz:lz + hb, ly:ly + hb, lx:lx + hb].contiguous())
    return l0, lod, lod_field


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def psnr(orig, rec):
    """PSNR = 20 log10(range / RMSE), range = max - min of the original
    (src/examples/tools/imgtools.h:137-143)."""
    orig = np.asarray(orig, dtype=np.float64)
    rec = np.asarray(rec, dtype=np.float64)
    rmse = np.sqrt(np.mean((orig - rec) ** 2))
    rng = orig.max() - orig.min()
    return float("inf") if rmse == 0 else float(20.0 * np.log10(rng / rmse))
