// Batched float bricks: compressVolumeFloat / decompressVolumeFloat (src/examples/
// CompressVolume.cpp:313-516) for many bricks of different sizes at once.  Each brick's
// bitstream is word for word what the single-volume call produces for it: the same streams in
// the same order (lowpass with encodeHuff, then the 7 highpass bands of every level, coarsest
// first, with encodeRLHuff), the same per-level quantization steps (q0 / 2^level, lowpass
// q0 / 2^(levels-1)), coded by the device-resident coder with one call per brick.
#include "bricks.hpp"

#include <algorithm>
#include <cstring>

#include "../bricks.hpp"
#include "batch.hpp"

namespace cca {

namespace {

struct Layout {
    std::vector<uint64_t> a, b, sym;            // per brick: work buffer and symbol offsets (elements)
    std::vector<std::vector<uint64_t>> lp;      // [level][brick] lowpass buffer offsets (floats)
    uint64_t work_floats = 0, sym_elems = 0;
    std::vector<uint64_t> level_elems;          // sum over bricks of the level's element count
};

uint64_t elems(const cc_brick& b) { return (uint64_t)b.sizeX * b.sizeY * b.sizeZ; }

Layout make_layout(const cc_brick* bricks, uint32_t n, uint32_t L) {
    Layout lay;
    lay.a.resize(n);
    lay.b.resize(n);
    lay.sym.resize(n);
    lay.lp.assign(L, std::vector<uint64_t>(n, 0));
    lay.level_elems.assign(L, 0);
    uint64_t w = 0, s = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t N = elems(bricks[i]);
        lay.a[i] = w;
        w += align_up(N, 64);
        lay.b[i] = w;
        w += align_up(N, 64);
        for (uint32_t l = 0; l + 1 < L; l++) {
            lay.lp[l][i] = w;
            w += align_up(N >> (3 * (l + 1)), 64);
        }
        for (uint32_t l = 0; l < L; l++) lay.level_elems[l] += N >> (3 * l);
        lay.sym[i] = s;
        s += align_up(N, 64);
    }
    lay.work_floats = w;
    lay.sym_elems = s;
    return lay;
}

void check_bricks(const cc_brick* bricks, uint32_t n, uint32_t L) {
    CCA_REQUIRE(L >= 1 && L <= 10, CC_ERR_INVALID_ARG, "numLevels must be in [1, 10]");
    CCA_REQUIRE(n == 0 || bricks, CC_ERR_INVALID_ARG, "null brick array");
    const uint32_t f = 1u << L;
    for (uint32_t i = 0; i < n; i++) {
        const cc_brick& b = bricks[i];
        CCA_REQUIRE(b.sizeX % f == 0 && b.sizeY % f == 0 && b.sizeZ % f == 0 && b.sizeX && b.sizeY && b.sizeZ > 1,
                    CC_ERR_INVALID_ARG, "brick sizes must be divisible by 2^numLevels (3D)");
        CCA_REQUIRE(elems(b) < (1ull << 31), CC_ERR_INVALID_ARG, "brick too large");
    }
}

// Coder items of brick i in bitstream order; ptr(i, k) gives the k-th stream's symbols.
template <class P>
void brick_items(const cc_brick& b, uint32_t L, uint32_t call, P ptr, std::vector<ItemSpec>& items) {
    const uint64_t N = elems(b);
    const uint32_t nlp = (uint32_t)(N >> (3 * L));
    items.push_back(ItemSpec{ptr(0), nlp, false, call});
    uint32_t k = 1;
    for (int l = (int)L - 1; l >= 0; l--) {
        const uint32_t nb = (uint32_t)(N >> (3 * (l + 1)));
        for (uint32_t band = 0; band < 7; band++) items.push_back(ItemSpec{ptr(k++), nb, true, call});
    }
}

// symbol offset (elements, from the brick's symbol base) of stream k in bitstream order
uint64_t stream_off(uint64_t N, uint32_t L, uint32_t k) {
    if (k == 0) return 0;
    uint64_t off = N >> (3 * L);
    uint32_t kk = 1;
    for (int l = (int)L - 1; l >= 0; l--) {
        const uint64_t nb = N >> (3 * (l + 1));
        for (uint32_t band = 0; band < 7; band++, kk++) {
            if (kk == k) return off;
            off += nb;
        }
    }
    return off;
}

// Uploads the per-level brick descriptors and tile prefixes; returns the device pointers.
struct LevelTables {
    const BrickLevel* levels;  // [L][n]
    const uint32_t* tiles;     // [L][3][n+1]
    std::vector<uint32_t> totals;  // [L][3]
};

LevelTables upload_levels(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, const Layout& lay,
                          float* work, unsigned char* sym, int symSize, bool inverse) {
    const size_t lv_bytes = sizeof(BrickLevel) * (size_t)L * n;
    const size_t tb_off = align_up(lv_bytes, 256);
    const size_t bytes = tb_off + 4ull * L * 3 * (n + 1);
    unsigned char* h = (unsigned char*)I->h_blevels.get(bytes);
    BrickLevel* hl = (BrickLevel*)h;
    uint32_t* ht = (uint32_t*)(h + tb_off);
    LevelTables t;
    t.totals.assign(3 * L, 0);
    for (uint32_t l = 0; l < L; l++) {
        uint32_t acc[3] = {0, 0, 0};
        for (uint32_t i = 0; i < n; i++) {
            const cc_brick& b = bricks[i];
            const uint64_t N = elems(b);
            BrickLevel& e = hl[(size_t)l * n + i];
            e.lx = (int)(b.sizeX >> l);
            e.ly = (int)(b.sizeY >> l);
            e.lz = (int)(b.sizeZ >> l);
            e.last = l == L - 1;
            e.in = l == 0 ? b.src : work + lay.lp[l - 1][i];
            e.a = work + lay.a[i];
            e.b = work + lay.b[i];
            e.lp = (l + 1 < L) ? work + lay.lp[l][i] : nullptr;
            // highpass bands of level l: stream index of its first band in bitstream order
            const uint32_t k = 1 + 7 * (L - 1 - l);
            e.hp = sym + (lay.sym[i] + stream_off(N, L, k)) * symSize;
            e.lpsym = sym + lay.sym[i] * symSize;
            e.out = l == 0 ? b.dst : work + lay.lp[l - 1][i];
            const float q0 = b.quantStep;
            e.step_hp = q0 / float(1u << l);      // CompressVolume.cpp:366 (q0 / (f/2), f = 2^(l+1))
            e.inv_hp = 1.0f / e.step_hp;          // Quantize.cu:94
            e.step_lp = q0 / float(1u << (L - 1));
            e.inv_lp = 1.0f / e.step_lp;
            for (int ax = 0; ax < 3; ax++) {
                ht[((size_t)l * 3 + ax) * (n + 1) + i] = acc[ax];
                acc[ax] += brick_tiles(ax, e.lx, e.ly, e.lz);
            }
        }
        for (int ax = 0; ax < 3; ax++) {
            ht[((size_t)l * 3 + ax) * (n + 1) + n] = acc[ax];
            t.totals[3 * l + ax] = acc[ax];
        }
    }
    (void)inverse;
    unsigned char* d = (unsigned char*)I->d_blevels.get(bytes);
    CCA_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, I->stream));
    t.levels = (const BrickLevel*)d;
    t.tiles = (const uint32_t*)(d + tb_off);
    return t;
}

}  // namespace

void compress_bricks(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, int symSize, uint32_t* d_out,
                     uint64_t cap, uint64_t* offsets, uint32_t* status) {
    DeviceGuard guard(I->device);
    check_bricks(bricks, n, L);
    for (uint32_t i = 0; i < n; i++) CCA_REQUIRE(bricks[i].src, CC_ERR_INVALID_ARG, "null brick source");
    const Layout lay = make_layout(bricks, n, L);
    float* work = I->d_bwork.as<float>(std::max<uint64_t>(lay.work_floats, 64));
    unsigned char* sym = (unsigned char*)I->d_bsym.get(std::max<uint64_t>(lay.sym_elems, 64) * symSize);
    const LevelTables t = upload_levels(I, bricks, n, L, lay, work, sym, symSize, false);
    for (uint32_t l = 0; l < L; l++)
        bricks_forward_level(t.levels + (size_t)l * n, t.tiles + (size_t)l * 3 * (n + 1), n, &t.totals[3 * l], symSize,
                             lay.level_elems[l], I->stream);
    std::vector<ItemSpec> items;
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t N = elems(bricks[i]);
        brick_items(bricks[i], L, i, [&](uint32_t k) { return (void*)(sym + (lay.sym[i] + stream_off(N, L, k)) * symSize); },
                    items);
    }
    std::vector<CallOut> res;
    const uint64_t total = device_encode_items(I, items, n, symSize, d_out, cap, &res, true);
    CCA_REQUIRE(total <= cap, CC_ERR_INVALID_ARG, "output buffer too small for the compressed bricks");
    bool ok = true;
    for (uint32_t i = 0; i < n; i++) {
        offsets[i] = res[i].word_begin;
        const uint32_t e = res[i].err;
        const cc_status s = e == 0 ? CC_OK
                            : e == kErrOverflow ? CC_ERR_SYMBOL_OVERFLOW
                            : e == kErrTooLong ? CC_ERR_CODEWORD_TOO_LONG
                            : e == kErrOutput ? CC_ERR_INVALID_ARG : CC_ERR_OUT_OF_MEMORY;
        if (status) status[i] = (uint32_t)s;
        ok = ok && s == CC_OK;
    }
    offsets[n] = total;
    CCA_REQUIRE(ok, CC_ERR_SYMBOL_OVERFLOW, "some bricks failed to encode (see the per-brick status)");
}

void decompress_bricks(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, int symSize,
                       const uint32_t* d_in, const uint64_t* offsets, uint32_t* status) {
    DeviceGuard guard(I->device);
    check_bricks(bricks, n, L);
    for (uint32_t i = 0; i < n; i++) {
        CCA_REQUIRE(bricks[i].dst, CC_ERR_INVALID_ARG, "null brick destination");
        CCA_REQUIRE(offsets[i + 1] >= offsets[i], CC_ERR_INVALID_ARG, "brick offsets must be nondecreasing");
    }
    const Layout lay = make_layout(bricks, n, L);
    float* work = I->d_bwork.as<float>(std::max<uint64_t>(lay.work_floats, 64));
    unsigned char* sym = (unsigned char*)I->d_bsym.get(std::max<uint64_t>(lay.sym_elems, 64) * symSize);
    std::vector<ItemSpec> items;
    std::vector<CallWords> calls(n);
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t N = elems(bricks[i]);
        calls[i] = CallWords{d_in + offsets[i], offsets[i + 1] - offsets[i]};
        brick_items(bricks[i], L, i, [&](uint32_t k) { return (void*)(sym + (lay.sym[i] + stream_off(N, L, k)) * symSize); },
                    items);
    }
    std::vector<DecCallOut> res;
    device_decode_items(I, items, calls, symSize, &res, false);
    const LevelTables t = upload_levels(I, bricks, n, L, lay, work, sym, symSize, true);
    for (int l = (int)L - 1; l >= 0; l--)
        bricks_inverse_level(t.levels + (size_t)l * n, t.tiles + (size_t)l * 3 * (n + 1), n, &t.totals[3 * l], symSize,
                             lay.level_elems[l], I->stream);
    // per-brick status from the decoder
    DeviceCoder& D = I->coder;
    DecCallOut* r = (DecCallOut*)D.h_dresult.get(sizeof(DecCallOut) * std::max<uint32_t>(n, 1));
    CCA_HIP(hipMemcpyAsync(r, D.d_dout.p, sizeof(DecCallOut) * n, hipMemcpyDeviceToHost, I->stream));
    CCA_HIP(hipStreamSynchronize(I->stream));
    bool ok = true;
    for (uint32_t i = 0; i < n; i++) {
        const cc_status s = r[i].err == 0 ? CC_OK : (r[i].err == kDecCorrupt ? CC_ERR_CORRUPT_STREAM : CC_ERR_OUT_OF_MEMORY);
        if (status) status[i] = (uint32_t)s;
        ok = ok && s == CC_OK;
    }
    CCA_REQUIRE(ok, CC_ERR_CORRUPT_STREAM, "some bricks failed to decode (see the per-brick status)");
}

}  // namespace cca
