// A C++ caller of the drop-in API, written the way the reference's own callers use it
// (src/examples/simpleExample.cpp:19-40 and the compressVolumeFloat / decompressVolumeFloat flow
// of src/examples/CompressVolume.cpp:313-516), linked against libcudacompress_amd.so and run on
// the GPU by tests/test_gpu_dropin.py.
//
//   dropin_volume <dir> <sizeX> <sizeY> <sizeZ> <numLevels> <quantStep>
//
// reads <dir>/volume.f32 (x fastest) and <dir>/streams.u16 (three streams of `count` symbols,
// count = sizeX*sizeY*sizeZ/8) and writes:
//   volume.words   compressVolumeFloat's bitstream (words of a BitStream over a std::vector)
//   volume.rec     decompressVolumeFloat's reconstruction
//   prefix.words   3 bits written by the caller, then encodeHuff of stream 0 (unaligned start)
//   alias1.words, alias2.words   encodeRLHuff(BitStream* b[] = {&b1, &b2, &b1}) of the 3 streams
//   alias.dec      decodeRLHuff of both bitstreams through the same aliased array (3 streams)
//   status.txt     "ok" or the failing step
#include <cudaCompress/BitStream.h>
#include <cudaCompress/Encode.h>
#include <cudaCompress/Instance.h>
#include <cudaCompress/util/DWT.h>
#include <cudaCompress/util/Quantize.h>

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace cudaCompress;

namespace {

std::string g_dir;

template <class T>
std::vector<T> read_file(const char* name, size_t count) {
    std::vector<T> v(count);
    FILE* f = std::fopen((g_dir + "/" + name).c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot read %s\n", name);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

template <class T>
void write_file(const char* name, const T* p, size_t count) {
    FILE* f = std::fopen((g_dir + "/" + name).c_str(), "wb");
    if (!f || std::fwrite(p, sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot write %s\n", name);
        std::exit(2);
    }
    std::fclose(f);
}

int fail(const char* step) {
    std::fprintf(stderr, "failed: %s (%s)\n", step, cc_last_error());
    FILE* f = std::fopen((g_dir + "/status.txt").c_str(), "w");
    if (f) {
        std::fprintf(f, "%s\n", step);
        std::fclose(f);
    }
    return 1;
}

#define HIPCHECK(x)                               \
    do {                                          \
        if ((x) != hipSuccess) return fail(#x);   \
    } while (0)

}  // namespace

int main(int argc, char** argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s dir sizeX sizeY sizeZ numLevels quantStep\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    const uint sizeX = (uint)std::atoi(argv[2]), sizeY = (uint)std::atoi(argv[3]), sizeZ = (uint)std::atoi(argv[4]);
    const uint numLevels = (uint)std::atoi(argv[5]);
    const float quantStep0 = (float)std::atof(argv[6]);
    const uint elemCount = sizeX * sizeY * sizeZ;
    const uint blockCount = 7;
    const uint elemCountPerBlockMax = elemCount / 8;

    std::vector<float> h_volume = read_file<float>("volume.f32", elemCount);
    std::vector<Symbol16> h_streams = read_file<Symbol16>("streams.u16", 3 * (size_t)elemCountPerBlockMax);

    Instance* inst = createInstance(0, blockCount, elemCountPerBlockMax, 128, 14);
    if (!inst) return fail("createInstance");

    float* dpImage = nullptr;
    float* dpDWT[2] = {nullptr, nullptr};
    float* dpOut = nullptr;
    HIPCHECK(hipMalloc(&dpImage, elemCount * sizeof(float)));
    HIPCHECK(hipMalloc(&dpDWT[0], elemCount * sizeof(float)));
    HIPCHECK(hipMalloc(&dpDWT[1], elemCount * sizeof(float)));
    HIPCHECK(hipMalloc(&dpOut, elemCount * sizeof(float)));
    HIPCHECK(hipMemcpy(dpImage, h_volume.data(), elemCount * sizeof(float), hipMemcpyHostToDevice));
    std::vector<Symbol16*> dpSymbolStreams(blockCount);
    for (uint i = 0; i < blockCount; i++) HIPCHECK(hipMalloc(&dpSymbolStreams[i], elemCountPerBlockMax * sizeof(Symbol16)));

    // ---- compressVolumeFloat (CompressVolume.cpp:313-409)
    std::vector<uint> bits;
    {
        int d = 0;
        util::dwtFloat3DForward(dpDWT[d], dpDWT[1 - d], dpDWT[d], dpImage, sizeX, sizeY, sizeZ);
        uint factor = 2;
        for (uint level = 1; level < numLevels; level++) {
            util::dwtFloat3DForward(dpDWT[1 - d], dpDWT[d], dpDWT[1 - d], dpDWT[d], sizeX / factor, sizeY / factor,
                                    sizeZ / factor, 1, sizeX, sizeX * sizeY, sizeX, sizeX * sizeY);
            d = 1 - d;
            factor *= 2;
        }
        BitStream bitStream(&bits);
        float q = quantStep0 / float(factor / 2);
        util::quantizeToSymbols(dpSymbolStreams[0], dpDWT[d], sizeX / factor, sizeY / factor, sizeZ / factor, q,
                                sizeX, sizeX * sizeY, util::QUANTIZE_DEADZONE);
        if (!encodeHuff(inst, bitStream, dpSymbolStreams.data(), 1, elemCount / (factor * factor * factor)))
            return fail("encodeHuff LP");
        for (int level = (int)numLevels - 1; level >= 0; level--) {
            q = quantStep0 / float(factor / 2);
            const uint bx = sizeX / factor, by = sizeY / factor, bz = sizeZ / factor;
            for (uint block = 1; block < 1 + blockCount; block++) {
                const uint x = block % 2, y = block / 2 % 2, z = block / 4;
                const uint offset = x * bx + y * by * sizeX + z * bz * sizeX * sizeY;
                util::quantizeToSymbols(dpSymbolStreams[block - 1], dpDWT[d] + offset, bx, by, bz, q, sizeX,
                                        sizeX * sizeY, util::QUANTIZE_DEADZONE);
            }
            d = 1 - d;
            if (!encodeRLHuff(inst, bitStream, dpSymbolStreams.data(), blockCount, bx * by * bz))
                return fail("encodeRLHuff HP");
            factor /= 2;
        }
        // bits.size() words now hold the stream (the BitStream writes into the vector)
    }
    write_file("volume.words", bits.data(), bits.size());

    // ---- decompressVolumeFloat (CompressVolume.cpp:412-516), WARP_SIZE * codingBlockSize padding
    {
        const uint pad = 32 * getInstanceCodingBlockSize(inst);
        Symbol16* dpSymbolStream = nullptr;
        const uint padded = (elemCountPerBlockMax + pad - 1) / pad * pad;
        HIPCHECK(hipMalloc(&dpSymbolStream, blockCount * padded * sizeof(Symbol16)));
        std::vector<Symbol16*> streams(blockCount * numLevels);
        int factor = 1 << numLevels;
        uint next = 0;
        for (int level = (int)numLevels - 1; level >= 0; level--) {
            const uint n = (sizeX / factor) * (sizeY / factor) * (sizeZ / factor);
            const uint np = (n + pad - 1) / pad * pad;
            for (uint i = 0; i < blockCount; i++) streams[next++] = dpSymbolStream + i * np;
            factor /= 2;
        }
        Symbol16** dppStreams = nullptr;
        HIPCHECK(hipMalloc(&dppStreams, streams.size() * sizeof(Symbol16*)));
        HIPCHECK(hipMemcpy(dppStreams, streams.data(), streams.size() * sizeof(Symbol16*), hipMemcpyHostToDevice));

        BitStreamReadOnly bitStream(bits.data(), (uint)bits.size() * 32);
        factor = 1 << numLevels;
        const uint elemCountLP = elemCount / (factor * factor * factor);
        if (!decodeHuff(inst, bitStream, &dpSymbolStream, 1, elemCountLP)) return fail("decodeHuff LP");
        util::unquantizeFromSymbols(dpOut, dpSymbolStream, sizeX / factor, sizeY / factor, sizeZ / factor,
                                    quantStep0 / float(factor / 2), sizeX, sizeX * sizeY);
        Symbol16** host_next = streams.data();
        Symbol16** dev_next = dppStreams;
        for (int level = (int)numLevels - 1; level >= 0; level--) {
            const uint n = (sizeX / factor) * (sizeY / factor) * (sizeZ / factor);
            if (!decodeRLHuff(inst, bitStream, host_next, blockCount, n)) return fail("decodeRLHuff HP");
            factor /= 2;
            util::dwtFloat3DInverseFromSymbols(dpOut, dpDWT[1], dpDWT[0], dpOut, (const ushort* const*)dev_next,
                                               quantStep0 / float(factor), sizeX / factor, sizeY / factor,
                                               sizeZ / factor, 1, sizeX, sizeX * sizeY, sizeX, sizeX * sizeY);
            host_next += blockCount;
            dev_next += blockCount;
        }
        std::vector<float> rec(elemCount);
        HIPCHECK(hipDeviceSynchronize());
        HIPCHECK(hipMemcpy(rec.data(), dpOut, elemCount * sizeof(float), hipMemcpyDeviceToHost));
        write_file("volume.rec", rec.data(), rec.size());
        HIPCHECK(hipFree(dppStreams));
        HIPCHECK(hipFree(dpSymbolStream));
    }

    // ---- entropy coder calls on caller streams: an unaligned start, then an aliased per-stream array
    const uint n = elemCountPerBlockMax;
    for (uint i = 0; i < 3; i++)
        HIPCHECK(hipMemcpy(dpSymbolStreams[i], h_streams.data() + (size_t)i * n, n * sizeof(Symbol16),
                           hipMemcpyHostToDevice));
    {
        BitStream bs;
        bs.writeBits(5u, 3);
        if (!encodeHuff(inst, bs, dpSymbolStreams.data(), 1, n)) return fail("encodeHuff unaligned");
        write_file("prefix.words", bs.getRaw(), bs.getRawSizeUInts());
        // read back: skip the 3 bits, decode, compare on the host
        BitStreamReadOnly rs(bs.getRaw(), bs.getBitSize());
        uint three = 0;
        rs.readBits(three, 3);
        Symbol16* dpDec = nullptr;
        HIPCHECK(hipMalloc(&dpDec, n * sizeof(Symbol16)));
        if (three != 5u || !decodeHuff(inst, rs, &dpDec, 1, n)) return fail("decodeHuff unaligned");
        std::vector<Symbol16> back(n);
        HIPCHECK(hipMemcpy(back.data(), dpDec, n * sizeof(Symbol16), hipMemcpyDeviceToHost));
        for (uint i = 0; i < n; i++)
            if (back[i] != h_streams[i]) return fail("unaligned round trip");
        HIPCHECK(hipFree(dpDec));
    }
    {
        BitStream b1, b2;
        BitStream* pbs[3] = {&b1, &b2, &b1};
        if (!encodeRLHuff(inst, pbs, dpSymbolStreams.data(), 3, n)) return fail("encodeRLHuff aliased");
        write_file("alias1.words", b1.getRaw(), b1.getRawSizeUInts());
        write_file("alias2.words", b2.getRaw(), b2.getRawSizeUInts());
        BitStreamReadOnly r1(b1.getRaw(), b1.getBitSize()), r2(b2.getRaw(), b2.getBitSize());
        BitStreamReadOnly* prs[3] = {&r1, &r2, &r1};
        std::vector<Symbol16*> dec(3);
        for (auto& p : dec) HIPCHECK(hipMalloc(&p, n * sizeof(Symbol16)));
        if (!decodeRLHuff(inst, prs, dec.data(), 3, n)) return fail("decodeRLHuff aliased");
        if (r1.getBitPosition() != b1.getBitSize() || r2.getBitPosition() != b2.getBitSize())
            return fail("aliased decode positions");
        std::vector<Symbol16> out(3 * (size_t)n);
        for (uint i = 0; i < 3; i++) {
            HIPCHECK(hipMemcpy(out.data() + (size_t)i * n, dec[i], n * sizeof(Symbol16), hipMemcpyDeviceToHost));
            HIPCHECK(hipFree(dec[i]));
        }
        write_file("alias.dec", out.data(), out.size());
    }

    for (auto* p : dpSymbolStreams) HIPCHECK(hipFree(p));
    HIPCHECK(hipFree(dpImage));
    HIPCHECK(hipFree(dpDWT[0]));
    HIPCHECK(hipFree(dpDWT[1]));
    HIPCHECK(hipFree(dpOut));
    destroyInstance(inst);
    FILE* f = std::fopen((g_dir + "/status.txt").c_str(), "w");
    if (f) {
        std::fprintf(f, "ok\n");
        std::fclose(f);
    }
    return 0;
}
