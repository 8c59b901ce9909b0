// Differential trace of the host bit stream API (include/cudaCompress/BitStream.h).
// TEST INFRASTRUCTURE ONLY.  The same source is compiled twice: against this repository's
// drop-in header (tests/cpp/_bin/bitstream_trace) and, when the reference tree is present,
// against the reference's own header where it lies (oracle/_ref/bitstream_trace).  Each build
// runs the same seeded sequence of writes and reads and prints every returned value, position,
// size and the raw words, so the two traces must be identical line for line.
//
// The sequences cover: writeBits of 0..32 bits at every word offset, writeAligned of 1/2/4/8-byte
// elements at unaligned positions, writeAlign past the end, BitStream over a caller vector with
// existing words, setBitSize shrinking and growing, readBit/readBits/skipBits at and past the end,
// align() at the end of the stream (the clamped case), readAligned/skipAligned/getAlignedPointer,
// setBitPosition beyond the size.
#include <cudaCompress/BitStream.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace cudaCompress;

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint rnd(uint n) {  // xorshift64*, deterministic across builds
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return (uint)((g_state * 2685821657736338717ull) >> 33) % n;
}

void dump(const char* tag, const BitStreamReadOnly& bs) {
    std::printf("%s pos=%u size=%u words=%u bytes=%u :", tag, bs.getBitPosition(), bs.getBitSize(),
                bs.getRawSizeUInts(), bs.getRawSizeBytes());
    for (uint i = 0; i < bs.getRawSizeUInts(); i++) std::printf(" %08x", bs.getRaw()[i]);
    std::printf("\n");
}

void write_sequence(BitStream& bs, int round) {
    const int ops = 40 + round % 17;
    for (int i = 0; i < ops; i++) {
        const uint op = rnd(10);
        if (op < 5) {
            const uint n = rnd(33);
            const uint v = rnd(0x7fffffff) * 2u + rnd(2);
            bs.writeBits(v, n);
            std::printf("writeBits %u %08x -> %u\n", n, v, bs.getBitPosition());
        } else if (op < 7) {
            const uint k = rnd(4);
            const uint cnt = rnd(6);
            if (k == 0) {
                uchar a[8];
                for (uint j = 0; j < cnt; j++) a[j] = (uchar)rnd(256);
                bs.writeAligned(a, cnt);
            } else if (k == 1) {
                ushort a[8];
                for (uint j = 0; j < cnt; j++) a[j] = (ushort)rnd(65536);
                bs.writeAligned(a, cnt);
            } else if (k == 2) {
                uint a[8];
                for (uint j = 0; j < cnt; j++) a[j] = rnd(0x7fffffff) * 3u;
                bs.writeAligned(a, cnt);
            } else {
                uint64 a[8];
                for (uint j = 0; j < cnt; j++) a[j] = ((uint64)rnd(0x7fffffff) << 33) ^ rnd(0x7fffffff);
                bs.writeAligned(a, cnt);
            }
            std::printf("writeAligned k%u x%u -> %u\n", k, cnt, bs.getBitPosition());
        } else if (op < 8) {
            const uint b = 1u << rnd(3);
            bs.writeAlign(b);
            std::printf("writeAlign %u -> %u size %u\n", b, bs.getBitPosition(), bs.getBitSize());
        } else if (op < 9) {
            const uint p = rnd(bs.getBitSize() + 1);
            const bool ok = bs.setBitPosition(p);
            std::printf("setBitPosition %u -> %d %u\n", p, (int)ok, bs.getBitPosition());
        } else {
            const uint p = bs.getBitSize() + rnd(3);
            const bool ok = bs.setBitPosition(p);
            std::printf("setBitPosition past %u -> %d %u\n", p, (int)ok, bs.getBitPosition());
        }
    }
}

void read_sequence(BitStreamReadOnly& rs) {
    for (int i = 0; i < 60; i++) {
        const uint op = rnd(12);
        if (op < 4) {
            uint d = 0xA5A5A5A5u;
            const uint n = rnd(33);
            const uint got = rs.readBits(d, n);
            std::printf("readBits %u -> %u %08x %u\n", n, got, d, rs.getBitPosition());
        } else if (op < 5) {
            uint d = 0xFFFFFFFEu;
            const uint got = rs.readBit(d);
            std::printf("readBit -> %u %08x %u\n", got, d, rs.getBitPosition());
        } else if (op < 6) {
            const uint n = rnd(40);
            const uint got = rs.skipBits(n);
            std::printf("skipBits %u -> %u %u\n", n, got, rs.getBitPosition());
        } else if (op < 7) {
            const uint b = 1u << rnd(3);
            const bool ok = rs.align(b);
            std::printf("align %u -> %d %u\n", b, (int)ok, rs.getBitPosition());
        } else if (op < 9) {
            const uint k = rnd(3), cnt = rnd(5);
            uint got = 0;
            uint64 sum = 0;
            if (k == 0) {
                uchar a[8] = {0};
                got = rs.readAligned(a, cnt);
                for (uint j = 0; j < got; j++) sum = sum * 131 + a[j];
            } else if (k == 1) {
                ushort a[8] = {0};
                got = rs.readAligned(a, cnt);
                for (uint j = 0; j < got; j++) sum = sum * 131 + a[j];
            } else {
                uint a[8] = {0};
                got = rs.readAligned(a, cnt);
                for (uint j = 0; j < got; j++) sum = sum * 131 + a[j];
            }
            std::printf("readAligned k%u x%u -> %u %llu %u\n", k, cnt, got, (unsigned long long)sum,
                        rs.getBitPosition());
        } else if (op < 10) {
            const uint cnt = rnd(5);
            const uint got = rs.skipAligned<ushort>(cnt);
            std::printf("skipAligned x%u -> %u %u\n", cnt, got, rs.getBitPosition());
        } else if (op < 11) {
            const ushort* p = rs.getAlignedPointer<ushort>();
            std::printf("getAlignedPointer -> %ld %u\n", (long)((const byte*)p - (const byte*)rs.getRaw()),
                        rs.getBitPosition());
        } else {
            const uint p = rnd(rs.getBitSize() + 40);
            const bool ok = rs.setBitPosition(p);
            std::printf("setBitPosition %u -> %d %u\n", p, (int)ok, rs.getBitPosition());
        }
        // align() of a position near the end may step past it (the clamped case); reads from
        // there would leave the buffer, so the trace returns to the end before going on
        if (rs.getBitPosition() > rs.getBitSize()) {
            std::printf("beyond end %u > %u\n", rs.getBitPosition(), rs.getBitSize());
            rs.setBitPosition(rs.getBitSize());
        }
    }
}

}  // namespace

int main() {
    for (int round = 0; round < 64; round++) {
        std::printf("round %d\n", round);
        std::vector<uint> storage;
        const uint pre = rnd(3);
        for (uint i = 0; i < pre; i++) storage.push_back(0x01020304u * (i + 1));
        BitStream bs(round % 2 ? &storage : nullptr);
        dump("new", bs);
        write_sequence(bs, round);
        dump("written", bs);
        if (round % 5 == 0) {
            const uint s = rnd(bs.getBitSize() + 70);
            bs.setBitSize(s);
            dump("setBitSize", bs);
        }
        BitStreamReadOnly rs(bs.getRaw(), bs.getBitSize());
        read_sequence(rs);
        dump("read", rs);
        // a full-width round trip at every offset
        BitStream w;
        for (uint k = 0; k <= 32; k++) w.writeBits(0x80000001u ^ (k * 0x01010101u), k);
        BitStreamReadOnly r(w.getRaw(), w.getBitSize());
        for (uint k = 0; k <= 32; k++) {
            uint d = 0;
            r.readBits(d, k);
            std::printf("%08x ", d);
        }
        std::printf("\n");
        dump("fullwidth", w);
    }
    return 0;
}
