// Drop-in replacement of include/cudaCompress/BitStream.h (m0bl0/cudaCompress): a host-side bit
// stream over 32-bit words.  Bits are stored most-significant first within each word; aligned
// element reads/writes first round the position up to a 32-bit word and to the element size,
// copy little-endian bytes, and round up to a 32-bit word again.  Same class names, methods and
// semantics as the reference (BitStream.h:14-122); the implementation is this project's.
#ifndef CUDACOMPRESS_AMD_DROPIN_BITSTREAM_H
#define CUDACOMPRESS_AMD_DROPIN_BITSTREAM_H

#include <cudaCompress/global.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace cudaCompress {

class BitStreamReadOnly {
public:
    BitStreamReadOnly(const uint* pBuffer, uint bitSize) : m_pBuffer(pBuffer), m_bitSize(bitSize), m_pos(0) {}
    BitStreamReadOnly(BitStreamReadOnly&& other)
        : m_pBuffer(other.m_pBuffer), m_bitSize(other.m_bitSize), m_pos(other.m_pos) {
        other.m_pBuffer = nullptr;
        other.m_bitSize = 0;
        other.m_pos = 0;
    }
    virtual ~BitStreamReadOnly() {}

    uint getBitPosition() const { return m_pos; }
    bool setBitPosition(uint bitPos) {
        if (bitPos > m_bitSize) {
            m_pos = m_bitSize;
            return false;
        }
        m_pos = bitPos;
        return true;
    }
    uint getBitSize() const { return m_bitSize; }

    // next bit into the least significant bit of data; returns the number of bits read
    uint readBit(uint& data) {
        if (m_pos == m_bitSize) return 0;
        data = (data & ~1u) | bitAt(m_pos);
        m_pos++;
        return 1;
    }
    // next bitCount bits into the bitCount least significant bits of data (rest unchanged);
    // returns the number of bits actually read (fewer at the end of the stream)
    uint readBits(uint& data, uint bitCount) {
        const uint n = std::min(bitCount, m_bitSize - m_pos);
        uint v = 0;
        for (uint i = 0; i < n; i++) v = (v << 1) | bitAt(m_pos + i);
        const uint mask = n >= 32 ? ~0u : ((1u << n) - 1u);
        data = (data & ~mask) | (v & mask);
        m_pos += n;
        return n;
    }
    uint skipBits(uint bitCount) {
        const uint n = std::min(bitCount, m_bitSize - m_pos);
        m_pos += n;
        return n;
    }

    template <typename T>
    uint readAligned(T* pTarget, uint count) {
        align<uint>();
        align<T>();
        const uint avail = (m_bitSize - m_pos) / 8 / (uint)sizeof(T);
        const uint n = std::min(avail, count);
        std::memcpy(pTarget, reinterpret_cast<const byte*>(m_pBuffer) + m_pos / 8, n * sizeof(T));
        m_pos += n * (uint)sizeof(T) * 8;
        align<uint>();
        return n;
    }
    template <typename T>
    uint skipAligned(uint count) {
        align<uint>();
        align<T>();
        const uint avail = (m_bitSize - m_pos) / 8 / (uint)sizeof(T);
        const uint n = std::min(avail, count);
        m_pos += n * (uint)sizeof(T) * 8;
        align<uint>();
        return n;
    }
    template <typename T>
    const T* getAlignedPointer() {
        align<uint>();
        align<T>();
        const T* p = reinterpret_cast<const T*>(getRaw()) + m_pos / (8 * sizeof(T));
        align<uint>();
        return p;
    }

    template <typename T>
    bool align() { return align((uint)sizeof(T)); }
    // advances to the next multiple of byteCount bytes within the current 32-bit word;
    // returns false (and stops at the end) when that would leave the stream
    bool align(uint byteCount) {
        const uint bits = byteCount * 8;
        uint off = (m_pos % 32) % bits;
        const uint room = m_bitSize - m_pos;
        bool ok = true;
        if (off > room) {
            off = room;
            ok = false;
        }
        if (off != 0) m_pos += bits - off;
        return ok;
    }

    const uint* getRaw() const { return m_pBuffer; }
    uint getRawSizeUInts() const { return (m_bitSize + 31) / 32; }
    uint getRawSizeBytes() const { return (m_bitSize + 7) / 8; }

protected:
    uint bitAt(uint p) const { return (m_pBuffer[p / 32] >> (31 - p % 32)) & 1u; }

    const uint* m_pBuffer;
    uint m_bitSize;
    uint m_pos;

private:
    BitStreamReadOnly(const BitStreamReadOnly&);
    BitStreamReadOnly& operator=(const BitStreamReadOnly&);
};

class BitStream : public BitStreamReadOnly {
public:
    // nullptr: the stream owns its storage; otherwise it writes into *pVector (whose current
    // contents count as written bits)
    explicit BitStream(std::vector<uint>* pVector = nullptr) : BitStreamReadOnly(nullptr, 0) {
        if (pVector) {
            m_own = false;
            m_vec = pVector;
            m_bitSize = (uint)m_vec->size() * 32;
        } else {
            m_own = true;
            m_vec = new std::vector<uint>;
        }
        m_pBuffer = m_vec->data();
    }
    BitStream(BitStream&& other) : BitStreamReadOnly(std::move(other)), m_vec(other.m_vec), m_own(other.m_own) {
        other.m_vec = nullptr;
        other.m_own = false;
    }
    virtual ~BitStream() {
        if (m_own) delete m_vec;
    }

    void setBitSize(uint bitSize) {
        m_bitSize = bitSize;
        resizeWords((bitSize + 31) / 32);
        if (m_pos > m_bitSize) m_pos = m_bitSize;
    }
    void reserveBitSize(uint bitSize) { m_vec->reserve((bitSize + 31) / 32); }

    // writes the bitCount least significant bits of data, most significant first
    void writeBits(uint data, uint bitCount) {
        growTo(m_pos + bitCount);
        uint* w = m_vec->data();
        for (uint i = 0; i < bitCount; i++) {
            const uint p = m_pos + i;
            const uint bit = (data >> (bitCount - 1 - i)) & 1u;
            const uint sh = 31 - p % 32;
            w[p / 32] = (w[p / 32] & ~(1u << sh)) | (bit << sh);
        }
        m_pos += bitCount;
    }
    template <typename T>
    void writeAligned(const T* pSource, uint count) {
        writeAlign<uint>();
        writeAlign<T>();
        if (count == 0) return;
        const uint start = m_pos;
        m_pos += count * (uint)sizeof(T) * 8;
        growTo(m_pos);
        std::memcpy(reinterpret_cast<byte*>(m_vec->data()) + start / 8, pSource, count * sizeof(T));
        writeAlign<uint>();
    }
    template <typename T>
    void writeAlign() { writeAlign((uint)sizeof(T)); }
    void writeAlign(uint byteCount) {
        const uint bits = byteCount * 8;
        const uint off = (m_pos % 32) % bits;
        if (off != 0) m_pos += bits - off;
        growTo(m_pos);
    }

    uint* getRaw() { return m_vec->data(); }
    using BitStreamReadOnly::getRaw;

    // Not in the reference API: places a segment produced by the encoder (cc_encode_* in
    // cudacompress_amd.h) at the current position.  segBits counts from bit 0 of words[0]; the
    // first startBit bits are placeholders for bits already in this stream.
    void appendEncodedSegment(const uint* words, uint64 segBits, uint startBit) {
        const uint nw = (uint)((segBits + 31) / 32);
        if (nw == 0) return;
        if (startBit != 0) {
            writeBits(words[0], 32 - startBit);
            writeAligned(words + 1, nw - 1);
        } else {
            writeAligned(words, nw);
        }
    }

private:
    void growTo(uint bits) {
        if (bits > m_bitSize) {
            m_bitSize = bits;
            const uint need = (bits + 31) / 32;
            if (need > m_vec->size()) resizeWords(need);
        }
    }
    void resizeWords(uint n) {
        m_vec->resize(n, 0u);
        m_pBuffer = m_vec->data();
    }

    std::vector<uint>* m_vec;
    bool m_own;

    BitStream(const BitStream&);
    BitStream& operator=(const BitStream&);
};

}  // namespace cudaCompress

#endif
