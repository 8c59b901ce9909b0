// Host bit writer/reader with the semantics of include/cudaCompress/BitStream.h: bits are
// written MSB-first into 32-bit words (writeBits, BitStream.h:377-416); aligned writes pad to
// the next 32-bit word and store little-endian bytes (writeAligned, :421-456).
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

namespace cca {

class BitWriter {
public:
    explicit BitWriter(std::vector<uint32_t>& words) : w_(words) {}
    uint64_t pos() const { return pos_; }
    void set_pos(uint64_t p) {
        pos_ = p;
        grow(pos_);
    }
    // write the low n bits of v (n <= 32), most significant first
    void bits(uint32_t v, uint32_t n) {
        if (n == 0) return;
        if (n < 32) v &= (1u << n) - 1u;
        grow(pos_ + n);
        uint32_t wi = (uint32_t)(pos_ >> 5), o = (uint32_t)(pos_ & 31);
        uint32_t room = 32 - o;
        if (n <= room) {
            uint32_t sh = room - n;
            uint32_t mask = (n == 32 ? 0xFFFFFFFFu : ((1u << n) - 1u)) << sh;
            w_[wi] = (w_[wi] & ~mask) | (v << sh);
        } else {
            uint32_t hi = n - room;  // bits spilling into the next word
            uint32_t mask = (1u << room) - 1u;
            w_[wi] = (w_[wi] & ~mask) | (v >> hi);
            uint32_t sh = 32 - hi;
            uint32_t mask2 = ((1u << hi) - 1u) << sh;
            w_[wi + 1] = (w_[wi + 1] & ~mask2) | (v << sh);
        }
        pos_ += n;
    }
    void align32() {
        pos_ = (pos_ + 31) & ~uint64_t(31);
        grow(pos_);
    }
    void aligned_u32(const uint32_t* v, uint64_t count) {
        align32();
        if (count == 0) return;
        grow(pos_ + 32 * count);
        memcpy(&w_[pos_ >> 5], v, 4 * count);
        pos_ += 32 * count;
    }
    void aligned_u16(const uint16_t* v, uint64_t count) {
        align32();
        if (count == 0) return;
        grow(pos_ + 16 * count);
        memcpy(reinterpret_cast<unsigned char*>(w_.data()) + (pos_ >> 3), v, 2 * count);
        pos_ += 16 * count;
        align32();
    }

private:
    void grow(uint64_t bits) {
        size_t need = (size_t)((bits + 31) >> 5);
        if (w_.size() < need) w_.resize(need, 0u);
    }
    std::vector<uint32_t>& w_;
    uint64_t pos_ = 0;
};

class BitReader {
public:
    BitReader(const uint32_t* w, uint64_t size, uint64_t pos) : w_(w), size_(size), pos_(pos) {}
    uint64_t pos() const { return pos_; }
    bool ok() const { return ok_; }
    // read n <= 32 bits; reading past the end yields zero bits and clears ok()
    uint32_t bits(uint32_t n) {
        if (n == 0) return 0;
        if (pos_ + n > size_) {
            ok_ = false;
            uint32_t v = 0;
            for (uint32_t i = 0; i < n; i++) v = (v << 1) | bit();
            return v;
        }
        uint32_t wi = (uint32_t)(pos_ >> 5), o = (uint32_t)(pos_ & 31);
        uint64_t win = (uint64_t)w_[wi] << 32;
        if (o + n > 32) win |= w_[wi + 1];
        pos_ += n;
        return (uint32_t)((win << o) >> (64 - n));
    }
    void align32() {
        uint64_t p = (pos_ + 31) & ~uint64_t(31);
        if (p > size_) { ok_ = false; p = size_; }
        pos_ = p;
    }
    uint32_t aligned_u32() {
        align32();
        if (pos_ + 32 > size_) { ok_ = false; return 0; }
        uint32_t v = w_[pos_ >> 5];
        pos_ += 32;
        return v;
    }
    const uint32_t* word_ptr() const { return w_ + (pos_ >> 5); }
    // skip `count` aligned elements of `bytes` bytes (BitStream.h:214-244 skipAligned)
    void skip_aligned(uint64_t count, uint32_t bytes) {
        align32();
        uint64_t need = count * bytes * 8;
        if (pos_ + need > size_) { ok_ = false; need = size_ - pos_; }
        pos_ += need;
        align32();
    }
    void skip_bits(uint64_t n) {
        if (pos_ + n > size_) { ok_ = false; n = size_ - pos_; }
        pos_ += n;
    }

private:
    uint32_t bit() {
        if (pos_ >= size_) { pos_++; return 0; }
        uint32_t b = (w_[pos_ >> 5] >> (31 - (pos_ & 31))) & 1u;
        pos_++;
        return b;
    }
    const uint32_t* w_;
    uint64_t size_;
    uint64_t pos_;
    bool ok_ = true;
};

}  // namespace cca
