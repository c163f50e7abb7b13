// Golomb-Rice descriptor stream of a RecSplit forest, bit-for-bit the layout of the reference's
// shockhash::RiceBitVector (include/shockhash/RiceBitVector.h:98-266):
//  * fixed parts are appended LSB-first into 64-bit words (appendFixed :116-133);
//  * per bucket, after all fixed parts, one unary code per node: `u` zero bits then a one
//    (appendUnaryAll :135-149);
//  * the word count after the last append is words_for(bits) = ((ceil(bits/8) + 14) / 8), which leaves
//    the padding the reader's 16-byte fixed-part loads rely on.
#pragma once
#include <cstdint>
#include <cstring>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <vector>

namespace shockhash {

class RiceBitVector {
  public:
    static size_t words_for(uint64_t bits) { return (((bits + 7) / 8) + 7 + 7) / 8; }

    class Builder {
        std::vector<uint64_t> data;
        uint64_t bit_count = 0;
        void grow(uint64_t bits) {
            const size_t w = words_for(bits);
            if (data.size() < w) data.resize(w, 0);
        }

      public:
        void appendFixed(uint64_t v, int log2golomb) {
            if (log2golomb == 0) { grow(bit_count); return; }
            grow(bit_count + log2golomb);
            const uint64_t low = log2golomb >= 64 ? v : (v & ((uint64_t(1) << log2golomb) - 1));
            const unsigned used = bit_count & 63;
            uint64_t *w = &data[bit_count / 64];
            w[0] |= low << used;
            if (used + log2golomb > 64) w[1] |= low >> (64 - used);
            bit_count += log2golomb;
        }
        void appendUnaryAll(const std::vector<uint32_t> &unary) {
            uint64_t inc = 0;
            for (uint32_t u : unary) inc += (uint64_t)u + 1;
            grow(bit_count + inc);
            for (uint32_t u : unary) {
                bit_count += u;
                data[bit_count / 64] |= uint64_t(1) << (bit_count & 63);
                ++bit_count;
            }
        }
        // Bitwise concatenation of another builder (appendRiceBitVector :151-169).
        void appendRiceBitVector(const Builder &o) {
            grow(bit_count + o.bit_count);
            const unsigned used = bit_count & 63;
            const size_t nw = (o.bit_count + 63) / 64;
            uint64_t *w = &data[bit_count / 64];
            for (size_t i = 0; i < nw; i++) {
                const uint64_t v = o.data[i];
                w[i] |= v << used;
                if (used && (i + 1) < data.size() - bit_count / 64) w[i + 1] |= v >> (64 - used);
            }
            bit_count += o.bit_count;
        }
        uint64_t getBits() const { return bit_count; }
        std::vector<uint64_t> build() {
            data.resize(words_for(bit_count), 0);
            return std::move(data);
        }
    };

  private:
    std::vector<uint64_t> data;

  public:
    RiceBitVector() = default;
    explicit RiceBitVector(std::vector<uint64_t> d) : data(std::move(d)) {}
    size_t getBits() const { return data.size() * 64; }
    const std::vector<uint64_t> &words() const { return data; }
    // Serialized form: word count, then the words (little-endian u64).
    void save(std::ostream &os) const {
        const uint64_t n = data.size();
        os.write((const char *)&n, 8);
        os.write((const char *)data.data(), (std::streamsize)(n * 8));
    }
    void load(std::istream &is) {
        uint64_t n = 0;
        is.read((char *)&n, 8);
        if (!is || n > (1ull << 40)) throw std::runtime_error("RiceBitVector: bad word count");
        data.assign(n, 0);
        is.read((char *)data.data(), (std::streamsize)(n * 8));
        if (!is) throw std::runtime_error("RiceBitVector: truncated stream");
    }
    // Stream operators of the reference's RiceBitVector (RiceBitVector.h:133-141), in the form above.
    friend std::ostream &operator<<(std::ostream &os, const RiceBitVector &rbv) {
        rbv.save(os);
        return os;
    }
    friend std::istream &operator>>(std::istream &is, RiceBitVector &rbv) {
        rbv.load(is);
        return is;
    }

    // Sequential reader (Reader :198-263): fixed parts from bit_pos, unary parts from
    // bit_pos + unary_offset.
    class Reader {
        const uint64_t *w;
        size_t nw;
        uint64_t fixed = 0;
        uint64_t upos = 0;  // next unary bit

        uint64_t word_at(size_t i) const { return i < nw ? w[i] : 0; }

      public:
        explicit Reader(const std::vector<uint64_t> &d) : w(d.data()), nw(d.size()) {}
        void readReset(uint64_t bit_pos, uint64_t unary_offset) {
            fixed = bit_pos;
            upos = bit_pos + unary_offset;
        }
        uint64_t readNext(int log2golomb) {
            // unary: count zero bits up to the next one
            uint64_t q = 0;
            for (;;) {
                const uint64_t cur = word_at(upos / 64) >> (upos & 63);
                if (cur) {
                    const unsigned z = (unsigned)__builtin_ctzll(cur);
                    q += z;
                    upos += z + 1;
                    break;
                }
                const unsigned skip = 64 - (unsigned)(upos & 63);
                q += skip;
                upos += skip;
                if (upos / 64 >= nw) break;  // malformed stream: stop at the end
            }
            uint64_t r = 0;
            if (log2golomb) {
                const unsigned sh = fixed & 63;
                unsigned __int128 v = ((unsigned __int128)word_at(fixed / 64 + 1) << 64) | word_at(fixed / 64);
                v >>= sh;
                r = (uint64_t)v & (log2golomb >= 64 ? ~0ull : ((uint64_t(1) << log2golomb) - 1));
                fixed += log2golomb;
            }
            return (q << log2golomb) | r;
        }
        void skipSubtree(uint64_t nodes, uint64_t fixed_len) {
            while (nodes) {
                uint64_t cur = word_at(upos / 64) >> (upos & 63);
                const unsigned avail = 64 - (unsigned)(upos & 63);
                const uint64_t ones = (uint64_t)__builtin_popcountll(cur);
                if (ones < nodes) {
                    nodes -= ones;
                    upos += avail;
                    if (upos / 64 >= nw) break;
                    continue;
                }
                // position of the nodes-th one in cur
                for (uint64_t k = 1; k < nodes; k++) cur &= cur - 1;
                upos += (unsigned)__builtin_ctzll(cur) + 1;
                nodes = 0;
            }
            fixed += fixed_len;
        }
    };
    Reader reader() const { return Reader(data); }
};

}  // namespace shockhash
