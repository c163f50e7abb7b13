// Elias-Fano storage of the two per-bucket cumulative sequences a RecSplit query needs: keys before
// bucket i and the descriptor bit position of bucket i. It answers the same question as sux's
// DoubleEF::get(i, cum_keys, cum_keys_next, bit_pos) (include/ShockHash.h:202); the bit layout is this
// library's own (sux is not part of the drop-in; query answers depend only on the stored values).
#pragma once
#include <cstdint>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <vector>

namespace shockhash {

class EliasFano {
    uint64_t n = 0;
    unsigned l = 0;
    std::vector<uint64_t> lower, upper;
    std::vector<uint64_t> samples;  // position in `upper` of every 256th one
    static constexpr unsigned SAMPLE = 256;

    static uint64_t get_bits(const std::vector<uint64_t> &v, uint64_t pos, unsigned len) {
        if (!len) return 0;
        const uint64_t w = pos / 64, b = pos % 64;
        uint64_t r = v[w] >> b;
        if (b + len > 64) r |= v[w + 1] << (64 - b);
        return len == 64 ? r : (r & ((uint64_t(1) << len) - 1));
    }
    static void set_bits(std::vector<uint64_t> &v, uint64_t pos, unsigned len, uint64_t x) {
        if (!len) return;
        const uint64_t w = pos / 64, b = pos % 64;
        v[w] |= x << b;
        if (b + len > 64) v[w + 1] |= x >> (64 - b);
    }

  public:
    EliasFano() = default;
    explicit EliasFano(const std::vector<uint64_t> &vals) {
        n = vals.size();
        const uint64_t u = n ? vals.back() + 1 : 1;
        l = 0;
        while (n && (u >> (l + 1)) >= n) l++;  // l = floor(log2(u / n))
        lower.assign((n * l + 127) / 64 + 1, 0);
        const uint64_t ubits = n + (u >> l) + 2;
        upper.assign(ubits / 64 + 2, 0);
        for (uint64_t i = 0; i < n; i++) {
            set_bits(lower, i * l, l, l ? (vals[i] & ((uint64_t(1) << l) - 1)) : 0);
            const uint64_t p = (vals[i] >> l) + i;
            upper[p / 64] |= uint64_t(1) << (p % 64);
            if (i % SAMPLE == 0) samples.push_back(p);
        }
    }
    uint64_t get(uint64_t i) const {
        // select the i-th one of `upper`, starting from the sample below it
        uint64_t pos = samples[i / SAMPLE];
        uint64_t left = i % SAMPLE;
        uint64_t w = pos / 64;
        uint64_t cur = upper[w] & (~uint64_t(0) << (pos % 64));
        for (;;) {
            const uint64_t c = (uint64_t)__builtin_popcountll(cur);
            if (left < c) break;
            left -= c;
            cur = upper[++w];
        }
        for (uint64_t k = 0; k < left; k++) cur &= cur - 1;
        const uint64_t p = w * 64 + (uint64_t)__builtin_ctzll(cur);
        return ((p - i) << l) | get_bits(lower, i * l, l);
    }
    uint64_t size() const { return n; }
    uint64_t bitCount() const { return 64 * (lower.size() + upper.size() + samples.size()) + 128; }
    // Serialized form: n, l, then the lower, upper and sample arrays, each as a count and its words.
    void save(std::ostream &os) const {
        const uint64_t hdr[2] = {n, l};
        os.write((const char *)hdr, 16);
        for (const auto *v : {&lower, &upper, &samples}) {
            const uint64_t c = v->size();
            os.write((const char *)&c, 8);
            os.write((const char *)v->data(), (std::streamsize)(c * 8));
        }
    }
    void load(std::istream &is) {
        uint64_t hdr[2] = {0, 0};
        is.read((char *)hdr, 16);
        if (!is || hdr[1] > 64) throw std::runtime_error("EliasFano: bad header");
        n = hdr[0];
        l = (unsigned)hdr[1];
        for (auto *v : {&lower, &upper, &samples}) {
            uint64_t c = 0;
            is.read((char *)&c, 8);
            if (!is || c > (1ull << 40)) throw std::runtime_error("EliasFano: bad array size");
            v->assign(c, 0);
            is.read((char *)v->data(), (std::streamsize)(c * 8));
        }
        if (!is) throw std::runtime_error("EliasFano: truncated stream");
    }
};

// The pair of sequences (bucket_size_acc, bucket_pos_acc), include/ShockHash.h:654.
class DoubleEF {
    EliasFano cum_keys, position;

  public:
    DoubleEF() = default;
    DoubleEF(const std::vector<uint64_t> &cum, const std::vector<uint64_t> &pos) : cum_keys(cum), position(pos) {}
    void get(uint64_t i, uint64_t &cum_keys_i, uint64_t &cum_keys_next, uint64_t &bit_pos) const {
        cum_keys_i = cum_keys.get(i);
        cum_keys_next = cum_keys.get(i + 1);
        bit_pos = position.get(i);
    }
    uint64_t bitCountCumKeys() const { return cum_keys.bitCount(); }
    uint64_t bitCountPosition() const { return position.bitCount(); }
    void save(std::ostream &os) const {
        cum_keys.save(os);
        position.save(os);
    }
    void load(std::istream &is) {
        cum_keys.load(is);
        position.load(is);
    }
};

}  // namespace shockhash
