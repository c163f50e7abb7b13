// Host data structures of the drop-in classes (test infrastructure): Rice bit vector round trips
// against a plain bit list, Elias-Fano random access, the ribbon retrieval's query on structures solved here
// (an independent sequential solver of the documented layout), serialization round trips.
#include <cstdio>
#include <random>
#include <vector>

#include <sstream>

#include "shockhash/RibbonRetrieval.h"
#include "shockhash/EliasFano.h"
#include "shockhash/RiceBitVector.h"

using namespace shockhash;

static int fails = 0;
#define CHECK(c) do { if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); fails++; } } while (0)

int main() {
    std::mt19937_64 rng(7);
    // Rice: fixed parts of random widths then unary parts; compare with an explicit bit list.
    for (int trial = 0; trial < 200; trial++) {
        RiceBitVector::Builder a, b;
        std::vector<int> bits;
        std::vector<std::pair<uint64_t, int>> fixed;
        std::vector<uint32_t> unary;
        const int nf = rng() % 40 + 1;
        for (int i = 0; i < nf; i++) {
            const int k = rng() % 40;
            const uint64_t v = rng();
            fixed.push_back({v, k});
            a.appendFixed(v, k);
            for (int j = 0; j < k; j++) bits.push_back((v >> j) & 1);
            unary.push_back(rng() % 70);
        }
        a.appendUnaryAll(unary);
        for (uint32_t u : unary) { for (uint32_t j = 0; j < u; j++) bits.push_back(0); bits.push_back(1); }
        // concatenation at a non-byte-aligned offset
        b.appendFixed(5, 3);
        b.appendRiceBitVector(a);
        CHECK(b.getBits() == bits.size() + 3);
        std::vector<uint64_t> w = b.build();
        CHECK(w.size() == RiceBitVector::words_for(bits.size() + 3));
        for (size_t i = 0; i < bits.size(); i++) CHECK((int)((w[(i + 3) / 64] >> ((i + 3) % 64)) & 1) == bits[i]);
        for (size_t i = bits.size() + 3; i < w.size() * 64; i++) CHECK(((w[i / 64] >> (i % 64)) & 1) == 0);
        // reading back: fixed from bit 3, unary after all fixed parts
        RiceBitVector rv(std::move(w));
        auto r = rv.reader();
        uint64_t fixed_len = 0;
        for (auto &f : fixed) fixed_len += f.second;
        r.readReset(3, fixed_len);
        for (size_t i = 0; i < fixed.size(); i++) {
            const int k = fixed[i].second;
            const uint64_t lowmask = k >= 64 ? ~0ull : ((1ull << k) - 1);
            const uint64_t want = ((uint64_t)unary[i] << k) | (fixed[i].first & lowmask);
            CHECK(r.readNext(k) == want);
        }
    }
    // skipSubtree skips whole unary codes and fixed bits
    {
        RiceBitVector::Builder b;
        for (int i = 0; i < 10; i++) b.appendFixed(i, 4);
        std::vector<uint32_t> u = {3, 0, 130, 2, 7, 1, 64, 0, 5, 9};
        b.appendUnaryAll(u);
        RiceBitVector rv(b.build());
        auto r = rv.reader();
        r.readReset(0, 40);
        CHECK(r.readNext(4) == ((3ull << 4) | 0));
        r.skipSubtree(3, 12);  // nodes 1..3
        CHECK(r.readNext(4) == ((7ull << 4) | 4));
        r.skipSubtree(4, 16);
        CHECK(r.readNext(4) == ((9ull << 4) | 9));
    }
    // Elias-Fano
    for (int trial = 0; trial < 50; trial++) {
        std::vector<uint64_t> v(rng() % 3000 + 1);
        uint64_t x = 0;
        for (auto &e : v) { x += rng() % (trial % 2 ? 5000 : 3); e = x; }
        EliasFano ef(v);
        for (size_t i = 0; i < v.size(); i++) CHECK(ef.get(i) == v[i]);
    }
    // ribbon retrieval: a sequential solver of the layout documented in include/shockhash_amd.h (shards of ~2048
    // keys, n + ceil(n/20) slots, rows of 64 slots), then every key's query through RibbonRetrieval
    for (size_t n : {size_t(0), size_t(1), size_t(5), size_t(100), size_t(2047), size_t(20000), size_t(100000)}) {
        std::vector<std::pair<uint64_t, uint8_t>> in;
        for (size_t i = 0; i < n; i++) in.push_back({rng(), (uint8_t)(rng() & 1)});
        const uint64_t S = n == 0 ? 1 : (n + 2047) / 2048;
        std::vector<std::vector<size_t>> shard_keys(S);
        for (size_t i = 0; i < n; i++)
            shard_keys[fastrange64(remix(in[i].first + SHGPU_RIBBON_SALT_SHARD), S)].push_back(i);
        std::vector<uint32_t> offs(S + 1, 0);
        for (uint64_t s = 0; s < S; s++) {
            const uint32_t ns = (uint32_t)shard_keys[s].size();
            offs[s + 1] = offs[s] + (ns == 0 ? 0 : std::max<uint32_t>(64, ns + (ns + 19) / 20));
        }
        std::vector<uint8_t> seeds(S, 0);
        std::vector<uint64_t> words(offs[S] / 64 + 2, 0);
        for (uint64_t s = 0; s < S; s++) {
            const uint32_t m = offs[s + 1] - offs[s];
            if (!m) continue;
            for (;; seeds[s]++) {
                std::vector<uint64_t> rows(m, 0);
                std::vector<uint8_t> bb(m, 0);
                bool ok = true;
                for (size_t i : shard_keys[s]) {
                    const uint64_t g = remix(remix(in[i].first + SHGPU_RIBBON_SALT_SHARD) + (seeds[s] + 1ull) * SHGPU_RIBBON_SEED_STEP);
                    uint64_t st = fastrange64(g, m - 63), row = remix(g ^ SHGPU_RIBBON_SALT_ROW) | 1;
                    uint8_t b = in[i].second;
                    while (true) {
                        if (!rows[st]) { rows[st] = row; bb[st] = b; break; }
                        row ^= rows[st]; b ^= bb[st];
                        if (!row) { ok = b == 0; break; }
                        const int t = __builtin_ctzll(row);
                        row >>= t; st += t;
                    }
                    if (!ok) break;
                }
                if (!ok) { CHECK(seeds[s] < 255); continue; }
                std::vector<uint8_t> z(m + 64, 0);
                for (int64_t p = (int64_t)m - 1; p >= 0; p--) {
                    if (!rows[p]) continue;
                    uint8_t v = bb[p];
                    for (int j = 1; j < 64; j++) if ((rows[p] >> j) & 1) v ^= z[p + j];
                    z[p] = v;
                }
                for (uint32_t p = 0; p < m; p++)
                    if (z[p]) words[(offs[s] + p) / 64] |= 1ull << ((offs[s] + p) % 64);
                break;
            }
        }
        RibbonRetrieval rr(offs, seeds, words);
        for (auto &p : in) CHECK(rr.retrieve(p.first) == p.second);
        if (n >= 20000) CHECK(rr.sizeBytes() * 8 < 1.08 * n + 256);
        std::stringstream ss;
        rr.save(ss);
        RibbonRetrieval back;
        back.load(ss);
        for (auto &p : in) CHECK(back.retrieve(p.first) == p.second);
        CHECK(back.sizeBytes() == rr.sizeBytes());
    }
    // Rice bit vector and Elias-Fano serialization round trips
    {
        RiceBitVector::Builder b;
        for (int i = 0; i < 1000; i++) b.appendFixed(rng(), (int)(rng() % 30));
        RiceBitVector rv(b.build()), back;
        std::stringstream ss;
        rv.save(ss);
        back.load(ss);
        CHECK(back.words() == rv.words());
        std::stringstream s3;
        RiceBitVector viaop;
        s3 << rv;
        s3 >> viaop;
        CHECK(viaop.words() == rv.words());
        std::vector<uint64_t> v(5000), w(5000);
        uint64_t x = 0, y = 0;
        for (size_t i = 0; i < v.size(); i++) { x += rng() % 4000; y += rng() % 50000; v[i] = x; w[i] = y; }
        DoubleEF ef(v, w), ef2;
        std::stringstream s2;
        ef.save(s2);
        ef2.load(s2);
        for (size_t i = 0; i + 1 < v.size(); i++) {
            uint64_t a, bb, c, a2, b2, c2;
            ef.get(i, a, bb, c);
            ef2.get(i, a2, b2, c2);
            CHECK(a == a2 && bb == b2 && c == c2 && a == v[i] && c == w[i]);
        }
    }
    std::printf("host structures: %d failures\n", fails);
    return fails != 0;
}
