// Host data structures of the drop-in classes (test infrastructure): Rice bit vector round trips
// against a plain bit list, Elias-Fano random access, 1-bit retrieval exactness.
#include <cstdio>
#include <random>
#include <vector>

#include "shockhash/BitRetrieval.h"
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
    // retrieval
    {
        std::vector<std::pair<uint64_t, uint8_t>> in;
        for (int i = 0; i < 100000; i++) in.push_back({rng(), (uint8_t)(rng() & 1)});
        BitRetrieval br(in);
        for (auto &p : in) CHECK(br.retrieve(p.first) == p.second);
        CHECK(br.sizeBytes() * 8 < 1.3 * in.size() + 256);
    }
    std::printf("host structures: %d failures\n", fails);
    return fails != 0;
}
