// Exhaustive check (test infrastructure): the oracle's three sequential pseudoforest predicates
// (CuckooUnionFind, UnionFind with tree flags, TinyBinaryCuckooHashTable insertion) must reject an
// edge sequence at exactly the first prefix that is not a pseudoforest, where "pseudoforest" is the
// textbook definition: every connected component has at most as many edges as vertices.
// Covers every sequence of up to K edges (self-loops and parallel edges included) on N vertices.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" size_t orc_pseudoforest_prefix(int pred, const uint32_t *c1, const uint32_t *c2, size_t n, size_t n_cells);

static size_t textbook_prefix(const uint32_t *a, const uint32_t *b, size_t n, size_t nc) {
    for (size_t k = 1; k <= n; k++) {
        std::vector<int> comp(nc);
        for (size_t v = 0; v < nc; v++) comp[v] = (int)v;
        auto find = [&](int x) { while (comp[x] != x) x = comp[x]; return x; };
        for (size_t i = 0; i < k; i++) {
            int r1 = find(a[i]), r2 = find(b[i]);
            if (r1 != r2) comp[r1] = r2;
        }
        std::vector<int> V(nc, 0), E(nc, 0);
        for (size_t v = 0; v < nc; v++) V[find(v)]++;
        for (size_t i = 0; i < k; i++) E[find(a[i])]++;
        for (size_t v = 0; v < nc; v++) if (E[v] > V[v]) return k - 1;
    }
    return n;
}

int main(int argc, char **argv) {
    const int maxn = argc > 1 ? atoi(argv[1]) : 4, maxk = argc > 2 ? atoi(argv[2]) : 5;
    long total = 0, bad = 0;
    for (int nc = 1; nc <= maxn; nc++) {
        for (int k = 1; k <= maxk; k++) {
            long combos = 1;
            for (int i = 0; i < k; i++) combos *= nc * nc;
            std::vector<uint32_t> a(k), b(k);
            for (long c = 0; c < combos; c++) {
                long t = c;
                for (int i = 0; i < k; i++) { int e = (int)(t % (nc * nc)); t /= nc * nc; a[i] = e / nc; b[i] = e % nc; }
                const size_t want = textbook_prefix(a.data(), b.data(), k, nc);
                for (int p = 0; p < 3; p++)
                    if (orc_pseudoforest_prefix(p, a.data(), b.data(), k, nc) != want) bad++;
                total++;
            }
        }
    }
    printf("sequences %ld mismatches %ld\n", total, bad);
    return bad != 0;
}
