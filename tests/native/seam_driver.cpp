// Test driver of libShockHash2Precompiled.so through the reference's C++ signatures
// (include/shockhash/ShockHash2-precompiled.h:8-10): reads work from stdin, writes results to stdout, both as
// little-endian binary.
//   seam_driver construct : per leaf  m (u32) + m keys (u64)  ->  code (u64), then m (key u64, bit u8) appended
//                           to ribbonInput, then m cells from shockhash2query(m, code, key, bit) (u64)
//   seam_driver query     : per tuple param, seed, key, retrieved (u64 each)  ->  shockhash2query (u64)
//   seam_driver trivial   : construct for param 0, 1 and 129 (no GPU call): prints the codes and how many
//                           ribbon entries were appended
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#include "shockhash/ShockHash2-precompiled.h"

template <class T> static bool rd(T &v) { return std::fread(&v, sizeof(T), 1, stdin) == 1; }
template <class T> static void wr(const T &v) { std::fwrite(&v, sizeof(T), 1, stdout); }

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    if (!std::strcmp(argv[1], "query")) {
        uint64_t p, s, k, r;
        while (rd(p) && rd(s) && rd(k) && rd(r)) wr((uint64_t)shockhash::shockhash2query(p, s, k, r));
        return 0;
    }
    if (!std::strcmp(argv[1], "trivial")) {
        for (size_t param : {0, 1, 129}) {
            std::vector<uint64_t> keys(param, 0x1234);
            std::vector<std::pair<uint64_t, uint8_t>> ribbon;
            const size_t code = shockhash::shockhash2construct(param, keys, ribbon);
            std::printf("%zu %zu %zu\n", param, code, ribbon.size());
        }
        return 0;
    }
    if (!std::strcmp(argv[1], "construct")) {
        uint32_t m;
        std::vector<std::pair<uint64_t, uint8_t>> ribbon;  // one vector for all leaves: entries are appended
        while (rd(m)) {
            std::vector<uint64_t> keys(m);
            for (auto &k : keys)
                if (!rd(k)) return 3;
            const size_t before = ribbon.size();
            const size_t code = shockhash::shockhash2construct(m, keys, ribbon);
            wr((uint64_t)code);
            for (size_t i = before; i < ribbon.size(); i++) {
                wr(ribbon[i].first);
                wr(ribbon[i].second);
            }
            for (size_t i = before; i < ribbon.size(); i++)
                wr((uint64_t)shockhash::shockhash2query(m, code, ribbon[i].first, ribbon[i].second));
        }
        return 0;
    }
    return 2;
}
