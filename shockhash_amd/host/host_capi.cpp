// libshockhash_host.so: extern "C" handles on instantiations of the drop-in class templates
// (ShockHash, SIMDShockHash, ShockHash2), used by the Python tests and the benchmark.
#include <cstdio>
#include <chrono>
#include <fstream>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../include/SIMDShockHash.hpp"
#include "../include/ShockHash.h"
#include "../include/ShockHash2.h"
#include "../include/ShockHash2Flat.h"

using namespace shockhash;

namespace {

struct Handle {
    virtual ~Handle() = default;
    virtual size_t query(const hash128_t &h) = 0;
    virtual size_t query_string(const std::string &s) = 0;
    virtual const std::vector<uint64_t> &words() const = 0;
    virtual const std::vector<uint64_t> &ribbon_keys() const = 0;
    virtual const std::vector<uint8_t> &ribbon_bits() const = 0;
    virtual const gpu::Forest &forest() const = 0;
    virtual size_t bits() const = 0;
    virtual double host_ms() const = 0;
    virtual const double *host_phases() const = 0;
    virtual void save(std::ostream &os) const = 0;
    double ctor_ms = 0;  // wall time of the class constructor
};

template <class T>
struct HandleT : Handle {
    std::unique_ptr<T> obj;
    explicit HandleT(T *o) : obj(o) {}
    size_t query(const hash128_t &h) override { return (*obj)(h); }
    size_t query_string(const std::string &s) override { return (*obj)(s); }
    const std::vector<uint64_t> &words() const override { return obj->descriptorWords(); }
    const std::vector<uint64_t> &ribbon_keys() const override { return obj->ribbonKeys(); }
    const std::vector<uint8_t> &ribbon_bits() const override { return obj->ribbonBits(); }
    const gpu::Forest &forest() const override { return obj->forest(); }
    size_t bits() const override { return obj->getBits(); }
    double host_ms() const override { return obj->hostEncodeMs(); }
    const double *host_phases() const override { return obj->hostPhaseMs(); }
    void save(std::ostream &os) const override { obj->save(os); }
};

template <class T>
Handle *make_loaded(std::istream &is) {
    std::unique_ptr<T> o(new T());
    o->load(is);
    return new HandleT<T>(o.release());
}

template <class T>
Handle *make_h(std::vector<hash128_t> &h, size_t bucket, size_t threads) {
    return new HandleT<T>(new T(h, bucket, threads));
}
template <class T>
Handle *make_s(const std::vector<std::string> &s, size_t bucket, size_t threads) {
    return new HandleT<T>(new T(s, bucket, threads));
}

#define SHH_LEAVES_SH1(X) X(2) X(4) X(8) X(16) X(24) X(30)
#define SHH_LEAVES_SIMD(X) X(8) X(16) X(24) X(58)
#define SHH_LEAVES_SH2(X) X(4) X(8) X(16) X(24) X(32) X(64) X(100) X(128)
// ROTATION_FITTING instantiations (variant | SHH_ROTATION)
#define SHH_LEAVES_SH1_ROT(X) X(8) X(16) X(30) X(40)
#define SHH_LEAVES_SIMD_ROT(X) X(16) X(58)
constexpr int SHH_ROTATION = 0x100;

template <class Keys, class MakeSH1, class MakeSIMD, class MakeSH2>
Handle *dispatch(int variant, int leaf, Keys &keys, size_t bucket, size_t threads) {
    if (variant == (SHGPU_SHOCKHASH | SHH_ROTATION)) {
#define C(L) if (leaf == L) return MakeSH1::template make<ShockHash<L, true>>(keys, bucket, threads);
        SHH_LEAVES_SH1_ROT(C)
#undef C
    } else if (variant == (SHGPU_SIMD_SHOCKHASH | SHH_ROTATION)) {
#define C(L) if (leaf == L) return MakeSIMD::template make<SIMDShockHash<L, true>>(keys, bucket, threads);
        SHH_LEAVES_SIMD_ROT(C)
#undef C
    } else if (variant == SHGPU_SHOCKHASH) {
#define C(L) if (leaf == L) return MakeSH1::template make<ShockHash<L, false>>(keys, bucket, threads);
        SHH_LEAVES_SH1(C)
#undef C
    } else if (variant == SHGPU_SIMD_SHOCKHASH) {
#define C(L) if (leaf == L) return MakeSIMD::template make<SIMDShockHash<L, false>>(keys, bucket, threads);
        SHH_LEAVES_SIMD(C)
#undef C
    } else if (variant == SHGPU_SHOCKHASH2) {
#define C(L) if (leaf == L) return MakeSH2::template make<ShockHash2<L>>(keys, bucket, threads);
        SHH_LEAVES_SH2(C)
#undef C
    }
    throw gpu::Error(SHGPU_ERR_UNSUPPORTED, "leaf size " + std::to_string(leaf) + " not instantiated in libshockhash_host");
}

struct ByHash {
    template <class T> static Handle *make(std::vector<hash128_t> &h, size_t b, size_t t) {
        if constexpr (std::is_constructible_v<T, std::vector<hash128_t> &, size_t, size_t>) return make_h<T>(h, b, t);
        else { T *o = new T(std::span<hash128_t>(h.data(), h.size()), b, t); return new HandleT<T>(o); }
    }
};
struct ByString {
    template <class T> static Handle *make(std::vector<std::string> &s, size_t b, size_t t) {
        if constexpr (std::is_constructible_v<T, const std::vector<std::string> &, size_t, size_t>) return make_s<T>(s, b, t);
        else { T *o = new T(std::span<const std::string>(s.data(), s.size()), b, t); return new HandleT<T>(o); }
    }
};

void set_err(char *err, size_t len, const char *msg) {
    if (err && len) std::snprintf(err, len, "%s", msg);
}

// Status of the calling thread's last failed construction: the engine's status code for a gpu::Error, -1 for
// any other exception.
thread_local int g_status = 0;
void set_status(const std::exception &e) {
    const auto *ge = dynamic_cast<const gpu::Error *>(&e);
    g_status = ge ? ge->status : -1;
}

// ShockHash2Flat<k> over 64-bit key hashes (the mhc values the reference computes from string keys).
#define SHH_FLAT_K(X) X(16) X(32) X(64) X(100) X(128)
struct FlatHandle {
    virtual ~FlatHandle() = default;
    virtual size_t query(uint64_t mhc) const = 0;
    virtual size_t bits() const = 0;
    virtual void info(uint64_t *out) const = 0;
    virtual const std::vector<uint64_t> &codes() const = 0;
    virtual const double *phases() const = 0;
    double ctor_ms = 0;
};
template <size_t K>
struct FlatHandleT : FlatHandle {
    ShockHash2Flat<K> f;
    explicit FlatHandleT(std::span<const uint64_t> m) : f(m) {}
    size_t query(uint64_t mhc) const override { return f(mhc); }
    size_t bits() const override { return f.getBits(); }
    void info(uint64_t *out) const override {
        out[0] = f.bucketCount(); out[1] = f.layerCount(); out[2] = f.fallbackSeedCount(); out[3] = f.getBits();
    }
    const std::vector<uint64_t> &codes() const override { return f.baseCaseCodes(); }
    const double *phases() const override { return f.phaseTimings(); }
};

}  // namespace

extern "C" {

void *shh_build(int variant, int leaf, const uint64_t *hashes, size_t n, size_t bucket, int threads, char *err,
                size_t errlen) {
    try {
        std::vector<hash128_t> h(n);
        std::memcpy(h.data(), hashes, n * 16);
        const auto t0 = std::chrono::steady_clock::now();
        Handle *hd = dispatch<std::vector<hash128_t>, ByHash, ByHash, ByHash>(variant, leaf, h, bucket, (size_t)threads);
        hd->ctor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return hd;
    } catch (const std::exception &e) {
        set_err(err, errlen, e.what());
        set_status(e);
        return nullptr;
    }
}

void *shh_build_strings(int variant, int leaf, const char *data, const uint64_t *offsets, size_t n, size_t bucket,
                        int threads, char *err, size_t errlen) {
    try {
        std::vector<std::string> s(n);
        for (size_t i = 0; i < n; i++) s[i].assign(data + offsets[i], offsets[i + 1] - offsets[i]);
        const auto t0 = std::chrono::steady_clock::now();
        Handle *hd = dispatch<std::vector<std::string>, ByString, ByString, ByString>(variant, leaf, s, bucket, (size_t)threads);
        hd->ctor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return hd;
    } catch (const std::exception &e) {
        set_err(err, errlen, e.what());
        set_status(e);
        return nullptr;
    }
}

void shh_free(void *h) { delete (Handle *)h; }

// Status of the calling thread's last failed shh_build / shh_build_strings (SHGPU_* code, -1 otherwise).
int shh_last_status(void) { return g_status; }

// Writes the function to a file (the class's save()); 0 on success.
int shh_save(void *h, const char *path, char *err, size_t errlen) {
    try {
        std::ofstream os(path, std::ios::binary);
        ((Handle *)h)->save(os);
        os.close();
        if (!os) throw std::runtime_error("write failed");
        return 0;
    } catch (const std::exception &e) {
        set_err(err, errlen, e.what());
        return 1;
    }
}

// A function of the given class and leaf size read from a file (the class's load()); queries only.
void *shh_load(int variant, int leaf, const char *path, char *err, size_t errlen) {
    try {
        std::ifstream is(path, std::ios::binary);
        if (!is) throw std::runtime_error(std::string("cannot open ") + path);
        if (variant == (SHGPU_SHOCKHASH | SHH_ROTATION)) {
#define C(L) if (leaf == L) return make_loaded<ShockHash<L, true>>(is);
            SHH_LEAVES_SH1_ROT(C)
#undef C
        } else if (variant == (SHGPU_SIMD_SHOCKHASH | SHH_ROTATION)) {
#define C(L) if (leaf == L) return make_loaded<SIMDShockHash<L, true>>(is);
            SHH_LEAVES_SIMD_ROT(C)
#undef C
        } else if (variant == SHGPU_SHOCKHASH) {
#define C(L) if (leaf == L) return make_loaded<ShockHash<L, false>>(is);
            SHH_LEAVES_SH1(C)
#undef C
        } else if (variant == SHGPU_SIMD_SHOCKHASH) {
#define C(L) if (leaf == L) return make_loaded<SIMDShockHash<L, false>>(is);
            SHH_LEAVES_SIMD(C)
#undef C
        } else if (variant == SHGPU_SHOCKHASH2) {
#define C(L) if (leaf == L) return make_loaded<ShockHash2<L>>(is);
            SHH_LEAVES_SH2(C)
#undef C
        }
        throw std::runtime_error("leaf size " + std::to_string(leaf) + " not instantiated in libshockhash_host");
    } catch (const std::exception &e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}

void shh_query(void *h, const uint64_t *hashes, size_t n, uint64_t *out) {
    Handle *x = (Handle *)h;
    for (size_t i = 0; i < n; i++) out[i] = x->query(hash128_t(hashes[2 * i], hashes[2 * i + 1]));
}

void shh_query_strings(void *h, const char *data, const uint64_t *offsets, size_t n, uint64_t *out) {
    Handle *x = (Handle *)h;
    for (size_t i = 0; i < n; i++) out[i] = x->query_string(std::string(data + offsets[i], offsets[i + 1] - offsets[i]));
}

size_t shh_nwords(void *h) { return ((Handle *)h)->words().size(); }
void shh_words(void *h, uint64_t *out) {
    const auto &w = ((Handle *)h)->words();
    std::memcpy(out, w.data(), w.size() * 8);
}
size_t shh_ribbon_size(void *h) { return ((Handle *)h)->ribbon_keys().size(); }
void shh_ribbon(void *h, uint64_t *keys, uint8_t *bits) {
    const auto &k = ((Handle *)h)->ribbon_keys();
    const auto &b = ((Handle *)h)->ribbon_bits();
    std::memcpy(keys, k.data(), k.size() * 8);
    std::memcpy(bits, b.data(), b.size());
}
// Keep the retrieval input of later builds (for shh_ribbon); off by default.
void shh_keep_retrieval_input(int keep) { gpu::keep_retrieval_input = keep != 0; }
size_t shh_nnodes(void *h) { return ((Handle *)h)->forest().codes.size(); }
void shh_codes(void *h, uint64_t *out) {
    const auto &c = ((Handle *)h)->forest().codes;
    std::memcpy(out, c.data(), c.size() * 8);
}
size_t shh_bits(void *h) { return ((Handle *)h)->bits(); }
// After the search: Golomb-Rice + Elias-Fano encoding, leaf orientation and retrieval construction; then the
// whole constructor (ms).
void shh_phase_timings(void *h, double *out4) {
    const double *p = ((Handle *)h)->host_phases();
    out4[0] = p[0]; out4[1] = p[1]; out4[2] = p[2];
    out4[3] = ((Handle *)h)->ctor_ms;
}
// Sort + bucketing, tree setup, seed search (host clock), host phases after the search, and the persistent
// search kernel alone (HIP events), in ms.
void shh_timings(void *h, double *out5) {
    const auto &f = ((Handle *)h)->forest();
    out5[0] = f.ms[0]; out5[1] = f.ms[1]; out5[2] = f.ms[2];
    out5[3] = ((Handle *)h)->host_ms();
    out5[4] = f.ms[3];
}

// ShockHash2Flat<k> over 64-bit key hashes (the mhc values the reference computes from string keys).
void *shh_flat_build(int k, const uint64_t *mhcs, size_t n, char *err, size_t errlen) {
    try {
        const auto t0 = std::chrono::steady_clock::now();
        FlatHandle *h = nullptr;
        std::span<const uint64_t> m(mhcs, n);
#define C(K) if (k == K) h = new FlatHandleT<K>(m);
        SHH_FLAT_K(C)
#undef C
        if (!h) throw gpu::Error(SHGPU_ERR_UNSUPPORTED, "ShockHash2Flat k " + std::to_string(k) + " not instantiated in libshockhash_host");
        h->ctor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return h;
    } catch (const std::exception &e) {
        set_err(err, errlen, e.what());
        set_status(e);
        return nullptr;
    }
}
void shh_flat_free(void *h) { delete (FlatHandle *)h; }
// out[i] = the function's value of mhcs[i]; keys outside the set may raise inside (they are not queried by tests)
int shh_flat_query(void *h, const uint64_t *mhcs, size_t n, uint64_t *out, char *err, size_t errlen) {
    try {
        for (size_t i = 0; i < n; i++) out[i] = ((FlatHandle *)h)->query(mhcs[i]);
        return 0;
    } catch (const std::exception &e) {
        set_err(err, errlen, e.what());
        return 1;
    }
}
// nbuckets, layers, base-case seeds kept in the fallback map, getBits(), constructor ms
void shh_flat_info(void *h, double *out5) {
    uint64_t v[4];
    ((FlatHandle *)h)->info(v);
    for (int i = 0; i < 4; i++) out5[i] = (double)v[i];
    out5[4] = ((FlatHandle *)h)->ctor_ms;
}
// construction phases in ms: threshold layers and bumping, fallback function, placement of every key, device
// base-case searches + orientation + retrieval, and the batched base-case search alone
void shh_flat_phases(void *h, double *out5) {
    const double *p = ((FlatHandle *)h)->phases();
    for (int i = 0; i < 5; i++) out5[i] = p[i];
}
size_t shh_flat_nbuckets(void *h) { return ((FlatHandle *)h)->codes().size(); }
void shh_flat_codes(void *h, uint64_t *out) {
    const auto &c = ((FlatHandle *)h)->codes();
    std::memcpy(out, c.data(), c.size() * 8);
}

// ShockHash2Flat's string hash (MurmurHash64A with a seed; the class uses seed 0), for checks against its
// published verification value.
uint64_t shh_murmur64a(const void *key, size_t len, uint64_t seed) { return flat_detail::murmurHash64A(key, len, seed); }

// Compile-time Golomb-Rice memo of a class (for comparison with the oracle's run-time formula).
int shh_memo(int variant, int leaf, uint64_t *out) {
    if (variant == SHGPU_SHOCKHASH || variant == SHGPU_SIMD_SHOCKHASH) {
#define C(L) if (leaf == L) { auto m = golomb_memo<L, false>(); std::memcpy(out, m.data(), m.size() * 8); return 0; }
        SHH_LEAVES_SH1(C) SHH_LEAVES_SIMD(C)
#undef C
    } else {
#define C(L) if (leaf == L) { auto m = golomb_memo<L, true>(); std::memcpy(out, m.data(), m.size() * 8); return 0; }
        SHH_LEAVES_SH2(C)
#undef C
    }
    return 1;
}

}  // extern "C"
