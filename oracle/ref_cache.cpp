// Test infrastructure only: the reference's calibration-cache building blocks
// built from the third-party headers it vendors (include/xxhash.h, XXH3 with
// the streaming Hash128 use of dummest-calibration.hh:106-146; include/cereal,
// the BinaryArchive of unordered_map<string, CacheEntry>,
// dummest-calibration.hh:148-181 and :267-336), so the engine's own hash and
// cache-file code can be checked against them. CacheEntry itself sits in
// dummest-calibration.hh, which needs a configured Kokkos; the entry here is
// 27 doubles archived one after another, as CacheEntry::serialize does with
// FILTER_PASS on.
#define XXH_INLINE_ALL
#include "xxhash.h"

#include <cereal/archives/binary.hpp>
#include <cereal/types/string.hpp>
#include <cereal/types/unordered_map.hpp>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_map>

namespace {

struct Entry27 {
    double v[27];
    template <class Archive>
    void serialize(Archive &ar) {
        for (double &x : v) ar(x);
    }
};

}  // namespace

extern "C" {

// XXH3_128bits over the concatenation of n pieces fed one update each.
void ref_xxh3_128_pieces(const char *data, const int64_t *sizes, int n, uint64_t out[2]) {
    XXH3_state_t *st = XXH3_createState();
    XXH3_128bits_reset(st);
    for (int i = 0; i < n; i++) {
        XXH3_128bits_update(st, data, (size_t)sizes[i]);
        data += sizes[i];
    }
    const XXH128_hash_t h = XXH3_128bits_digest(st);
    XXH3_freeState(st);
    out[0] = h.low64;
    out[1] = h.high64;
}

// Writes n entries (keys NUL-terminated at key_stride) through cereal.
int ref_cache_write(const char *path, int64_t n, const char *keys, int key_stride, const double *values) {
    std::unordered_map<std::string, Entry27> m;
    for (int64_t i = 0; i < n; i++) {
        Entry27 e;
        std::memcpy(e.v, values + 27 * i, sizeof e.v);
        m[std::string(keys + i * key_stride)] = e;
    }
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) return -1;
    cereal::BinaryOutputArchive ar(out);
    ar(m);
    return 0;
}

// Reads a cache file through cereal; returns the entry count (copies up to
// cap entries in the map's iteration order) or -1 when cereal throws.
int64_t ref_cache_read(const char *path, int64_t cap, char *keys, int key_stride, double *values) {
    std::unordered_map<std::string, Entry27> m;
    try {
        std::ifstream in(path, std::ios::binary);
        cereal::BinaryInputArchive ar(in);
        ar(m);
    } catch (const std::exception &) {
        return -1;
    }
    int64_t i = 0;
    for (const auto &kv : m) {
        if (i >= cap) break;
        std::strncpy(keys + i * key_stride, kv.first.c_str(), (size_t)key_stride - 1);
        keys[i * key_stride + key_stride - 1] = 0;
        std::memcpy(values + 27 * i, kv.second.v, sizeof kv.second.v);
        i++;
    }
    return (int64_t)m.size();
}

}  // extern "C"
