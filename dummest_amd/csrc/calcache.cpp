// Calibration cache compatible with the reference engine's (ProfileCache,
// dummest-calibration.hh:106-336), shared with bin/dummest's cache directory.
//
// * Key: the XXH3 128-bit hash (seed 0, default secret) of the byte sequence
//   the reference feeds its Hash128 (computeCacheKey, :201-226), printed as
//   32 hex digits, high half first (Hash128::to_string, :134-143).
// * File: cache.bin in the first usable directory of $DUMMEST_CACHE_DIR,
//   $XDG_CACHE_HOME/.dummest-cache, $HOME/.dummest-cache, ./.dummest-cache,
//   $TMPDIR/.dummest-cache, /tmp/.dummest-cache (getCacheFilePath, :233-265),
//   in the layout of cereal's BinaryArchive for unordered_map<string,
//   CacheEntry> (:259-336): a 64-bit entry count, then per entry a 64-bit key
//   length, the key bytes and 27 doubles (CacheEntry with FILTER_PASS, :148-181).
//   Native byte order, as the reference writes it.
// * XXH3: written here from the published algorithm (xxHash 0.8 XXH3_128bits:
//   short inputs 0-16, 17-128, 129-240 bytes, and the striped long loop with
//   1 KiB blocks); only the 192-byte default secret is the spec's table.
#include <dlfcn.h>
#include <unistd.h>

#include <algorithm>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "dm_host.h"
#include "../../include/dummest_amd.h"

namespace dm {

namespace {

constexpr uint64_t kP32_1 = 0x9E3779B1u, kP32_2 = 0x85EBCA77u, kP32_3 = 0xC2B2AE3Du;
constexpr uint64_t kP64_1 = 0x9E3779B185EBCA87ull, kP64_2 = 0xC2B2AE3D27D4EB4Full, kP64_3 = 0x165667B19E3779F9ull;
constexpr uint64_t kP64_4 = 0x85EBCA77C2B2AE63ull, kP64_5 = 0x27D4EB2F165667C5ull;
constexpr uint64_t kMX1 = 0x165667919E3779F9ull, kMX2 = 0x9FB21C651E98DF25ull;

// XXH3's default secret (the xxHash specification's kSecret table).
const uint8_t kSecret[192] = {
    0xb8, 0xfe, 0x6c, 0x39, 0x23, 0xa4, 0x4b, 0xbe, 0x7c, 0x01, 0x81, 0x2c, 0xf7, 0x21, 0xad, 0x1c,
    0xde, 0xd4, 0x6d, 0xe9, 0x83, 0x90, 0x97, 0xdb, 0x72, 0x40, 0xa4, 0xa4, 0xb7, 0xb3, 0x67, 0x1f,
    0xcb, 0x79, 0xe6, 0x4e, 0xcc, 0xc0, 0xe5, 0x78, 0x82, 0x5a, 0xd0, 0x7d, 0xcc, 0xff, 0x72, 0x21,
    0xb8, 0x08, 0x46, 0x74, 0xf7, 0x43, 0x24, 0x8e, 0xe0, 0x35, 0x90, 0xe6, 0x81, 0x3a, 0x26, 0x4c,
    0x3c, 0x28, 0x52, 0xbb, 0x91, 0xc3, 0x00, 0xcb, 0x88, 0xd0, 0x65, 0x8b, 0x1b, 0x53, 0x2e, 0xa3,
    0x71, 0x64, 0x48, 0x97, 0xa2, 0x0d, 0xf9, 0x4e, 0x38, 0x19, 0xef, 0x46, 0xa9, 0xde, 0xac, 0xd8,
    0xa8, 0xfa, 0x76, 0x3f, 0xe3, 0x9c, 0x34, 0x3f, 0xf9, 0xdc, 0xbb, 0xc7, 0xc7, 0x0b, 0x4f, 0x1d,
    0x8a, 0x51, 0xe0, 0x4b, 0xcd, 0xb4, 0x59, 0x31, 0xc8, 0x9f, 0x7e, 0xc9, 0xd9, 0x78, 0x73, 0x64,
    0xea, 0xc5, 0xac, 0x83, 0x34, 0xd3, 0xeb, 0xc3, 0xc5, 0x81, 0xa0, 0xff, 0xfa, 0x13, 0x63, 0xeb,
    0x17, 0x0d, 0xdd, 0x51, 0xb7, 0xf0, 0xda, 0x49, 0xd3, 0x16, 0x55, 0x26, 0x29, 0xd4, 0x68, 0x9e,
    0x2b, 0x16, 0xbe, 0x58, 0x7d, 0x47, 0xa1, 0xfc, 0x8f, 0xf8, 0xb8, 0xd1, 0x7a, 0xd0, 0x31, 0xce,
    0x45, 0xcb, 0x3a, 0x8f, 0x95, 0x16, 0x04, 0x28, 0xaf, 0xd7, 0xfb, 0xca, 0xbb, 0x4b, 0x40, 0x7e,
};

inline uint64_t rd64(const uint8_t *p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}
inline uint32_t rd32(const uint8_t *p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
inline Hash128 mul128(uint64_t a, uint64_t b) {
    const unsigned __int128 p = (unsigned __int128)a * b;
    return Hash128{(uint64_t)p, (uint64_t)(p >> 64)};
}
inline uint64_t mulFold(uint64_t a, uint64_t b) {
    const Hash128 p = mul128(a, b);
    return p.lo ^ p.hi;
}
inline uint64_t xorShift(uint64_t v, int s) { return v ^ (v >> s); }
inline uint64_t avalancheXXH64(uint64_t h) {
    h ^= h >> 33;
    h *= kP64_2;
    h ^= h >> 29;
    h *= kP64_3;
    h ^= h >> 32;
    return h;
}
inline uint64_t avalanche(uint64_t h) { return xorShift(xorShift(h, 37) * kMX1, 32); }
inline uint64_t mix16(const uint8_t *in, const uint8_t *sec) {
    return mulFold(rd64(in) ^ rd64(sec), rd64(in + 8) ^ rd64(sec + 8));
}
inline void mix32(Hash128 &acc, const uint8_t *a, const uint8_t *b, const uint8_t *sec) {
    acc.lo += mix16(a, sec);
    acc.lo ^= rd64(b) + rd64(b + 8);
    acc.hi += mix16(b, sec + 16);
    acc.hi ^= rd64(a) + rd64(a + 8);
}
inline Hash128 finish(const Hash128 &acc, uint64_t len) {
    const uint64_t lo = acc.lo + acc.hi;
    const uint64_t hi = acc.lo * kP64_1 + acc.hi * kP64_4 + len * kP64_2;
    return Hash128{avalanche(lo), 0 - avalanche(hi)};
}
// one 64-byte stripe into the eight accumulators
inline void stripe(uint64_t acc[8], const uint8_t *in, const uint8_t *sec) {
    for (int l = 0; l < 8; l++) {
        const uint64_t v = rd64(in + 8 * l);
        const uint64_t k = v ^ rd64(sec + 8 * l);
        acc[l ^ 1] += v;
        acc[l] += (uint64_t)(uint32_t)k * (k >> 32);
    }
}
inline void scramble(uint64_t acc[8], const uint8_t *sec) {
    for (int l = 0; l < 8; l++) acc[l] = (xorShift(acc[l], 47) ^ rd64(sec + 8 * l)) * kP32_1;
}
inline uint64_t merge(const uint64_t acc[8], const uint8_t *sec, uint64_t start) {
    uint64_t r = start;
    for (int i = 0; i < 4; i++) r += mulFold(acc[2 * i] ^ rd64(sec + 16 * i), acc[2 * i + 1] ^ rd64(sec + 16 * i + 8));
    return avalanche(r);
}

}  // namespace

Hash128 xxh3_128(const void *data, size_t len) {
    const uint8_t *in = static_cast<const uint8_t *>(data);
    const uint8_t *s = kSecret;
    if (len <= 16) {
        if (len > 8) {
            const uint64_t flipLo = rd64(s + 32) ^ rd64(s + 40), flipHi = rd64(s + 48) ^ rd64(s + 56);
            uint64_t lo = rd64(in), hi = rd64(in + len - 8);
            Hash128 m = mul128(lo ^ hi ^ flipLo, kP64_1);
            m.lo += (uint64_t)(len - 1) << 54;
            hi ^= flipHi;
            m.hi += hi + (uint64_t)(uint32_t)hi * (kP32_2 - 1);
            m.lo ^= __builtin_bswap64(m.hi);
            Hash128 h = mul128(m.lo, kP64_2);
            h.hi += m.hi * kP64_2;
            return Hash128{avalanche(h.lo), avalanche(h.hi)};
        }
        if (len >= 4) {
            const uint64_t v = (uint64_t)rd32(in) + ((uint64_t)rd32(in + len - 4) << 32);
            Hash128 m = mul128(v ^ (rd64(s + 16) ^ rd64(s + 24)), kP64_1 + ((uint64_t)len << 2));
            m.hi += m.lo << 1;
            m.lo ^= m.hi >> 3;
            m.lo = xorShift(xorShift(m.lo, 35) * kMX2, 28);
            m.hi = avalanche(m.hi);
            return m;
        }
        if (len > 0) {
            const uint32_t c1 = in[0], c2 = in[len >> 1], c3 = in[len - 1];
            const uint32_t lo = (c1 << 16) | (c2 << 24) | c3 | ((uint32_t)len << 8);
            const uint32_t sw = __builtin_bswap32(lo);
            const uint32_t hi = (sw << 13) | (sw >> 19);
            return Hash128{avalancheXXH64(lo ^ (uint64_t)(rd32(s) ^ rd32(s + 4))),
                           avalancheXXH64(hi ^ (uint64_t)(rd32(s + 8) ^ rd32(s + 12)))};
        }
        return Hash128{avalancheXXH64(rd64(s + 64) ^ rd64(s + 72)), avalancheXXH64(rd64(s + 80) ^ rd64(s + 88))};
    }
    if (len <= 128) {
        Hash128 acc{len * kP64_1, 0};
        if (len > 32) {
            if (len > 64) {
                if (len > 96) mix32(acc, in + 48, in + len - 64, s + 96);
                mix32(acc, in + 32, in + len - 48, s + 64);
            }
            mix32(acc, in + 16, in + len - 32, s + 32);
        }
        mix32(acc, in, in + len - 16, s);
        return finish(acc, len);
    }
    if (len <= 240) {
        Hash128 acc{len * kP64_1, 0};
        for (size_t i = 32; i < 160; i += 32) mix32(acc, in + i - 32, in + i - 16, s + i - 32);
        acc.lo = avalanche(acc.lo);
        acc.hi = avalanche(acc.hi);
        for (size_t i = 160; i <= len; i += 32) mix32(acc, in + i - 32, in + i - 16, s + 3 + i - 160);
        mix32(acc, in + len - 16, in + len - 32, s + 136 - 17 - 16);
        return finish(acc, len);
    }
    // long inputs: 64-byte stripes, 16 per 1 KiB block, scrambled between blocks
    uint64_t acc[8] = {kP32_3, kP64_1, kP64_2, kP64_3, kP64_4, kP32_2, kP64_5, kP32_1};
    const size_t perBlock = (sizeof(kSecret) - 64) / 8, blockLen = 64 * perBlock;
    const size_t blocks = (len - 1) / blockLen;
    for (size_t b = 0; b < blocks; b++) {
        for (size_t k = 0; k < perBlock; k++) stripe(acc, in + b * blockLen + 64 * k, s + 8 * k);
        scramble(acc, s + sizeof(kSecret) - 64);
    }
    const size_t rest = ((len - 1) - blockLen * blocks) / 64;
    for (size_t k = 0; k < rest; k++) stripe(acc, in + blocks * blockLen + 64 * k, s + 8 * k);
    stripe(acc, in + len - 64, s + sizeof(kSecret) - 64 - 7);
    return Hash128{merge(acc, s + 11, len * kP64_1), merge(acc, s + sizeof(kSecret) - 64 - 11, ~(len * kP64_2))};
}

std::string hashHex(const Hash128 &h) {
    std::ostringstream ss;
    ss << std::hex << std::setfill('0') << std::setw(16) << h.hi << std::setw(16) << h.lo;
    return ss.str();
}

// The reference hashes its own executable (getBinaryHash, :148-159), which
// holds its whole engine; here the engine is the library, so both files are
// hashed (executable, then library).
std::string binaryHash() {
    std::string bytes;
    auto slurp = [&](const std::string &path) {
        std::ifstream f(path, std::ios::binary);
        if (f) bytes.append(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    };
    slurp("/proc/self/exe");
    Dl_info info{};
    if (dladdr((void *)&binaryHash, &info) && info.dli_fname) slurp(info.dli_fname);
    return hashHex(xxh3_128(bytes.data(), bytes.size()));
}

std::string calibrationKey(const std::string &binHash, const HostProfile &p, int seqLen, int border, int trials,
                           const double rates[6], float tantanThr) {
    std::string b;
    auto add = [&](const void *d, size_t n) { b.append(static_cast<const char *>(d), n); };
    auto addF = [&](float v) { add(&v, sizeof v); };
    auto addI = [&](int v) { add(&v, sizeof v); };
    add(binHash.data(), binHash.size());
    add(p.name.data(), p.name.size());
    addI(kWidth);
    addI(p.M);
    add(p.values.data(), sizeof(float) * (size_t)kWidth * (p.M + 1));
    add(p.values.data() + (size_t)kWidth * p.M + 4, sizeof(float) * 20);  // the background letter frequencies
    addI(seqLen);
    addI(border);
    addI(trials);
    const float i1 = (float)rates[0], i2 = (float)rates[1], d1 = (float)rates[2], d2 = (float)rates[3];
    addF(i1);
    addF(i2);
    addF(d1);
    addF(d2);
    addF(i1 + d2);  // BACKGROUND_FRAMESHIFT_RATE
    addF(i2 + d1);  // BACKGROUND_FRAMESHIFT_RATE_2
    addF((float)rates[4]);
    addF((float)rates[5]);
    addF(tantanThr);
    return hashHex(xxh3_128(b.data(), b.size()));
}

std::string calibrationCachePath() {
    namespace fs = std::filesystem;
    std::vector<fs::path> cand;
    auto env = [](const char *n) -> const char * {
        const char *v = std::getenv(n);
        return (v && *v) ? v : nullptr;
    };
    if (const char *d = env("DUMMEST_CACHE_DIR")) cand.emplace_back(d);
    if (const char *d = env("XDG_CACHE_HOME")) cand.emplace_back(fs::path(d) / ".dummest-cache");
    if (const char *d = env("HOME")) cand.emplace_back(fs::path(d) / ".dummest-cache");
    std::error_code ec0;
    const fs::path cwd = fs::current_path(ec0);
    cand.emplace_back(cwd / ".dummest-cache");
    if (const char *d = env("TMPDIR")) cand.emplace_back(fs::path(d) / ".dummest-cache");
    cand.emplace_back("/tmp/.dummest-cache");
    for (const auto &dir : cand) {
        std::error_code ec;
        fs::create_directories(dir, ec);
        if (!ec && fs::is_directory(dir, ec)) return (dir / "cache.bin").string();
    }
    std::cerr << "# Warning: Could not create cache directory; using current directory\n";
    return (cwd / "cache.bin").string();
}

bool readCalibrationCache(const std::string &path, std::vector<std::pair<std::string, CacheEntry>> &out,
                          std::string &err) {
    out.clear();
    std::ifstream in(path, std::ios::binary);
    if (!in || in.peek() == std::ifstream::traits_type::eof()) return true;  // no file or empty: no entries
    in.seekg(0, std::ios::end);
    const uint64_t size = (uint64_t)in.tellg();
    in.seekg(0);
    uint64_t count = 0, pos = 8;
    if (size < 8 || !in.read(reinterpret_cast<char *>(&count), 8)) {
        err = "truncated entry count";
        return false;
    }
    for (uint64_t k = 0; k < count; k++) {
        uint64_t klen = 0;
        if (pos + 8 > size || !in.read(reinterpret_cast<char *>(&klen), 8) || klen > size - pos - 8) {
            err = "truncated key";
            return false;
        }
        pos += 8;
        std::string key(klen, '\0');
        CacheEntry e{};
        if (!in.read(&key[0], (std::streamsize)klen) || pos + klen + sizeof(e.v) > size ||
            !in.read(reinterpret_cast<char *>(e.v), sizeof(e.v))) {
            err = "truncated entry";
            return false;
        }
        pos += klen + sizeof(e.v);
        out.emplace_back(std::move(key), e);
    }
    return true;
}

bool writeCalibrationCache(const std::string &path, const std::vector<std::pair<std::string, CacheEntry>> &entries) {
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) return false;
    const uint64_t count = entries.size();
    out.write(reinterpret_cast<const char *>(&count), 8);
    for (const auto &kv : entries) {
        const uint64_t klen = kv.first.size();
        out.write(reinterpret_cast<const char *>(&klen), 8);
        out.write(kv.first.data(), (std::streamsize)klen);
        out.write(reinterpret_cast<const char *>(kv.second.v), sizeof(kv.second.v));
    }
    return (bool)out;
}

}  // namespace dm

// ------------------------------------------------------------------ C ABI
extern "C" {

int dm_xxh3_128(const void *data, int64_t len, uint64_t out[2]) {
    if (len < 0 || (len > 0 && !data) || !out) return DM_EINVAL;
    const dm::Hash128 h = dm::xxh3_128(data, (size_t)len);
    out[0] = h.lo;
    out[1] = h.hi;
    return DM_OK;
}

int dm_calibration_key(const char *name, int32_t M, const float *values, const char *binary_hash, int32_t seq_len,
                       int32_t border, int32_t trials, const double rates[6], double tantan_threshold, char out[33]) {
    if (!name || M < 1 || !values || !rates || !out) return DM_EINVAL;
    dm::HostProfile p;
    p.name = name;
    p.M = M;
    p.values.assign(values, values + (size_t)dm::kWidth * (M + 1));
    const std::string key = dm::calibrationKey(binary_hash ? binary_hash : "", p, seq_len, border, trials, rates,
                                               (float)tantan_threshold);
    std::memcpy(out, key.c_str(), 33);
    return DM_OK;
}

int dm_cache_file_path(char *out, int32_t cap) {
    if (!out || cap < 1) return DM_EINVAL;
    const std::string path = dm::calibrationCachePath();
    if ((int64_t)path.size() + 1 > cap) return DM_EINVAL;
    std::memcpy(out, path.c_str(), path.size() + 1);
    return DM_OK;
}

int64_t dm_cache_read(const char *path, int64_t cap, char *keys, int32_t key_stride, double *values) {
    if (!path || cap < 0 || (cap > 0 && (!keys || key_stride < 1 || !values))) return DM_EINVAL;
    std::vector<std::pair<std::string, dm::CacheEntry>> entries;
    std::string err;
    if (!dm::readCalibrationCache(path, entries, err)) return DM_EFORMAT;
    for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)entries.size()); i++) {
        const std::string &k = entries[i].first;
        if ((int64_t)k.size() + 1 > key_stride) return DM_EINVAL;
        std::memcpy(keys + i * key_stride, k.c_str(), k.size() + 1);
        std::memcpy(values + 27 * i, entries[i].second.v, sizeof(entries[i].second.v));
    }
    return (int64_t)entries.size();
}

int dm_cache_write(const char *path, int64_t n, const char *keys, int32_t key_stride, const double *values) {
    if (!path || n < 0 || (n > 0 && (!keys || key_stride < 1 || !values))) return DM_EINVAL;
    std::vector<std::pair<std::string, dm::CacheEntry>> entries((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        entries[i].first.assign(keys + i * key_stride, strnlen(keys + i * key_stride, (size_t)key_stride));
        std::memcpy(entries[i].second.v, values + 27 * i, sizeof(entries[i].second.v));
    }
    return dm::writeCalibrationCache(path, entries) ? DM_OK : DM_EIO;
}

}  // extern "C"
