// On-disk tensor cache (tensor_cache.hpp).  Replaces util::CacheManager (cache_manager.cpp:49-114)
// and the cache branch of get_Nfft (demag_field.cpp:185-204).
#include "tensor_cache.hpp"

#include "../../include/mafx.h"

#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <system_error>
#include <unistd.h>

namespace mafx {

namespace fs = std::filesystem;

namespace {

constexpr char kMagic[8] = {'M', 'A', 'F', 'X', 'N', 'F', 'T', '1'};
constexpr uint32_t kVersion = 1;

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    TensorCacheKey key;
    uint64_t payload_bytes;
    uint64_t checksum;
};

// FNV-1a over 64-bit little-endian words (the tail bytes zero-extended into a last word)
uint64_t checksum(const void* data, uint64_t bytes) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    uint64_t h = 0xcbf29ce484222325ull;
    const uint64_t prime = 0x100000001b3ull;
    uint64_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h = (h ^ w) * prime;
    }
    if (i < bytes) {
        uint64_t w = 0;
        std::memcpy(&w, p + i, bytes - i);
        h = (h ^ w) * prime;
    }
    return h;
}

bool same_key(const TensorCacheKey& a, const TensorCacheKey& b) {
    return a.n0 == b.n0 && a.n1 == b.n1 && a.n2 == b.n2 && a.hxs == b.hxs && a.precision == b.precision &&
           a.dipole == b.dipole && std::memcmp(&a.dx, &b.dx, sizeof(double)) == 0 &&
           std::memcmp(&a.dy, &b.dy, sizeof(double)) == 0 && std::memcmp(&a.dz, &b.dz, sizeof(double)) == 0;
}

uint64_t max_bytes() {
    if (const char* e = std::getenv("MAFX_CACHE_MAX_BYTES")) {
        char* end = nullptr;
        const unsigned long long v = std::strtoull(e, &end, 10);
        if (end != e) return v;
    }
    return 2000000000ull; // cache_manager.hpp default max_size_in_byte
}

std::string hex_bits(double v) {
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    char s[17];
    std::snprintf(s, sizeof s, "%016" PRIx64, b);
    return s;
}

} // namespace

fs::path tensor_cache_dir() {
    if (const char* e = std::getenv("MAFX_CACHE_DIR"))
        if (*e) return fs::path(e);
    std::error_code ec;
    fs::path t = fs::temp_directory_path(ec);
    if (ec) t = "/tmp";
    return t / "magnumaf.cache";
}

int32_t tensor_cache_layout(int precision, int term_flags) {
    if (precision != MAFX_F64 && (term_flags & MAFX_TERM_FULL_TENSOR))
        return precision | kCacheFullTensor | ((term_flags & MAFX_TERM_DEVICE_TENSOR) ? kCacheDeviceTensor : 0);
    if (term_flags & MAFX_TERM_DEVICE_TENSOR) return precision | kCacheDeviceTensor;
    if (precision == MAFX_F64) return (term_flags & MAFX_TERM_OCTANT_TENSOR) && !(term_flags & MAFX_TERM_HOST_TENSOR)
                                          ? kCacheOctantTensor : 0;
    return precision | ((term_flags & MAFX_TERM_HOST_TENSOR) ? kCacheHostTensor : 0);
}

std::string tensor_cache_name(const TensorCacheKey& k) {
    // demag_field.cpp:185-190: "Nfft_n0exp_<n0>_n1exp_<n1>_n2exp_<n2>_dx_<1e9 dx>_dy_.._dz_.."
    std::string s = "Nfft_n0exp_" + std::to_string(k.n0) + "_n1exp_" + std::to_string(k.n1) + "_n2exp_" +
                    std::to_string(k.n2) + "_dx_" + std::to_string(1e9 * k.dx) + "_dy_" + std::to_string(1e9 * k.dy) +
                    "_dz_" + std::to_string(1e9 * k.dz);
    const int32_t layout = k.precision & 0xff;
    // FP32 layouts: the parity-extended host octant (default), the whole host tensor, the GPU tensor
    s += layout == 0 ? ((k.precision & kCacheDeviceTensor)   ? "_mafx_c64_dev"
                        : (k.precision & kCacheOctantTensor) ? "_mafx_c64_oct"
                                                             : "_mafx_c64")
                     : ((k.precision & kCacheFullTensor)     ? ((k.precision & kCacheDeviceTensor) ? "_mafx_f32res_dev"
                                                                                                   : "_mafx_f32res_host")
                        : (k.precision & kCacheHostTensor)     ? "_mafx_f32sym_host"
                        : (k.precision & kCacheDeviceTensor) ? "_mafx_f32sym_dev"
                                                             : "_mafx_f32oct");
    if (k.dipole) s += "_dipole";
    s += "_" + hex_bits(k.dx) + hex_bits(k.dy) + hex_bits(k.dz);
    return s;
}

bool tensor_cache_read(const TensorCacheKey& k, uint64_t bytes, std::vector<char>& out, bool verbose) {
    const fs::path p = tensor_cache_dir() / tensor_cache_name(k);
    std::error_code ec;
    if (!fs::exists(p, ec)) return false;
    std::ifstream f(p, std::ios::binary);
    Header h{};
    bool ok = static_cast<bool>(f.read(reinterpret_cast<char*>(&h), sizeof h));
    ok = ok && std::memcmp(h.magic, kMagic, sizeof kMagic) == 0 && h.version == kVersion &&
         h.header_bytes == sizeof(Header) && same_key(h.key, k) && h.payload_bytes == bytes;
    if (ok) {
        out.resize(bytes);
        ok = static_cast<bool>(f.read(out.data(), static_cast<std::streamsize>(bytes))) &&
             checksum(out.data(), bytes) == h.checksum;
    }
    if (!ok) {
        std::cout << "Warning: demag tensor cache file " << p << " is invalid; computing the tensor instead."
                  << std::endl;
        out.clear();
        return false;
    }
    if (verbose) std::cout << "Info: Reading demag tensor from " << p << std::endl;
    return true;
}

void tensor_cache_write(const TensorCacheKey& k, const void* data, uint64_t bytes, bool verbose) {
    const fs::path dir = tensor_cache_dir();
    const fs::path p = dir / tensor_cache_name(k);
    std::error_code ec;
    fs::create_directories(dir, ec);
    if (ec) {
        std::cout << "Warning: cannot create the demag tensor cache folder " << dir << "; not caching." << std::endl;
        return;
    }
    if (verbose) std::cout << "Info: Saving demag tensor to " << p << std::endl;
    const fs::path tmp = p.string() + ".tmp" + std::to_string(static_cast<long>(::getpid()));
    Header h{};
    std::memcpy(h.magic, kMagic, sizeof kMagic);
    h.version = kVersion;
    h.header_bytes = sizeof(Header);
    h.key = k;
    h.payload_bytes = bytes;
    h.checksum = checksum(data, bytes);
    bool ok;
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        ok = static_cast<bool>(f.write(reinterpret_cast<const char*>(&h), sizeof h)) &&
             static_cast<bool>(f.write(static_cast<const char*>(data), static_cast<std::streamsize>(bytes)));
        f.close();
        ok = ok && !f.fail();
    }
    if (ok) fs::rename(tmp, p, ec);
    if (!ok || ec) {
        fs::remove(tmp, ec);
        std::cout << "Warning: writing the demag tensor cache file " << p << " failed; not caching." << std::endl;
    }
}

void tensor_cache_shrink(bool verbose) {
    const fs::path dir = tensor_cache_dir();
    std::error_code ec;
    if (!fs::is_directory(dir, ec)) return;
    struct F {
        fs::path p;
        uint64_t size;
        fs::file_time_type t;
    };
    std::vector<F> files;
    uint64_t total = 0;
    for (const auto& e : fs::directory_iterator(dir, ec)) {
        if (!e.is_regular_file(ec)) continue;
        const uint64_t sz = e.file_size(ec);
        files.push_back({e.path(), sz, e.last_write_time(ec)});
        total += sz;
    }
    const uint64_t cap = max_bytes(), target = cap / 10;
    if (total <= cap) return;
    // newest first; drop from the back (cache_manager.cpp:29-48)
    std::sort(files.begin(), files.end(), [](const F& a, const F& b) { return a.t > b.t; });
    while (total > target && !files.empty()) {
        if (verbose) std::cout << "removing file: " << files.back().p << std::endl;
        fs::remove(files.back().p, ec);
        total -= files.back().size;
        files.pop_back();
    }
}

} // namespace mafx
