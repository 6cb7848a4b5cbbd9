// On-disk cache of transformed demag / dipole tensors (host code, no HIP).
//
// Mirrors DemagField's caching (demag_field.cpp:185-204) and CacheManager
// (cache_manager.cpp:49-114): one file per mesh in <temp dir>/magnumaf.cache, named after the mesh,
// read when present, written after a computation, and the folder shrunk, oldest files first, when
// it outgrows its size limit.  Each file carries a header with the full key and a checksum of the
// payload, so a file of another mesh, layout or an interrupted write is never used.
#pragma once
#include <cstdint>
#include <filesystem>
#include <string>
#include <vector>

namespace mafx {

constexpr int32_t kCacheHostTensor = 0x100;   // FP32 layout from the whole host-built real-space tensor
constexpr int32_t kCacheDeviceTensor = 0x200; // either layout from the GPU-built real-space tensor
constexpr int32_t kCacheOctantTensor = 0x400; // FP64 layout from the parity-extended host octant
constexpr int32_t kCacheFullTensor = 0x800;   // FP32 symmetric quarter + FP16 residual of the whole tensor's spectrum

struct TensorCacheKey {
    int64_t n0, n1, n2, hxs; // expanded lengths and the padded row stride of the stored planes
    int32_t precision;       // MAFX_F64 (complex half spectrum) or MAFX_F32_FIELD (real quarter),
                             // | kCacheHostTensor when the FP32 tensor comes from the whole host
                             // tensor (default: the host octant), | kCacheDeviceTensor when the
                             // tensor comes from the GPU
    int32_t dipole;          // 1 for the atomistic point-dipole tensor
    double dx, dy, dz;
};

// The key's layout field of a context of this precision and DEMAG term flags.
int32_t tensor_cache_layout(int precision, int term_flags);

// $MAFX_CACHE_DIR, else std::filesystem::temp_directory_path() / "magnumaf.cache"
std::filesystem::path tensor_cache_dir();
// The reference's mesh string (demag_field.cpp:185-190), then the layout and the exact bits of
// dx, dy, dz.
std::string tensor_cache_name(const TensorCacheKey& k);
// Reads the payload of a valid file for k of exactly `bytes` bytes into out; false otherwise.
bool tensor_cache_read(const TensorCacheKey& k, uint64_t bytes, std::vector<char>& out, bool verbose);
// Writes the payload (through a temporary file renamed into place); failures only warn.
void tensor_cache_write(const TensorCacheKey& k, const void* data, uint64_t bytes, bool verbose);
// Removes the oldest files while the folder holds more than $MAFX_CACHE_MAX_BYTES (default 2e9),
// down to a tenth of that.
void tensor_cache_shrink(bool verbose);

} // namespace mafx
