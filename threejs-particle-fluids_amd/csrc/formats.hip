// On-disk formats: the PSDF signed-distance-field binary (src/sdf/
// writeBinary.ts:21-106). Host code only.
//
// Layout (little endian): "PSDF" magic, u32 version 1, u32 resolution x, y, z,
// u32 reserved 0, f32 origin x, y, z, f32 voxel size x, y, z (48 bytes), then
// f32 voxels with x fastest, then y, then z.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "runtime.hpp"

using namespace pf;

namespace {

constexpr uint32_t kSdfMagic = 0x46445350u;  // "PSDF" read as a little-endian u32
constexpr uint32_t kSdfVersion = 1u;
constexpr size_t kSdfHeader = 48;

uint32_t rd_u32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
float rd_f32(const uint8_t* p) {
  const uint32_t u = rd_u32(p);
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
void wr_u32(uint8_t* p, uint32_t v) {
  p[0] = (uint8_t)v;
  p[1] = (uint8_t)(v >> 8);
  p[2] = (uint8_t)(v >> 16);
  p[3] = (uint8_t)(v >> 24);
}
void wr_f32(uint8_t* p, float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  wr_u32(p, u);
}

}  // namespace

#define PF_API_BEGIN try {
#define PF_API_END                                    \
  }                                                   \
  catch (const pf::PfError& e) { return e.code; }     \
  catch (const std::bad_alloc&) {                     \
    pf::set_error("host allocation failed");          \
    return PF_ENOMEM;                                 \
  }                                                   \
  catch (const std::exception& e) {                   \
    pf::set_error("%s", e.what());                    \
    return PF_EINVAL;                                 \
  }                                                   \
  return PF_OK;

extern "C" {

int pf_sdf_decode(const void* bytes, size_t size, uint32_t resolution[3], double origin[3], double voxel_size[3],
                  float* data, size_t data_capacity, size_t* voxel_count) {
  PF_API_BEGIN
  if (!bytes && size) fail(PF_EINVAL, "decodeSdfBinary: null buffer");
  if (size < kSdfHeader) fail(PF_EINVAL, "decodeSdfBinary: buffer too small for header (%zu < %zu)", size, kSdfHeader);
  const uint8_t* b = static_cast<const uint8_t*>(bytes);
  const uint32_t magic = rd_u32(b);
  if (magic != kSdfMagic)
    fail(PF_EINVAL, "decodeSdfBinary: bad magic 0x%x (expected 0x%x \"PSDF\")", magic, kSdfMagic);
  const uint32_t version = rd_u32(b + 4);
  if (version != kSdfVersion)
    fail(PF_EINVAL, "decodeSdfBinary: unsupported version %u (expected %u)", version, kSdfVersion);
  const uint32_t rx = rd_u32(b + 8), ry = rd_u32(b + 12), rz = rd_u32(b + 16);
  const float o[3] = {rd_f32(b + 24), rd_f32(b + 28), rd_f32(b + 32)};
  const float v[3] = {rd_f32(b + 36), rd_f32(b + 40), rd_f32(b + 44)};
  if (rx < 2 || ry < 2 || rz < 2)
    fail(PF_EINVAL, "decodeSdfBinary: resolution must be ≥ 2, got [%u, %u, %u]", rx, ry, rz);
  for (int a = 0; a < 3; ++a)
    if (!(v[a] > 0.0f) || !std::isfinite(v[a]))
      fail(PF_EINVAL, "decodeSdfBinary: voxelSize must be positive, got [%g, %g, %g]", (double)v[0], (double)v[1],
           (double)v[2]);
  const size_t count = (size_t)rx * ry * rz;
  const size_t expected = kSdfHeader + count * 4;
  if (size < expected)
    fail(PF_EINVAL, "decodeSdfBinary: buffer too small (%zu < %zu for %u·%u·%u voxels)", size, expected, rx, ry, rz);
  if (resolution) { resolution[0] = rx; resolution[1] = ry; resolution[2] = rz; }
  for (int a = 0; a < 3; ++a) {
    if (origin) origin[a] = (double)o[a];
    if (voxel_size) voxel_size[a] = (double)v[a];
  }
  if (voxel_count) *voxel_count = count;
  // Voxels only when the caller's buffer holds them all (first call: header).
  if (data && data_capacity >= count)
    for (size_t i = 0; i < count; ++i) data[i] = rd_f32(b + kSdfHeader + 4 * i);
  PF_API_END
}

int pf_sdf_encode(const float* data, size_t data_count, const uint32_t resolution[3], const double origin[3],
                  const double voxel_size[3], void* out, size_t capacity, size_t* size) {
  PF_API_BEGIN
  if (!resolution || !origin || !voxel_size) fail(PF_EINVAL, "encodeSdfBinary: null argument");
  const size_t count = (size_t)resolution[0] * resolution[1] * resolution[2];
  if (count != data_count)
    fail(PF_EINVAL, "encodeSdfBinary: resolution %u·%u·%u=%zu does not match data length %zu", resolution[0],
         resolution[1], resolution[2], count, data_count);
  const size_t total = kSdfHeader + count * 4;
  if (size) *size = total;
  if (!out) return PF_OK;  // size query
  if (capacity < total) fail(PF_EINVAL, "encodeSdfBinary: output holds %zu bytes, %zu needed", capacity, total);
  if (count && !data) fail(PF_EINVAL, "encodeSdfBinary: null data");
  uint8_t* b = static_cast<uint8_t*>(out);
  wr_u32(b, kSdfMagic);
  wr_u32(b + 4, kSdfVersion);
  wr_u32(b + 8, resolution[0]);
  wr_u32(b + 12, resolution[1]);
  wr_u32(b + 16, resolution[2]);
  wr_u32(b + 20, 0u);
  for (int a = 0; a < 3; ++a) {
    wr_f32(b + 24 + 4 * a, (float)origin[a]);
    wr_f32(b + 36 + 4 * a, (float)voxel_size[a]);
  }
  for (size_t i = 0; i < count; ++i) wr_f32(b + kSdfHeader + 4 * i, data[i]);
  PF_API_END
}

}  // extern "C"
