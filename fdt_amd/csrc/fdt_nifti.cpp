// fdt_nifti.cpp -- NIfTI-1 reader/writer and mask maps (see fdt_nifti.h).
#include "fdt_nifti.h"

#include <zlib.h>
#include <sys/stat.h>

#include <atomic>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

namespace fdt {

namespace {

bool exists(const std::string& p) { struct stat sb; return stat(p.c_str(), &sb) == 0 && S_ISREG(sb.st_mode); }
bool ends_with(const std::string& s, const char* suf) {
  size_t n = std::strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

void swap_bytes(void* p, int size) {
  uint8_t* b = static_cast<uint8_t*>(p);
  for (int i = 0; i < size / 2; ++i) std::swap(b[i], b[size - 1 - i]);
}
template <class T>
T get(const uint8_t* h, int off, bool sw) {
  T v;
  std::memcpy(&v, h + off, sizeof(T));
  if (sw) swap_bytes(&v, sizeof(T));
  return v;
}
template <class T>
void put(uint8_t* h, int off, T v) { std::memcpy(h + off, &v, sizeof(T)); }

// header offsets of the NIfTI-1 specification
enum {
  H_SIZEOF = 0, H_DIM = 40, H_DATATYPE = 70, H_BITPIX = 72, H_PIXDIM = 76, H_VOXOFF = 108, H_SLOPE = 112,
  H_INTER = 116, H_XYZT = 123, H_CALMAX = 124, H_CALMIN = 128, H_DESCRIP = 148, H_QFORM = 252, H_SFORM = 254,
  H_QUATERN = 256, H_SROW = 280, H_MAGIC = 344, H_SIZE = 348
};

struct GzIn {
  gzFile f = nullptr;
  ~GzIn() { if (f) gzclose(f); }
  bool read(void* dst, size_t n) {
    uint8_t* p = static_cast<uint8_t*>(dst);
    while (n) {
      unsigned chunk = n > (1u << 30) ? (1u << 30) : (unsigned)n;
      int got = gzread(f, p, chunk);
      if (got <= 0) return false;
      p += got;
      n -= (size_t)got;
    }
    return true;
  }
};

// gzip of `n` bytes as one member into `out`: level 1 with run-length
// matching only (Z_RLE), which on float images with a zero background is
// 2.6x faster than the default strategy at a slightly smaller size
bool gzip_member(const uint8_t* src, size_t n, std::vector<uint8_t>& out) {
  z_stream zs;
  std::memset(&zs, 0, sizeof zs);
  if (deflateInit2(&zs, 1, Z_DEFLATED, 15 + 16, 8, Z_RLE) != Z_OK) return false;
  out.resize(deflateBound(&zs, (uLong)n) + 64);
  zs.next_in = const_cast<Bytef*>(src);
  zs.avail_in = (uInt)n;
  zs.next_out = out.data();
  zs.avail_out = (uInt)out.size();
  int rc = deflate(&zs, Z_FINISH);
  out.resize(zs.total_out);
  deflateEnd(&zs);
  return rc == Z_STREAM_END;
}

template <class T>
void convert(const uint8_t* src, size_t n, bool sw, float* dst) {
  for (size_t i = 0; i < n; ++i) {
    T v;
    std::memcpy(&v, src + i * sizeof(T), sizeof(T));
    if (sw) swap_bytes(&v, sizeof(T));
    dst[i] = (float)v;
  }
}

}  // namespace

std::string image_path(const std::string& base) {
  if (exists(base)) return base;
  if (exists(base + ".nii.gz")) return base + ".nii.gz";
  if (exists(base + ".nii")) return base + ".nii";
  return "";
}

bool read_nifti(const std::string& path, Nifti& img, std::string& err) {
  GzIn in;
  in.f = gzopen(path.c_str(), "rb");  // reads plain files too
  if (!in.f) { err = "cannot open image " + path; return false; }
  gzbuffer(in.f, 1 << 20);
  uint8_t h[H_SIZE];
  if (!in.read(h, H_SIZE)) { err = path + ": short header"; return false; }
  bool sw = false;
  int32_t sz = get<int32_t>(h, H_SIZEOF, false);
  if (sz != 348) {
    sw = true;
    if (get<int32_t>(h, H_SIZEOF, true) != 348) { err = path + ": not a NIfTI-1 / ANALYZE header"; return false; }
  }
  if (std::memcmp(h + H_MAGIC, "n+1", 4) != 0) { err = path + ": not a single-file NIfTI-1 image (magic n+1)"; return false; }
  for (int i = 0; i < 8; ++i) img.dim[i] = get<int16_t>(h, H_DIM + 2 * i, sw);
  for (int i = 0; i < 8; ++i) img.pixdim[i] = get<float>(h, H_PIXDIM + 4 * i, sw);
  if (img.dim[0] < 1 || img.dim[0] > 7) { err = path + ": bad dim[0]"; return false; }
  for (int i = img.dim[0] + 1; i < 8; ++i) img.dim[i] = 1;
  img.datatype = get<int16_t>(h, H_DATATYPE, sw);
  float voxoff = get<float>(h, H_VOXOFF, sw);
  img.scl_slope = get<float>(h, H_SLOPE, sw);
  img.scl_inter = get<float>(h, H_INTER, sw);
  img.xyzt_units = h[H_XYZT];
  img.cal_max = get<float>(h, H_CALMAX, sw);
  img.cal_min = get<float>(h, H_CALMIN, sw);
  img.qform_code = get<int16_t>(h, H_QFORM, sw);
  img.sform_code = get<int16_t>(h, H_SFORM, sw);
  for (int i = 0; i < 6; ++i) img.quatern[i] = get<float>(h, H_QUATERN + 4 * i, sw);
  for (int i = 0; i < 12; ++i) img.srow[i] = get<float>(h, H_SROW + 4 * i, sw);
  // at most 2^38 values (1 TiB as float): larger products come from corrupt headers, and the
  // check keeps the product from wrapping
  constexpr size_t kMaxValues = (size_t)1 << 38;
  size_t n = 1;
  for (int i = 1; i <= img.dim[0]; ++i) {
    if (img.dim[i] < 1) { err = path + ": bad dimension"; return false; }
    if (n > kMaxValues / (size_t)img.dim[i]) { err = path + ": image too large (corrupt dimensions?)"; return false; }
    n *= (size_t)img.dim[i];
  }
  int bytes = 0;
  switch (img.datatype) {
    case 2: case 256: bytes = 1; break;          // uint8, int8
    case 4: case 512: bytes = 2; break;          // int16, uint16
    case 8: case 16: case 768: bytes = 4; break; // int32, float32, uint32
    case 64: bytes = 8; break;                   // float64
    default: err = path + ": unsupported datatype " + std::to_string(img.datatype); return false;
  }
  // extensions up to vox_offset are skipped (vox_offset below the header size reads from the header's end)
  if (!(voxoff < 2147483648.0f)) { err = path + ": bad vox_offset"; return false; }
  size_t skip = voxoff > (float)H_SIZE ? (size_t)voxoff - H_SIZE : 0;
  uint8_t pad[4096];
  while (skip) {
    const size_t m = std::min(skip, sizeof pad);
    if (!in.read(pad, m)) { err = path + ": short file"; return false; }
    skip -= m;
  }
  std::vector<uint8_t> raw;
  try {
    raw.resize(n * (size_t)bytes);
    img.data.resize(n);
  } catch (const std::bad_alloc&) {
    err = path + ": not enough host memory for " + std::to_string(n) + " values";
    return false;
  }
  if (!in.read(raw.data(), raw.size())) { err = path + ": fewer voxels than the header declares"; return false; }
  switch (img.datatype) {
    case 2: convert<uint8_t>(raw.data(), n, false, img.data.data()); break;
    case 256: convert<int8_t>(raw.data(), n, false, img.data.data()); break;
    case 4: convert<int16_t>(raw.data(), n, sw, img.data.data()); break;
    case 512: convert<uint16_t>(raw.data(), n, sw, img.data.data()); break;
    case 8: convert<int32_t>(raw.data(), n, sw, img.data.data()); break;
    case 768: convert<uint32_t>(raw.data(), n, sw, img.data.data()); break;
    case 16: convert<float>(raw.data(), n, sw, img.data.data()); break;
    case 64: convert<double>(raw.data(), n, sw, img.data.data()); break;
  }
  if (img.scl_slope != 0.0f && std::isfinite(img.scl_slope) && !(img.scl_slope == 1.0f && img.scl_inter == 0.0f))
    for (auto& v : img.data) v = v * img.scl_slope + img.scl_inter;
  return true;
}

bool write_nifti(const std::string& path, const Nifti& like, int nt, const float* data, float cal_min, float cal_max,
                 std::string& err, int threads) {
  uint8_t h[H_SIZE + 4];
  std::memset(h, 0, sizeof h);
  put<int32_t>(h, H_SIZEOF, 348);
  int16_t dim[8] = {(int16_t)(nt > 1 ? 4 : 3), (int16_t)like.dim[1], (int16_t)like.dim[2], (int16_t)like.dim[3],
                    (int16_t)nt, 1, 1, 1};
  for (int i = 0; i < 8; ++i) put<int16_t>(h, H_DIM + 2 * i, dim[i]);
  put<int16_t>(h, H_DATATYPE, 16);
  put<int16_t>(h, H_BITPIX, 32);
  float pd[8];
  for (int i = 0; i < 8; ++i) pd[i] = like.pixdim[i];
  pd[4] = 1.0f;
  for (int i = 0; i < 8; ++i) put<float>(h, H_PIXDIM + 4 * i, pd[i]);
  put<float>(h, H_VOXOFF, 352.0f);
  put<float>(h, H_SLOPE, 1.0f);
  put<float>(h, H_INTER, 0.0f);
  h[H_XYZT] = like.xyzt_units;
  put<float>(h, H_CALMAX, cal_max);
  put<float>(h, H_CALMIN, cal_min);
  std::strncpy(reinterpret_cast<char*>(h + H_DESCRIP), "fdt-mi355x bedpostx_gpu", 79);
  put<int16_t>(h, H_QFORM, like.qform_code);
  put<int16_t>(h, H_SFORM, like.sform_code);
  for (int i = 0; i < 6; ++i) put<float>(h, H_QUATERN + 4 * i, like.quatern[i]);
  for (int i = 0; i < 12; ++i) put<float>(h, H_SROW + 4 * i, like.srow[i]);
  std::memcpy(h + H_MAGIC, "n+1", 4);
  const size_t bytes = like.nxyz() * (size_t)nt * sizeof(float);
  if (ends_with(path, ".gz")) {
    // a multi-member gzip file (members of 4 MB compressed in parallel, as
    // pigz does): any gzip reader, zlib's gzread included, reads it as one stream
    const size_t CH = (size_t)4 << 20;
    const size_t nch = (bytes + CH - 1) / CH;
    std::vector<std::vector<uint8_t>> mem(nch + 1);
    bool ok = gzip_member(h, sizeof h, mem[0]);
    std::atomic<size_t> next{0};
    std::atomic<bool> good{true};
    auto work = [&]() {
      for (size_t j; (j = next++) < nch;) {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(data) + j * CH;
        if (!gzip_member(p, std::min(CH, bytes - j * CH), mem[j + 1])) good = false;
      }
    };
    const int nt_workers = (int)std::min<size_t>(nch, (size_t)std::max(1, threads));
    std::vector<std::thread> pool;
    for (int w = 1; w < nt_workers; ++w) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (!ok || !good) { err = "compression failed for " + path; return false; }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { err = "cannot write " + path; return false; }
    for (auto& m : mem) ok = ok && std::fwrite(m.data(), 1, m.size(), f) == m.size();
    if (std::fclose(f) != 0) ok = false;
    if (!ok) { err = "error writing " + path; return false; }
    return true;
  }
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { err = "cannot write " + path; return false; }
  bool ok = std::fwrite(h, 1, sizeof h, f) == sizeof h && std::fwrite(data, 1, bytes, f) == bytes;
  if (std::fclose(f) != 0) ok = false;
  if (!ok) { err = "error writing " + path; return false; }
  return true;
}

std::vector<size_t> mask_indices(const Nifti& mask) {
  std::vector<size_t> idx;
  const size_t n = mask.nxyz();
  for (size_t v = 0; v < n; ++v)
    if (mask.data[v] != 0.0f) idx.push_back(v);
  return idx;
}

void mask_matrix(const Nifti& img, const std::vector<size_t>& idx, float* out) {
  const size_t nxyz = img.nxyz();
  const int nt = img.nt();
  for (size_t i = 0; i < idx.size(); ++i)
    for (int t = 0; t < nt; ++t) out[i * nt + t] = img.data[(size_t)t * nxyz + idx[i]];
}

void scatter_volumes(const std::vector<size_t>& idx, size_t nxyz, int nt, const float* m, float* vol) {
  std::memset(vol, 0, nxyz * (size_t)nt * sizeof(float));
  const size_t nv = idx.size();
  for (int t = 0; t < nt; ++t)
    for (size_t i = 0; i < nv; ++i) vol[(size_t)t * nxyz + idx[i]] = m[(size_t)t * nv + i];
}

}  // namespace fdt
