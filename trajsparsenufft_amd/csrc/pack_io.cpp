// Pack directory format: header + five arrays per component, error state.
// Format and checks follow the reference reader/writer:
//   header "<Q11I"             src/export_gpu_case.py:242-256, nufft_spmma_bench.cu:42-59
//   header validation          nufft_spmma_bench.cu:88-98
//   real/imag pair validation  nufft_integrated_bench.cu:135-194
//   production file set        src/export_torchkbnufft_gpu_case.py:407-436
#include <sys/stat.h>

#include <cerrno>
#include <cstdarg>
#include <cstdio>

#include "tsn_internal.h"

namespace tsn {

static thread_local char g_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  std::vsnprintf(g_error, sizeof(g_error), fmt, args);
  va_end(args);
}

int fail(int code, const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  std::vsnprintf(g_error, sizeof(g_error), fmt, args);
  va_end(args);
  return code;
}

int validate_header(const tsn_pack_header& h) {
  if (h.magic != kMagic || h.version != kVersion || h.n_channels != static_cast<uint32_t>(kN) ||
      h.m_tile != static_cast<uint32_t>(kM) || h.k_tile != static_cast<uint32_t>(kK))
    return fail(TSN_ERR_RUNTIME, "unsupported input header (magic/version/n_channels/m/k)");
  if (h.residual_nnz != 0)
    return fail(TSN_ERR_INCOMPATIBLE, "pack has a CSR residual (residual_nnz=%u)", h.residual_nnz);
  if (h.groups == 0 || h.tiles == 0 || h.output_size == 0 || h.n_cols == 0)
    return fail(TSN_ERR_RUNTIME, "empty pack (groups=%u tiles=%u)", h.groups, h.tiles);
  if (static_cast<uint64_t>(h.groups) * kM < h.active_rows)
    return fail(TSN_ERR_RUNTIME, "groups cannot hold active_rows");
  return TSN_OK;
}

// Every index a kernel will dereference is checked here, on the host.
int validate_pack(const HostPack& p) {
  TSN_CHECK(validate_header(p.h));
  const tsn_pack_header& h = p.h;
  if (p.group_offsets.size() != h.groups + 1ull || p.group_row_ids.size() != h.groups * 16ull ||
      p.tile_col_ids.size() != h.tiles * 32ull || p.real_comp.size() != h.tiles * 256ull ||
      p.imag_comp.size() != h.tiles * 256ull || p.real_meta.size() != h.tiles * 16ull ||
      p.imag_meta.size() != h.tiles * 16ull)
    return fail(TSN_ERR_RUNTIME, "pack array sizes do not match the header");
  if (p.group_offsets[0] != 0 || p.group_offsets[h.groups] != h.tiles)
    return fail(TSN_ERR_RUNTIME, "group_offsets must start at 0 and end at tiles");
  for (uint32_t g = 0; g < h.groups; ++g)
    if (p.group_offsets[g + 1] < p.group_offsets[g])
      return fail(TSN_ERR_RUNTIME, "group_offsets not monotone at group %u", g);
  for (int32_t r : p.group_row_ids)
    if (r < -1 || r >= static_cast<int64_t>(h.output_size))
      return fail(TSN_ERR_RUNTIME, "group_row_ids entry %d outside [-1, %u)", r, h.output_size);
  for (int32_t c : p.tile_col_ids)
    if (c < 0 || c >= static_cast<int64_t>(h.n_cols))
      return fail(TSN_ERR_RUNTIME, "tile_col_ids entry %d outside [0, %u)", c, h.n_cols);
  // each 4-bit chunk code must name two distinct positions p0 < p1
  for (const std::vector<uint32_t>* meta : {&p.real_meta, &p.imag_meta}) {
    for (size_t t = 0; t < h.tiles; ++t) {
      uint32_t rows[16];
      decode_row_meta(meta->data() + t * 16, rows);
      for (int r = 0; r < 16; ++r)
        for (int c = 0; c < 8; ++c) {
          const uint32_t code = (rows[r] >> (4 * c)) & 0xF;
          if ((code & 3) >= (code >> 2))
            return fail(TSN_ERR_RUNTIME, "tile %zu row %d chunk %d: invalid 2:4 code 0x%x", t, r, c,
                        code);
        }
    }
  }
  return TSN_OK;
}

}  // namespace tsn

using namespace tsn;

namespace {

int read_file(const std::string& path, void* dst, size_t bytes) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return fail(TSN_ERR_RUNTIME, "cannot open %s", path.c_str());
  const size_t got = bytes ? std::fread(dst, 1, bytes, f) : 0;
  const int extra = std::fgetc(f);
  std::fclose(f);
  if (got != bytes) return fail(TSN_ERR_RUNTIME, "short read %s (%zu of %zu bytes)", path.c_str(), got, bytes);
  if (extra != EOF) return fail(TSN_ERR_RUNTIME, "file %s is longer than the header implies", path.c_str());
  return TSN_OK;
}

int write_file(const std::string& path, const void* src, size_t bytes) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return fail(TSN_ERR_RUNTIME, "cannot create %s", path.c_str());
  const size_t put = bytes ? std::fwrite(src, 1, bytes, f) : 0;
  const int closed = std::fclose(f);
  if (put != bytes || closed != 0) return fail(TSN_ERR_RUNTIME, "failed to write %s", path.c_str());
  return TSN_OK;
}

int make_dirs(const std::string& path) {
  std::string partial;
  for (size_t i = 0; i <= path.size(); ++i) {
    if (i == path.size() || path[i] == '/') {
      if (!partial.empty() && mkdir(partial.c_str(), 0755) != 0 && errno != EEXIST)
        return fail(TSN_ERR_RUNTIME, "cannot create directory %s", partial.c_str());
    }
    if (i < path.size()) partial.push_back(path[i]);
  }
  return TSN_OK;
}

template <typename T>
int read_array(const std::string& dir, const char* name, std::vector<T>& out, size_t count) {
  out.resize(count);
  return read_file(dir + "/" + name, out.data(), count * sizeof(T));
}

template <typename T>
int write_array(const std::string& dir, const char* name, const std::vector<T>& v) {
  return write_file(dir + "/" + name, v.data(), v.size() * sizeof(T));
}

int read_component(const std::string& dir, tsn_pack_header* h, std::vector<uint32_t>& offsets,
                   std::vector<int32_t>& rows, std::vector<int32_t>& cols, std::vector<uint16_t>& comp,
                   std::vector<uint32_t>& meta) {
  TSN_CHECK(read_file(dir + "/header.bin", h, sizeof(*h)));
  TSN_CHECK(validate_header(*h));
  TSN_CHECK(read_array(dir, "group_offsets.bin", offsets, h->groups + 1ull));
  TSN_CHECK(read_array(dir, "group_row_ids.bin", rows, h->groups * 16ull));
  TSN_CHECK(read_array(dir, "tile_col_ids.bin", cols, h->tiles * 32ull));
  TSN_CHECK(read_array(dir, "tile_a_comp.bin", comp, h->tiles * 256ull));
  TSN_CHECK(read_array(dir, "tile_pair_meta.bin", meta, h->tiles * 16ull));
  return TSN_OK;
}

}  // namespace

extern "C" {

const char* tsn_last_error(void) { return g_error; }

const char* tsn_version(void) { return "trajsparsenufft_amd 0.1 gfx950"; }

int tsn_host_pack_read(const char* real_dir, const char* imag_dir, tsn_host_pack** out) {
  if (!real_dir || !imag_dir || !out) return fail(TSN_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  auto* p = new tsn_host_pack();
  tsn_pack_header ih{};
  std::vector<uint32_t> io;
  std::vector<int32_t> ir, ic;
  int rc = read_component(real_dir, &p->h, p->group_offsets, p->group_row_ids, p->tile_col_ids,
                          p->real_comp, p->real_meta);
  if (rc == TSN_OK)
    rc = read_component(imag_dir, &ih, io, ir, ic, p->imag_comp, p->imag_meta);
  if (rc == TSN_OK && std::memcmp(&p->h, &ih, sizeof(ih)) != 0)
    rc = fail(TSN_ERR_INCOMPATIBLE, "incompatible real/imag packed direction");
  if (rc == TSN_OK && (io != p->group_offsets || ir != p->group_row_ids || ic != p->tile_col_ids))
    rc = fail(TSN_ERR_INCOMPATIBLE, "real/imag structures differ");
  if (rc == TSN_OK) rc = validate_pack(*p);
  if (rc != TSN_OK) {
    delete p;
    return rc;
  }
  *out = p;
  return TSN_OK;
}

int tsn_host_pack_write(const tsn_host_pack* p, const char* real_dir, const char* imag_dir) {
  if (!p || !real_dir || !imag_dir) return fail(TSN_ERR_ARGUMENT, "null argument");
  const std::string dirs[2] = {real_dir, imag_dir};
  for (int c = 0; c < 2; ++c) {
    TSN_CHECK(make_dirs(dirs[c]));
    TSN_CHECK(write_file(dirs[c] + "/header.bin", &p->h, sizeof(p->h)));
    TSN_CHECK(write_array(dirs[c], "group_offsets.bin", p->group_offsets));
    TSN_CHECK(write_array(dirs[c], "group_row_ids.bin", p->group_row_ids));
    TSN_CHECK(write_array(dirs[c], "tile_col_ids.bin", p->tile_col_ids));
    TSN_CHECK(write_array(dirs[c], "tile_a_comp.bin", c == 0 ? p->real_comp : p->imag_comp));
    TSN_CHECK(write_array(dirs[c], "tile_pair_meta.bin", c == 0 ? p->real_meta : p->imag_meta));
  }
  return TSN_OK;
}

int tsn_host_pack_create(const tsn_pack_header* h, const uint32_t* group_offsets,
                         const int32_t* group_row_ids, const int32_t* tile_col_ids,
                         const uint16_t* real_comp, const uint32_t* real_meta,
                         const uint16_t* imag_comp, const uint32_t* imag_meta, tsn_host_pack** out) {
  if (!h || !group_offsets || !group_row_ids || !tile_col_ids || !real_comp || !real_meta ||
      !imag_comp || !imag_meta || !out)
    return fail(TSN_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  TSN_CHECK(validate_header(*h));
  auto* p = new tsn_host_pack();
  p->h = *h;
  p->group_offsets.assign(group_offsets, group_offsets + h->groups + 1);
  p->group_row_ids.assign(group_row_ids, group_row_ids + h->groups * 16ull);
  p->tile_col_ids.assign(tile_col_ids, tile_col_ids + h->tiles * 32ull);
  p->real_comp.assign(real_comp, real_comp + h->tiles * 256ull);
  p->real_meta.assign(real_meta, real_meta + h->tiles * 16ull);
  p->imag_comp.assign(imag_comp, imag_comp + h->tiles * 256ull);
  p->imag_meta.assign(imag_meta, imag_meta + h->tiles * 16ull);
  const int rc = validate_pack(*p);
  if (rc != TSN_OK) {
    delete p;
    return rc;
  }
  *out = p;
  return TSN_OK;
}

int tsn_host_pack_header(const tsn_host_pack* p, tsn_pack_header* out) {
  if (!p || !out) return fail(TSN_ERR_ARGUMENT, "null argument");
  *out = p->h;
  return TSN_OK;
}

int tsn_host_pack_arrays(const tsn_host_pack* p, const uint32_t** group_offsets,
                         const int32_t** group_row_ids, const int32_t** tile_col_ids,
                         const uint16_t** real_comp, const uint32_t** real_meta,
                         const uint16_t** imag_comp, const uint32_t** imag_meta) {
  if (!p) return fail(TSN_ERR_ARGUMENT, "null pack");
  if (group_offsets) *group_offsets = p->group_offsets.data();
  if (group_row_ids) *group_row_ids = p->group_row_ids.data();
  if (tile_col_ids) *tile_col_ids = p->tile_col_ids.data();
  if (real_comp) *real_comp = p->real_comp.data();
  if (real_meta) *real_meta = p->real_meta.data();
  if (imag_comp) *imag_comp = p->imag_comp.data();
  if (imag_meta) *imag_meta = p->imag_meta.data();
  return TSN_OK;
}

void tsn_host_pack_free(tsn_host_pack* p) { delete p; }

}  // extern "C"
