// grmhd.hip - GRMHD packed-volume ingestion (nubhlight_pack metadata and
// RGBA32F binaries, single frames and time series), device volumes over the
// data (the handles of the render launch's GRMHD slots), and the render
// launch's GRMHD sample exposed for checks (include/bh_grmhd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bh_grmhd.h"
#include "grmhd_device.hpp"

namespace bhk {
int fail(const std::string& msg);

namespace gj {

// A small JSON reader: objects, arrays, strings (with escapes), numbers,
// true / false / null -- what the packer writes.
struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;
  const Value* get(const std::string& k) const {
    for (const auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct Parser {
  const std::string& s;
  size_t i = 0;
  bool ok = true;
  explicit Parser(const std::string& text) : s(text) {}
  void ws() {
    while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
  }
  bool lit(const char* w) {
    const size_t n = std::strlen(w);
    if (s.compare(i, n, w) == 0) {
      i += n;
      return true;
    }
    return false;
  }
  std::string string_body() {
    std::string out;
    ++i;  // opening quote
    while (i < s.size() && s[i] != '"') {
      char c = s[i++];
      if (c == '\\' && i < s.size()) {
        const char e = s[i++];
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'u': {  // keep BMP code points below 0x80, others as '?'
            unsigned cp = 0;
            for (int k = 0; k < 4 && i < s.size(); ++k) cp = cp * 16 + std::stoul(std::string(1, s[i++]), nullptr, 16);
            c = cp < 0x80 ? static_cast<char>(cp) : '?';
            break;
          }
          default: c = e;
        }
      }
      out.push_back(c);
    }
    if (i >= s.size()) ok = false;
    ++i;  // closing quote
    return out;
  }
  Value parse() {
    Value v;
    ws();
    if (i >= s.size()) {
      ok = false;
      return v;
    }
    const char c = s[i];
    if (c == '{') {
      v.kind = Value::Object;
      ++i;
      ws();
      if (i < s.size() && s[i] == '}') {
        ++i;
        return v;
      }
      while (ok) {
        ws();
        if (i >= s.size() || s[i] != '"') {
          ok = false;
          break;
        }
        std::string k = string_body();
        ws();
        if (i >= s.size() || s[i] != ':') {
          ok = false;
          break;
        }
        ++i;
        Value item = parse();
        v.obj.emplace_back(std::move(k), std::move(item));
        ws();
        if (i < s.size() && s[i] == ',') {
          ++i;
          continue;
        }
        if (i < s.size() && s[i] == '}') {
          ++i;
          break;
        }
        ok = false;
      }
    } else if (c == '[') {
      v.kind = Value::Array;
      ++i;
      ws();
      if (i < s.size() && s[i] == ']') {
        ++i;
        return v;
      }
      while (ok) {
        v.arr.push_back(parse());
        ws();
        if (i < s.size() && s[i] == ',') {
          ++i;
          continue;
        }
        if (i < s.size() && s[i] == ']') {
          ++i;
          break;
        }
        ok = false;
      }
    } else if (c == '"') {
      v.kind = Value::String;
      v.str = string_body();
    } else if (lit("true")) {
      v.kind = Value::Bool;
      v.b = true;
    } else if (lit("false")) {
      v.kind = Value::Bool;
    } else if (lit("null")) {
      v.kind = Value::Null;
    } else {
      const char* start = s.c_str() + i;
      char* end = nullptr;
      v.num = std::strtod(start, &end);
      if (end == start) {
        ok = false;
      } else {
        v.kind = Value::Number;
        i += static_cast<size_t>(end - start);
      }
    }
    return v;
  }
};

bool read_text(const std::string& path, std::string* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  out->assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

std::string dir_of(const std::string& path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

std::string lower(std::string v) {
  for (char& c : v) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return v;
}

uint64_t fnv1a64(const unsigned char* p, size_t n) {
  uint64_t h = 14695981039346656037ull;
  for (size_t k = 0; k < n; ++k) {
    h ^= p[k];
    h *= 1099511628211ull;
  }
  return h;
}

int parse_file(const char* path, Value* v, const char* who) {
  std::string text;
  if (!read_text(path, &text)) return fail(std::string(who) + ": cannot read metadata " + path);
  if (text.empty()) return fail(std::string(who) + ": empty metadata file");
  Parser ps(text);
  *v = ps.parse();
  if (!ps.ok || v->kind != Value::Object) return fail(std::string(who) + ": invalid JSON in " + path);
  return 0;
}

}  // namespace gj

// Volumes made here: descriptor address -> (descriptor, voxels).
namespace {
struct VolAlloc {
  GrmhdTex* desc;
  float4* data;
};
std::mutex g_mu;
std::map<unsigned long long, VolAlloc> g_vols;
}  // namespace

bool grmhd_volume_known(unsigned long long h) {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_vols.count(h) != 0;
}

// A volume like bh_grmhd_texture_create's, uploaded on stream s (host
// texels should be pinned for the copy to be asynchronous); returns once the
// upload has completed.  Released with bh_grmhd_texture_destroy.
int grmhd_volume_create_on_stream(const float* host_rgba, int w, int h, int d, hipStream_t s,
                                  unsigned long long* tex) {
  const size_t cells = static_cast<size_t>(w) * h * d;
  VolAlloc va{nullptr, nullptr};
  if (hipMalloc(&va.data, cells * sizeof(float4)) != hipSuccess) return fail("grmhd volume: hipMalloc failed");
  if (hipMalloc(&va.desc, sizeof(GrmhdTex)) != hipSuccess) {
    (void)hipFree(va.data);
    return fail("grmhd volume: hipMalloc failed");
  }
  const GrmhdTex dd{va.data, w, h, d, 0};
  if (hipMemcpyAsync(va.data, host_rgba, cells * sizeof(float4), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(va.desc, &dd, sizeof(dd), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipFree(va.data);
    (void)hipFree(va.desc);
    return fail("grmhd volume: upload failed");
  }
  const unsigned long long hd = reinterpret_cast<unsigned long long>(va.desc);
  std::lock_guard<std::mutex> lk(g_mu);
  g_vols[hd] = va;
  *tex = hd;
  return 0;
}

__global__ void k_grmhd_sample(unsigned long long left, unsigned long long right, float alpha, float r_min,
                               float r_max, const float* __restrict__ xyz, int n, float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = grmhd_sample(left, right, alpha, r_min, r_max, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

}  // namespace bhk

using namespace bhk;

extern "C" {

uint64_t bh_fnv1a64(const void* data, int64_t bytes) {
  if (!data || bytes <= 0) return gj::fnv1a64(nullptr, 0);
  return gj::fnv1a64(static_cast<const unsigned char*>(data), static_cast<size_t>(bytes));
}

int bh_grmhd_load_packed(const char* json_path, BH_GrmhdPacked* meta, float* rgba, int64_t rgba_floats) {
  const char* who = "bh_grmhd_load_packed";
  if (!json_path || !meta) return fail(std::string(who) + ": NULL argument");
  gj::Value js;
  if (gj::parse_file(json_path, &js, who) != 0) return -1;
  std::string format = "RGBA32F";
  if (const gj::Value* f = js.get("format"); f && f->kind == gj::Value::String) format = f->str;
  if (gj::lower(format) != "rgba32f") return fail(std::string(who) + ": unsupported texture format " + format);
  const gj::Value* bin = js.get("bin");
  if (!bin) bin = js.get("binary_file");
  if (!bin || bin->kind != gj::Value::String || bin->str.empty())
    return fail(std::string(who) + ": missing 'bin' (or 'binary_file') field");
  const gj::Value* gd = js.get("grid_dims");
  if (!gd || gd->kind != gj::Value::Array || gd->arr.size() != 3)
    return fail(std::string(who) + ": invalid or missing 'grid_dims' (expected [w, h, d])");
  int dims[3];
  for (int k = 0; k < 3; ++k) {
    const gj::Value& d = gd->arr[k];
    if (d.kind != gj::Value::Number || !(d.num >= 1.0) || d.num != std::floor(d.num) || d.num > 65536.0)
      return fail(std::string(who) + ": grid_dims must be positive integers");
    dims[k] = static_cast<int>(d.num);
  }
  const int64_t expected = static_cast<int64_t>(dims[0]) * dims[1] * dims[2] * 4;
  const std::string bin_path = gj::dir_of(json_path) + "/" + bin->str;
  std::ifstream bf(bin_path, std::ios::binary | std::ios::ate);
  if (!bf) return fail(std::string(who) + ": failed to read binary file " + bin_path);
  const std::streamsize size = bf.tellg();
  bf.seekg(0, std::ios::beg);
  if (size <= 0 || size % 4 != 0 || size / 4 != expected)
    return fail(std::string(who) + ": data size mismatch: expected " + std::to_string(expected) + " floats, got " +
                std::to_string(static_cast<long long>(size) / 4));
  std::vector<unsigned char> blob(static_cast<size_t>(size));
  if (!bf.read(reinterpret_cast<char*>(blob.data()), size)) return fail(std::string(who) + ": short read of " + bin_path);
  BH_GrmhdPacked m;
  std::memset(&m, 0, sizeof(m));
  m.width = dims[0];
  m.height = dims[1];
  m.depth = dims[2];
  m.floats = expected;
  m.checksum = gj::fnv1a64(blob.data(), blob.size());
  if (const gj::Value* cs = js.get("checksum_fnv1a64"); cs && cs->kind == gj::Value::String) {
    char hex[17];
    std::snprintf(hex, sizeof(hex), "%016llx", static_cast<unsigned long long>(m.checksum));
    if (gj::lower(cs->str) != hex)
      return fail(std::string(who) + ": checksum mismatch: expected " + cs->str + ", calculated " + hex);
    m.checksum_verified = 1;
  }
  const gj::Value* ch = js.get("channels");
  if (!ch || ch->kind != gj::Value::Array || ch->arr.size() != 4)
    return fail(std::string(who) + ": expected 4 channels, found " +
                std::to_string(ch && ch->kind == gj::Value::Array ? ch->arr.size() : 0));
  const gj::Value* jmin = js.get("min");
  const gj::Value* jmax = js.get("max");
  for (int c = 0; c < 4; ++c) {
    const gj::Value& e = ch->arr[c];
    std::string name = "unknown";
    float lo = 0.0f, hi = 1.0f;
    if (e.kind == gj::Value::Object) {  // legacy: {name, min, max}
      if (const gj::Value* n = e.get("name"); n && n->kind == gj::Value::String) name = n->str;
      if (const gj::Value* v = e.get("min"); v && v->kind == gj::Value::Number) lo = static_cast<float>(v->num);
      if (const gj::Value* v = e.get("max"); v && v->kind == gj::Value::Number) hi = static_cast<float>(v->num);
    } else if (e.kind == gj::Value::String) {
      name = e.str;
      if (jmin && jmin->kind == gj::Value::Array && c < static_cast<int>(jmin->arr.size()))
        lo = static_cast<float>(jmin->arr[c].num);
      if (jmax && jmax->kind == gj::Value::Array && c < static_cast<int>(jmax->arr.size()))
        hi = static_cast<float>(jmax->arr[c].num);
    } else {
      return fail(std::string(who) + ": channels must be strings or objects");
    }
    std::strncpy(m.channels[c], name.c_str(), sizeof(m.channels[c]) - 1);
    m.min[c] = lo;
    m.max[c] = hi;
  }
  m.r_min = 1.0f;
  m.r_max = 50.0f;
  if (const gj::Value* v = js.get("r_min"); v && v->kind == gj::Value::Number) m.r_min = static_cast<float>(v->num);
  if (const gj::Value* v = js.get("r_max"); v && v->kind == gj::Value::Number) m.r_max = static_cast<float>(v->num);
  if (rgba) {
    if (rgba_floats < expected) return fail(std::string(who) + ": rgba buffer too small");
    std::memcpy(rgba, blob.data(), blob.size());
  }
  *meta = m;
  return 0;
}

int bh_grmhd_timeseries_info(const char* json_path, int32_t* frame_count, int64_t* frame_bytes, int32_t* grid_dims,
                             double* times, int64_t* offsets, int32_t max_frames) {
  const char* who = "bh_grmhd_timeseries_info";
  if (!json_path || !frame_count || !frame_bytes || !grid_dims) return fail(std::string(who) + ": NULL argument");
  gj::Value js;
  if (gj::parse_file(json_path, &js, who) != 0) return -1;
  const gj::Value* fc = js.get("frameCount");
  const gj::Value* fs = js.get("frameSize");
  const gj::Value* fr = js.get("frames");
  const gj::Value* fm = js.get("frameMetadata");
  if (!fc || fc->kind != gj::Value::Number || !fs || fs->kind != gj::Value::Number || !fr ||
      fr->kind != gj::Value::Array || !fm || fm->kind != gj::Value::Object)
    return fail(std::string(who) + ": not a multi-frame metadata file (frameCount, frameSize, frames, frameMetadata)");
  const int n = static_cast<int>(fc->num);
  if (n <= 0 || static_cast<size_t>(n) != fr->arr.size())
    return fail(std::string(who) + ": frameCount does not match the frames array");
  const gj::Value* gd = fm->get("gridDims");
  if (!gd || gd->kind != gj::Value::Array || gd->arr.size() != 3)
    return fail(std::string(who) + ": frameMetadata.gridDims missing");
  int64_t cells = 4;
  for (int k = 0; k < 3; ++k) {
    grid_dims[k] = static_cast<int32_t>(gd->arr[k].num);
    if (grid_dims[k] <= 0) return fail(std::string(who) + ": gridDims must be positive");
    cells *= grid_dims[k];
  }
  const int64_t fsz = static_cast<int64_t>(fs->num);
  if (fsz != cells * 4) return fail(std::string(who) + ": frameSize is not gridDims x 4 channels x 4 bytes");
  for (int k = 0; k < n; ++k) {
    const gj::Value& f = fr->arr[k];
    const gj::Value* t = f.get("time");
    const gj::Value* o = f.get("offset");
    if (!t || !o || t->kind != gj::Value::Number || o->kind != gj::Value::Number)
      return fail(std::string(who) + ": frame entries need time and offset");
    if (k < max_frames) {
      if (times) times[k] = t->num;
      if (offsets) offsets[k] = static_cast<int64_t>(o->num);
    }
  }
  *frame_count = n;
  *frame_bytes = fsz;
  return 0;
}

int bh_grmhd_load_frame(const char* json_path, const char* bin_path, int32_t frame, float* rgba, int64_t rgba_floats) {
  const char* who = "bh_grmhd_load_frame";
  if (!json_path || !bin_path || !rgba) return fail(std::string(who) + ": NULL argument");
  int32_t n = 0, dims[3];
  int64_t fb = 0;
  std::vector<double> times(65536);
  std::vector<int64_t> offs(65536);
  if (bh_grmhd_timeseries_info(json_path, &n, &fb, dims, times.data(), offs.data(), 65536) != 0) return -1;
  if (frame < 0 || frame >= n || frame >= 65536) return fail(std::string(who) + ": frame out of range");
  if (rgba_floats < fb / 4) return fail(std::string(who) + ": rgba buffer too small");
  std::ifstream bf(bin_path, std::ios::binary);
  if (!bf) return fail(std::string(who) + ": cannot read " + std::string(bin_path));
  bf.seekg(offs[frame], std::ios::beg);
  if (!bf.read(reinterpret_cast<char*>(rgba), fb)) return fail(std::string(who) + ": short read of frame");
  return 0;
}

int bh_grmhd_frame_pair(const double* times, int32_t n, double t, int32_t* i0, float* alpha) {
  if (!times || n <= 0 || !i0 || !alpha) return fail("bh_grmhd_frame_pair: invalid arguments");
  if (n == 1) {
    *i0 = 0;
    *alpha = 0.0f;
    return 0;
  }
  int k = 0;
  while (k + 1 < n - 1 && times[k + 1] <= t) ++k;
  const double dt = times[k + 1] - times[k];
  double a = dt > 0.0 ? (t - times[k]) / dt : 0.0;
  a = std::min(std::max(a, 0.0), 1.0);
  *i0 = k;
  *alpha = static_cast<float>(a);
  return 0;
}

int bh_grmhd_texture_create(const float* host_rgba, int32_t width, int32_t height, int32_t depth,
                            unsigned long long* tex) {
  const char* who = "bh_grmhd_texture_create";
  if (!host_rgba || !tex || width <= 0 || height <= 0 || depth <= 0) return fail(std::string(who) + ": invalid arguments");
  const size_t cells = static_cast<size_t>(width) * height * depth;
  VolAlloc va{nullptr, nullptr};
  if (hipMalloc(&va.data, cells * sizeof(float4)) != hipSuccess) return fail(std::string(who) + ": hipMalloc failed");
  if (hipMalloc(&va.desc, sizeof(GrmhdTex)) != hipSuccess) {
    (void)hipFree(va.data);
    return fail(std::string(who) + ": hipMalloc failed");
  }
  GrmhdTex d{va.data, width, height, depth, 0};
  if (hipMemcpy(va.data, host_rgba, cells * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(va.desc, &d, sizeof(d), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(va.data);
    (void)hipFree(va.desc);
    return fail(std::string(who) + ": upload failed");
  }
  const unsigned long long h = reinterpret_cast<unsigned long long>(va.desc);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_vols[h] = va;
  }
  *tex = h;
  return 0;
}

int bh_grmhd_texture_destroy(unsigned long long tex) {
  VolAlloc va{nullptr, nullptr};
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_vols.find(tex);
    if (it == g_vols.end()) return fail("bh_grmhd_texture_destroy: not a volume made by bh_grmhd_texture_create");
    va = it->second;
    g_vols.erase(it);
  }
  if (hipDeviceSynchronize() != hipSuccess) return fail("bh_grmhd_texture_destroy: device error");
  (void)hipFree(va.data);
  (void)hipFree(va.desc);
  return 0;
}

int bh_grmhd_sample(unsigned long long tex, unsigned long long tex_right, float alpha, float r_min, float r_max,
                    const float* xyz, int32_t n, float* out) {
  const char* who = "bh_grmhd_sample";
  if (!xyz || !out || n <= 0) return fail(std::string(who) + ": invalid arguments");
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (tex != 0 && !g_vols.count(tex)) return fail(std::string(who) + ": unknown left volume");
    if (tex_right != 0 && !g_vols.count(tex_right)) return fail(std::string(who) + ": unknown right volume");
  }
  float* d_xyz = nullptr;
  float4* d_out = nullptr;
  if (hipMalloc(&d_xyz, static_cast<size_t>(n) * 3 * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_out, static_cast<size_t>(n) * sizeof(float4)) != hipSuccess) {
    if (d_xyz) (void)hipFree(d_xyz);
    return fail(std::string(who) + ": hipMalloc failed");
  }
  int rc = 0;
  if (hipMemcpy(d_xyz, xyz, static_cast<size_t>(n) * 3 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail(std::string(who) + ": upload failed");
  } else {
    hipLaunchKernelGGL(k_grmhd_sample, dim3((n + 255) / 256), dim3(256), 0, 0, tex, tex_right, alpha, r_min, r_max,
                       d_xyz, n, d_out);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(out, d_out, static_cast<size_t>(n) * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(std::string(who) + ": sampling failed");
  }
  (void)hipFree(d_xyz);
  (void)hipFree(d_out);
  return rc;
}

}  // extern "C"
