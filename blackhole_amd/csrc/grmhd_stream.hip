// grmhd_stream.hip - GRMHD time-series streaming into device memory
// (bh_grmhd.h, bh_grmhd_streamer_*): the frame cache of grmhd_stream.hpp
// with a backend that reads a frame into pinned host memory and uploads it
// as a render-launch volume on the streamer's own non-blocking stream.
#include <hip/hip_runtime.h>

#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/bh_grmhd.h"
#include "grmhd_stream.hpp"

namespace bhk {
int fail(const std::string& msg);
int grmhd_volume_create_on_stream(const float* host_rgba, int w, int h, int d, hipStream_t s,
                                  unsigned long long* tex);
}  // namespace bhk

using bhk::fail;

namespace {

// Reads frames of one time series and makes them device volumes.  Used by
// the loader thread only.
struct DeviceFrames {
  std::string bin;
  int32_t dims[3] = {0, 0, 0};
  int64_t frame_bytes = 0;
  std::vector<int64_t> offsets;
  float* staging = nullptr;  // pinned, one frame
  hipStream_t stream = nullptr;
  int device = 0;

  unsigned long long load(int32_t f) {
    if (hipSetDevice(device) != hipSuccess) return 0;
    std::ifstream in(bin, std::ios::binary);
    if (!in) return 0;
    in.seekg(offsets[f], std::ios::beg);
    if (!in.read(reinterpret_cast<char*>(staging), frame_bytes)) return 0;
    unsigned long long h = 0;
    if (bhk::grmhd_volume_create_on_stream(staging, dims[0], dims[1], dims[2], stream, &h) != 0) return 0;
    return h;
  }
  void release(unsigned long long h) { (void)bh_grmhd_texture_destroy(h); }
};

}  // namespace

struct BH_GrmhdStreamer {
  DeviceFrames be;
  std::vector<double> times;
  std::unique_ptr<bhk::FrameStreamer<DeviceFrames>> st;
  ~BH_GrmhdStreamer() {
    st.reset();  // joins the loader and frees the frames
    if (be.staging) (void)hipHostFree(be.staging);
    if (be.stream) (void)hipStreamDestroy(be.stream);
  }
};

extern "C" {

BH_GrmhdStreamer* bh_grmhd_streamer_create(const char* json_path, const char* bin_path, int64_t cache_bytes,
                                           int32_t prefetch) {
  const std::string who = "bh_grmhd_streamer_create";
  if (!json_path || !bin_path || prefetch < 0) {
    fail(who + ": invalid arguments");
    return nullptr;
  }
  int32_t n = 0, dims[3] = {0, 0, 0};
  int64_t fb = 0;
  if (bh_grmhd_timeseries_info(json_path, &n, &fb, dims, nullptr, nullptr, 0) != 0) return nullptr;
  if (n <= 0 || n > (1 << 20)) {
    fail(who + ": frame count out of range");
    return nullptr;
  }
  const int64_t want = static_cast<int64_t>(dims[0]) * dims[1] * dims[2] * 16;
  if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0 || fb != want) {
    fail(who + ": frames must be RGBA32F volumes of grid_dims");
    return nullptr;
  }
  if (cache_bytes < fb) {
    fail(who + ": cache budget below one frame (" + std::to_string(fb) + " bytes)");
    return nullptr;
  }
  std::unique_ptr<BH_GrmhdStreamer> s(new BH_GrmhdStreamer());
  s->times.assign(n, 0.0);
  s->be.offsets.assign(n, 0);
  if (bh_grmhd_timeseries_info(json_path, &n, &fb, dims, s->times.data(), s->be.offsets.data(), n) != 0)
    return nullptr;
  {
    std::ifstream in(bin_path, std::ios::binary | std::ios::ate);
    if (!in) {
      fail(who + ": cannot read " + std::string(bin_path));
      return nullptr;
    }
    const int64_t size = static_cast<int64_t>(in.tellg());
    for (int32_t f = 0; f < n; ++f)
      if (s->be.offsets[f] < 0 || s->be.offsets[f] + fb > size) {
        fail(who + ": frame " + std::to_string(f) + " lies beyond the end of the binary");
        return nullptr;
      }
  }
  s->be.bin = bin_path;
  for (int i = 0; i < 3; ++i) s->be.dims[i] = dims[i];
  s->be.frame_bytes = fb;
  if (hipGetDevice(&s->be.device) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&s->be.staging), static_cast<size_t>(fb)) != hipSuccess ||
      hipStreamCreateWithFlags(&s->be.stream, hipStreamNonBlocking) != hipSuccess) {
    fail(who + ": device set-up failed");
    return nullptr;
  }
  s->st.reset(new bhk::FrameStreamer<DeviceFrames>(&s->be, n, fb, cache_bytes, prefetch));
  return s.release();
}

void bh_grmhd_streamer_destroy(BH_GrmhdStreamer* s) { delete s; }

int bh_grmhd_streamer_seek(BH_GrmhdStreamer* s, int32_t frame) {
  if (!s) return fail("bh_grmhd_streamer_seek: NULL streamer");
  s->st->seek(frame);
  return 0;
}

int bh_grmhd_streamer_get_frame(BH_GrmhdStreamer* s, int32_t offset, int32_t wait, unsigned long long* tex) {
  if (!s || !tex || (offset != 0 && offset != 1)) return fail("bh_grmhd_streamer_get_frame: invalid arguments");
  *tex = s->st->get(offset, wait != 0);
  return *tex ? 0 : 1;
}

int bh_grmhd_streamer_frame_pair(BH_GrmhdStreamer* s, double t, int32_t wait, unsigned long long* left,
                                 unsigned long long* right, float* alpha) {
  if (!s || !left || !right || !alpha) return fail("bh_grmhd_streamer_frame_pair: invalid arguments");
  int32_t i0 = 0;
  if (bh_grmhd_frame_pair(s->times.data(), static_cast<int32_t>(s->times.size()), t, &i0, alpha) != 0) return -1;
  s->st->seek(i0);
  *left = s->st->get(0, wait != 0);
  *right = s->times.size() > 1 ? s->st->get(1, wait != 0) : 0ull;
  return (*left && (*right || s->times.size() == 1)) ? 0 : 1;
}

int bh_grmhd_streamer_stats(const BH_GrmhdStreamer* s, BH_GrmhdStreamStats* out) {
  if (!s || !out) return fail("bh_grmhd_streamer_stats: invalid arguments");
  const auto st = s->st->stats();
  out->hits = st.hits;
  out->misses = st.misses;
  out->resident_bytes = st.resident_bytes;
  out->budget_bytes = st.budget_bytes;
  out->resident_frames = st.resident_frames;
  out->queue_depth = st.queue_depth;
  out->current_frame = st.current_frame;
  out->frame_count = st.frame_count;
  return 0;
}

}  // extern "C"
