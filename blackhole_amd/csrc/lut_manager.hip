// lut_manager.hip - the LUT slot registry (bh_lut_manager.h): device copies
// of the render path's lookup tables in the eight slots of
// src/cuda/lut_manager.h, with the GRMHD slots holding the volumes the
// render launch samples.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "../../include/bh_grmhd.h"
#include "../../include/bh_kernel_launch.h"
#include "../../include/bh_lut_manager.h"
#include "grmhd_device.hpp"

namespace bhk {
int fail(const std::string& msg);
}  // namespace bhk

using bhk::fail;

struct BH_LutManager {
  unsigned long long tex[BhLutCount] = {};
  unsigned long long background = 0;
};

namespace {

constexpr int kSlotBackground = BhLutCount;  // registry tag of the background table

// Live handles made here: geometry and the device texels, for bh_lut_read
// and for freeing.  GRMHD handles belong to bh_grmhd_texture_create.
struct Entry {
  BH_LutImage meta;
  float* d_data;           // table texels (NULL for a GRMHD volume)
  BH_LutTableDev* d_desc;  // table descriptor (NULL for a GRMHD volume)
  bool grmhd;
};
std::mutex g_mu;
std::map<unsigned long long, Entry> g_live;

bool is_grmhd_slot(int slot) { return slot == BhLutGrmhd || slot == BhLutGrmhdRight; }

// The table geometry each slot takes (lut_manager.cu registers GL_TEXTURE_2D
// for every slot but the galaxy cubemap; the GRMHD slots hold 3-D volumes).
int check_image(int slot, const BH_LutImage* img, const char* who) {
  const std::string w(who);
  if (!img || !img->host) return fail(w + ": NULL image");
  if (img->width <= 0 || img->height <= 0 || img->depth <= 0 || img->layers <= 0)
    return fail(w + ": empty image");
  if (img->channels != 1 && img->channels != 4) return fail(w + ": channels must be 1 or 4");
  if (is_grmhd_slot(slot)) {
    if (img->channels != 4 || img->layers != 1) return fail(w + ": a GRMHD slot takes one RGBA32F volume");
  } else if (slot == BhLutGalaxy) {
    if (img->layers != 6 || img->depth != 1) return fail(w + ": the galaxy slot takes a 6-layer cubemap");
  } else if (img->depth != 1 || img->layers != 1) {
    return fail(w + ": this slot takes a 2-D table");
  }
  const double n = static_cast<double>(img->width) * img->height * img->depth * img->channels * img->layers;
  if (n > 4.0e9) return fail(w + ": table too large");
  return 0;
}

int64_t image_floats(const BH_LutImage& m) {
  return static_cast<int64_t>(m.width) * m.height * m.depth * m.channels * m.layers;
}

// Device copy of a table; its handle in *h.
int make_table(int slot, const BH_LutImage& img, unsigned long long* h, const char* who) {
  const std::string w(who);
  if (is_grmhd_slot(slot)) {
    unsigned long long v = 0;
    if (bh_grmhd_texture_create(img.host, img.width, img.height, img.depth, &v) != 0) return -1;
    Entry e{img, nullptr, nullptr, true};
    e.meta.host = nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    g_live[v] = e;
    *h = v;
    return 0;
  }
  const size_t bytes = static_cast<size_t>(image_floats(img)) * sizeof(float);
  float* d = nullptr;
  BH_LutTableDev* desc = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) return fail(w + ": hipMalloc failed");
  if (hipMalloc(&desc, sizeof(BH_LutTableDev)) != hipSuccess) {
    (void)hipFree(d);
    return fail(w + ": hipMalloc failed");
  }
  const BH_LutTableDev hd{d, img.width, img.height, img.depth, img.channels, img.layers, slot};
  if (hipMemcpy(d, img.host, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(desc, &hd, sizeof(hd), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    (void)hipFree(desc);
    return fail(w + ": upload failed");
  }
  Entry e{img, d, desc, false};
  e.meta.host = nullptr;
  const unsigned long long v = reinterpret_cast<unsigned long long>(desc);
  std::lock_guard<std::mutex> lk(g_mu);
  g_live[v] = e;
  *h = v;
  return 0;
}

// Frees a handle made by make_table once the device is idle (a launch may
// still read the table).
void free_table(unsigned long long h) {
  if (!h) return;
  Entry e{};
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_live.find(h);
    if (it == g_live.end()) return;
    e = it->second;
    g_live.erase(it);
  }
  if (e.grmhd) {
    (void)bh_grmhd_texture_destroy(h);
    return;
  }
  (void)hipDeviceSynchronize();
  (void)hipFree(e.d_data);
  (void)hipFree(e.d_desc);
}

}  // namespace

extern "C" {

BH_LutManager* bh_lut_manager_create(void) { return new BH_LutManager(); }

int bh_lut_register(BH_LutManager* mgr, int slot, const BH_LutImage* img) {
  const char* who = "bh_lut_register";
  if (!mgr) return fail(std::string(who) + ": NULL manager");
  if (slot < 0 || slot >= BhLutCount) return fail(std::string(who) + ": invalid slot " + std::to_string(slot));
  if (check_image(slot, img, who) != 0) return -1;
  bh_lut_unregister(mgr, slot);
  unsigned long long h = 0;
  if (make_table(slot, *img, &h, who) != 0) return -1;
  mgr->tex[slot] = h;
  return 0;
}

void bh_lut_unregister(BH_LutManager* mgr, int slot) {
  if (!mgr || slot < 0 || slot >= BhLutCount) return;
  free_table(mgr->tex[slot]);
  mgr->tex[slot] = 0;
}

unsigned long long bh_lut_get_tex(const BH_LutManager* mgr, int slot) {
  if (!mgr || slot < 0 || slot >= BhLutCount) return 0ull;
  return mgr->tex[slot];
}

int bh_lut_register_background(BH_LutManager* mgr, const BH_LutImage* img) {
  const char* who = "bh_lut_register_background";
  if (!mgr) return fail(std::string(who) + ": NULL manager");
  if (check_image(BhLutEmissivity, img, who) != 0) return -1;  // a 2-D table, as the 2-D slots
  bh_lut_unregister_background(mgr);
  unsigned long long h = 0;
  if (make_table(kSlotBackground, *img, &h, who) != 0) return -1;
  mgr->background = h;
  return 0;
}

void bh_lut_unregister_background(BH_LutManager* mgr) {
  if (!mgr) return;
  free_table(mgr->background);
  mgr->background = 0;
}

unsigned long long bh_lut_get_background_tex(const BH_LutManager* mgr) { return mgr ? mgr->background : 0ull; }

void bh_lut_shutdown(BH_LutManager* mgr) {
  if (!mgr) return;
  for (int s = 0; s < BhLutCount; ++s) bh_lut_unregister(mgr, s);
  bh_lut_unregister_background(mgr);
}

void bh_lut_manager_destroy(BH_LutManager* mgr) {
  if (!mgr) return;
  bh_lut_shutdown(mgr);
  delete mgr;
}

int bh_lut_upload(const BH_LutManager* mgr) {
  if (!mgr) return fail("bh_lut_upload: NULL manager");
  const unsigned long long* t = mgr->tex;
  bh_upload_lut_textures(t[BhLutEmissivity], t[BhLutRedshift], t[BhLutSpectral], t[BhLutGrb], t[BhLutGalaxy],
                         t[BhLutGrmhd], t[BhLutSynchG], t[BhLutGrmhdRight]);
  bh_upload_bridge_background_texture(mgr->background);
  return 0;
}

int bh_lut_read(unsigned long long handle, BH_LutImage* meta, float* host_out, int64_t host_floats) {
  const char* who = "bh_lut_read";
  if (!meta) return fail(std::string(who) + ": NULL meta");
  Entry e{};
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_live.find(handle);
    if (it == g_live.end()) return fail(std::string(who) + ": not a handle of a registered table");
    e = it->second;
  }
  *meta = e.meta;
  if (!host_out) return 0;
  const int64_t n = image_floats(e.meta);
  if (host_floats < n) return fail(std::string(who) + ": host buffer too small");
  const float* src = e.d_data;
  if (e.grmhd) {  // the volume's texels sit behind its device descriptor
    bhk::GrmhdTex gd;
    if (hipMemcpy(&gd, reinterpret_cast<const void*>(handle), sizeof(gd), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(std::string(who) + ": descriptor read failed");
    src = reinterpret_cast<const float*>(gd.data);
  }
  if (hipMemcpy(host_out, src, static_cast<size_t>(n) * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(std::string(who) + ": read failed");
  return 0;
}

}  // extern "C"
