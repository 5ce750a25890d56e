// Host-side checks of the GRMHD frame cache (blackhole_amd/csrc/grmhd_stream.hpp)
// with a backend that keeps frames in host memory and counts loads: the
// behaviour of the reference's TileCache / GRMHDStreamer
// (src/grmhd_streaming.cpp): misses queue a load and return nothing, hits
// return the frame, a seek prefetches the next frames, least-recently-used
// eviction under the byte budget, the current and next frames kept.  Prints
// "ok" or the first failed check.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../blackhole_amd/csrc/grmhd_stream.hpp"

namespace {

struct HostFrames {
  int64_t floats;
  std::mutex mu;
  std::map<unsigned long long, std::vector<float>> live;
  std::atomic<int> loads{0}, releases{0};
  int fail_frame = -1;
  unsigned long long next = 1;
  unsigned long long load(int32_t f) {
    std::this_thread::sleep_for(std::chrono::milliseconds(2));  // a disk read
    if (f == fail_frame) return 0;
    std::lock_guard<std::mutex> lk(mu);
    const unsigned long long h = next++;
    live[h].assign(floats, static_cast<float>(f));
    ++loads;
    return h;
  }
  void release(unsigned long long h) {
    std::lock_guard<std::mutex> lk(mu);
    live.erase(h);
    ++releases;
  }
  float value(unsigned long long h) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = live.find(h);
    return it == live.end() ? -1.0f : it->second[0];
  }
};

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("failed line %d: %s\n", __LINE__, #c);           \
      return 1;                                                    \
    }                                                              \
  } while (0)

bool wait_until(const std::function<bool()>& f) {
  for (int i = 0; i < 2000; ++i) {
    if (f()) return true;
    std::this_thread::sleep_for(std::chrono::milliseconds(1));
  }
  return false;
}

}  // namespace

int main() {
  const int64_t fb = 4096;  // bytes per frame
  HostFrames be;
  be.floats = fb / 4;
  be.fail_frame = 9;
  {
    bhk::FrameStreamer<HostFrames> s(&be, 10, fb, 4 * fb, 3);
    // a miss queues the load and returns nothing; waiting returns the frame
    CHECK(s.get(0, false) == 0);
    CHECK(wait_until([&] { return s.resident(0); }));
    const auto h0 = s.get(0, false);
    CHECK(h0 != 0 && be.value(h0) == 0.0f);
    auto st = s.stats();
    CHECK(st.hits == 1 && st.misses == 1);
    // a seek prefetches the next three frames
    s.seek(2);
    CHECK(wait_until([&] { return s.resident(3) && s.resident(4) && s.resident(5); }));
    CHECK(s.get(1, false) != 0 && be.value(s.get(1, false)) == 3.0f);  // adjacent frame from the prefetch
    // the adjacent frame is only looked up: a miss does not queue it
    s.seek(8);
    CHECK(wait_until([&] { return s.stats().queue_depth == 0; }));
    CHECK(s.get(1, false) == 0);  // frame 9 fails to load
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    CHECK(!s.resident(9));
    const auto h8 = s.get(0, true);
    CHECK(h8 != 0 && be.value(h8) == 8.0f);
    // budget: at most four frames resident, least recently used evicted first
    st = s.stats();
    CHECK(st.resident_frames <= 4 && st.resident_bytes <= 4 * fb);
    // frames 2 and 8 recently used; 0, 3, 4, 5 older: 8 must stay (current)
    CHECK(s.resident(8));
    s.seek(0);
    const auto g0 = s.get(0, true);
    CHECK(g0 != 0 && be.value(g0) == 0.0f);
    CHECK(wait_until([&] { return s.stats().queue_depth == 0; }));
    st = s.stats();
    CHECK(st.resident_frames <= 4 && st.resident_bytes == st.resident_frames * fb);
    CHECK(s.resident(0) && s.resident(1));  // current and next are never evicted
    // waiting on a frame that cannot load returns nothing instead of hanging
    s.seek(9);
    CHECK(s.get(0, true) == 0);
    CHECK(s.stats().current_frame == 9);
    // seeks clamp to the series
    s.seek(100);
    CHECK(s.stats().current_frame == 9);
    s.seek(-5);
    CHECK(s.stats().current_frame == 0);
  }
  // every frame made resident was released at shutdown
  CHECK(be.loads.load() == be.releases.load());
  CHECK(be.live.empty());
  // a budget of one frame still keeps the current and next ones
  {
    HostFrames b2;
    b2.floats = fb / 4;
    bhk::FrameStreamer<HostFrames> s(&b2, 5, fb, fb, 3);
    s.seek(1);
    CHECK(s.get(0, true) != 0);
    CHECK(wait_until([&] { return s.stats().queue_depth == 0; }));
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    CHECK(s.resident(1) && s.resident(2));
    CHECK(s.stats().resident_frames == 2);
  }
  std::printf("ok\n");
  return 0;
}
