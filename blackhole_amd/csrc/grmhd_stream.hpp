// grmhd_stream.hpp - frame streaming of GRMHD time series: a least-recently-
// used cache of resident frames under a byte budget, filled by a background
// loader thread, with prefetch of the frames after the current one (the
// behaviour of the reference's GRMHDStreamer / TileCache,
// src/grmhd_streaming.h:125-301, whose tiles are whole frames,
// src/grmhd_streaming.cpp:355-411).
//
// Plain C++ (no HIP): the residency of a frame is delegated to a Backend,
//   Handle Backend::load(int frame)      read + make the frame resident (0 on failure)
//   void   Backend::release(Handle h)    free it
// so the same logic keeps frames in device memory in the library
// (grmhd_stream.hip) and in host memory in its CPU test.
//
// Eviction never takes the current frame or the next one: handles given out
// for them stay valid until a later seek moves the current frame away.
// When only protected frames remain, the cache may exceed its budget.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

namespace bhk {

template <class Backend>
class FrameStreamer {
 public:
  using Handle = unsigned long long;

  struct Stats {
    uint64_t hits = 0, misses = 0;
    int64_t resident_bytes = 0, budget_bytes = 0;
    int32_t resident_frames = 0, queue_depth = 0, current_frame = 0, frame_count = 0;
  };

  FrameStreamer(Backend* backend, int32_t frame_count, int64_t frame_bytes, int64_t budget_bytes, int32_t prefetch)
      : be_(backend), n_(frame_count), frame_bytes_(frame_bytes), budget_(budget_bytes), prefetch_(prefetch) {
    running_ = true;
    loader_ = std::thread([this] { loader_loop(); });
  }

  ~FrameStreamer() { shutdown(); }

  FrameStreamer(const FrameStreamer&) = delete;
  FrameStreamer& operator=(const FrameStreamer&) = delete;

  // Stops the loader (the queue is dropped) and releases every resident frame.
  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (!running_) return;
      running_ = false;
      queue_.clear();
    }
    cv_.notify_all();
    done_cv_.notify_all();
    if (loader_.joinable()) loader_.join();
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : cache_) be_->release(kv.second.h);
    cache_.clear();
    bytes_ = 0;
  }

  // seekFrame: the current frame becomes `frame` (clamped to the series);
  // the next `prefetch` frames that are not resident are queued.
  void seek(int32_t frame) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      cur_ = std::min(std::max(frame, 0), n_ - 1);
      for (int32_t i = 1; i <= prefetch_; ++i) {
        const int32_t f = cur_ + i;
        if (f < n_ && !cache_.count(f)) enqueue_locked(f);
      }
    }
    cv_.notify_all();
  }

  // getTile of the current frame (offset 0) or getAdjacentTile (offset 1):
  // the resident frame's handle (a hit), else 0 (a miss; for the current
  // frame its load is queued).  wait: block until the frame is resident
  // (or cannot be loaded).
  Handle get(int32_t offset, bool wait) {
    std::unique_lock<std::mutex> lk(mu_);
    const int32_t f = cur_ + offset;
    if (f < 0 || f >= n_) return 0;
    auto it = cache_.find(f);
    if (it != cache_.end()) {
      ++hits_;
      it->second.stamp = ++clock_;
      return it->second.h;
    }
    ++misses_;
    if (offset != 0 && !wait) return 0;  // the adjacent frame is only looked up
    failed_.erase(f);
    const uint64_t c0 = loads_[f];
    enqueue_locked(f);
    cv_.notify_all();
    if (!wait) return 0;
    // ends when the loader has finished (or failed) a load of f, or stops
    done_cv_.wait(lk, [&] { return loads_[f] != c0 || failed_.count(f) || !running_; });
    auto jt = cache_.find(f);
    if (jt == cache_.end()) return 0;  // failed, stopped, or already evicted again
    jt->second.stamp = ++clock_;
    return jt->second.h;
  }

  Stats stats() const {
    std::lock_guard<std::mutex> lk(mu_);
    Stats s;
    s.hits = hits_;
    s.misses = misses_;
    s.resident_bytes = bytes_;
    s.budget_bytes = budget_;
    s.resident_frames = static_cast<int32_t>(cache_.size());
    s.queue_depth = static_cast<int32_t>(queue_.size());
    s.current_frame = cur_;
    s.frame_count = n_;
    return s;
  }

  bool resident(int32_t f) const {
    std::lock_guard<std::mutex> lk(mu_);
    return cache_.count(f) != 0;
  }

 private:
  struct Entry {
    Handle h;
    uint64_t stamp;  // last access, for least-recently-used eviction
  };

  void enqueue_locked(int32_t f) {
    if (std::find(queue_.begin(), queue_.end(), f) == queue_.end()) queue_.push_back(f);
  }

  // Evicts least-recently-used frames other than the current and the next
  // until `need` more bytes fit the budget (TileCache::put, :67-84).
  void make_room_locked(int64_t need) {
    while (bytes_ + need > budget_) {
      auto victim = cache_.end();
      for (auto it = cache_.begin(); it != cache_.end(); ++it) {
        if (it->first == cur_ || it->first == cur_ + 1) continue;
        if (victim == cache_.end() || it->second.stamp < victim->second.stamp) victim = it;
      }
      if (victim == cache_.end()) return;  // only protected frames left
      be_->release(victim->second.h);
      bytes_ -= frame_bytes_;
      cache_.erase(victim);
    }
  }

  void loader_loop() {
    for (;;) {
      int32_t f;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !running_ || !queue_.empty(); });
        if (!running_) return;
        f = queue_.front();
        queue_.pop_front();
        if (cache_.count(f)) continue;  // loaded meanwhile
      }
      const Handle h = be_->load(f);  // outside the lock: disk read and upload
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (!h) {
          failed_[f] = true;
        } else if (!running_) {
          be_->release(h);
        } else {
          make_room_locked(frame_bytes_);
          cache_[f] = Entry{h, ++clock_};
          bytes_ += frame_bytes_;
          failed_.erase(f);
        }
        ++loads_[f];
      }
      done_cv_.notify_all();
    }
  }

  Backend* be_;
  const int32_t n_;
  const int64_t frame_bytes_, budget_;
  const int32_t prefetch_;
  mutable std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<int32_t> queue_;
  std::map<int32_t, Entry> cache_;
  std::map<int32_t, bool> failed_;
  std::map<int32_t, uint64_t> loads_;  // finished load attempts per frame
  int64_t bytes_ = 0;
  uint64_t clock_ = 0, hits_ = 0, misses_ = 0;
  int32_t cur_ = 0;
  bool running_ = false;
  std::thread loader_;
};

}  // namespace bhk
