// Autotuning of launch choices that change no result bit.
//
// A choice is tuned per device and per key (kernel family, dimension, grade,
// size class = ceil(log2 n)): the first call that needs it times every
// candidate on the device (HIP events on the context's stream, one warm-up
// launch, then the median of three), keeps the fastest in a process-wide
// table, and later calls take it from there. Every candidate writes the same
// outputs bit for bit (only which lane computes an entry, or how many launches
// do the work, differs), so the timing launches are real work and the outputs
// the call returns are the last candidate's, identical to the chosen one's.
// A stream under graph capture is never timed: it takes the table's choice,
// or candidate 0 before one exists.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "feec_internal.hpp"

namespace feec {
namespace {

struct Entry {
  int choice = 0;
  std::vector<float> ms;  // median time of each candidate
};
std::mutex g_mu;
std::map<std::tuple<int, std::string, int, int, int>, Entry> g_table;
std::atomic<int> g_mode{0};

int size_class(int64_t n) {
  int c = 0;
  while (c < 62 && (int64_t(1) << c) < n) ++c;
  return c;
}

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

}  // namespace

int tune_run(Ctx* ctx, const char* family, int d, int k, int64_t n, int ncand,
             const std::function<void(int)>& launch) {
  FEEC_REQUIRE(ncand >= 1, FEEC_E_ARG, "no candidates");
  const int mode = g_mode.load();
  if (mode > 0) {
    const int c = std::min(mode - 1, ncand - 1);
    launch(c);
    return c;
  }
  const auto key = std::make_tuple(ctx->device, std::string(family), d, k, size_class(n));
  {
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_table.find(key);
    if (it != g_table.end()) {
      launch(it->second.choice);
      return it->second.choice;
    }
  }
  if (ncand == 1 || capturing(ctx->stream)) {
    launch(0);
    return 0;
  }
  hipEvent_t e0, e1;
  FEEC_HIP(hipEventCreate(&e0));
  FEEC_HIP(hipEventCreate(&e1));
  Entry e;
  e.ms.assign(ncand, 0.f);
  try {
    for (int c = 0; c < ncand; ++c) {
      launch(c);  // warm-up (first-launch code loading, caches)
      float t[3];
      for (int r = 0; r < 3; ++r) {
        FEEC_HIP(hipEventRecord(e0, ctx->stream));
        launch(c);
        FEEC_HIP(hipEventRecord(e1, ctx->stream));
        FEEC_HIP(hipEventSynchronize(e1));
        FEEC_HIP(hipEventElapsedTime(&t[r], e0, e1));
      }
      std::sort(t, t + 3);
      e.ms[c] = t[1];
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  e.choice = (int)(std::min_element(e.ms.begin(), e.ms.end()) - e.ms.begin());
  {
    std::lock_guard<std::mutex> lock(g_mu);
    g_table.emplace(key, e);
  }
  return e.choice;
}

}  // namespace feec

using namespace feec;

extern "C" {

feec_status feec_tune_mode(int mode) {
  return guard([&] {
    FEEC_REQUIRE(mode >= 0 && mode <= 64, FEEC_E_ARG, "mode 0 (tune) or a candidate number 1..64");
    g_mode.store(mode);
  });
}

feec_status feec_tune_table(char* out, int64_t* len) {
  return guard([&] {
    FEEC_REQUIRE(len, FEEC_E_ARG, "null argument");
    std::string s;
    {
      std::lock_guard<std::mutex> lock(g_mu);
      char line[256];
      for (const auto& kv : g_table) {
        std::snprintf(line, sizeof line, "device %d %s d=%d k=%d n<=2^%d choice %d ms",
                      std::get<0>(kv.first), std::get<1>(kv.first).c_str(), std::get<2>(kv.first),
                      std::get<3>(kv.first), std::get<4>(kv.first), kv.second.choice);
        s += line;
        for (float m : kv.second.ms) {
          std::snprintf(line, sizeof line, " %.4f", m);
          s += line;
        }
        s += "\n";
      }
    }
    if (out) {
      FEEC_REQUIRE(*len >= (int64_t)s.size() + 1, FEEC_E_ARG, "output buffer too small");
      std::copy(s.begin(), s.end(), out);
      out[s.size()] = 0;
    }
    *len = (int64_t)s.size() + 1;
  });
}

}  // extern "C"
