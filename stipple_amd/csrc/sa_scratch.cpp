// sa_scratch.cpp -- device scratch blocks reused across calls.
//
// Every engine call needs scratch (staged inputs, memo tables, the octree
// build's pair and node arrays) for the duration of its work on a stream.
// Stream-ordered allocation (hipMallocAsync / hipFreeAsync) would serve, but
// on this runtime a hipFreeAsync of a large block can hold the host for
// hundreds of milliseconds once the blocks requested vary in size -- which is
// what a serving loop or a batch builder of varying batch sizes does.  So the
// library keeps its own cache: blocks come from hipMalloc, go back to the
// cache when a call is done with them (an event recorded on the call's
// stream marks the end of their last use), and are handed out again best-fit
// to the same stream, or to another one once that event has completed (a
// stream never waits for another stream's queued work through the cache; it
// gets a new block instead).  Free cached bytes above kKeep go back to the
// runtime, oldest first.  A stream under graph capture allocates stream-ordered instead (a
// captured graph owns such memory; hipMalloc is not allowed during capture).
//
// Host code only: no device work.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "sa_internal.h"

namespace sa {
namespace {

struct Block {
  void* p = nullptr;
  size_t bytes = 0;
  hipEvent_t ev = nullptr;  // recorded where the last user's work ends
  hipStream_t last = nullptr;
  bool busy = false;
  uint64_t stamp = 0;  // when it was last given back (oldest leave first)
};

std::mutex g_mu;
std::vector<Block> g_blocks;
size_t g_free = 0;
uint64_t g_clock = 0;
constexpr size_t kKeep = (size_t)32 << 30;
constexpr size_t kGrain = (size_t)2 << 20;

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// Frees the free block at `i` once its last use is over (caller holds g_mu).
void drop(size_t i) {
  Block& b = g_blocks[i];
  (void)hipEventSynchronize(b.ev);
  (void)hipEventDestroy(b.ev);
  (void)hipFree(b.p);
  g_free -= b.bytes;
  g_blocks.erase(g_blocks.begin() + (std::ptrdiff_t)i);
}

void trim(size_t keep) {
  while (g_free > keep) {
    size_t oldest = g_blocks.size();
    for (size_t i = 0; i < g_blocks.size(); ++i)
      if (!g_blocks[i].busy && (oldest == g_blocks.size() || g_blocks[i].stamp < g_blocks[oldest].stamp))
        oldest = i;
    if (oldest == g_blocks.size()) return;
    drop(oldest);
  }
}

}  // namespace

hipError_t scratch_alloc(void** out, size_t bytes, hipStream_t s) {
  *out = nullptr;
  if (bytes == 0) bytes = 256;
  if (capturing(s)) return hipMallocAsync(out, bytes, s);
  std::lock_guard<std::mutex> g(g_mu);
  // Best fit among the blocks this stream used last or whose last use is
  // already over: a block still in use on another stream would make this
  // stream wait for that stream's queued work (a palette build behind a walk
  // step, say), so such blocks are left alone and a new one is allocated.
  size_t best = g_blocks.size();
  for (size_t i = 0; i < g_blocks.size(); ++i) {
    const Block& b = g_blocks[i];
    if (b.busy || b.bytes < bytes || b.bytes > 2 * bytes + ((size_t)64 << 20)) continue;
    if (best != g_blocks.size() && b.bytes >= g_blocks[best].bytes) continue;
    if (b.last != s) {
      const hipError_t q = hipEventQuery(b.ev);
      if (q == hipErrorNotReady) (void)hipGetLastError();  // not an error: keep it off the launch checks
      if (q != hipSuccess) continue;
    }
    best = i;
  }
  if (best < g_blocks.size()) {
    Block& b = g_blocks[best];
    b.busy = true;
    b.last = s;
    g_free -= b.bytes;
    *out = b.p;
    return hipSuccess;
  }
  Block b;
  b.bytes = (bytes + kGrain - 1) & ~(kGrain - 1);
  hipError_t e = hipMalloc(&b.p, b.bytes);
  if (e == hipErrorOutOfMemory) {  // give the cached free blocks back and retry
    (void)hipGetLastError();
    trim(0);
    e = hipMalloc(&b.p, b.bytes);
  }
  if (e != hipSuccess) return e;
  e = hipEventCreateWithFlags(&b.ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    (void)hipFree(b.p);
    return e;
  }
  // An entry at this address can only be stale (its block was freed outside
  // the cache): drop it, so a later scratch_free finds the live block.
  for (size_t i = 0; i < g_blocks.size(); ++i) {
    if (g_blocks[i].p != b.p) continue;
    if (!g_blocks[i].busy) g_free -= g_blocks[i].bytes;
    (void)hipEventDestroy(g_blocks[i].ev);
    g_blocks.erase(g_blocks.begin() + (std::ptrdiff_t)i);
    break;
  }
  b.busy = true;
  b.last = s;
  g_blocks.push_back(b);
  *out = b.p;
  return hipSuccess;
}

hipError_t scratch_free(void* p, hipStream_t s) {
  if (!p) return hipSuccess;
  {
    std::lock_guard<std::mutex> g(g_mu);
    for (Block& b : g_blocks) {
      if (b.p != p) continue;
      const hipError_t e = hipEventRecord(b.ev, s);
      b.last = s;
      b.busy = false;
      b.stamp = ++g_clock;
      g_free += b.bytes;
      trim(kKeep);
      return e;
    }
  }
  return hipFreeAsync(p, s);  // stream-ordered (taken under capture)
}

}  // namespace sa
