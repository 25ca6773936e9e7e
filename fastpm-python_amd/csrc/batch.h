// Internal: the parts of a batched step (fpm_batch_*) that live in each
// translation unit, and the device-side helpers they share.
//
// A batched launch walks one flat item space over at most kBatchMax boxes:
// item i belongs to the box b with pre[b] <= i < pre[b + 1] and is that box's
// local block i - pre[b] of cnt[b].  Each box's range is padded to a multiple
// of 8 items (the padding items do nothing), so that a box's local block j
// runs on XCD j mod 8 as in the box's own launch, and its bodies' XCD-aware
// tile order holds.  One workgroup per item: a loop over items would let the
// compiler hoist the bodies' per-thread index arithmetic out of it and hold
// it in vector registers for the whole kernel (measured: 1.6-3x the VGPRs of
// the single-box kernels, half their occupancy).
#pragma once
#include <vector>

#include "fpm_common.h"

namespace fpm {

constexpr int kBatchMax = 63;  // boxes per launch: pre[] fits one wave

// The box of `item`: every lane of the wave tests one prefix entry and the
// ballot's population count is the box (one load instead of a dependent
// search).  Called by full waves in uniform control flow; the result is made
// wave-uniform explicitly, so the box's arguments load into scalar registers
// as a kernel's own arguments do.
__device__ __forceinline__ int batch_box(const int64_t* __restrict__ pre, int nb, int64_t item) {
  const int lane = threadIdx.x & 63;
  const int64_t v = lane <= nb ? pre[lane] : INT64_MAX;
  return __builtin_amdgcn_readfirstlane(__popcll(__ballot(v <= item)) - 1);
}

// Bump allocation of the plan's tables in the caller's device buffer; with a
// null base it only counts (the sizing pass of fpm_batch_device_bytes).
struct DevArena {
  char* base;
  int64_t size;
  int64_t used;
  // reserves `bytes` (256-aligned) and copies `host` there when the arena
  // has memory; returns the device address or null
  void* put(const void* host, int64_t bytes, int* rc);
};

// One launch of a batched kernel: the compiled kernel it selects (kind and
// template choices), its argument table and item prefix, and the item count.
struct BatchLaunch {
  int kind;
  int a, b, c, d;          // template choices, meaning per kind
  const void* tab;
  const int64_t* pre;      // padded item prefix, nb + 1 entries
  const int64_t* cnt;      // items per box
  int nb;
  int64_t total;
};

// Groups boxes by a key and uploads, per group of at most kBatchMax boxes,
// the boxes' argument structs, padded item prefix and item counts; appends
// one BatchLaunch per group.
template <typename A, typename KeyFn, typename ItemsFn>
int batch_groups(const std::vector<A>& args, int kind, KeyFn key, ItemsFn items, DevArena* arena,
                 std::vector<BatchLaunch>& out) {
  std::vector<int> done(args.size(), 0);
  for (size_t i = 0; i < args.size(); ++i) {
    if (done[i]) continue;
    const auto k = key(args[i]);
    std::vector<A> tab;
    std::vector<int64_t> pre(1, 0), cnt;
    for (size_t j = i; j < args.size() && (int)tab.size() < kBatchMax; ++j) {
      if (done[j] || !(key(args[j]) == k)) continue;
      done[j] = 1;
      const int64_t n = items(args[j]);
      if (n <= 0) continue;
      tab.push_back(args[j]);
      cnt.push_back(n);
      pre.push_back(pre.back() + (n + 7) / 8 * 8);
    }
    if (tab.empty()) continue;
    BatchLaunch L{};
    L.kind = kind;
    L.a = k.a;
    L.b = k.b;
    L.c = k.c;
    L.d = k.d;
    L.nb = (int)tab.size();
    L.total = pre.back();
    if (L.total >= ((int64_t)1 << 31)) return FPM_ERR_ARG;
    int rc = 0;
    L.tab = arena->put(tab.data(), (int64_t)(tab.size() * sizeof(A)), &rc);
    if (rc) return rc;
    L.pre = static_cast<const int64_t*>(
        arena->put(pre.data(), (int64_t)(pre.size() * sizeof(int64_t)), &rc));
    if (rc) return rc;
    L.cnt = static_cast<const int64_t*>(
        arena->put(cnt.data(), (int64_t)(cnt.size() * sizeof(int64_t)), &rc));
    if (rc) return rc;
    out.push_back(L);
  }
  return FPM_OK;
}

struct BatchKey {
  int a, b, c, d;
  bool operator==(const BatchKey& o) const { return a == o.a && b == o.b && c == o.c && d == o.d; }
};

// Workgroups of a batched launch: one per item (batch_groups keeps the
// total below 2^31).
inline unsigned batch_grid(int64_t total) { return (unsigned)total; }

// The per-unit parts.  Each builds its launches from the boxes into `out`
// (tables placed in `arena`) and enqueues them.
int tile_batch_build(int nbox, const fpm_batch_box* boxes, DevArena* arena,
                     std::vector<BatchLaunch>& bin, std::vector<BatchLaunch>& paint,
                     std::vector<BatchLaunch>& readout);
// counter: the step row of the count pass's kick-drift (NULL: none)
int tile_batch_run(const std::vector<BatchLaunch>& ls, hipStream_t st, const int32_t* counter);

int fft_batch_build(int nbox, const fpm_batch_box* boxes, DevArena* arena,
                    std::vector<BatchLaunch>& out);
int fft_batch_run(const std::vector<BatchLaunch>& ls, hipStream_t st);

}  // namespace fpm
