// Batched step of many boxes (fpm_batch_*): the plan holds, per pass, the
// launches the translation units built (binning with the step boundary's
// kick-drift / paint / readout in tiles.hip, the transform in fft.hip), each
// with its argument table in the caller's device buffer.
#include <cstring>
#include <new>

#include "batch.h"

namespace fpm {

void* DevArena::put(const void* host, int64_t bytes, int* rc) {
  const int64_t at = (used + 255) / 256 * 256;
  used = at + bytes;
  if (!base) return nullptr;
  if (used > size) {
    *rc = FPM_ERR_WORKSPACE;
    return nullptr;
  }
  if (hipMemcpy(base + at, host, (size_t)bytes, hipMemcpyHostToDevice) != hipSuccess) {
    *rc = FPM_ERR_HIP;
    return nullptr;
  }
  return base + at;
}

}  // namespace fpm

struct fpm_batch {
  std::vector<fpm::BatchLaunch> bin, paint, fft, readout;
};

using namespace fpm;

static int build_all(int nbox, const fpm_batch_box* boxes, DevArena* arena, fpm_batch* b) {
  if (nbox < 1 || !boxes) return FPM_ERR_ARG;
  for (int i = 0; i < nbox; ++i) {
    if (boxes[i].order != 2 && boxes[i].order != 3) return FPM_ERR_ARG;
    if (boxes[i].part.n > 0 && boxes[i].part.s == nullptr) return FPM_ERR_ARG;
  }
  int rc = tile_batch_build(nbox, boxes, arena, b->bin, b->paint, b->readout);
  if (rc) return rc;
  return fft_batch_build(nbox, boxes, arena, b->fft);
}

extern "C" {

int64_t fpm_batch_device_bytes(int nbox, const fpm_batch_box* boxes) {
  DevArena arena{nullptr, 0, 0};
  fpm_batch b;
  const int rc = build_all(nbox, boxes, &arena, &b);
  return rc ? (int64_t)rc : arena.used;
}

fpm_batch* fpm_batch_create(int nbox, const fpm_batch_box* boxes, void* dev, int64_t dev_bytes,
                            int* status) {
  int dummy = 0;
  int* st = status ? status : &dummy;
  if (!dev || dev_bytes <= 0) {
    *st = FPM_ERR_ARG;
    return nullptr;
  }
  fpm_batch* b = new (std::nothrow) fpm_batch();
  if (!b) {
    *st = FPM_ERR_ARG;
    return nullptr;
  }
  DevArena arena{static_cast<char*>(dev), dev_bytes, 0};
  const int rc = build_all(nbox, boxes, &arena, b);
  if (rc) {
    delete b;
    *st = rc;
    return nullptr;
  }
  *st = FPM_OK;
  return b;
}

int fpm_batch_step(const fpm_batch* b, const int32_t* counter, fpm_stream_t stream) {
  if (!b) return FPM_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  int rc = FPM_OK;
  // the count pass applies the step boundary (counter) before binning
  if ((rc = tile_batch_run(b->bin, st, counter))) return rc;
  if ((rc = tile_batch_run(b->paint, st, nullptr))) return rc;
  if ((rc = fft_batch_run(b->fft, st))) return rc;
  return tile_batch_run(b->readout, st, nullptr);
}

void fpm_batch_destroy(fpm_batch* b) { delete b; }

}  // extern "C"
