// Host-side helpers shared by the C-ABI entry points: error state, argument
// checks, FP64 camera preparation, launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "ngt_common.hpp"
#include "ngt_transport.h"

#define NGT_HOST_DEVICE_INLINE __host__ __device__ inline

namespace ngt {

// Thread-local message of the last failure (returned by ngt_last_error).
void set_error(const std::string& msg);
void clear_error();

struct Status {
  int code = NGT_OK;
  static Status ok() { return Status(); }
};

// Return an error code from an entry point with a message.
#define NGT_FAIL(code_, msg_)            \
  do {                                   \
    ::ngt::set_error(msg_);              \
    return (code_);                      \
  } while (0)

#define NGT_REQUIRE(cond_, code_, msg_) \
  do {                                  \
    if (!(cond_)) NGT_FAIL(code_, msg_); \
  } while (0)

// Checks hipGetLastError after launches; turns a HIP error into NGT_ERR_HIP.
int check_launch(const char* where);

inline hipStream_t as_stream(ngt_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

struct Shape5 {
  int64_t n, d, h, w, c;
  int64_t cells() const { return d * h * w; }
  int64_t elems() const { return n * d * h * w * c; }
  Dims dims() const { return Dims{static_cast<int>(w), static_cast<int>(h), static_cast<int>(d)}; }
};

inline Shape5 shape5(const int64_t* s) { return Shape5{s[0], s[1], s[2], s[3], s[4]}; }

// Validates an NDHWC input shape like the reference ops
// (resample_grid.cc:371-377): 5-D, C <= 4, C not 3, non-empty.
int check_grid_shape(const int64_t* s, const char* what, bool allow_c3 = false);

// (data entry, camera, output slot) triples of one op call.
struct Pair {
  int n, v, out;
};
int make_pairs(int N, int V, bool separate_camera_batch, std::vector<Pair>& pairs);

// FP64 camera parameters of one pair: MV = V*M, inverses in double, rounded
// to float (transformations.hpp:48-68 does the same in float).
CamParams make_cam(const float* m16, const float* v16, const float* p16, const float* f6, int data_index,
                   int out_index);

// Zero-fill with a kernel (not hipMemsetAsync, so graph capture stays clean).
void fill_zero(float* p, int64_t n, hipStream_t s);
void fill_zero2(float* a, int64_t na, float* b, int64_t nb, hipStream_t s);
void fill_zero_u32(uint32_t* p, int64_t n, hipStream_t s);

inline unsigned blocks_for(int64_t n, int threads) { return static_cast<unsigned>((n + threads - 1) / threads); }

// ------------------------------------------------------------ work lists ---
// Uneven batches run one persistent launch per kernel: a work-group takes
// tiles t = blockIdx.x, blockIdx.x + gridDim.x, ... of a flat tile list whose
// per-instance ranges are begin[i] .. begin[i+1].

// Instance owning tile t (begin is non-decreasing, begin[0] = 0).
NGT_DEVICE int find_item(const int* __restrict__ begin, int n, int t) {
  int lo = 0, hi = n;  // begin[lo] <= t < begin[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (begin[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

// Grid size of a persistent launch: blocks_per_cu work-groups per compute
// unit of the current device (0: one work-group per tile), at most the tiles.
unsigned persistent_grid(int64_t tiles, int blocks_per_cu);

// Byte layout of a plan blob (identical on the host and the device copy).
struct BlobLayout {
  size_t bytes = 0;
  template <typename T>
  size_t add(size_t count) {
    const size_t align = alignof(T) < 16 ? 16 : alignof(T);
    bytes = (bytes + align - 1) / align * align;
    const size_t off = bytes;
    bytes += count * sizeof(T);
    return off;
  }
};

// Read-only plan data seen through the constant address space, so uniform
// reads of it become scalar loads (as kernel arguments are) and the kernel
// bodies see it like a by-value argument.  Valid because a plan is never
// written while a kernel runs.
template <typename T>
NGT_DEVICE const T* as_constant(const void* p) {
  return (const T*)((const __attribute__((address_space(4))) char*)p);
}

// A device pointer stored in a plan blob, typed as global memory.  The
// compiler cannot tell the address space of a generic pointer loaded from
// memory and would access it with flat instructions, whose completion a
// wavefront can only wait for together with its LDS and scalar traffic; a
// pointer loaded with this type and converted to a generic one where it is
// used keeps its global accesses.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T>
NGT_HOST_DEVICE_INLINE T* blob_at(void* base, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(base) + off);
}
template <typename T>
NGT_HOST_DEVICE_INLINE const T* blob_at(const void* base, size_t off) {
  return reinterpret_cast<const T*>(static_cast<const char*>(base) + off);
}

}  // namespace ngt
