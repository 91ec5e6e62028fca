// Host helpers: error state, shape checks, FP64 camera preparation, fills.
#include "ngt_host.hpp"

#include <cmath>
#include <cstring>

namespace ngt {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int check_launch(const char* where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(where) + ": HIP error '" + hipGetErrorString(e) + "'");
    return NGT_ERR_HIP;
  }
  return NGT_OK;
}

int check_grid_shape(const int64_t* s, const char* what, bool allow_c3) {
  if (s == nullptr) {
    set_error(std::string("Invalid ") + what + " shape (NDHWC): null");
    return NGT_ERR_INVALID_ARGUMENT;
  }
  for (int i = 0; i < 5; ++i) {
    if (s[i] <= 0) {
      set_error(std::string("Empty ") + what + " (NDHWC): [" + std::to_string(s[0]) + "," + std::to_string(s[1]) +
                "," + std::to_string(s[2]) + "," + std::to_string(s[3]) + "," + std::to_string(s[4]) + "]");
      return NGT_ERR_INVALID_ARGUMENT;
    }
  }
  if (s[4] > 4 || (s[4] == 3 && !allow_c3)) {
    if (s[4] > 4) {
      set_error(std::string("Invalid ") + what + " shape (NDHWC): channel " + std::to_string(s[4]));
      return NGT_ERR_INVALID_ARGUMENT;
    }
    set_error("Only 1,2 and 4 Channel supported.");
    return NGT_ERR_UNIMPLEMENTED;
  }
  if (s[1] > (1 << 24) || s[2] > (1 << 24) || s[3] > (1 << 24) || s[1] * s[2] * s[3] * s[4] >= (int64_t(1) << 32)) {
    set_error(std::string(what) + " spatial size too large (a grid slice must hold fewer than 2^32 values)");
    return NGT_ERR_INVALID_ARGUMENT;
  }
  return NGT_OK;
}

int make_pairs(int N, int V, bool separate_camera_batch, std::vector<Pair>& pairs) {
  pairs.clear();
  if (separate_camera_batch) {
    for (int n = 0; n < N; ++n)
      for (int v = 0; v < V; ++v) pairs.push_back(Pair{n, v, n * V + v});
  } else {
    if (V != N) {
      set_error("camera batch must match data batch when not using global sampling.");
      return NGT_ERR_INVALID_ARGUMENT;
    }
    for (int n = 0; n < N; ++n) pairs.push_back(Pair{n, n, n});
  }
  return NGT_OK;
}

namespace {

// column-major flat (the op argument) -> row-major double math matrix
void load_math(const float* f, double m[16]) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) m[r * 4 + c] = static_cast<double>(f[c * 4 + r]);
}

void matmul(const double a[16], const double b[16], double o[16]) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a[r * 4 + k] * b[k * 4 + c];
      o[r * 4 + c] = s;
    }
}

// Gauss-Jordan with partial pivoting in double.
bool invert(const double a[16], double o[16]) {
  double m[4][8];
  for (int r = 0; r < 4; ++r) {
    for (int c = 0; c < 4; ++c) {
      m[r][c] = a[r * 4 + c];
      m[r][c + 4] = (r == c) ? 1.0 : 0.0;
    }
  }
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    for (int r = col + 1; r < 4; ++r)
      if (std::fabs(m[r][col]) > std::fabs(m[piv][col])) piv = r;
    if (std::fabs(m[piv][col]) < 1e-300) return false;
    if (piv != col)
      for (int c = 0; c < 8; ++c) std::swap(m[piv][c], m[col][c]);
    const double inv = 1.0 / m[col][col];
    for (int c = 0; c < 8; ++c) m[col][c] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == col) continue;
      const double f = m[r][col];
      if (f == 0.0) continue;
      for (int c = 0; c < 8; ++c) m[r][c] -= f * m[col][c];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) o[r * 4 + c] = m[r][c + 4];
  return true;
}

void to_float(const double a[16], float o[16]) {
  for (int i = 0; i < 16; ++i) o[i] = static_cast<float>(a[i]);
}

}  // namespace

CamParams make_cam(const float* m16, const float* v16, const float* p16, const float* f6, int data_index,
                   int out_index) {
  CamParams cp;
  std::memset(&cp, 0, sizeof(cp));
  double M[16], V[16], P[16], MV[16], MVi[16], Pi[16];
  load_math(m16, M);
  load_math(v16, V);
  load_math(p16, P);
  matmul(V, M, MV);
  if (!invert(MV, MVi)) std::memset(MVi, 0, sizeof(MVi));
  if (!invert(P, Pi)) std::memset(Pi, 0, sizeof(Pi));
  to_float(MV, cp.mv);
  to_float(MVi, cp.mv_inv);
  to_float(P, cp.p);
  to_float(Pi, cp.p_inv);
  cp.near_ = f6[0];
  cp.far_ = f6[1];
  cp.data_index = data_index;
  cp.out_index = out_index;
  return cp;
}

// 16-byte stores where the range starts 16-byte aligned (workspace regions
// and torch allocations do), else 4-byte stores.
__device__ __forceinline__ void fill_range(float* __restrict__ p, int64_t n, float value, int64_t tid,
                                           int64_t stride) {
  if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
    const int64_t n4 = n >> 2;
    float4* __restrict__ q = reinterpret_cast<float4*>(p);
    const float4 v = make_float4(value, value, value, value);
    for (int64_t i = tid; i < n4; i += stride) q[i] = v;
    for (int64_t i = (n4 << 2) + tid; i < n; i += stride) p[i] = value;
  } else {
    for (int64_t i = tid; i < n; i += stride) p[i] = value;
  }
}

inline unsigned fill_blocks(int64_t n) {
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + 1023) / 1024, 4096)));
}

__global__ void fill_kernel(float* __restrict__ p, int64_t n, float value) {
  fill_range(p, n, value, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x,
             static_cast<int64_t>(gridDim.x) * blockDim.x);
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t value) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = value;
}

void fill_zero(float* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(fill_blocks(n)), dim3(256), 0, s, p, n, 0.f);
}

// Two ranges in one launch: grid.y selects the range.
__global__ void fill2_kernel(float* __restrict__ a, int64_t na, float* __restrict__ b, int64_t nb) {
  fill_range(blockIdx.y == 0 ? a : b, blockIdx.y == 0 ? na : nb, 0.f,
             static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x, static_cast<int64_t>(gridDim.x) * blockDim.x);
}

void fill_zero2(float* a, int64_t na, float* b, int64_t nb, hipStream_t s) {
  const int64_t n = std::max(na, nb);
  if (n <= 0) return;
  hipLaunchKernelGGL(fill2_kernel, dim3(fill_blocks(n), 2), dim3(256), 0, s, a, na, b, nb);
}

void fill_zero_u32(uint32_t* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const unsigned b = static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(fill_u32_kernel, dim3(b), dim3(256), 0, s, p, n, 0u);
}

unsigned persistent_grid(int64_t tiles, int blocks_per_cu) {
  if (tiles <= 0) return 1;
  if (blocks_per_cu <= 0) return static_cast<unsigned>(tiles);
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  return static_cast<unsigned>(std::min<int64_t>(tiles, static_cast<int64_t>(cus) * blocks_per_cu));
}

}  // namespace ngt

extern "C" {

const char* ngt_last_error(void) { return ngt::g_last_error.c_str(); }

int ngt_version(void) { return 7; }

}  // extern "C"
