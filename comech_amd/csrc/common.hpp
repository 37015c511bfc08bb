// Shared HIP utilities: error handling, device buffers, wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace cm {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define CM_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::cm::Error(std::string(#call) + " failed: " + hipGetErrorString(e_) + " at " + \
                        __FILE__ + ":" + std::to_string(__LINE__));                        \
  } while (0)

#define CM_CHECK(cond, msg)                                     \
  do {                                                          \
    if (!(cond)) throw ::cm::Error(std::string("comech: ") + (msg)); \
  } while (0)

#define CM_LAUNCH_CHECK() CM_HIP(hipGetLastError())

// Owning device buffer (hipMalloc / hipFree).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) CM_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T* host, size_t count, hipStream_t s) {
    alloc(count);
    if (count) {
      CM_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
      CM_HIP(hipStreamSynchronize(s));
    }
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  size_t bytes() const { return n * sizeof(T); }
};

// ---- reductions: wavefront (64 lanes) shuffle, then LDS across the waves of a block
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block of BS threads; returns the total in thread 0.
template <int BS>
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[BS / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) t += red[w];
  }
  __syncthreads();
  return t;
}

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace cm
