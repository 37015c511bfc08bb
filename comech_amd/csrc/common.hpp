// Shared HIP utilities: error handling, device buffers, wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <mutex>
#include <string>
#include <unordered_map>
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

// Process-wide cache of device blocks.  Blocks come back here only from a destroyed system,
// after its stream was synchronised (nothing in flight can touch them), and are handed to later
// allocations of the same size class on the same device: repeated solves (one system per solve,
// as the engine does per action) skip hipMalloc/hipFree.  Size classes: 8 per octave (at most
// 12.5 % padding).  At most 16 GiB stay cached per process; a failed hipMalloc empties the cache
// and retries.
struct BlockCache {
  std::mutex mu;
  std::unordered_map<unsigned long long, std::vector<void*>> free;
  size_t cached = 0;
  static constexpr size_t kCap = size_t(16) << 30;
  static size_t size_class(size_t b) {
    if (b <= 4096) return 4096;
    size_t p = 4096;
    while (p < b) p <<= 1;
    const size_t q = p >> 4;      // p/2 < b <= p: steps of p/16 = 8 per octave
    return (b + q - 1) / q * q;
  }
  static unsigned long long key(int dev, size_t cls) { return ((unsigned long long)dev << 48) | cls; }
  void* take(int dev, size_t cls) {
    std::lock_guard<std::mutex> g(mu);
    auto it = free.find(key(dev, cls));
    if (it == free.end() || it->second.empty()) return nullptr;
    void* q = it->second.back();
    it->second.pop_back();
    cached -= cls;
    return q;
  }
  bool give(int dev, size_t cls, void* q) {
    std::lock_guard<std::mutex> g(mu);
    if (cached + cls > kCap) return false;
    free[key(dev, cls)].push_back(q);
    cached += cls;
    return true;
  }
  void flush() {
    std::lock_guard<std::mutex> g(mu);
    for (auto& kv : free) {
      const int dev = (int)(kv.first >> 48);
      int cur = 0;
      (void)hipGetDevice(&cur);
      (void)hipSetDevice(dev);
      for (void* q : kv.second) (void)hipFree(q);
      (void)hipSetDevice(cur);
    }
    free.clear();
    cached = 0;
  }
};
inline BlockCache& block_cache() {
  static BlockCache* c = new BlockCache;   // never destroyed: outlives the HIP runtime teardown
  return *c;
}
// Set while a synchronised system is being destroyed: its buffers go to the cache.
inline bool& release_to_cache() {
  static thread_local bool on = false;
  return on;
}

// Owning device buffer.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  size_t cls = 0;   // allocated bytes (size class)
  int dev = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), cls(o.cls), dev(o.dev) { o.p = nullptr; o.n = 0; o.cls = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; n = o.n; cls = o.cls; dev = o.dev;
      o.p = nullptr; o.n = 0; o.cls = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (!count) return;
    CM_HIP(hipGetDevice(&dev));
    cls = BlockCache::size_class(count * sizeof(T));
    void* q = block_cache().take(dev, cls);
    if (!q) {
      if (hipMalloc(&q, cls) != hipSuccess) {
        (void)hipGetLastError();
        block_cache().flush();
        CM_HIP(hipMalloc(&q, cls));
      }
    }
    p = static_cast<T*>(q);
  }
  void release() {
    if (p) {
      if (!(release_to_cache() && block_cache().give(dev, cls, p))) (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    cls = 0;
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

// Level-0 smoother blocks over split-copy groups: a grid node and its split copies, coupled by
// interface springs, take one dense inverse of their principal submatrix of A in the smoother
// (block Jacobi over the group) instead of one 3x3 inverse per node.  Member nodes carry a zero
// 3x3 inverse in Dinv, so the per-node smoother steps leave them to the group pass.
constexpr int GROUP_MAX = 8;   // members per group (larger groups keep per-node blocks)

template <typename T>
struct CopyGroups {
  int ng = 0;                        // groups (0: none)
  const int* gptr = nullptr;         // [ng + 1] member ranges
  const int* gslot = nullptr;        // members as slots of the touched-node list
  const int* tnode = nullptr;        // touched-node list (slot -> node)
  const uint8_t* gmem = nullptr;     // [ntouch] 1 if the slot is a group member
  const int* goff = nullptr;         // [ng + 1] offsets of the dense (3m)^2 inverses
  const T* ginv = nullptr;           // dense symmetric inverses, row-major per group
};

// A zero diagonal marks a group member's 3x3 slot in Dinv (a node's own inverse is never zero:
// singular blocks become the identity).
template <typename T>
__device__ __forceinline__ bool group_member(const T* m) {
  return m[0] == T(0) && m[3] == T(0) && m[5] == T(0);
}

// z = c Ginv r over the members of group g, gathered from and scattered to node vectors.
template <typename T>
__device__ __forceinline__ void group_apply(const CopyGroups<T>& ga, int g, const T* __restrict__ r,
                                            T* __restrict__ z, T c) {
  const int b = ga.gptr[g], m = ga.gptr[g + 1] - b, N = 3 * m;
  const T* G = ga.ginv + ga.goff[g];
  T rg[3 * GROUP_MAX];
  for (int j = 0; j < m; ++j) {
    const long long n = ga.tnode[ga.gslot[b + j]];
    for (int q = 0; q < 3; ++q) rg[3 * j + q] = r[3 * n + q];
  }
  for (int j = 0; j < m; ++j) {
    const long long n = ga.tnode[ga.gslot[b + j]];
    for (int q = 0; q < 3; ++q) {
      T acc = T(0);
      const T* row = G + (size_t)(3 * j + q) * N;
      for (int k = 0; k < N; ++k) acc += row[k] * rg[k];
      z[3 * n + q] = c * acc;
    }
  }
}

}  // namespace cm
