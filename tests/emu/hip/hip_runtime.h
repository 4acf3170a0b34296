// Host emulation of the HIP subset libquax_hip uses -- TEST INFRASTRUCTURE.
//
// Lets the real runtime (qx_runtime.hip, qx_kernels.hip) and the generated
// per-class kernels compile with g++ for the CPU, so that index arithmetic,
// work decomposition and memory accesses of every kernel can be checked under
// AddressSanitizer and compared with the CPU oracle without a GPU.
//
// Execution model: blocks run one after another; the lanes of a block run on a
// persistent pool of OS threads (one per lane), so __syncthreads and
// wavefront shuffles are real rendezvous.  A shuffle reached by only some
// lanes deadlocks -- which is itself a finding (divergent cross-lane op).
// Device memory is host heap memory (ASan-checked).
#pragma once
#include <atomic>
#include <barrier>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static

typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

struct EmuIdx {
  unsigned x = 0, y = 0, z = 0;
};
extern thread_local EmuIdx threadIdx, blockIdx;
extern thread_local dim3 blockDim, gridDim;

inline const char* hipGetErrorString(hipError_t e) { return e ? "emulated HIP error" : "no error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) {
  *n = 1;
  return hipSuccess;
}
inline hipError_t hipMalloc(void** p, size_t n) {
  *p = std::malloc(n);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
  *e = std::malloc(8);
  return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
  std::free(e);
  return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2 };
// kernels run synchronously at launch, so streams carry no ordering
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = std::malloc(8);
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
  std::free(s);
  return hipSuccess;
}
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* t, hipEvent_t, hipEvent_t) {
  *t = 0.f;
  return hipSuccess;
}

// ---- lane pool -----------------------------------------------------------
namespace emu {
void run_grid(dim3 grid, dim3 block, const std::function<void()>& body);
double shfl_xor(double v, int off);
int shfl_xor(int v, int off);
unsigned long long ballot(bool p);
void syncthreads();
int syncthreads_or(int p);
}  // namespace emu

#define __syncthreads() emu::syncthreads()
inline int __syncthreads_or(int p) { return emu::syncthreads_or(p); }
inline double __shfl_xor(double v, int off, int = 64) { return emu::shfl_xor(v, off); }
inline int __shfl_xor(int v, int off, int = 64) { return emu::shfl_xor(v, off); }
inline unsigned long long __ballot(bool p) { return emu::ballot(p); }
inline unsigned __lane_id() { return threadIdx.x % 64; }
inline int __ffsll(long long m) { return m ? __builtin_ffsll(m) : 0; }

inline double unsafeAtomicAdd(double* p, double v) {
  return std::atomic_ref<double>(*p).fetch_add(v);
}
inline double atomicAdd(double* p, double v) { return std::atomic_ref<double>(*p).fetch_add(v); }
inline float atomicAdd(float* p, float v) { return std::atomic_ref<float>(*p).fetch_add(v); }
inline int atomicMin(int* p, int v) {
  std::atomic_ref<int> a(*p);
  int o = a.load();
  while (v < o && !a.compare_exchange_weak(o, v)) {
  }
  return o;
}
inline int atomicMax(int* p, int v) {
  std::atomic_ref<int> a(*p);
  int o = a.load();
  while (v > o && !a.compare_exchange_weak(o, v)) {
  }
  return o;
}

namespace emu {
// dynamic shared memory of the running grid: exactly the launch's byte
// count, so AddressSanitizer flags any access past it
inline std::vector<double*>& dyn_pool() {
  static std::vector<double*> p(1, nullptr);
  return p;
}
inline void* dyn_shared() { return dyn_pool()[0]; }
}  // namespace emu
#define HIP_DYNAMIC_SHARED(type, var) type* var = reinterpret_cast<type*>(emu::dyn_shared());

template <class K, class... A>
inline void emu_launch(K kernel, dim3 grid, dim3 block, size_t shmem, A... args) {
  double*& d = emu::dyn_pool()[0];
  delete[] d;
  d = shmem ? new double[(shmem + 7) / 8] : nullptr;
  if (shmem && std::getenv("EMU_TRACE_SHMEM")) std::fprintf(stderr, "emu: launch with %zu B dynamic LDS\n", shmem);
  emu::run_grid(grid, block, [=]() { kernel(args...); });
  delete[] d;
  d = nullptr;
}
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  emu_launch(kernel, grid, block, (size_t)(shmem), __VA_ARGS__)

using std::max;
using std::min;
