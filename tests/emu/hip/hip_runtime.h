/*
 * CPU emulation of the HIP subset the library uses -- test infrastructure.
 *
 * The library's .hip sources are rewritten (launch syntax only, by
 * translate.py) and compiled with g++ and AddressSanitizer against this
 * header, so that every kernel runs on the host with each device buffer a
 * separate heap allocation: an out-of-bounds index in any kernel is
 * reported with its stack instead of faulting a GPU.  Blocks run one after
 * another; the threads of a block run concurrently on a pool of host
 * threads, with __syncthreads a real barrier and wavefront shuffles
 * exchanged through a per-block buffer.  Kernel pointer arguments are
 * checked against the registry of device allocations, so passing a host
 * pointer to a kernel is reported too.
 *
 * Nothing here is part of the product; the product is the HIP build.
 */
#ifndef BB_EMU_HIP_RUNTIME_H
#define BB_EMU_HIP_RUNTIME_H

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <functional>
#include <type_traits>
#include <typeinfo>
#include <utility>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) x

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1)
    : x(a), y(b), z(c) {}
};

struct double2 { double x, y; };
inline double2 make_double2(double x, double y) { return double2{x, y}; }

extern thread_local dim3 threadIdx;
extern thread_local dim3 blockIdx;
extern dim3 blockDim;
extern dim3 gridDim;

typedef int hipError_t;
enum {
  hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2,
  hipErrorNoDevice = 100, hipErrorLaunchFailure = 719
};
typedef struct bb_emu_stream *hipStream_t;
typedef struct bb_emu_event *hipEvent_t;
enum hipMemcpyKind {
  hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
};
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0

struct hipDeviceProp_t {
  char name[256];
  char gcnArchName[256];
  int multiProcessorCount;
};

hipError_t hipGetDeviceCount(int *n);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
/* the address of a file-scope __device__ array: registered as a device
 * allocation of `bytes` on its first lookup (the emulated device's copy is
 * the host array itself) */
hipError_t bb_emu_symbol_address(void **p, const void *sym, size_t bytes);
template <typename T>
inline hipError_t hipGetSymbolAddress(void **p, const T &sym)
{
  return bb_emu_symbol_address(p, (const void *) &sym, sizeof(T));
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize(void);
hipError_t hipMalloc(void **p, size_t n);
hipError_t hipFree(void *p);
template<typename T>
inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **) p, n); }
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k,
  hipStream_t st = 0);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st = 0);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = 0);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipGetLastError(void);
const char *hipGetErrorString(hipError_t e);

/* ---- device-side intrinsics ---- */
namespace bb_emu {
void syncthreads(void);
double shfl_down_d(double v, unsigned delta, int width);
float shfl_down_f(float v, unsigned delta, int width);
int shfl_down_i(int v, unsigned delta, int width);
double shfl_idx_d(double v, int src, int width);
int shfl_idx_i(int v, int src, int width);
/* bit l set when lane l of the caller's 64-lane wavefront passes p != 0 */
unsigned long long ballot(int p);
/* true when [p, p+bytes) lies inside one device (or pinned host)
 * allocation; null passes */
bool is_device_range(const void *p, size_t bytes);
void report_bad_pointer(const char *kernel, int arg, const void *p);
void run_grid(const char *name, dim3 g, dim3 b, const std::function<void()> &body);
/* Launch capture for the machine-code replay (tests/isa): with
 * BB_EMU_ISA_DUMP=<dir> set, the first launches of each kernel are written
 * out as (kernel, launch shape, explicit argument bytes in kernarg layout,
 * every device allocation before and after the launch). */
bool isa_capture_wanted(const char *name);
void isa_capture(const char *name, const char *ftype, dim3 g, dim3 b,
  const std::vector<unsigned char> &args, int phase);
}

#define __syncthreads() bb_emu::syncthreads()

inline double __shfl_down(double v, unsigned d, int w = 64) { return bb_emu::shfl_down_d(v, d, w); }
inline float __shfl_down(float v, unsigned d, int w = 64) { return bb_emu::shfl_down_f(v, d, w); }
inline int __shfl_down(int v, unsigned d, int w = 64) { return bb_emu::shfl_down_i(v, d, w); }
inline double __shfl(double v, int src, int w = 64) { return bb_emu::shfl_idx_d(v, src, w); }
inline int __shfl(int v, int src, int w = 64) { return bb_emu::shfl_idx_i(v, src, w); }
inline unsigned long long __ballot(int p) { return bb_emu::ballot(p); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long) x); }

inline int atomicAdd(int *a, int v) { return __atomic_fetch_add(a, v, __ATOMIC_SEQ_CST); }
inline int atomicExch(int *a, int v) { return __atomic_exchange_n(a, v, __ATOMIC_SEQ_CST); }
inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
inline unsigned atomicAdd(unsigned *a, unsigned v) { return __atomic_fetch_add(a, v, __ATOMIC_SEQ_CST); }
inline double atomicAdd(double *a, double v)
{
  double old, want;
  __atomic_load(a, &old, __ATOMIC_SEQ_CST);
  do {
    want = old + v;
  } while(!__atomic_compare_exchange(a, &old, &want, false, __ATOMIC_SEQ_CST,
    __ATOMIC_SEQ_CST));
  return old;
}
inline int atomicMax(int *a, int v)
{
  int old = __atomic_load_n(a, __ATOMIC_SEQ_CST);
  while(old < v && !__atomic_compare_exchange_n(a, &old, v, false,
    __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
  return old;
}
inline int atomicMin(int *a, int v)
{
  int old = __atomic_load_n(a, __ATOMIC_SEQ_CST);
  while(old > v && !__atomic_compare_exchange_n(a, &old, v, false,
    __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
  return old;
}

/* HIP's device min/max overloads */
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline long min(long a, long b) { return a < b ? a : b; }
inline long max(long a, long b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline double min(double a, double b) { return fmin(a, b); }
inline double max(double a, double b) { return fmax(a, b); }

/* ---- kernel launch: kern<<<g, b, sh, st>>>(args) is rewritten by
 * translate.py into bb_emu::launch(#kern, kern, g, b, sh, st)(args) ---- */
namespace bb_emu {
template<typename A>
inline void check_arg(const char *name, int idx, const A &a)
{
  if constexpr(std::is_pointer<A>::value) {
    if(a != nullptr && !is_device_range((const void *) a, 1))
      report_bad_pointer(name, idx, (const void *) a);
  }
}

/* one explicit argument converted to its parameter type P and laid out
 * at the next multiple of P's alignment, as in the kernarg segment */
template<typename P, typename A>
inline void put_param(std::vector<unsigned char> &kb, const A &a)
{
  P v = static_cast<P>(a);
  size_t off = (kb.size() + alignof(P) - 1) / alignof(P) * alignof(P);
  kb.resize(off + sizeof(P));
  memcpy(kb.data() + off, &v, sizeof(P));
}

template<typename... P, typename... Args>
inline void pack_params(std::vector<unsigned char> &kb, void (*)(P...),
  const Args &...args)
{
  static_assert(sizeof...(P) == sizeof...(Args), "kernel argument count");
  (put_param<P>(kb, args), ...);
}

template<typename F>
struct Launch {
  const char *name;
  F f;
  dim3 g, b;
  template<typename... Args>
  void operator()(Args... args) const
  {
    int idx = 0;
    (check_arg(name, idx++, args), ...);
    F fn = f;
    const bool cap = isa_capture_wanted(name);
    std::vector<unsigned char> kb;
    if(cap) {
      pack_params(kb, fn, args...);
      isa_capture(name, typeid(F).name(), g, b, kb, 0);
    }
    run_grid(name, g, b, [&]() { fn(args...); });
    if(cap) isa_capture(name, typeid(F).name(), g, b, kb, 1);
  }
};

template<typename F>
inline Launch<F> launch(const char *name, F f, dim3 g, dim3 b, size_t = 0,
  hipStream_t = 0)
{
  return Launch<F>{name, f, g, b};
}
}

#endif
