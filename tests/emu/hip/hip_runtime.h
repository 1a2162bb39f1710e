// hip_runtime.h (host emulation) -- lets the library's HIP sources compile with g++ and run on
// the CPU, for tests only.  A kernel launch runs every block in turn; each block runs one host
// thread per work-item.  Wavefront operations (shuffles, readfirstlane) exchange values through
// a per-wavefront slot array between two barriers of the wavefront's live threads, which gives
// SIMT lock-step semantics for operations reached in uniform control flow (the only place the
// library uses them).  Every state load and store goes through JFA_ACCESS, which checks it
// against the ranges the test registered.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ inline __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__ static

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_uint3 {
    unsigned x, y, z;
};
extern thread_local emu_uint3 threadIdx, blockIdx;
extern thread_local dim3 blockDim, gridDim;

typedef int hipError_t;
enum { hipSuccess = 0 };
typedef void* hipStream_t;
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 16 };

hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipGetDevice(int*);
hipError_t hipDeviceGetAttribute(int*, hipDeviceAttribute_t, int);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int*, const void*, int, size_t);
hipError_t hipGetSymbolAddress(void** dev_ptr, const void* symbol);
#define HIP_SYMBOL(x) ((const void*)&(x))
#define __align__(n) alignas(n)

namespace emu {
void launch(dim3 grid, dim3 block, const std::function<void()>& body);
uint64_t exchange(uint64_t v, int mode, int arg);  // mode 0 down, 1 xor, 2 index, 3 first live, 4 up
void syncthreads();
void access(const void* p, size_t bytes);
double rcp(double x);
double rsq(double x);

template <class T>
inline uint64_t to_bits(T v) {
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    return b;
}
template <class T>
inline T from_bits(uint64_t b) {
    T v;
    std::memcpy(&v, &b, sizeof(T));
    return v;
}
}  // namespace emu

#define hipLaunchKernelGGL(K, G, B, S, ST, ...) emu::launch(dim3(G), dim3(B), [=]() { K(__VA_ARGS__); })
#define JFA_ACCESS(p, bytes) emu::access((const void*)(p), (bytes))
#define JFA_OPAQUE(x) ((void)0)
#define JFA_WAVES_PER_EU(n)

template <class T>
inline T __shfl_down(T v, unsigned d, int = 64) {
    return emu::from_bits<T>(emu::exchange(emu::to_bits(v), 0, (int)d));
}
template <class T>
inline T __shfl_up(T v, unsigned d, int = 64) {
    return emu::from_bits<T>(emu::exchange(emu::to_bits(v), 4, (int)d));
}
template <class T>
inline T __shfl_xor(T v, int m, int = 64) {
    return emu::from_bits<T>(emu::exchange(emu::to_bits(v), 1, m));
}
template <class T>
inline T __shfl(T v, int src, int = 64) {
    return emu::from_bits<T>(emu::exchange(emu::to_bits(v), 2, src));
}
inline unsigned __builtin_amdgcn_readfirstlane_emu(unsigned v) { return (unsigned)emu::exchange(v, 3, 0); }
#define __builtin_amdgcn_readfirstlane(v) __builtin_amdgcn_readfirstlane_emu((unsigned)(v))
#define __builtin_amdgcn_rcp(x) emu::rcp(x)
#define __builtin_amdgcn_rsq(x) emu::rsq(x)
inline void __syncthreads() { emu::syncthreads(); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
    unsigned long long old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
using std::max;  // HIP provides integer min/max in the global namespace
using std::min;
inline long long __double_as_longlong(double x) { return emu::from_bits<long long>(emu::to_bits(x)); }
inline double __longlong_as_double(long long x) { return emu::from_bits<double>(emu::to_bits(x)); }
