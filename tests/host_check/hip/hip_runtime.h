// Host stand-ins for the HIP names the device headers use, so that the lane
// kernel body (dg_valu_body.h) compiles as plain C++ for a CPU bounds check
// (tests/test_lane_kernel_host.py), and the batch's work-list layout
// (dg_worklist.h) for tests/test_worklist_host.py.  One "lane": no
// cross-lane operations.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#define __device__
#define __host__
#define __global__
#define __constant__
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#define HIP_SYMBOL(x) x
typedef int hipError_t;
typedef void* hipStream_t;
enum { hipSuccess = 0 };
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
struct Idx { unsigned x = 0, y = 0, z = 0; };
struct int2 { int x, y; };
static Idx threadIdx, blockIdx, blockDim, gridDim;
template <typename T> inline hipError_t hipMemcpyToSymbol(T&, const void*, size_t) { return 0; }
inline long long __double_as_longlong(double d) { long long x; memcpy(&x, &d, 8); return x; }
inline double __longlong_as_double(long long x) { double d; memcpy(&d, &x, 8); return d; }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) { auto o = *p; if (v > o) *p = v; return o; }
inline double __shfl_xor(double v, int, int) { return v; }
inline void __syncthreads() {}
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
inline float __builtin_amdgcn_sqrtf(float x) { return std::sqrt(x); }
inline int __builtin_amdgcn_readfirstlane(int x) { return x; }
