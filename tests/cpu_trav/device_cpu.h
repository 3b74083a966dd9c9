// CPU stand-in for the device intrinsics rtg_device.h uses, so that its traversal functions
// (trav_begin, trav_step, the stack helpers) can be compiled with g++ and checked on the CPU,
// under AddressSanitizer, against brute force (tests/test_traversal_cpu.py). One lane per "wave".
// Test infrastructure only.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#define __forceinline__ inline
using std::min;
using std::max;
inline float __uint_as_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t __float_as_uint(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float __int_as_float(int32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline uint32_t __lane_id() { return 0; }
inline unsigned long long __ballot(bool p) { return p ? 1ull : 0ull; }
template <class T> inline T __shfl(T v, int) { return v; }
template <class T> inline T __shfl_xor(T v, int) { return v; }
template <class T> inline T __shfl_up(T v, unsigned) { return v; }
template <class T, class U> inline T atomicAdd(T* p, U v) { T o = *p; *p += (T)v; return o; }
struct CpuThreadIdx { uint32_t x = 0, y = 0, z = 0; };
static CpuThreadIdx threadIdx;
inline void __syncthreads() {}
inline bool __any(bool p) { return p; }
