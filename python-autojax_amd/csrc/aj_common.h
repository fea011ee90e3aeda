// Shared definitions for the gfx950 kernels of libautojax_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <type_traits>

#include "../../include/autojax_amd.h"

namespace aj {

// Thread-local error text returned by aj_last_error().
void set_error(const char* fmt, ...);

// Optional HIP-event timing of selected launches (aj_timing_enable / aj_timing_read).
void timer_begin(int id, hipStream_t st);
void timer_end(int id, hipStream_t st);

#define AJ_CHECK_HIP(expr)                                                              \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess) {                                                             \
      ::aj::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__,  \
                      __LINE__);                                                        \
      return AJ_EHIP;                                                                   \
    }                                                                                   \
  } while (0)

#define AJ_REQUIRE(cond, ...)          \
  do {                                 \
    if (!(cond)) {                     \
      ::aj::set_error(__VA_ARGS__);    \
      return AJ_EINVAL;                \
    }                                  \
  } while (0)

#define AJ_LAUNCH_CHECK()                                                               \
  do {                                                                                  \
    hipError_t _e = hipGetLastError();                                                  \
    if (_e != hipSuccess) {                                                             \
      ::aj::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e),        \
                      __FILE__, __LINE__);                                              \
      return AJ_EHIP;                                                                   \
    }                                                                                   \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Bump allocator over a caller-provided workspace.
struct Arena {
  char* base;
  size_t cap;
  size_t off = 0;
  Arena(void* p, size_t n) : base(static_cast<char*>(p)), cap(n) {}
  template <class T>
  T* take(size_t count) {
    off = align_up(off, 256);
    T* p = reinterpret_cast<T*>(base + off);
    off += count * sizeof(T);
    return p;
  }
  bool ok() const { return off <= cap; }
};

// Wave-level sum of a double over the 64 lanes of a wavefront.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Direct global -> LDS loads (raw buffer addressing): lane L of the wave writes its `size` bytes at
// LDS address lds_ptr + size * L (lds_ptr wave-uniform); offsets at or past the resource's byte count
// read as zeros.
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) uint32_t lds_u32;

__device__ void raw_buffer_load_lds(i32x4 rsrc, __attribute__((address_space(3))) uint32_t* lds_ptr, int size,
                                    int voffset, int soffset, int offset, int aux)
    __asm("llvm.amdgcn.raw.buffer.load.lds");

// One dword or double by a raw buffer load (vector offset + scalar offset, in bytes): offsets at or past the
// resource's byte count read as zero (no branch).
__device__ int raw_buffer_load_i32(i32x4 rsrc, int voffset, int soffset, int aux) __asm("llvm.amdgcn.raw.buffer.load.i32");
__device__ double raw_buffer_load_f64(i32x4 rsrc, int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.raw.buffer.load.f64");

// Buffer resource for byte offsets below 2^32 from base (stride 0, raw addressing).
__device__ inline i32x4 make_rsrc(const void* base, uint32_t bytes) {
  const uint64_t a = reinterpret_cast<uint64_t>(base);
  i32x4 r;
  r[0] = (int)(uint32_t)a;
  r[1] = (int)(uint32_t)(a >> 32) & 0xFFFF;
  r[2] = (int)bytes;
  r[3] = 0x00020000;
  return r;
}

// s_waitcnt immediate that waits for vmcnt <= n and leaves expcnt and lgkmcnt unconstrained
// (gfx9 encoding: vmcnt[3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]).
constexpr unsigned vmcnt_imm(unsigned n) { return (n & 0xF) | ((n >> 4) << 14) | (0x7u << 4) | (0xFu << 8); }

}  // namespace aj
