// Device characterisation, no reference counterpart: register-only MFMA loops that measure the matrix
// units' achievable rate on this device (the ceiling that the SYRK kernels' rooflines are read against
// next to the published peaks).  Each wave keeps many independent accumulators of the MFMA instruction
// the SYRK of that arithmetic uses (32 for int8, as the residue SYRK's 8 x 4 blocks per wave; 16 for
// FP64) in workgroups of eight waves, one per CU as the SYRK runs, with lane-dependent small integer
// operands (built with a run-time zero, so nothing folds).  Every product and sum is exact, so a launch
// of n iterations returns exactly n times the one-iteration result: a test checks that.

#include "aj_common.h"

namespace aj {
namespace probe {

typedef int i4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int PROBE_THREADS = 512;

// operand bytes in [-8, 7] (a lane- and register-dependent pattern): |acc word| <= iters * 64 * 64
__device__ inline int probe_bytes(int lane, int r, int zero) {
  int w = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) w |= (int)(((unsigned)(((lane * 5 + r * 3 + k * 7) & 15) - 8 + zero) & 0xFFu) << (8 * k));
  return w;
}

// v_mfma_i32_16x16x64_i8 (the residue SYRK's instruction) on 8 x 4 accumulators per wave; per lane the
// sum of its 32 accumulators' four words.  zero is 0 at run time.
__global__ void __launch_bounds__(PROBE_THREADS, 1) mfma_i8_kernel(int iters, int zero, int64_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  i4 fa[8], fb[4], acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; ++a)
    fa[a] = i4{probe_bytes(lane, a, zero), probe_bytes(lane, a + 1, zero), probe_bytes(lane, a + 2, zero),
               probe_bytes(lane, a + 3, zero)};
#pragma unroll
  for (int b = 0; b < 4; ++b)
    fb[b] = i4{probe_bytes(lane, b + 11, zero), probe_bytes(lane, b + 12, zero), probe_bytes(lane, b + 13, zero),
               probe_bytes(lane, b + 14, zero)};
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = i4{0, 0, 0, 0};
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[a], fb[b], acc[a][b], 0, 0, 0);
  }
  int64_t s = 0;
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) s += (int64_t)acc[a][b][0] + acc[a][b][1] + acc[a][b][2] + acc[a][b][3];
  out[(int64_t)blockIdx.x * PROBE_THREADS + threadIdx.x] = s;
}

// v_mfma_f64_16x16x4f64 (the FP64 SYRK's and the Cholesky's instruction) on 4 x 4 accumulators per
// wave, operands small integers (|v| <= 8): every sum stays an exact integer far below 2^53.
__global__ void __launch_bounds__(PROBE_THREADS, 1) mfma_f64_kernel(int iters, double zero, double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  double fa[4], fb[4];
  d4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a) fa[a] = (double)(((lane * 5 + a * 3) & 15) - 8) + zero * (a + 1);
#pragma unroll
  for (int b = 0; b < 4; ++b) fb[b] = (double)(((lane * 7 + b * 5) & 15) - 8) + zero * (b + 5);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[b], acc[a][b], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) s += (acc[a][b][0] + acc[a][b][1]) + (acc[a][b][2] + acc[a][b][3]);
  out[(int64_t)blockIdx.x * PROBE_THREADS + threadIdx.x] = s;
}

}  // namespace probe
}  // namespace aj

using namespace aj;

// kind 0: int8 MFMA (16x16x64, 32 accumulators per wave), 1: FP64 MFMA (16x16x4, 16 accumulators per
// wave).  workgroups x 512 threads, iters loop iterations; out holds one value per thread (int64 for
// kind 0, double for kind 1).  MFMAs per launch: workgroups * 8 waves * iters * (32 or 16).
extern "C" int aj_mfma_peak_probe(int kind, int64_t workgroups, int64_t iters, void* out, void* stream) {
  AJ_REQUIRE(kind == 0 || kind == 1, "probe kind %d (0 int8, 1 fp64)", kind);
  AJ_REQUIRE(workgroups > 0 && workgroups <= (1 << 20) && iters > 0 && iters <= (1 << 18) && out,
             "bad probe size workgroups=%lld iters=%lld", (long long)workgroups, (long long)iters);
  if (kind == 0) {
    // each accumulator word stays within iters * 64 * 64 <= 2^30 (iters <= 2^18)
    hipLaunchKernelGGL(probe::mfma_i8_kernel, dim3((unsigned)workgroups), dim3(probe::PROBE_THREADS), 0,
                       as_stream(stream), (int)iters, 0, static_cast<int64_t*>(out));
  } else {
    hipLaunchKernelGGL(probe::mfma_f64_kernel, dim3((unsigned)workgroups), dim3(probe::PROBE_THREADS), 0,
                       as_stream(stream), (int)iters, 0.0, static_cast<double*>(out));
  }
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}
