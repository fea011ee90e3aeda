// Lattice (compact) form of W-tilde for a regular pixel grid.
//
// Reference: src/autojax/numba.py
//   w_compact_curvature_interferometer_from :210-234
//     w_compact[m, n] = sum_k n_k^-2 cos(2 pi d [(n - (N-1)) u_k0 - m u_k1]),  m in [0, N), n in [0, 2N-1)
//     with d = pixel_scale in radians (the reference's TWOPI_D = 2 pi d)
//   w_tilde_via_compact_from                :237-255
//     W[i, j] = w_compact[|r_i - r_j|, (c_i - c_j if r_i >= r_j else c_j - c_i) + N - 1]
//
// On the device the sum is a real FP64 GEMM over 2K terms: with per-visibility phasors
//   Q_m = exp(-2 pi i d u_k1 m),  P_n = exp(2 pi i d u_k0 (n - N + 1)),
//   w_compact[m, n] = sum_k w_k Re(P_n Q_m) = sum_k w_k (cos q cos p + sin q sin p),
// so A[m, (k, 0|1)] = w_k (cos q, sin q) and B[n, (k, 0|1)] = (cos p, sin p) and
// w_compact = A B^T.  Each workgroup owns a 32 x 64 output tile and a slice of
// the visibilities, builds the tile's phasors for 32 visibilities at a time in
// LDS (FP64 sincospi of the phase reduced to [-1/2, 1/2] revolution), runs
// v_mfma_f64_16x16x4 on them, and writes its partial tile; a second kernel sums
// the slices in a fixed order (deterministic).  The work is 2 N (2N-1) 2K flops
// plus (N + 2N - 1) K sincos per covering tile row/column, against N (2N-1) K
// cosines in the reference's loop.

#include "aj_common.h"

namespace aj {
namespace cp {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int CT_M = 32;   // tile rows (m)
constexpr int CT_N = 64;   // tile columns (n)
constexpr int CT_KV = 32;  // visibilities per LDS chunk
constexpr int CT_TARGET_BLOCKS = 2048;

__global__ void __launch_bounds__(256) compact_gemm_kernel(const double* __restrict__ noise,
                                                           const double* __restrict__ uv, int64_t K, int64_t N,
                                                           double d_rad, int64_t k_per_split,
                                                           double* __restrict__ partial) {
  __shared__ double Qs[CT_M][2 * CT_KV + 1];
  __shared__ double Ps[CT_N][2 * CT_KV + 1];
  const int64_t ND = 2 * N - 1, O = N - 1;
  const int64_t n0 = (int64_t)blockIdx.x * CT_N, m0 = (int64_t)blockIdx.y * CT_M;
  const int64_t z = blockIdx.z;
  const int64_t kb = z * k_per_split, ke = std::min<int64_t>(K, kb + k_per_split);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int br = wave & 1, bc0 = (wave >> 1) * 2;  // wave: block row br, block columns bc0, bc0 + 1
  d4 acc[2] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};

  for (int64_t k0 = kb; k0 < ke; k0 += CT_KV) {
    // phasors of the tile's rows and columns for visibilities k0 .. k0 + 31 (zero outside)
    for (int e = tid; e < (CT_M + CT_N) * CT_KV; e += 256) {
      const int j = e % CT_KV, r = e / CT_KV;
      const int64_t k = k0 + j;
      double c = 0.0, s = 0.0;
      if (r < CT_M) {
        const int64_t m = m0 + r;
        if (k < ke && m < N) {
          const double w = 1.0 / (noise[k] * noise[k]);
          double rv = d_rad * (double)m * uv[2 * k + 1];
          rv -= rint(rv);
          sincospi(2.0 * rv, &s, &c);
          c *= w;
          s *= w;
        }
        Qs[r][2 * j] = c;
        Qs[r][2 * j + 1] = s;
      } else {
        const int64_t n = n0 + (r - CT_M);
        if (k < ke && n < ND) {
          double rv = d_rad * (double)(n - O) * uv[2 * k];
          rv -= rint(rv);
          sincospi(2.0 * rv, &s, &c);
        }
        Ps[r - CT_M][2 * j] = c;
        Ps[r - CT_M][2 * j + 1] = s;
      }
    }
    __syncthreads();
    // v_mfma_f64_16x16x4: A lane l holds A[l%16][l/16], B lane l holds B[l/16][l%16]
#pragma unroll
    for (int ks = 0; ks < 2 * CT_KV / 4; ++ks) {
      const int kk = ks * 4 + lk;
      const double fa = Qs[br * 16 + lr][kk];
      const double fb0 = Ps[bc0 * 16 + lr][kk];
      const double fb1 = Ps[(bc0 + 1) * 16 + lr][kk];
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa, fb0, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa, fb1, acc[1], 0, 0, 0);
    }
    __syncthreads();
  }
  // accumulator register r of lane l holds row 4r + l/16, column l%16
  double* out = partial + z * N * ND;
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t m = m0 + br * 16 + 4 * r + lk, n = n0 + (bc0 + b) * 16 + lr;
      if (m < N && n < ND) out[m * ND + n] = acc[b][r];
    }
}

// w_compact[e] = sum_z partial[z][e], z in order.
__global__ void compact_reduce_kernel(const double* __restrict__ partial, int64_t nsplit, int64_t total,
                                      double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  double s = 0.0;
  for (int64_t z = 0; z < nsplit; ++z) s += partial[z * total + e];
  out[e] = s;
}

// W[i, j] from the lattice table (numba.py:246-254).
__global__ void compact_expand_kernel(const double* __restrict__ C, int64_t N, const int64_t* __restrict__ native,
                                      int64_t M, double* __restrict__ W) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const int64_t ND = 2 * N - 1, O = N - 1;
  const int64_t rj = native[2 * j], cj = native[2 * j + 1];
  for (int64_t i = blockIdx.y; i < M; i += gridDim.y) {
    const int64_t ri = native[2 * i], ci = native[2 * i + 1];
    const int64_t m = ri - rj;
    const int64_t n = (m >= 0 ? ci - cj : cj - ci) + O;
    W[i * M + j] = C[(m >= 0 ? m : -m) * ND + n];
  }
}

// W[i, j] = preload[y_j - y_i, x_j - x_i] for every pair (i, j), negative offsets wrapping as numpy
// indexes (numba.py:282-291: each entry is read from the table, the lower triangle included, so W is
// symmetric only when the table is point-symmetric).
__global__ void preload_expand_kernel(const double* __restrict__ P, int64_t Py, int64_t Px,
                                      const int64_t* __restrict__ native, int64_t M, double* __restrict__ W) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const int64_t yj = native[2 * j], xj = native[2 * j + 1];
  for (int64_t i = blockIdx.y; i < M; i += gridDim.y) {
    int64_t dy = yj - native[2 * i], dx = xj - native[2 * i + 1];
    dy += dy < 0 ? Py : 0;
    dx += dx < 0 ? Px : 0;
    W[i * M + j] = P[dy * Px + dx];
  }
}

// ---------------------------------------------------------------------------
// Preload table of original.py:230-403 (w_tilde_curvature_preload_interferometer_from) on the Y x X
// corner of the grid g (f8[Gy, Gx, 2], (y, x) in radians): a 2Y x 2X table whose four quadrants hold
//   [i, j]          offset g[0, 0]         - g[i, j]                    (i < Y, j < X)
//   [i, 2X - j]     offset g[0, X-1]       - g[i, X-1-j]                (j >= 1)
//   [2Y - i, j]     offset g[Y-1, 0]       - g[Y-1-i, j]                (i >= 1)
//   [2Y - i, 2X - j] offset g[Y-1, X-1]    - g[Y-1-i, X-1-j]            (i, j >= 1)
// each sum_k n_k^-2 cos(2 pi (dx u_k0 + dy u_k1)); row Y and column X stay zero.  Thread = one table
// entry; the workgroup stages PRE_KV visibilities at a time in LDS; blockIdx.y is a slice of the
// visibilities whose partial table compact_reduce_kernel sums in slice order.
// ---------------------------------------------------------------------------
constexpr int PRE_KV = 256;
constexpr int PRE_TARGET_BLOCKS = 2048;

__device__ inline bool preload_offset(const double* __restrict__ g, int64_t Gx, int64_t Y, int64_t X, int64_t r,
                                      int64_t c, double& dy, double& dx) {
  // corner (cy, cx) and pixel (py, px) of the entry's quadrant; false for row Y / column X
  int64_t cy, cx, py, px;
  if (r < Y) {
    cy = 0;
    py = r;
  } else if (r > Y) {
    const int64_t i = 2 * Y - r;
    cy = Y - 1;
    py = Y - 1 - i;
  } else {
    return false;
  }
  if (c < X) {
    cx = 0;
    px = c;
  } else if (c > X) {
    const int64_t j = 2 * X - c;
    cx = X - 1;
    px = X - 1 - j;
  } else {
    return false;
  }
  dy = g[2 * (cy * Gx + cx)] - g[2 * (py * Gx + px)];
  dx = g[2 * (cy * Gx + cx) + 1] - g[2 * (py * Gx + px) + 1];
  return true;
}

__global__ void __launch_bounds__(256) preload_build_kernel(const double* __restrict__ noise,
                                                            const double* __restrict__ uv, int64_t K,
                                                            const double* __restrict__ g, int64_t Gx, int64_t Y,
                                                            int64_t X, int64_t k_per_split,
                                                            double* __restrict__ partial) {
  __shared__ double sw[PRE_KV], su0[PRE_KV], su1[PRE_KV];
  const int64_t E = 4 * Y * X;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t kb = (int64_t)blockIdx.y * k_per_split, ke = std::min<int64_t>(K, kb + k_per_split);
  double dy = 0.0, dx = 0.0;
  const bool live = e < E && preload_offset(g, Gx, Y, X, e / (2 * X), e % (2 * X), dy, dx);
  double acc = 0.0;
  for (int64_t k0 = kb; k0 < ke; k0 += PRE_KV) {
    const int n = (int)std::min<int64_t>(PRE_KV, ke - k0);
    if (threadIdx.x < n) {
      const double s = noise[k0 + threadIdx.x];
      sw[threadIdx.x] = 1.0 / (s * s);
      su0[threadIdx.x] = uv[2 * (k0 + threadIdx.x)];
      su1[threadIdx.x] = uv[2 * (k0 + threadIdx.x) + 1];
    }
    __syncthreads();
    if (live) {
      for (int k = 0; k < n; ++k) {
        double rv = fma(dx, su0[k], dy * su1[k]);
        rv -= rint(rv);
        acc = fma(sw[k], cospi(2.0 * rv), acc);
      }
    }
    __syncthreads();
  }
  if (e < E) partial[(int64_t)blockIdx.y * E + e] = acc;
}

// Visibility slices of the preload build: enough workgroups to fill the device, at least PRE_KV
// visibilities per slice.
struct PrePlan {
  int64_t nblk, nsplit, k_per_split;
  size_t bytes;
};

inline PrePlan make_pre_plan(int64_t K, int64_t Y, int64_t X) {
  PrePlan p{};
  const int64_t E = 4 * Y * X;
  p.nblk = ceil_div(E, (int64_t)256);
  const int64_t chunks = std::max<int64_t>(1, ceil_div(K, (int64_t)PRE_KV));
  p.nsplit = std::min<int64_t>(chunks, std::max<int64_t>(1, ceil_div((int64_t)PRE_TARGET_BLOCKS, p.nblk)));
  p.nsplit = std::min<int64_t>(p.nsplit, 65535);
  p.k_per_split = ceil_div(chunks, p.nsplit) * PRE_KV;
  p.nsplit = std::max<int64_t>(1, ceil_div(std::max<int64_t>(K, 1), p.k_per_split));
  p.bytes = align_up(sizeof(double) * p.nsplit * E, 256);
  return p;
}

struct Plan {
  int64_t ntn, ntm, nsplit, k_per_split;
  size_t bytes;
};

inline Plan make_plan(int64_t K, int64_t N) {
  Plan p{};
  const int64_t ND = 2 * N - 1;
  p.ntn = ceil_div(ND, (int64_t)CT_N);
  p.ntm = ceil_div(N, (int64_t)CT_M);
  const int64_t chunks = std::max<int64_t>(1, ceil_div(K, (int64_t)CT_KV));
  p.nsplit = std::min<int64_t>(chunks, std::max<int64_t>(1, ceil_div((int64_t)CT_TARGET_BLOCKS, p.ntn * p.ntm)));
  p.k_per_split = ceil_div(chunks, p.nsplit) * CT_KV;
  p.nsplit = std::max<int64_t>(1, ceil_div(std::max<int64_t>(K, 1), p.k_per_split));
  p.bytes = align_up(sizeof(double) * p.nsplit * N * ND, 256);
  return p;
}

}  // namespace cp
}  // namespace aj

using namespace aj;

extern "C" size_t aj_w_compact_workspace_bytes(int64_t K, int64_t grid_size) {
  if (K < 0 || grid_size <= 0) return 0;
  return cp::make_plan(K, grid_size).bytes;
}

extern "C" int aj_w_compact_curvature_interferometer_from(const double* noise_map_real, const double* uv_wavelengths,
                                                         int64_t K, int64_t grid_size, double pixel_scale,
                                                         double* w_compact, void* workspace, size_t workspace_bytes,
                                                         void* stream) {
  AJ_REQUIRE(K >= 0 && grid_size > 0 && grid_size < (1 << 20), "bad sizes K=%lld grid_size=%lld", (long long)K,
             (long long)grid_size);
  AJ_REQUIRE(w_compact && (K == 0 || (noise_map_real && uv_wavelengths)), "null pointer");
  const int64_t N = grid_size, ND = 2 * N - 1;
  hipStream_t st = as_stream(stream);
  if (K == 0) {
    AJ_CHECK_HIP(hipMemsetAsync(w_compact, 0, sizeof(double) * N * ND, st));
    return AJ_OK;
  }
  cp::Plan p = cp::make_plan(K, N);
  AJ_REQUIRE(workspace && workspace_bytes >= p.bytes, "workspace too small: need %zu bytes, got %zu", p.bytes,
             workspace_bytes);
  double* partial = static_cast<double*>(workspace);
  const double d_rad = pixel_scale * M_PI / 648000.0;  // arcsec -> rad (numba.py:221)
  hipLaunchKernelGGL(cp::compact_gemm_kernel, dim3((unsigned)p.ntn, (unsigned)p.ntm, (unsigned)p.nsplit), dim3(256),
                     0, st, noise_map_real, uv_wavelengths, K, N, d_rad, p.k_per_split, partial);
  AJ_LAUNCH_CHECK();
  const int64_t total = N * ND;
  hipLaunchKernelGGL(cp::compact_reduce_kernel, dim3((unsigned)ceil_div(total, (int64_t)256)), dim3(256), 0, st,
                     partial, p.nsplit, total, w_compact);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" int aj_w_tilde_via_compact_from(const double* w_compact, int64_t grid_size,
                                           const int64_t* native_index_for_slim_index, int64_t M, double* w_tilde,
                                           void* stream) {
  AJ_REQUIRE(grid_size > 0 && M > 0 && M < (1ll << 31), "bad sizes");
  AJ_REQUIRE(w_compact && native_index_for_slim_index && w_tilde, "null pointer");
  hipLaunchKernelGGL(cp::compact_expand_kernel,
                     dim3((unsigned)ceil_div(M, (int64_t)256), (unsigned)std::min<int64_t>(M, 32768)), dim3(256), 0,
                     as_stream(stream), w_compact, grid_size, native_index_for_slim_index, M, w_tilde);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" int aj_w_tilde_via_preload_from(const double* w_tilde_preload, int64_t preload_rows, int64_t preload_cols,
                                           const int64_t* native_index_for_slim_index, int64_t M, double* w_tilde,
                                           void* stream) {
  AJ_REQUIRE(preload_rows > 0 && preload_cols > 0 && M > 0 && M < (1ll << 31), "bad sizes");
  AJ_REQUIRE(w_tilde_preload && native_index_for_slim_index && w_tilde, "null pointer");
  hipLaunchKernelGGL(cp::preload_expand_kernel,
                     dim3((unsigned)ceil_div(M, (int64_t)256), (unsigned)std::min<int64_t>(M, 32768)), dim3(256), 0,
                     as_stream(stream), w_tilde_preload, preload_rows, preload_cols, native_index_for_slim_index, M,
                     w_tilde);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" size_t aj_w_tilde_preload_workspace_bytes(int64_t K, int64_t y_shape, int64_t x_shape) {
  if (K < 0 || y_shape <= 0 || x_shape <= 0) return 0;
  return cp::make_pre_plan(K, y_shape, x_shape).bytes;
}

extern "C" int aj_w_tilde_curvature_preload_interferometer_from(const double* noise_map_real,
                                                                const double* uv_wavelengths, int64_t K,
                                                                int64_t y_shape, int64_t x_shape,
                                                                const double* grid_radians_2d, int64_t grid_rows,
                                                                int64_t grid_cols, double* curvature_preload,
                                                                void* workspace, size_t workspace_bytes,
                                                                void* stream) {
  AJ_REQUIRE(K >= 0 && y_shape > 0 && x_shape > 0 && y_shape <= grid_rows && x_shape <= grid_cols &&
                 4 * y_shape * x_shape < (1ll << 40),
             "bad sizes K=%lld shape (%lld, %lld) grid (%lld, %lld)", (long long)K, (long long)y_shape,
             (long long)x_shape, (long long)grid_rows, (long long)grid_cols);
  AJ_REQUIRE(curvature_preload && grid_radians_2d && (K == 0 || (noise_map_real && uv_wavelengths)), "null pointer");
  hipStream_t st = as_stream(stream);
  const int64_t E = 4 * y_shape * x_shape;
  if (K == 0) {
    AJ_CHECK_HIP(hipMemsetAsync(curvature_preload, 0, sizeof(double) * E, st));
    return AJ_OK;
  }
  const cp::PrePlan p = cp::make_pre_plan(K, y_shape, x_shape);
  AJ_REQUIRE(workspace && workspace_bytes >= p.bytes, "workspace too small: need %zu bytes, got %zu", p.bytes,
             workspace_bytes);
  AJ_REQUIRE(p.nblk < (1ll << 31), "preload table too large");
  double* partial = static_cast<double*>(workspace);
  hipLaunchKernelGGL(cp::preload_build_kernel, dim3((unsigned)p.nblk, (unsigned)p.nsplit), dim3(256), 0, st,
                     noise_map_real, uv_wavelengths, K, grid_radians_2d, grid_cols, y_shape, x_shape, p.k_per_split,
                     partial);
  AJ_LAUNCH_CHECK();
  hipLaunchKernelGGL(cp::compact_reduce_kernel, dim3((unsigned)ceil_div(E, (int64_t)256)), dim3(256), 0, st, partial,
                     p.nsplit, E, curvature_preload);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}
