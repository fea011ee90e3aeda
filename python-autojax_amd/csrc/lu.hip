// Dense LU factorisation with partial pivoting of one general S x S matrix, spread over the device, and
// the solve and log|det| that follow from it: what np.linalg.solve and np.linalg.slogdet do in the
// reference (LAPACK dgesv / dgetrf: numba.py:731 `reconstruction_positive_negative_from`, numba.py:849-850
// the two log-determinants).  Any nonsingular matrix is solved (symmetric or not, definite or not); a zero
// pivot is reported as the 1-based column, as dgetrf's info.
//
// Blocked right-looking LU, panel width LU_NB:
//   panel_kernel   one workgroup: for each column of the panel, the pivot (first row of the largest |a|,
//                  as idamax), the row swap inside the panel, the multipliers, the rank-1 update of the
//                  panel's remaining columns (the panel stays in global memory / L2);
//   swap_kernel    the panel's row swaps applied to every column outside the panel (one thread per column,
//                  swaps in pivot order);
//   trsm_kernel    U12 = L11^-1 A12 (one thread per column, L11 and the column block in LDS);
//   update_kernel  A22 -= L21 U12 on v_mfma_f64_16x16x4 in 64 x 64 tiles (L21 and U12 tiles staged in LDS).
// Then solve_kernel (one workgroup): row swaps of the right-hand side, unit-lower and upper substitutions
// in 64-row blocks (a block's off-diagonal part as wave dot products over the rows, its diagonal block by
// one wave with lane shuffles), and log|det| = sum log|U_jj|.
#include "aj_common.h"

#include <cmath>
#include <mutex>

namespace aj {
namespace lu {

constexpr int LU_NB = 32;
constexpr int PANEL_THREADS = 1024;
constexpr int SOLVE_THREADS = 1024;
constexpr int64_t SOLVE_MAX = 16384;  // the solve keeps its vector in LDS (128 KB)

typedef double d4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(PANEL_THREADS) panel_kernel(double* __restrict__ A, int64_t n, int64_t k0, int nb,
                                                              int64_t* __restrict__ piv, int64_t* __restrict__ info) {
  __shared__ double red_v[PANEL_THREADS / 64];
  __shared__ int64_t red_i[PANEL_THREADS / 64];
  __shared__ int64_t p_s;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int jj = 0; jj < nb; ++jj) {
    const int64_t j = k0 + jj;
    // pivot: the first row of the largest |A[i, j]|, i >= j (each thread scans rows in increasing order,
    // so a strict > keeps its first maximum; ties across threads resolve to the smaller row)
    double best = -1.0;
    int64_t bi = j;
    for (int64_t i = j + tid; i < n; i += PANEL_THREADS) {
      const double v = fabs(A[i * n + j]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o, 64);
      const int64_t oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) {
        best = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      red_v[wave] = best;
      red_i[wave] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double bv = red_v[0];
      int64_t bix = red_i[0];
      for (int w = 1; w < PANEL_THREADS / 64; ++w)
        if (red_v[w] > bv || (red_v[w] == bv && red_i[w] < bix)) {
          bv = red_v[w];
          bix = red_i[w];
        }
      if (!(bv > 0.0) && *info == 0) *info = j + 1;  // exact zero pivot: singular (dgetrf info)
      p_s = bix;
      piv[j] = bix;
    }
    __syncthreads();
    const int64_t p = p_s;
    if (p != j)
      for (int c = tid; c < nb; c += PANEL_THREADS) {
        const double t = A[j * n + k0 + c];
        A[j * n + k0 + c] = A[p * n + k0 + c];
        A[p * n + k0 + c] = t;
      }
    __syncthreads();
    const double pivot = A[j * n + j];
    if (pivot != 0.0)
      for (int64_t i = j + 1 + tid; i < n; i += PANEL_THREADS) A[i * n + j] /= pivot;
    __syncthreads();
    // rank-1 update of the panel's columns right of j
    const int rc = nb - jj - 1;
    if (rc > 0) {
      const int64_t rows = n - j - 1;
      for (int64_t e = tid; e < rows * rc; e += PANEL_THREADS) {
        const int64_t i = j + 1 + e / rc;
        const int64_t c = j + 1 + e % rc;
        A[i * n + c] = fma(-A[i * n + j], A[j * n + c], A[i * n + c]);
      }
    }
    __syncthreads();
  }
}

// Row swaps of the panel [k0, k0 + nb) applied to the columns outside it, in pivot order.
__global__ void swap_kernel(double* __restrict__ A, int64_t n, int64_t k0, int nb, const int64_t* __restrict__ piv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ncol = n - nb;
  if (t >= ncol) return;
  const int64_t c = t < k0 ? t : t + nb;
  for (int jj = 0; jj < nb; ++jj) {
    const int64_t j = k0 + jj, p = piv[j];
    if (p != j) {
      const double v = A[j * n + c];
      A[j * n + c] = A[p * n + c];
      A[p * n + c] = v;
    }
  }
}

// U12 = L11^-1 A12 for the rows of the panel and every column right of it (L11 unit lower triangular):
// TRSM_COLS columns per workgroup, one thread per column substituting down its own column of the block,
// which sits in LDS beside L11 (no barrier between rows: a thread reads only its own column).
constexpr int TRSM_COLS = 128;

__global__ void __launch_bounds__(TRSM_COLS) trsm_kernel(double* __restrict__ A, int64_t n, int64_t k0, int nb) {
  __shared__ double L[LU_NB][LU_NB + 1];
  __shared__ double X[LU_NB][TRSM_COLS + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < LU_NB * LU_NB; e += TRSM_COLS) {
    const int r = e / LU_NB, c = e % LU_NB;
    L[r][c] = (r < nb && c < r) ? A[(k0 + r) * n + k0 + c] : 0.0;
  }
  const int64_t c = k0 + nb + (int64_t)blockIdx.x * TRSM_COLS + tid;
  const bool live = c < n;
  for (int r = 0; r < nb; ++r) X[r][tid] = live ? A[(k0 + r) * n + c] : 0.0;
  __syncthreads();
  if (!live) return;
  for (int r = 1; r < nb; ++r) {
    double v = X[r][tid];
    for (int l = 0; l < r; ++l) v = fma(-L[r][l], X[l][tid], v);
    X[r][tid] = v;
  }
  for (int r = 1; r < nb; ++r) A[(k0 + r) * n + c] = X[r][tid];
}

// A22 -= L21 U12 over rows and columns [k0 + nb, n): one 64 x 64 tile per workgroup of four waves, each
// wave a 32 x 32 quarter as 2 x 2 MFMA blocks; the L21 (64 x nb) and U12 (nb x 64) tiles in LDS.
__global__ void __launch_bounds__(256) update_kernel(double* __restrict__ A, int64_t n, int64_t k0, int nb) {
  __shared__ double Ls[64][LU_NB + 1];
  __shared__ double Us[LU_NB][64 + 1];
  const int64_t r0 = k0 + nb;
  const int64_t R0 = r0 + (int64_t)blockIdx.y * 64, C0 = r0 + (int64_t)blockIdx.x * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < 64 * LU_NB; e += 256) {
    const int r = e / LU_NB, k = e % LU_NB;
    const int64_t gi = R0 + r;
    Ls[r][k] = (gi < n && k < nb) ? A[gi * n + k0 + k] : 0.0;
  }
  for (int e = tid; e < LU_NB * 64; e += 256) {
    const int k = e / 64, c = e % 64;
    const int64_t gj = C0 + c;
    Us[k][c] = (gj < n && k < nb) ? A[(k0 + k) * n + gj] : 0.0;
  }
  __syncthreads();
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = R0 + wr + a * 16 + lk + 4 * r, gj = C0 + wc + b * 16 + lr;
        acc[a][b][r] = (gi < n && gj < n) ? A[gi * n + gj] : 0.0;
      }
#pragma unroll
  for (int ks = 0; ks < LU_NB / 4; ++ks) {
    const int k = ks * 4 + lk;
    const double fa0 = -Ls[wr + lr][k], fa1 = -Ls[wr + 16 + lr][k];
    const double fb0 = Us[k][wc + lr], fb1 = Us[k][wc + 16 + lr];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa0, fb0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa0, fb1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1, fb0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1, fb1, acc[1][1], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = R0 + wr + a * 16 + lk + 4 * r, gj = C0 + wc + b * 16 + lr;
        if (gi < n && gj < n) A[gi * n + gj] = acc[a][b][r];
      }
}

// x = A^-1 b from the factors (P A = L U): swaps, L y = P b, U x = y; logabsdet = sum log|U_jj|.
__global__ void __launch_bounds__(SOLVE_THREADS) solve_kernel(const double* __restrict__ LUm, int64_t n,
                                                              const int64_t* __restrict__ piv,
                                                              const double* __restrict__ b, double* __restrict__ x,
                                                              double* __restrict__ logabsdet) {
  extern __shared__ double y[];
  __shared__ double part[64];
  __shared__ double red[SOLVE_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = SOLVE_THREADS / 64;
  for (int64_t i = tid; i < n; i += SOLVE_THREADS) y[i] = b ? b[i] : 0.0;
  __syncthreads();
  if (tid == 0)
    for (int64_t j = 0; j < n; ++j) {
      const int64_t p = piv[j];
      if (p != j) {
        const double t = y[j];
        y[j] = y[p];
        y[p] = t;
      }
    }
  __syncthreads();
  // forward, unit lower: 64-row blocks
  for (int64_t r0 = 0; r0 < n; r0 += 64) {
    const int rows = (int)std::min<int64_t>(64, n - r0);
    // part[r] = sum_{c < r0} L[r0 + r, c] y[c]: wave w takes rows w, w + NW, ...
    for (int r = wave; r < rows; r += NW) {
      double acc = 0.0;
      const double* row = LUm + (r0 + r) * n;
      for (int64_t c = lane; c < r0; c += 64) acc = fma(row[c], y[c], acc);
      acc = wave_sum(acc);
      if (lane == 0) part[r] = acc;
    }
    __syncthreads();
    if (wave == 0) {
      double v = lane < rows ? y[r0 + lane] - part[lane] : 0.0;
      for (int l = 0; l < rows; ++l) {
        const double xl = __shfl(v, l, 64);
        if (lane > l && lane < rows) v = fma(-LUm[(r0 + lane) * n + r0 + l], xl, v);
      }
      if (lane < rows) y[r0 + lane] = v;
    }
    __syncthreads();
  }
  // backward, upper with its diagonal: 64-row blocks from the bottom
  for (int64_t r1 = n; r1 > 0; r1 -= 64) {
    const int64_t r0 = std::max<int64_t>(0, r1 - 64);
    const int rows = (int)(r1 - r0);
    for (int r = wave; r < rows; r += NW) {
      double acc = 0.0;
      const double* row = LUm + (r0 + r) * n;
      for (int64_t c = r1 + lane; c < n; c += 64) acc = fma(row[c], y[c], acc);
      acc = wave_sum(acc);
      if (lane == 0) part[r] = acc;
    }
    __syncthreads();
    if (wave == 0) {
      double v = lane < rows ? y[r0 + lane] - part[lane] : 0.0;
      for (int l = rows - 1; l >= 0; --l) {
        if (lane == l) v = v / LUm[(r0 + l) * n + r0 + l];
        const double xl = __shfl(v, l, 64);
        if (lane < l) v = fma(-LUm[(r0 + lane) * n + r0 + l], xl, v);
      }
      if (lane < rows) y[r0 + lane] = v;
    }
    __syncthreads();
  }
  if (x)
    for (int64_t i = tid; i < n; i += SOLVE_THREADS) x[i] = y[i];
  double ld = 0.0;
  for (int64_t i = tid; i < n; i += SOLVE_THREADS) ld += log(fabs(LUm[i * n + i]));
  ld = wave_sum(ld);
  if (lane == 0) red[wave] = ld;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int w = 0; w < NW; ++w) s += red[w];
    if (logabsdet) *logabsdet = s;
  }
}

__global__ void copy_kernel(const double* __restrict__ src, int64_t count, double* __restrict__ dst,
                            int64_t* __restrict__ info) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
    dst[e] = src[e];
  if (blockIdx.x == 0 && threadIdx.x == 0) *info = 0;
}

}  // namespace lu
}  // namespace aj

using namespace aj;
using namespace aj::lu;

extern "C" size_t aj_solve_workspace_bytes(int64_t S) {
  return S > 0 ? align_up(sizeof(double) * S * S, 256) + align_up(sizeof(int64_t) * S, 256) + 256 : 0;
}

namespace {

// The factorisation and solve on a workspace of at least aj_solve_workspace_bytes(S) (checked by the
// callers).  Nothing in the workspace needs to start at any value: the factors are a copy of the input and
// every pivot is written before it is read.
int lu_solve(const double* data_vector, const double* curvature_reg_matrix, int64_t S, double* reconstruction,
             double* logdet, int64_t* info, void* workspace, size_t workspace_bytes, hipStream_t st) {
  Arena ar(workspace, workspace_bytes);
  double* LUm = ar.take<double>(S * S);
  int64_t* piv = ar.take<int64_t>(S);
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(S * S, 256), 4096)), dim3(256), 0, st,
                     curvature_reg_matrix, S * S, LUm, info);
  for (int64_t k0 = 0; k0 < S; k0 += LU_NB) {
    const int nb = (int)std::min<int64_t>(LU_NB, S - k0);
    hipLaunchKernelGGL(panel_kernel, dim3(1), dim3(PANEL_THREADS), 0, st, LUm, S, k0, nb, piv, info);
    if (S - nb > 0)
      hipLaunchKernelGGL(swap_kernel, dim3((unsigned)ceil_div(S - nb, 256)), dim3(256), 0, st, LUm, S, k0, nb, piv);
    const int64_t rest = S - k0 - nb;
    if (rest > 0) {
      hipLaunchKernelGGL(trsm_kernel, dim3((unsigned)ceil_div(rest, TRSM_COLS)), dim3(TRSM_COLS), 0, st, LUm, S, k0,
                         nb);
      const unsigned t = (unsigned)ceil_div(rest, 64);
      hipLaunchKernelGGL(update_kernel, dim3(t, t), dim3(256), 0, st, LUm, S, k0, nb);
    }
  }
  AJ_CHECK_HIP(hipFuncSetAttribute((const void*)solve_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(double) * SOLVE_MAX)));
  hipLaunchKernelGGL(solve_kernel, dim3(1), dim3(SOLVE_THREADS), sizeof(double) * S, st, LUm, S, piv,
                     reconstruction ? data_vector : nullptr, reconstruction, logdet);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

// Scratch of the allocation-free variant: the workspace of the largest S it takes, one copy per device
// (a file-scope device array: each GPU that loads the library holds its own until the process ends).
constexpr int64_t NOALLOC_S_MAX = 2048;
constexpr size_t NOALLOC_BYTES = ((sizeof(double) * NOALLOC_S_MAX * NOALLOC_S_MAX + 255) / 256) * 256 +
                                 ((sizeof(int64_t) * NOALLOC_S_MAX + 255) / 256) * 256 + 256;
__device__ __align__(256) char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
void* scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per device (the array's name in
// host code is only a handle, never a device address).
int scratch_on_current_device(void** out) {
  int dev = 0;
  AJ_CHECK_HIP(hipGetDevice(&dev));
  AJ_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "device %d beyond the %d devices this library tracks", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    AJ_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    AJ_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the solve scratch");
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return AJ_OK;
}

}  // namespace

extern "C" int aj_reconstruction_positive_negative_from(const double* data_vector, const double* curvature_reg_matrix,
                                                        int64_t S, double* reconstruction, double* logdet,
                                                        int64_t* info, void* workspace, size_t workspace_bytes,
                                                        void* stream) {
  AJ_REQUIRE(S > 0 && curvature_reg_matrix && info && (reconstruction == nullptr || data_vector != nullptr),
             "bad arguments");
  AJ_REQUIRE(S <= SOLVE_MAX, "S = %lld exceeds the solve's limit of %lld", (long long)S, (long long)SOLVE_MAX);
  AJ_REQUIRE(workspace && workspace_bytes >= aj_solve_workspace_bytes(S), "workspace too small");
  return lu_solve(data_vector, curvature_reg_matrix, S, reconstruction, logdet, info, workspace, workspace_bytes,
                  as_stream(stream));
}

extern "C" int aj_reconstruction_positive_negative_from_noalloc(const double* data_vector,
                                                                const double* curvature_reg_matrix, int64_t S,
                                                                double* reconstruction, double* logdet,
                                                                int64_t* info, void* stream) {
  AJ_REQUIRE(S > 0 && curvature_reg_matrix && info && (reconstruction == nullptr || data_vector != nullptr),
             "bad arguments");
  AJ_REQUIRE(S <= NOALLOC_S_MAX, "S = %lld exceeds the allocation-free solve's limit of %lld (its scratch is %zu "
             "bytes); use aj_reconstruction_positive_negative_from with a workspace", (long long)S,
             (long long)NOALLOC_S_MAX, NOALLOC_BYTES);
  void* ws = nullptr;
  const int rc = scratch_on_current_device(&ws);
  if (rc != AJ_OK) return rc;
  return lu_solve(data_vector, curvature_reg_matrix, S, reconstruction, logdet, info, ws, NOALLOC_BYTES,
                  as_stream(stream));
}
