// W-tilde curvature build (the NUFFT Gram matrix) for gfx950.
//
// Reference: src/autojax/numba.py:155-207 (w_tilde_curvature_interferometer_from)
//   W_ij = sum_k n_k^-2 cos(2 pi [(g_i1 - g_j1) u_k0 + (g_i0 - g_j0) u_k1])
//
// Rewritten as a Gram matrix of cos/sin features so that it runs on MFMA:
//   r_ik = g_i1 u_k0 + g_i0 u_k1            (revolutions, FP64, reduced to [-1/2, 1/2])
//   C_ik = sqrt(w_k) cos(2 pi r_ik),  S_ik = sqrt(w_k) sin(2 pi r_ik)
//   W    = [C S] [C S]^T                    (lower-triangle SYRK, inner dim 2K)
// The grid is centred first (W depends on differences only), which keeps the
// phases small.  K is processed in chunks of KC visibilities; a launch covers
// a group of chunks whose features sit in the caller's workspace.
//
// Precision modes (feature element / MFMA / accumulation):
//   FP64    : double features (sincospi), v_mfma_f64_16x16x4, FP64 accumulate.
//   FP32    : float features, v_mfma_f32_16x16x4, FP32 within a chunk, FP64 across.
//   BF16X3  : float features split hi+lo bf16, hi*hi + hi*lo + lo*hi on
//             v_mfma_f32_16x16x32_bf16, FP32 within a chunk, FP64 across.
//   INT8XN  : float features scaled per chunk to N signed 8-bit digits,
//             all N^2 digit products on v_mfma_i32_16x16x64_i8 with exact int32
//             accumulation per digit-order group, FP64 across chunks.  The
//             result is the exact Gram matrix of the quantised features.
#include "aj_common.h"
#include "data_sums.h"

#include <vector>

namespace aj {
namespace wt {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef int i16v __attribute__((ext_vector_type(16)));

constexpr int KC = 4096;        // visibilities per chunk
constexpr int FC = 2 * KC;      // features per chunk (cos block then sin block)
constexpr int TILE = 64;        // output tile edge per workgroup
constexpr int ROWB = 128;       // bytes per row per slice per LDS stage
constexpr int PITCH = ROWB + 16;  // padded LDS row pitch (bytes)

template <int MODE> struct Mode;
template <> struct Mode<AJ_PREC_FP64> {
  using E = double; static constexpr int NS = 1, MK = 4, NG = 1;
};
template <> struct Mode<AJ_PREC_FP32> {
  using E = float; static constexpr int NS = 1, MK = 4, NG = 1;
};
template <> struct Mode<AJ_PREC_BF16X3> {
  using E = uint16_t; static constexpr int NS = 2, MK = 32, NG = 1;
};
template <> struct Mode<AJ_PREC_INT8X3> {
  using E = int8_t; static constexpr int NS = 3, MK = 64, NG = 5;
};
template <> struct Mode<AJ_PREC_INT8X4> {
  using E = int8_t; static constexpr int NS = 4, MK = 64, NG = 6;
};
template <> struct Mode<AJ_PREC_INT8X5> {
  using E = int8_t; static constexpr int NS = 5, MK = 64, NG = 7;
};
__host__ __device__ constexpr bool is_i8(int mode) {
  return mode == AJ_PREC_INT8X3 || mode == AJ_PREC_INT8X4 || mode == AJ_PREC_INT8X5;
}

// Digit-order groups kept by the integer SYRK kernels (Mode<>::NG: products a_p*b_q with p+q < NG).
//   int8x3  : all orders 0..4 of 3 digits (9 products).
//   int8x4  : orders 0..5 of 4 digits (15 products); the single order-6 product a3*b3 (weight 256^-6
//             of the leading term, about 1e-17 relative) is below the FP64 rounding of the sum.
//   int8x5  : orders 0..6 of 5 digits (22 products); orders 7 and 8 (weight 256^-7 and less) are
//             below the FP64 rounding.
// Dropping whole orders above the FP64 rounding (e.g. keeping orders 0..3 of 4 digits) leaves W-tilde
// accurate to ~1e-9 but no longer the Gram matrix of any feature matrix: it can lose positive
// semi-definiteness, and F + H then fails to factor at small regularization coefficients.

inline int slices_of(int mode) {
  switch (mode) {
    case AJ_PREC_BF16X3: return 2;
    case AJ_PREC_INT8X3: return 3;
    case AJ_PREC_INT8X4: return 4;
    case AJ_PREC_INT8X5: return 5;
    default: return 1;
  }
}
inline int esize_of(int mode) {
  switch (mode) {
    case AJ_PREC_FP64: return 8;
    case AJ_PREC_FP32: return 4;
    case AJ_PREC_BF16X3: return 2;
    default: return 1;
  }
}

// s_waitcnt immediate that waits for vmcnt <= n and leaves expcnt and lgkmcnt unconstrained
// (gfx9 encoding: vmcnt[3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]).
constexpr unsigned aj_vmcnt_imm(unsigned n) { return (n & 0xF) | ((n >> 4) << 14) | (0x7u << 4) | (0xFu << 8); }
// s_waitcnt immediate that waits for lgkmcnt == 0 only (vmcnt 63 and expcnt 7 unconstrained).
constexpr unsigned aj_lgkmcnt0_imm() { return 0xFu | (0x3u << 14) | (0x7u << 4); }

// Largest integer digit pattern for N signed 8-bit digits: 127 * 256^(N-1).
__host__ __device__ inline double qmax_of(int ns) {
  double q = 127.0;
  for (int i = 1; i < ns; ++i) q *= 256.0;
  return q;
}

// ---------------------------------------------------------------------------
// Grid centring: mean of the grid, FP64, one block.
// ---------------------------------------------------------------------------
__global__ void grid_mean_kernel(const double* __restrict__ grid, int64_t M, double* __restrict__ out) {
  __shared__ double s0[256], s1[256];
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < M; i += 256) {
    a += grid[2 * i];
    b += grid[2 * i + 1];
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s0[threadIdx.x] += s0[threadIdx.x + o];
      s1[threadIdx.x] += s1[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s0[0] / (double)M;
    out[1] = s1[0] / (double)M;
  }
}

// Per-chunk scale for the integer modes: max_k 1/|n_k| over the chunk.
__global__ void chunk_scale_kernel(const double* __restrict__ noise, int64_t K, int64_t k_begin, int nchunks,
                                   double* __restrict__ sc) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  double m = 0.0;
  for (int k = threadIdx.x; k < KC; k += 256) {
    int64_t kg = k_begin + (int64_t)c * KC + k;
    if (kg < K) m = fmax(m, 1.0 / fabs(noise[kg]));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) sc[c] = red[0] > 0.0 ? red[0] : 1.0;
}

// ---------------------------------------------------------------------------
// Feature kernels.  Thread = one visibility of the group, loops over ROWS
// image pixels.  Layout: X[slice][i][c*FC + k] = cos part, [.. + KC + k] = sin.
// Rows are padded to a multiple of TILE; padded rows hold zeros.
// ---------------------------------------------------------------------------
constexpr int FEAT_ROWS = 16;

__device__ inline double reduced_revs(double gx, double gy, double u0, double u1) {
  double r = fma(gx, u0, gy * u1);
  return r - rint(r);
}

// With DS (FP64 only) the kernel also accumulates the data-side sums of its 256 visibilities and 16 rows
// from the same cos/sin values (data_sums.h); workgroup (x, y) writes block blk0 + x of the row sums.
// Every thread runs all 16 rows: the grid covers whole chunks (KC is a multiple of 256) and Mpad is a
// multiple of FEAT_ROWS.
static_assert(KC % 256 == 0 && TILE % FEAT_ROWS == 0, "feature grid must cover whole workgroups");

template <int MODE, bool DS>
__global__ void __launch_bounds__(256) features_kernel(
    const double* __restrict__ noise, const double* __restrict__ uv, const double* __restrict__ grid,
    const double* __restrict__ centre, int64_t K, int64_t M, int64_t Mpad, int64_t k_begin, int nchunks,
    const double* __restrict__ chunk_sc, void* __restrict__ Xv, int64_t ldx, DataSums dsum, int64_t blk0) {
  static_assert(!DS || MODE == AJ_PREC_FP64, "data-side sums need FP64 features");
  using E = typename Mode<MODE>::E;
  constexpr int NS = Mode<MODE>::NS;
  E* X = reinterpret_cast<E*>(Xv);
  const int64_t kl = (int64_t)blockIdx.x * 256 + threadIdx.x;  // local visibility index in group
  if (!DS && kl >= (int64_t)nchunks * KC) return;
  const int c = (int)(kl / KC);
  const int kc = (int)(kl % KC);
  const int64_t kg = k_begin + kl;
  double u0 = 0.0, u1 = 0.0, sw = 0.0;
  if (kg < K) {
    u0 = uv[2 * kg];
    u1 = uv[2 * kg + 1];
    sw = 1.0 / fabs(noise[kg]);
  }
  const double cy = centre[0], cx = centre[1];
  const int64_t col_c = (int64_t)c * FC + kc;
  const int64_t col_s = col_c + KC;
  const int64_t sstride = Mpad * ldx;
  double qscale = 0.0;
  if constexpr (is_i8(MODE)) qscale = sw / chunk_sc[c] * qmax_of(NS);
  double ca = 0.0, cb = 0.0, cc = 0.0, cd = 0.0;  // data-side coefficients of this visibility
  if constexpr (DS) {
    if (kg < K) {
      ca = dsum.coef[4 * kg];
      cb = dsum.coef[4 * kg + 1];
      cc = dsum.coef[4 * kg + 2];
      cd = dsum.coef[4 * kg + 3];
    }
  }
  RowSum16 rs0, rs1;
  double s0 = 0.0, s1 = 0.0;
  const int64_t i0 = (int64_t)blockIdx.y * FEAT_ROWS;
  for (int r = 0; r < FEAT_ROWS; ++r) {
    const int64_t i = i0 + r;
    if (!DS && i >= Mpad) break;
    double cv = 0.0, sv = 0.0;
    float cf = 0.f, sf = 0.f;
    const bool live = (i < M) && (kg < K);
    if constexpr (MODE == AJ_PREC_FP64) {
      double d0 = 0.0, d1 = 0.0;
      if (live) {
        double rv = reduced_revs(grid[2 * i + 1] - cx, grid[2 * i] - cy, u0, u1);
        sincospi(2.0 * rv, &sv, &cv);
        if constexpr (DS) {
          d0 = fma(ca, cv, cb * sv);
          d1 = fma(cc, cv, cd * sv);
        }
        cv *= sw;
        sv *= sw;
      }
      X[i * ldx + col_c] = cv;
      X[i * ldx + col_s] = sv;
      if constexpr (DS) {
        rs0.push(r, d0, threadIdx.x & 63, s0);
        rs1.push(r, d1, threadIdx.x & 63, s1);
      }
    } else if constexpr (is_i8(MODE)) {
      // FP64 cos/sin, quantised to NS balanced 8-bit digits of a fixed-point value
      if (live) {
        double rv = reduced_revs(grid[2 * i + 1] - cx, grid[2 * i] - cy, u0, u1);
        sincospi(2.0 * rv, &sv, &cv);
      }
      long long qc = live ? llrint(cv * qscale) : 0;
      long long qs = live ? llrint(sv * qscale) : 0;
#pragma unroll
      for (int p = NS - 1; p >= 0; --p) {
        int8_t dc = (int8_t)(qc & 0xFF);
        int8_t ds = (int8_t)(qs & 0xFF);
        qc = (qc - dc) >> 8;
        qs = (qs - ds) >> 8;
        X[p * sstride + i * ldx + col_c] = dc;
        X[p * sstride + i * ldx + col_s] = ds;
      }
    } else {
      if (live) {
        double rv = reduced_revs(grid[2 * i + 1] - cx, grid[2 * i] - cy, u0, u1);
        sincospif(2.0f * (float)rv, &sf, &cf);
      }
      if constexpr (MODE == AJ_PREC_FP32) {
        X[i * ldx + col_c] = cf * (float)sw;
        X[i * ldx + col_s] = sf * (float)sw;
      } else if constexpr (MODE == AJ_PREC_BF16X3) {
        float a = cf * (float)sw, b = sf * (float)sw;
        __bf16 ah = (__bf16)a, bh = (__bf16)b;
        __bf16 al = (__bf16)(a - (float)ah), bl = (__bf16)(b - (float)bh);
        X[i * ldx + col_c] = __builtin_bit_cast(uint16_t, ah);
        X[i * ldx + col_s] = __builtin_bit_cast(uint16_t, bh);
        X[sstride + i * ldx + col_c] = __builtin_bit_cast(uint16_t, al);
        X[sstride + i * ldx + col_s] = __builtin_bit_cast(uint16_t, bl);
      }
    }
  }
  if constexpr (DS) block_row_sums_out(s0, s1, dsum, blk0 + blockIdx.x, i0);
}

// ---------------------------------------------------------------------------
// SYRK over the lower triangle of 64x64 tiles.  4 waves as 2x2, each wave a
// 32x32 block = 2x2 MFMA 16x16 blocks.  Per LDS stage each slice contributes
// 64 rows x 128 bytes of features for both operands.
// ---------------------------------------------------------------------------
__device__ inline void tile_of(int64_t t, int64_t& I, int64_t& J) {
  int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (i * (i + 1) / 2 > t) --i;
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  I = i;
  J = t - i * (i + 1) / 2;
}

template <int MODE>
__global__ void __launch_bounds__(256) syrk_kernel(const void* __restrict__ Xv, int64_t M, int64_t Mpad,
                                                   int64_t ldx, int nchunks, const double* __restrict__ chunk_sc,
                                                   double* __restrict__ W, int first) {
  using E = typename Mode<MODE>::E;
  constexpr int NS = Mode<MODE>::NS;
  constexpr int MK = Mode<MODE>::MK;
  constexpr int NG = Mode<MODE>::NG;
  constexpr int ES = (int)sizeof(E);
  constexpr int KT = ROWB / ES;          // features per stage
  constexpr int NSTEP = KT / MK;         // MFMA k-steps per stage
  constexpr bool INT8 = is_i8(MODE);

  __shared__ __attribute__((aligned(16))) char lds[2][NS][TILE * PITCH];

  int64_t TI, TJ;
  tile_of(blockIdx.x, TI, TJ);
  const int64_t I0 = TI * TILE, J0 = TJ * TILE;
  const bool diag = (TI == TJ);
  const char* X = reinterpret_cast<const char*>(Xv);
  const int64_t sstride = Mpad * ldx * ES;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = (wave >> 1) * 32;  // wave row offset in tile
  const int wc = (wave & 1) * 32;   // wave col offset in tile
  const int lr = lane & 15;         // fragment row within 16-block
  const int lk = lane >> 4;         // k-group of the lane

  double accd[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) accd[a][b][r] = 0.0;

  using AccT = typename std::conditional<MODE == AJ_PREC_FP64, d4,
               typename std::conditional<INT8, i4, f4>::type>::type;
  AccT acc[NG][2][2];

  const int64_t chunk_cols = FC;
  for (int c = 0; c < nchunks; ++c) {
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[g][a][b] = AccT{0, 0, 0, 0};

    for (int64_t k0 = (int64_t)c * chunk_cols; k0 < (int64_t)(c + 1) * chunk_cols; k0 += KT) {
      // Global -> LDS: 64 rows x 128 B per slice per operand; 8 threads per row.
      __syncthreads();
#pragma unroll
      for (int s = 0; s < NS; ++s) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int idx = tid + h * 256;
          const int row = idx >> 3, seg = idx & 7;
          const int4* srcA = reinterpret_cast<const int4*>(X + s * sstride + ((I0 + row) * ldx + k0) * ES) + seg;
          *reinterpret_cast<int4*>(&lds[0][s][row * PITCH + seg * 16]) = *srcA;
          if (!diag) {
            const int4* srcB = reinterpret_cast<const int4*>(X + s * sstride + ((J0 + row) * ldx + k0) * ES) + seg;
            *reinterpret_cast<int4*>(&lds[1][s][row * PITCH + seg * 16]) = *srcB;
          }
        }
      }
      __syncthreads();
      const int bsel = diag ? 0 : 1;
#pragma unroll
      for (int st = 0; st < NSTEP; ++st) {
        if constexpr (MODE == AJ_PREC_FP64 || MODE == AJ_PREC_FP32) {
          E fa[2], fb[2];
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            fa[a] = *reinterpret_cast<const E*>(&lds[0][0][(wr + a * 16 + lr) * PITCH + (st * MK + lk) * ES]);
            fb[a] = *reinterpret_cast<const E*>(&lds[bsel][0][(wc + a * 16 + lr) * PITCH + (st * MK + lk) * ES]);
          }
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
              if constexpr (MODE == AJ_PREC_FP64)
                acc[0][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[b], acc[0][a][b], 0, 0, 0);
              else
                acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a], fb[b], acc[0][a][b], 0, 0, 0);
            }
        } else if constexpr (MODE == AJ_PREC_BF16X3) {
          bf8 fa[2][2], fb[2][2];  // [slice][block]
#pragma unroll
          for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
              fa[s][a] = *reinterpret_cast<const bf8*>(&lds[0][s][(wr + a * 16 + lr) * PITCH + (st * MK + lk * 8) * ES]);
              fb[s][a] = *reinterpret_cast<const bf8*>(&lds[bsel][s][(wc + a * 16 + lr) * PITCH + (st * MK + lk * 8) * ES]);
            }
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
              acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[1][a], fb[0][b], acc[0][a][b], 0, 0, 0);
              acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0][a], fb[1][b], acc[0][a][b], 0, 0, 0);
              acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0][a], fb[0][b], acc[0][a][b], 0, 0, 0);
            }
        } else {
          i4 fa[NS][2], fb[NS][2];
#pragma unroll
          for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
              fa[s][a] = *reinterpret_cast<const i4*>(&lds[0][s][(wr + a * 16 + lr) * PITCH + st * MK + lk * 16]);
              fb[s][a] = *reinterpret_cast<const i4*>(&lds[bsel][s][(wc + a * 16 + lr) * PITCH + st * MK + lk * 16]);
            }
#pragma unroll
          for (int p = 0; p < NS; ++p)
#pragma unroll
            for (int q = 0; q < NS; ++q)
#pragma unroll
              for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                  acc[p + q][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[p][a], fb[q][b], acc[p + q][a][b], 0, 0, 0);
        }
      }
    }
    // Chunk flush into FP64.
    if constexpr (MODE == AJ_PREC_FP64) {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) accd[a][b][r] += acc[0][a][b][r];
    } else if constexpr (INT8) {
      const double q = qmax_of(NS);
      const double s2 = chunk_sc[c] * chunk_sc[c] / (q * q);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            double v = 0.0;
#pragma unroll
            for (int g = 0; g < NG; ++g) v = v * 256.0 + (double)acc[g][a][b][r];
            accd[a][b][r] += v * s2;
          }
    } else {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) accd[a][b][r] += (double)acc[0][a][b][r];
    }
  }

  // Write / accumulate the tile: lane holds rows 4*(lane/16)+r, column lane%16.
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // v_mfma_f64_16x16x4 holds row 4r + lane/16; the f32/bf16/i8 forms hold row 4*(lane/16) + r
        const int64_t row = I0 + wr + a * 16 + (MODE == AJ_PREC_FP64 ? lk + 4 * r : lk * 4 + r);
        const int64_t col = J0 + wc + b * 16 + lr;
        if (row < M && col < M) {
          double* p = W + row * M + col;
          *p = first ? accd[a][b][r] : *p + accd[a][b][r];
        }
      }
}

// ---------------------------------------------------------------------------
// FP64 SYRK (the plain path): 128 x 128 lower-triangle tiles, eight waves as 4 x 2, each a 32 x 64 block
// (2 x 4 v_mfma_f64_16x16x4 blocks).  Each launch's K (its chunks' features) is cut into F64_SPLIT parts;
// unit u is part u / ntiles of tile u % ntiles, so there are F64_SPLIT times as many units as tiles (a
// last, partial round costs a smaller share of the launch) and the units in flight read the same part
// of K (shared row panels in L2).  Part 0 accumulates into W, the other parts into their own partial
// tiles P (tile-major), added to W once after the last launch.  Per 16-double stage the next stage's
// panels are loaded into registers while the MFMAs read the current one from LDS: one barrier per stage.
// Two workgroups fit a CU (74 KB LDS, 127 VGPRs each).  At M = 5,000 (820 tiles) the SYRK ran at 49, 57,
// 60, 63, 60, 61 and 63 TFLOP/s with 1, 2, 3, 4, 5, 6 and 8 parts; 4 is the default (compile-time
// AJ_F64_SPLIT).
// ---------------------------------------------------------------------------
#ifndef AJ_F64_SPLIT
#define AJ_F64_SPLIT 4
#endif
constexpr int F64_TILE = 128;
constexpr int F64_KS = 16;             // doubles of a row per stage
constexpr int F64_PITCH = F64_KS + 2;  // LDS row pitch in doubles
constexpr int F64_SPLIT = AJ_F64_SPLIT;
static_assert(F64_SPLIT >= 1 && FC % F64_KS == 0 && FC / F64_KS >= F64_SPLIT, "every part of K holds a stage");

__global__ void __launch_bounds__(512) syrk_f64_kernel(const double* __restrict__ X, int64_t M, int64_t ldx,
                                                       int64_t kfeat, int ntiles, double* __restrict__ W,
                                                       double* __restrict__ P, int first) {
  __shared__ __attribute__((aligned(16))) double lds[2][2][F64_TILE * F64_PITCH];
  const int h = (int)blockIdx.x / ntiles, t = (int)blockIdx.x % ntiles;
  int64_t TI, TJ;
  tile_of(t, TI, TJ);
  const int64_t I0 = TI * F64_TILE, J0 = TJ * F64_TILE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 64, lr = lane & 15, lk = lane >> 4;
  const int64_t nst = kfeat / F64_KS;
  const int64_t s_beg = h * nst / F64_SPLIT, s_end = (h + 1) * nst / F64_SPLIT;
  // staging: per operand 128 rows x 8 16-byte segments; thread tid loads (row, seg) = (tid >> 3, tid & 7) and
  // (row + 64, seg) of both operands (a diagonal tile loads its one panel twice: no branch in the loop).
  // The next stage's loads are issued before this stage's MFMAs and land in registers while they run; the
  // clamp reloads the last stage instead of branching.
  static_assert(F64_KS == 16, "the staging below moves 8 segments of 16 bytes per row");
  const int row0 = tid >> 3, sg = tid & 7;
  const double* pa0 = X + (I0 + row0) * ldx + 2 * sg;
  const double* pa1 = pa0 + 64 * ldx;
  const double* pb0 = X + (J0 + row0) * ldx + 2 * sg;
  const double* pb1 = pb0 + 64 * ldx;
  const int o0 = row0 * F64_PITCH + 2 * sg, o1 = o0 + 64 * F64_PITCH;
  d4 acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  int64_t k = s_beg * F64_KS;
  double2 ra0 = *reinterpret_cast<const double2*>(pa0 + k), ra1 = *reinterpret_cast<const double2*>(pa1 + k);
  double2 rb0 = *reinterpret_cast<const double2*>(pb0 + k), rb1 = *reinterpret_cast<const double2*>(pb1 + k);
  *reinterpret_cast<double2*>(lds[0][0] + o0) = ra0;
  *reinterpret_cast<double2*>(lds[0][0] + o1) = ra1;
  *reinterpret_cast<double2*>(lds[0][1] + o0) = rb0;
  *reinterpret_cast<double2*>(lds[0][1] + o1) = rb1;
  __syncthreads();
  int buf = 0;
  for (int64_t s = s_beg; s < s_end; ++s) {
    // issue the next stage's loads first; they are written to the other buffer after this stage's MFMAs
    // (that buffer was last read before the previous barrier)
    k = std::min<int64_t>(s + 1, s_end - 1) * F64_KS;
    ra0 = *reinterpret_cast<const double2*>(pa0 + k);
    ra1 = *reinterpret_cast<const double2*>(pa1 + k);
    rb0 = *reinterpret_cast<const double2*>(pb0 + k);
    rb1 = *reinterpret_cast<const double2*>(pb1 + k);
    const double* La = lds[buf][0];
    const double* Lb = lds[buf][1];
#pragma unroll
    for (int kk = 0; kk < F64_KS / 4; ++kk) {
      double fa[2], fb[4];
#pragma unroll
      for (int a = 0; a < 2; ++a) fa[a] = La[(wr + a * 16 + lr) * F64_PITCH + kk * 4 + lk];
#pragma unroll
      for (int b = 0; b < 4; ++b) fb[b] = Lb[(wc + b * 16 + lr) * F64_PITCH + kk * 4 + lk];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
    buf ^= 1;
    *reinterpret_cast<double2*>(lds[buf][0] + o0) = ra0;
    *reinterpret_cast<double2*>(lds[buf][0] + o1) = ra1;
    *reinterpret_cast<double2*>(lds[buf][1] + o0) = rb0;
    *reinterpret_cast<double2*>(lds[buf][1] + o1) = rb1;
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int lrow = wr + a * 16 + lk + 4 * r, lcol = wc + b * 16 + lr;
        if (h == 0) {
          const int64_t row = I0 + lrow, col = J0 + lcol;
          if (row < M && col < M) {
            double* q = W + row * M + col;
            *q = first ? acc[a][b][r] : *q + acc[a][b][r];
          }
        } else {
          double* q = P + ((int64_t)(h - 1) * ntiles + t) * (F64_TILE * F64_TILE) + lrow * F64_TILE + lcol;
          *q = first ? acc[a][b][r] : *q + acc[a][b][r];
        }
      }
}

// W += the partial tiles of parts 1 .. F64_SPLIT - 1 (in part order), over the tiles' elements inside W.
__global__ void syrk_f64_partials_kernel(const double* __restrict__ P, int ntiles, int64_t M, double* __restrict__ W) {
  const int t = blockIdx.y;
  int64_t TI, TJ;
  tile_of(t, TI, TJ);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < F64_TILE * F64_TILE; e += gridDim.x * blockDim.x) {
    const int64_t row = TI * F64_TILE + e / F64_TILE, col = TJ * F64_TILE + e % F64_TILE;
    if (row >= M || col >= M) continue;
    double v = W[row * M + col];
    for (int h = 1; h < F64_SPLIT; ++h) v += P[((int64_t)(h - 1) * ntiles + t) * (F64_TILE * F64_TILE) + e];
    W[row * M + col] = v;
  }
}

// ---------------------------------------------------------------------------
// Integer-digit SYRK: W += sum_c scale_c^2/Q^2 * sum_{p,q} 256^(2NS-2-p-q) A_p A_q^T.
// Workgroup tile 64 rows x 128 columns (4 waves as 2x2, each 32x64 = 2x4
// MFMA 16x16 blocks); one int32 accumulator set per digit order p+q, exact
// within a 4096-visibility chunk, flushed in FP64 straight into W (zeroed by
// the caller).  LDS double-buffered by 64-feature stages with a register
// prefetch of the next stage.  Tiles: row blocks I of 64, column blocks
// J <= I/2 of 128 (the lower triangle), placed so that consecutive tiles,
// which share row panels, land on the same XCD.
// ---------------------------------------------------------------------------
constexpr int I8_TM = 64, I8_TN = 128, I8_KS = 64;
constexpr int I8_NLAYOUT = 9;  // 0-4: register prefetch with the LDS layouts below; 5: direct-to-LDS; 6: 32x32x32 MFMA; 7: 8 waves; 8: 8 waves, LDS ring

// LDS placement of 16-byte unit u (0..3) of tile row r within one 64-byte stage row.
//   0: rows padded to 80 bytes;  1: 64-byte rows, unit ^= (r/2)%4;  2: rows padded to 72 bytes;
//   3: 64-byte rows, unit ^= ((r/2) ^ (r/8))%4;  4: rows padded to 96 bytes.
template <int L>
struct I8Layout {
  static constexpr int pitch = L == 0 ? 80 : L == 2 ? 72 : L == 4 ? 96 : 64;
  __device__ static inline int off(int r, int u) {
    if constexpr (L == 1) return r * 64 + 16 * (u ^ ((r >> 1) & 3));
    else if constexpr (L == 3) return r * 64 + 16 * (u ^ (((r >> 1) ^ (r >> 3)) & 3));
    else return r * pitch + 16 * u;
  }
};

// Tile order: 512 x 512 super-tiles (8 row blocks x 4 column blocks), lower triangle, so that the
// 32 workgroups an XCD runs at a time share their row and column panels in its L2.
inline void build_i8_tiles(int64_t Mpad, std::vector<int>& out) {
  const int64_t nI = Mpad / I8_TM, nJ = Mpad / I8_TN;
  const int64_t nS = (nI + 7) / 8;
  out.clear();
  for (int64_t SI = 0; SI < nS; ++SI)
    for (int64_t SJ = 0; SJ <= SI; ++SJ)
      for (int64_t i = 0; i < 8; ++i)
        for (int64_t j = 0; j < 4; ++j) {
          const int64_t I = 8 * SI + i, J = 4 * SJ + j;
          if (I < nI && J < nJ && J <= I / 2) {
            out.push_back((int)I);
            out.push_back((int)J);
          }
        }
}

inline int64_t i8_tiles(int64_t Mpad) {
  const int64_t nI = Mpad / I8_TM;  // even (Mpad is a multiple of 128)
  return (nI / 2) * (nI / 2 + 1);
}

template <int NS, int NG, int LAYOUT>
__global__ void __launch_bounds__(256, 1) syrk_i8_kernel(const int8_t* __restrict__ X, int64_t M, int64_t Mpad,
                                                         int64_t ldx, int nchunks, const double* __restrict__ chunk_sc,
                                                         double* __restrict__ W, int64_t ntiles,
                                                         const int* __restrict__ tiles) {
  constexpr int NFULL = 2 * NS - 1;  // orders 0 .. 2NS-2
  constexpr int ROWS = I8_TM + I8_TN;
  using LY = I8Layout<LAYOUT>;
  __shared__ __attribute__((aligned(16))) int8_t lds[2][NS][ROWS * LY::pitch];

  // XCD-aware placement: block b runs on XCD b % 8; give each XCD a contiguous run of tiles.
  const int64_t per = (ntiles + 7) / 8;
  const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (t >= ntiles) return;
  const int64_t I0 = (int64_t)tiles[2 * t] * I8_TM, J0 = (int64_t)tiles[2 * t + 1] * I8_TN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t sstride = Mpad * ldx;

  // global -> register prefetch: thread loads 16 B of row (tid/4) [A], rows (tid/4) and 64+(tid/4) [B]
  const int grow = tid >> 2, gunit = tid & 3, gseg = gunit * 16;
  const int8_t* srcA = X + (I0 + grow) * ldx + gseg;
  const int8_t* srcB0 = X + (J0 + grow) * ldx + gseg;
  const int8_t* srcB1 = X + (J0 + 64 + grow) * ldx + gseg;
  // the next stage is prefetched in two halves of the digit slices, each overlapped with half of
  // the MFMA work of the current stage, so only half of the slices sit in registers at a time.
  // (All indices below are compile-time constants after unrolling, so pf stays in registers.)
  constexpr int H0 = (NS + 1) / 2;
  // named registers, up to three slices per half (no array: an array here is turned into LDS
  // storage by the compiler)
  static_assert(H0 <= 3, "prefetch holds at most three digit slices per half stage");
  int4 pA0, pB0, pC0, pA1, pB1, pC1, pA2, pB2, pC2;
#define AJ_I8_LD1(V0, V1, V2, SS, K0)                                              \
  if ((SS) < NS) {                                                                 \
    V0 = *reinterpret_cast<const int4*>(srcA + (int64_t)(SS) * sstride + (K0));    \
    V1 = *reinterpret_cast<const int4*>(srcB0 + (int64_t)(SS) * sstride + (K0));   \
    V2 = *reinterpret_cast<const int4*>(srcB1 + (int64_t)(SS) * sstride + (K0));   \
  }
#define AJ_I8_ST1(V0, V1, V2, SS, BUF)                                                             \
  if ((SS) < NS) {                                                                                 \
    *reinterpret_cast<int4*>(&lds[BUF][SS][LY::off(grow, gunit)]) = V0;                            \
    *reinterpret_cast<int4*>(&lds[BUF][SS][LY::off(I8_TM + grow, gunit)]) = V1;                    \
    *reinterpret_cast<int4*>(&lds[BUF][SS][LY::off(I8_TM + 64 + grow, gunit)]) = V2;               \
  }
#define AJ_I8_GLOAD(HALF, K0)                                 \
  {                                                           \
    AJ_I8_LD1(pA0, pB0, pC0, (HALF) * H0 + 0, K0)             \
    if (H0 > 1) { AJ_I8_LD1(pA1, pB1, pC1, (HALF) * H0 + 1, K0) } \
    if (H0 > 2) { AJ_I8_LD1(pA2, pB2, pC2, (HALF) * H0 + 2, K0) } \
  }
#define AJ_I8_LSTORE(HALF, BUF)                               \
  {                                                           \
    AJ_I8_ST1(pA0, pB0, pC0, (HALF) * H0 + 0, BUF)            \
    if (H0 > 1) { AJ_I8_ST1(pA1, pB1, pC1, (HALF) * H0 + 1, BUF) } \
    if (H0 > 2) { AJ_I8_ST1(pA2, pB2, pC2, (HALF) * H0 + 2, BUF) } \
  }

  i4 acc[NG][2][4];
  const int64_t nstage = (int64_t)nchunks * (FC / I8_KS);
  constexpr int SPC = FC / I8_KS;  // stages per chunk
  AJ_I8_GLOAD(0, 0)
  AJ_I8_LSTORE(0, 0)
  AJ_I8_GLOAD(1, 0)
  AJ_I8_LSTORE(1, 0)
  __syncthreads();
  for (int64_t st = 0; st < nstage; ++st) {
    if (st % SPC == 0) {
#pragma unroll
      for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[g][a][b] = i4{0, 0, 0, 0};
    }
    const int buf = (int)(st & 1);
    const bool more = st + 1 < nstage;
    const int64_t knext = (st + 1) * I8_KS;
    i4 fa[NS][2];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int a = 0; a < 2; ++a)
        fa[s][a] = *reinterpret_cast<const i4*>(&lds[buf][s][LY::off(wr + a * 16 + lr, lk)]);
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      if (more) { AJ_I8_GLOAD(half, knext) }
      // keep the order: loads issued, MFMA work, then the LDS stores that wait for the loads
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int b = 2 * half; b < 2 * half + 2; ++b) {
        i4 fb[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s)
          fb[s] = *reinterpret_cast<const i4*>(&lds[buf][s][LY::off(I8_TM + wc + b * 16 + lr, lk)]);
#pragma unroll
        for (int p = 0; p < NS; ++p)
#pragma unroll
          for (int q = 0; q < NS; ++q)
            if (p + q < NG)
#pragma unroll
              for (int a = 0; a < 2; ++a)
                acc[p + q][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[p][a], fb[q], acc[p + q][a][b], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
      if (more) {
        if (buf) { AJ_I8_LSTORE(half, 0) } else { AJ_I8_LSTORE(half, 1) }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#undef AJ_I8_GLOAD
#undef AJ_I8_LSTORE
#undef AJ_I8_LD1
#undef AJ_I8_ST1
    if (st % SPC == SPC - 1) {
      const double q = qmax_of(NS);
      const int c = (int)(st / SPC);
      const double s2 = chunk_sc[c] * chunk_sc[c] / (q * q);
      // an opaque copy of M keeps the 32 output addresses from being hoisted out of the
      // stage loop (where they would occupy registers the prefetch needs)
      int64_t Mo = M;
      asm volatile("" : "+s"(Mo));
      double* Wt = W + (I0 + wr + lk * 4) * Mo + (J0 + wc + lr);
      const int64_t rlim = M - (I0 + wr + lk * 4), clim = M - (J0 + wc + lr);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            double v = 0.0;
#pragma unroll
            for (int g = 0; g < NG; ++g) v = v * 256.0 + (double)acc[g][a][b][r];
#pragma unroll
            for (int g = NG; g < NFULL; ++g) v *= 256.0;
            if (a * 16 + r < rlim && b * 16 < clim) Wt[(a * 16 + r) * Mo + b * 16] += v * s2;
            // one output at a time: keeps the scheduler from reading all accumulators at once
            __builtin_amdgcn_sched_barrier(0);
          }
    }
    __syncthreads();
  }
}

template <int NS, int NG>
__global__ void __launch_bounds__(256, 1) syrk_i8_k32_kernel(const int8_t* __restrict__ X, int64_t M, int64_t Mpad,
                                                         int64_t ldx, int nchunks, const double* __restrict__ chunk_sc,
                                                         double* __restrict__ W, int64_t ntiles,
                                                         const int* __restrict__ tiles) {
  constexpr int NFULL = 2 * NS - 1;  // orders 0 .. 2NS-2
  constexpr int ROWS = I8_TM + I8_TN;
  using LY = I8Layout<1>;
  __shared__ __attribute__((aligned(16))) int8_t lds[2][NS][ROWS * LY::pitch];

  // XCD-aware placement: block b runs on XCD b % 8; give each XCD a contiguous run of tiles.
  const int64_t per = (ntiles + 7) / 8;
  const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (t >= ntiles) return;
  const int64_t I0 = (int64_t)tiles[2 * t] * I8_TM, J0 = (int64_t)tiles[2 * t + 1] * I8_TN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t sstride = Mpad * ldx;

  // global -> register prefetch: thread loads 16 B of row (tid/4) [A], rows (tid/4) and 64+(tid/4) [B]
  const int grow = tid >> 2, gunit = tid & 3, gseg = gunit * 16;
  const int8_t* srcA = X + (I0 + grow) * ldx + gseg;
  const int8_t* srcB0 = X + (J0 + grow) * ldx + gseg;
  const int8_t* srcB1 = X + (J0 + 64 + grow) * ldx + gseg;
  // the next stage is prefetched in two halves of the digit slices, each overlapped with half of
  // the MFMA work of the current stage, so only half of the slices sit in registers at a time.
  // (All indices below are compile-time constants after unrolling, so pf stays in registers.)
  constexpr int H0 = (NS + 1) / 2;
  // named registers, up to three slices per half (no array: an array here is turned into LDS
  // storage by the compiler)
  static_assert(H0 <= 3, "prefetch holds at most three digit slices per half stage");
  int4 pA0, pB0, pC0, pA1, pB1, pC1, pA2, pB2, pC2;
#define AJ_I8_LD1(V0, V1, V2, SS, K0)                                              \
  if ((SS) < NS) {                                                                 \
    V0 = *reinterpret_cast<const int4*>(srcA + (int64_t)(SS) * sstride + (K0));    \
    V1 = *reinterpret_cast<const int4*>(srcB0 + (int64_t)(SS) * sstride + (K0));   \
    V2 = *reinterpret_cast<const int4*>(srcB1 + (int64_t)(SS) * sstride + (K0));   \
  }
#define AJ_I8_ST1(V0, V1, V2, SS, BUF)                                                             \
  if ((SS) < NS) {                                                                                 \
    *reinterpret_cast<int4*>(&lds[BUF][SS][LY::off(grow, gunit)]) = V0;                            \
    *reinterpret_cast<int4*>(&lds[BUF][SS][LY::off(I8_TM + grow, gunit)]) = V1;                    \
    *reinterpret_cast<int4*>(&lds[BUF][SS][LY::off(I8_TM + 64 + grow, gunit)]) = V2;               \
  }
#define AJ_I8_GLOAD(HALF, K0)                                 \
  {                                                           \
    AJ_I8_LD1(pA0, pB0, pC0, (HALF) * H0 + 0, K0)             \
    if (H0 > 1) { AJ_I8_LD1(pA1, pB1, pC1, (HALF) * H0 + 1, K0) } \
    if (H0 > 2) { AJ_I8_LD1(pA2, pB2, pC2, (HALF) * H0 + 2, K0) } \
  }
#define AJ_I8_LSTORE(HALF, BUF)                               \
  {                                                           \
    AJ_I8_ST1(pA0, pB0, pC0, (HALF) * H0 + 0, BUF)            \
    if (H0 > 1) { AJ_I8_ST1(pA1, pB1, pC1, (HALF) * H0 + 1, BUF) } \
    if (H0 > 2) { AJ_I8_ST1(pA2, pB2, pC2, (HALF) * H0 + 2, BUF) } \
  }

  i16v acc[NG][2];
  const int64_t nstage = (int64_t)nchunks * (FC / I8_KS);
  constexpr int SPC = FC / I8_KS;  // stages per chunk
  AJ_I8_GLOAD(0, 0)
  AJ_I8_LSTORE(0, 0)
  AJ_I8_GLOAD(1, 0)
  AJ_I8_LSTORE(1, 0)
  __syncthreads();
  for (int64_t st = 0; st < nstage; ++st) {
    if (st % SPC == 0) {
#pragma unroll
      for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[g][b] = i16v{};
    }
    const int buf = (int)(st & 1);
    const bool more = st + 1 < nstage;
    const int64_t knext = (st + 1) * I8_KS;
    // v_mfma_i32_32x32x32_i8: lane l holds A[row l%32][k 16*(l/32) .. +16) and B likewise; the
    // wave's 32x64 block is two 32x32 blocks; a 64-byte stage row is two k-steps of 32.
    const int lr32 = lane & 31, lh = lane >> 5;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      if (more) { AJ_I8_GLOAD(half, knext) }
      __builtin_amdgcn_sched_barrier(0);
      const int ks = half;  // k-step of this half
      i4 fa[NS], fb[2][NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        fa[s] = *reinterpret_cast<const i4*>(&lds[buf][s][LY::off(wr + lr32, 2 * ks + lh)]);
#pragma unroll
        for (int b = 0; b < 2; ++b)
          fb[b][s] = *reinterpret_cast<const i4*>(&lds[buf][s][LY::off(I8_TM + wc + 32 * b + lr32, 2 * ks + lh)]);
      }
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int p = 0; p < NS; ++p)
#pragma unroll
          for (int q = 0; q < NS; ++q)
            if (p + q < NG) acc[p + q][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[p], fb[b][q], acc[p + q][b], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if (more) {
        if (buf) { AJ_I8_LSTORE(half, 0) } else { AJ_I8_LSTORE(half, 1) }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#undef AJ_I8_GLOAD
#undef AJ_I8_LSTORE
#undef AJ_I8_LD1
#undef AJ_I8_ST1
    if (st % SPC == SPC - 1) {
      const double q = qmax_of(NS);
      const int c = (int)(st / SPC);
      const double s2 = chunk_sc[c] * chunk_sc[c] / (q * q);
      // an opaque copy of M keeps the 32 output addresses from being hoisted out of the
      // stage loop (where they would occupy registers the prefetch needs)
      int64_t Mo = M;
      asm volatile("" : "+s"(Mo));
      // accumulator register r of lane l holds row 8(r/4) + 4(l/32) + r%4, column l%32
      double* Wt = W + (I0 + wr + 4 * lh) * Mo + (J0 + wc + lr32);
      const int64_t rlim = M - (I0 + wr + 4 * lh), clim = M - (J0 + wc + lr32);
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int ro = 8 * (r >> 2) + (r & 3);
          double v = 0.0;
#pragma unroll
          for (int g = 0; g < NG; ++g) v = v * 256.0 + (double)acc[g][b][r];
#pragma unroll
          for (int g = NG; g < NFULL; ++g) v *= 256.0;
          if (ro < rlim && 32 * b < clim) Wt[ro * Mo + 32 * b] += v * s2;
          __builtin_amdgcn_sched_barrier(0);
        }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Integer SYRK, two waves per SIMD: the same 64x128 workgroup tile split over
// 8 waves of 32x32 (2x2 MFMA 16x16 blocks), so each wave holds half the
// accumulators of syrk_i8_kernel and a second wave on the SIMD covers the
// LDS-read latency of the first.  Register prefetch of the next stage in two
// halves, LDS rows of 64 bytes with 16-byte units XOR-swizzled by (row/2)%4.
// ---------------------------------------------------------------------------
template <int NS, int NG>
__global__ void __launch_bounds__(512, 1) syrk_i8_w8_kernel(const int8_t* __restrict__ X, int64_t M, int64_t Mpad,
                                                            int64_t ldx, int nchunks,
                                                            const double* __restrict__ chunk_sc,
                                                            double* __restrict__ W, int64_t ntiles,
                                                            const int* __restrict__ tiles) {
  constexpr int NFULL = 2 * NS - 1;
  constexpr int ROWS = I8_TM + I8_TN;
  using LY = I8Layout<1>;
  __shared__ __attribute__((aligned(16))) int8_t lds[2][NS][ROWS * LY::pitch];

  const int64_t per = (ntiles + 7) / 8;
  const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (t >= ntiles) return;
  const int64_t I0 = (int64_t)tiles[2 * t] * I8_TM, J0 = (int64_t)tiles[2 * t + 1] * I8_TN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 2) * 32, wc = (wave & 3) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t sstride = Mpad * ldx;

  // thread t moves 16 bytes of B row t/4 and, for t < 256, of A row t/4
  const int grow = tid >> 2, gunit = tid & 3, gseg = gunit * 16;
  const bool hasA = tid < 256;
  const int8_t* srcA = X + (I0 + (grow & 63)) * ldx + gseg;
  const int8_t* srcB = X + (J0 + grow) * ldx + gseg;
  constexpr int H0 = (NS + 1) / 2;
  static_assert(H0 <= 3, "prefetch holds at most three digit slices per half stage");
  int4 pa0, pb0, pa1, pb1, pa2, pb2;
#define AJ_W8_LD1(VA, VB, SS, K0)                                                   \
  if ((SS) < NS) {                                                                  \
    if (hasA) VA = *reinterpret_cast<const int4*>(srcA + (int64_t)(SS) * sstride + (K0)); \
    VB = *reinterpret_cast<const int4*>(srcB + (int64_t)(SS) * sstride + (K0));    \
  }
#define AJ_W8_ST1(VA, VB, SS, BUF)                                                          \
  if ((SS) < NS) {                                                                          \
    if (hasA) *reinterpret_cast<int4*>(&lds[BUF][SS][LY::off(grow, gunit)]) = VA;           \
    *reinterpret_cast<int4*>(&lds[BUF][SS][LY::off(I8_TM + grow, gunit)]) = VB;             \
  }
#define AJ_W8_GLOAD(HALF, K0)                                       \
  {                                                                 \
    AJ_W8_LD1(pa0, pb0, (HALF) * H0 + 0, K0)                        \
    if (H0 > 1) { AJ_W8_LD1(pa1, pb1, (HALF) * H0 + 1, K0) }        \
    if (H0 > 2) { AJ_W8_LD1(pa2, pb2, (HALF) * H0 + 2, K0) }        \
  }
#define AJ_W8_LSTORE(HALF, BUF)                                     \
  {                                                                 \
    AJ_W8_ST1(pa0, pb0, (HALF) * H0 + 0, BUF)                       \
    if (H0 > 1) { AJ_W8_ST1(pa1, pb1, (HALF) * H0 + 1, BUF) }       \
    if (H0 > 2) { AJ_W8_ST1(pa2, pb2, (HALF) * H0 + 2, BUF) }       \
  }

  i4 acc[NG][2][2];
  const int64_t nstage = (int64_t)nchunks * (FC / I8_KS);
  constexpr int SPC = FC / I8_KS;
  AJ_W8_GLOAD(0, 0)
  AJ_W8_LSTORE(0, 0)
  AJ_W8_GLOAD(1, 0)
  AJ_W8_LSTORE(1, 0)
  __syncthreads();
  for (int64_t st = 0; st < nstage; ++st) {
    if (st % SPC == 0) {
#pragma unroll
      for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) acc[g][a][b] = i4{0, 0, 0, 0};
    }
    const int buf = (int)(st & 1);
    const bool more = st + 1 < nstage;
    const int64_t knext = (st + 1) * I8_KS;
    i4 fa[NS][2];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int a = 0; a < 2; ++a)
        fa[s][a] = *reinterpret_cast<const i4*>(&lds[buf][s][LY::off(wr + a * 16 + lr, lk)]);
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      if (more) { AJ_W8_GLOAD(half, knext) }
      __builtin_amdgcn_sched_barrier(0);
      const int b = half;
      i4 fb[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s)
        fb[s] = *reinterpret_cast<const i4*>(&lds[buf][s][LY::off(I8_TM + wc + b * 16 + lr, lk)]);
#pragma unroll
      for (int p = 0; p < NS; ++p)
#pragma unroll
        for (int q = 0; q < NS; ++q)
          if (p + q < NG)
#pragma unroll
            for (int a = 0; a < 2; ++a)
              acc[p + q][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[p][a], fb[q], acc[p + q][a][b], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if (more) {
        if (buf) { AJ_W8_LSTORE(half, 0) } else { AJ_W8_LSTORE(half, 1) }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    if (st % SPC == SPC - 1) {
      const double qm = qmax_of(NS);
      const int c = (int)(st / SPC);
      const double s2 = chunk_sc[c] * chunk_sc[c] / (qm * qm);
      int64_t Mo = M;
      asm volatile("" : "+s"(Mo));
      double* Wt = W + (I0 + wr + lk * 4) * Mo + (J0 + wc + lr);
      const int64_t rlim = M - (I0 + wr + lk * 4), clim = M - (J0 + wc + lr);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            double v = 0.0;
#pragma unroll
            for (int g = 0; g < NG; ++g) v = v * 256.0 + (double)acc[g][a][b][r];
#pragma unroll
            for (int g = NG; g < NFULL; ++g) v *= 256.0;
            if (a * 16 + r < rlim && b * 16 < clim) Wt[(a * 16 + r) * Mo + b * 16] += v * s2;
            __builtin_amdgcn_sched_barrier(0);
          }
    }
    __syncthreads();
  }
#undef AJ_W8_GLOAD
#undef AJ_W8_LSTORE
#undef AJ_W8_LD1
#undef AJ_W8_ST1
}

// ---------------------------------------------------------------------------
// Integer SYRK, pipelined variant: the same tiles and arithmetic as
// syrk_i8_kernel, with the feature tiles moved by direct global->LDS loads
// (global_load_lds, 16 bytes per lane, lane i of a wave writes base + 16 i)
// into two LDS stages, the next stage issued under the MFMA work of the
// current one, so no registers hold prefetched data.  64-byte LDS rows with the 16-byte
// units XOR-swizzled by (row/2)%4 (free of bank conflicts for the fragment
// reads); the swizzle is applied by choosing each lane's source address.
// ---------------------------------------------------------------------------
__device__ inline void dma16(const int8_t* g, int8_t* lds_wave_base) {
  __builtin_amdgcn_global_load_lds((const void*)g,
                                   (void __attribute__((address_space(3)))*)(
                                       (__attribute__((address_space(3))) int8_t*)lds_wave_base),
                                   16, 0, 0);
}

template <int NS, int NG>
__global__ void __launch_bounds__(256, 1) syrk_i8_dma_kernel(const int8_t* __restrict__ X, int64_t M, int64_t Mpad,
                                                             int64_t ldx, int nchunks,
                                                             const double* __restrict__ chunk_sc,
                                                             double* __restrict__ W, int64_t ntiles,
                                                             const int* __restrict__ tiles) {
  constexpr int NFULL = 2 * NS - 1;  // orders 0 .. 2NS-2
  constexpr int SLICE = (I8_TM + I8_TN) * 64;  // bytes of one slice of one stage
  __shared__ __attribute__((aligned(1024))) int8_t L0[NS * SLICE];
  __shared__ __attribute__((aligned(1024))) int8_t L1[NS * SLICE];

  const int64_t per = (ntiles + 7) / 8;
  const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (t >= ntiles) return;
  const int64_t I0 = (int64_t)tiles[2 * t] * I8_TM, J0 = (int64_t)tiles[2 * t + 1] * I8_TN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t sstride = Mpad * ldx;

  // DMA assignment: the 192 stage rows form 12 blocks of 16 rows (1 KB per slice); wave w moves
  // blocks w, w+4, w+8.  Lane L fills LDS slot L of its block: row (L/4), stored unit L%4, which
  // holds global unit (L%4) ^ ((row/2)%4).
  const int8_t* src[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int rb = wave + 4 * j;
    const int row = rb * 16 + (lane >> 2);
    const int64_t grow = rb < 4 ? I0 + row : J0 + (row - I8_TM);
    const int u = (lane & 3) ^ ((row >> 1) & 3);
    src[j] = X + grow * ldx + 16 * u;
  }
#define AJ_DMA_ISSUE(BUF, K0)                                                                  \
  {                                                                                           \
    const int64_t dk_ = (K0); /* evaluated once, outside the loops below */                   \
    _Pragma("unroll") for (int dsl_ = 0; dsl_ < NS; ++dsl_)                                   \
    _Pragma("unroll") for (int dj_ = 0; dj_ < 3; ++dj_)                                       \
      dma16(src[dj_] + dsl_ * sstride + dk_, BUF + dsl_ * SLICE + (wave + 4 * dj_) * 1024);   \
  }

  i4 acc[NG][2][4];
  const int64_t nstage = (int64_t)nchunks * (FC / I8_KS);
  constexpr int SPC = FC / I8_KS;

  AJ_DMA_ISSUE(L0, 0)
  __builtin_amdgcn_s_waitcnt(0);  // all counters: own DMA landed in LDS

  // Stage S from CUR; the DMA of stage S+1 into NXT is issued after all fragment reads of CUR and
  // runs under the MFMA work.  Each wave waits for its own DMA before the next barrier, so after the
  // barrier every wave's part of NXT is in LDS.
#define AJ_DMA_STEP(CUR, NXT, S)                                                                           \
  if ((S) < nstage) {                                                                                      \
    const int64_t s_ = (S);                                                                                \
    /* nothing is in flight here (each wave waited for its own DMA), so the fence costs nothing */         \
    __syncthreads();                                                                                       \
    if (s_ % SPC == 0) {                                                                                   \
      _Pragma("unroll") for (int g = 0; g < NG; ++g)                                                       \
      _Pragma("unroll") for (int a = 0; a < 2; ++a)                                                        \
      _Pragma("unroll") for (int b = 0; b < 4; ++b) acc[g][a][b] = i4{0, 0, 0, 0};                         \
    }                                                                                                      \
    i4 fa[NS][2], fb[4][NS];                                                                               \
    _Pragma("unroll") for (int s = 0; s < NS; ++s)                                                         \
    _Pragma("unroll") for (int a = 0; a < 2; ++a) {                                                        \
      const int r = wr + a * 16 + lr;                                                                      \
      fa[s][a] = *reinterpret_cast<const i4*>(&CUR[s * SLICE + r * 64 + 16 * (lk ^ ((r >> 1) & 3))]);     \
    }                                                                                                      \
    _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                          \
    _Pragma("unroll") for (int s = 0; s < NS; ++s) {                                                       \
      const int r = I8_TM + wc + b * 16 + lr;                                                              \
      fb[b][s] = *reinterpret_cast<const i4*>(&CUR[s * SLICE + r * 64 + 16 * (lk ^ ((r >> 1) & 3))]);     \
    }                                                                                                      \
    __builtin_amdgcn_sched_barrier(0);                                                                     \
    if (s_ + 1 < nstage) { AJ_DMA_ISSUE(NXT, (s_ + 1) * I8_KS) }                                           \
    __builtin_amdgcn_sched_barrier(0);                                                                     \
    _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                          \
    _Pragma("unroll") for (int p = 0; p < NS; ++p)                                                         \
    _Pragma("unroll") for (int q = 0; q < NS; ++q)                                                         \
    if (p + q < NG)                                                                                        \
    _Pragma("unroll") for (int a = 0; a < 2; ++a)                                                          \
      acc[p + q][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[p][a], fb[b][q], acc[p + q][a][b], 0, 0, 0); \
    if (s_ % SPC == SPC - 1) {                                                                             \
      const double qm = qmax_of(NS);                                                                       \
      const int c = (int)(s_ / SPC);                                                                       \
      const double s2 = chunk_sc[c] * chunk_sc[c] / (qm * qm);                                             \
      int64_t Mo = M;                                                                                      \
      asm volatile("" : "+s"(Mo));                                                                         \
      double* Wt = W + (I0 + wr + lk * 4) * Mo + (J0 + wc + lr);                                           \
      const int64_t rlim = M - (I0 + wr + lk * 4), clim = M - (J0 + wc + lr);                              \
      _Pragma("unroll") for (int a = 0; a < 2; ++a)                                                        \
      _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                        \
      _Pragma("unroll") for (int r = 0; r < 4; ++r) {                                                      \
        double v = 0.0;                                                                                    \
        _Pragma("unroll") for (int g = 0; g < NG; ++g) v = v * 256.0 + (double)acc[g][a][b][r];            \
        _Pragma("unroll") for (int g = NG; g < NFULL; ++g) v *= 256.0;                                     \
        if (a * 16 + r < rlim && b * 16 < clim) Wt[(a * 16 + r) * Mo + b * 16] += v * s2;                  \
        __builtin_amdgcn_sched_barrier(0);                                                                 \
      }                                                                                                    \
    }                                                                                                      \
    __builtin_amdgcn_sched_barrier(0);                                                                     \
    __builtin_amdgcn_s_waitcnt(0); /* all counters: own DMA of stage S+1 landed in LDS */                   \
  }

  for (int64_t st = 0; st < nstage; st += 2) {
    AJ_DMA_STEP(L0, L1, st)
    AJ_DMA_STEP(L1, L0, st + 1)
  }
#undef AJ_DMA_STEP
#undef AJ_DMA_ISSUE
}

// ---------------------------------------------------------------------------
// Integer SYRK, software-pipelined: the 64x128 workgroup tile as 8 waves of
// 32x32 (one v_mfma_i32_32x32x32_i8 block each, two waves per SIMD), K in
// stages of 32 bytes, a ring of four LDS stages filled by direct global->LDS
// loads (global_load_lds, 16 bytes per lane), and two register sets of
// fragments: while the MFMAs of stage S run on one set, the LDS reads of
// stage S+1 fill the other, so the LDS read burst that follows each barrier
// overlaps MFMA work instead of idling the matrix units.  Stage S+3 is loaded
// during stage S into the buffer of stage S-1, whose reads completed before
// the MFMAs of stage S-1.  Each LDS stage holds, per digit slice, 192 rows x
// 32 bytes as 16-byte units at position 2 r + (u ^ ((r >> 3) & 1)); the
// swizzle is applied through each lane's source address.  The four buffers
// are distinct LDS objects so that the compiler can tell the loads in flight
// from the fragment reads.
// ---------------------------------------------------------------------------
constexpr int RG_KS = 32;                          // bytes of K per stage
constexpr int RG_UNITS = (I8_TM + I8_TN) * 2;      // 16-byte units per slice per stage
constexpr int RG_BLOCKS = RG_UNITS / 64;           // 1 KiB blocks per slice per stage (6)
constexpr int RG_SLICE = RG_UNITS * 16;            // bytes per slice per stage

__device__ inline int rg_pos(int r, int u) { return 2 * r + (u ^ ((r >> 3) & 1)); }

template <int NS, int NG>
__global__ void __launch_bounds__(512, 1) syrk_i8_ring_kernel(const int8_t* __restrict__ X, int64_t M, int64_t Mpad,
                                                              int64_t ldx, int nchunks,
                                                              const double* __restrict__ chunk_sc,
                                                              double* __restrict__ W, int64_t ntiles,
                                                              const int* __restrict__ tiles) {
  constexpr int NFULL = 2 * NS - 1;
  constexpr int NBLK = RG_BLOCKS * NS;            // 1 KiB loads per stage
  constexpr int BPW = (NBLK + 7) / 8;             // loads per wave per stage (padded to equal counts)
  constexpr int SPC = FC / RG_KS;                 // stages per chunk
  static_assert(SPC % 4 == 0, "the stage loop is unrolled by the four ring buffers");
  __shared__ __attribute__((aligned(1024))) int8_t R0[NS * RG_SLICE];
  __shared__ __attribute__((aligned(1024))) int8_t R1[NS * RG_SLICE];
  __shared__ __attribute__((aligned(1024))) int8_t R2[NS * RG_SLICE];
  __shared__ __attribute__((aligned(1024))) int8_t R3[NS * RG_SLICE];
  __shared__ __attribute__((aligned(1024))) int8_t RD[1024];  // target of the padding loads, never read

  const int64_t per = (ntiles + 7) / 8;
  const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (t >= ntiles) return;
  const int64_t I0 = (int64_t)tiles[2 * t] * I8_TM, J0 = (int64_t)tiles[2 * t + 1] * I8_TN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 2) * 32, wc = (wave & 3) * 32;
  const int lr32 = lane & 31, lh = lane >> 5;
  const int64_t sstride = Mpad * ldx;
  const int offa = 16 * rg_pos(wr + lr32, lh), offb = 16 * rg_pos(I8_TM + wc + lr32, lh);

  // Load q of wave w moves block b = w + 8q (slice b / 6, 1 KiB block b % 6 of the slice); lane i
  // fills LDS unit 64 (b % 6) + i of the slice, i.e. row r, unit u with rg_pos(r, u) = that unit.
  // A wave with fewer real blocks repeats its first source into RD, so every wave issues BPW loads.
  const int8_t* src[BPW];
  int ldo[BPW];
  bool real[BPW];
#pragma unroll
  for (int q = 0; q < BPW; ++q) {
    const int b = wave + 8 * q;
    real[q] = b < NBLK;
    const int bb = real[q] ? b : wave;
    const int sl = bb / RG_BLOCKS, pos = (bb % RG_BLOCKS) * 64 + lane;
    const int r = pos >> 1;
    const int u = (pos & 1) ^ ((r >> 3) & 1);
    const int64_t grow = r < I8_TM ? I0 + r : J0 + (r - I8_TM);
    src[q] = X + (int64_t)sl * sstride + grow * ldx + 16 * u;
    ldo[q] = sl * RG_SLICE + (bb % RG_BLOCKS) * 1024;
  }
  const int64_t nstage = (int64_t)nchunks * SPC;
  const int64_t klast = (nstage - 1) * RG_KS;

#define AJ_RG_ISSUE(BUF, KOFF)                                                                            \
  {                                                                                                       \
    const int64_t ko_ = (KOFF) < klast ? (KOFF) : klast; /* past the end: a harmless in-range load */     \
    _Pragma("unroll") for (int q_ = 0; q_ < BPW; ++q_) {                                                  \
      int8_t* dst_ = real[q_] ? (BUF + ldo[q_]) : RD;                                                     \
      __builtin_amdgcn_global_load_lds((const void*)(src[q_] + ko_),                                      \
                                       (void __attribute__((address_space(3)))*)(                         \
                                           (__attribute__((address_space(3))) int8_t*)dst_),              \
                                       16, 0, 0);                                                         \
    }                                                                                                     \
  }
#define AJ_RG_READ(BUF, FA, FB)                                                                           \
  {                                                                                                       \
    _Pragma("unroll") for (int s = 0; s < NS; ++s) {                                                      \
      FA[s] = *reinterpret_cast<const i4*>(&BUF[s * RG_SLICE + offa]);                                    \
      FB[s] = *reinterpret_cast<const i4*>(&BUF[s * RG_SLICE + offb]);                                    \
    }                                                                                                     \
  }
  // Barrier without a memory fence (a fence would also wait for the loads still in flight).  Ordering
  // comes from the explicit waits: a wave's loads of the stage about to be read have landed (vmcnt)
  // before it arrives, and its reads of the buffer refilled after the barrier completed (lgkmcnt
  // before its MFMAs) a stage earlier; the empty asm statements keep the compiler from moving
  // memory accesses across the barrier.
#define AJ_RG_BARRIER()                \
  {                                    \
    asm volatile("" ::: "memory");     \
    __builtin_amdgcn_s_barrier();      \
    asm volatile("" ::: "memory");     \
  }

  i16v acc[NG];
  i4 fa0[NS], fb0[NS], fa1[NS], fb1[NS];
  // prologue: loads of stages 0, 1, 2 into R0, R1, R2; fragments of stage 0 into set 0
  AJ_RG_ISSUE(R0, 0)
  AJ_RG_ISSUE(R1, RG_KS)
  AJ_RG_ISSUE(R2, 2 * RG_KS)
  __builtin_amdgcn_s_waitcnt(aj_vmcnt_imm(2 * BPW));
  AJ_RG_BARRIER()
  AJ_RG_READ(R0, fa0, fb0)

  // Step S: own loads of stage S+1 landed (only stage S+2 may stay in flight), barrier, load stage
  // S+3 into DMA_BUF (the buffer of stage S-1), read stage S+1 from NXT_BUF into set NXT, MFMA on
  // set CUR (stage S).
#define AJ_RG_STEP(S, NXT_BUF, DMA_BUF, FAC, FBC, FAN, FBN)                                             \
  {                                                                                                    \
    const int64_t s_ = (S);                                                                            \
    __builtin_amdgcn_s_waitcnt(aj_vmcnt_imm(BPW));                                                     \
    AJ_RG_BARRIER()                                                                                    \
    /* set CUR was read a stage ago: wait for it now, so that only the reads of set NXT are in    */  \
    /* flight at the MFMAs (two sets at once exceed what the LDS counter can track)                */  \
    __builtin_amdgcn_s_waitcnt(aj_lgkmcnt0_imm());                                                     \
    AJ_RG_ISSUE(DMA_BUF, (s_ + 3) * RG_KS)                                                             \
    AJ_RG_READ(NXT_BUF, FAN, FBN)                                                                      \
    __builtin_amdgcn_sched_barrier(0);                                                                 \
    if (s_ % SPC == 0) {                                                                               \
      _Pragma("unroll") for (int g = 0; g < NG; ++g) acc[g] = i16v{};                                  \
    }                                                                                                  \
    _Pragma("unroll") for (int p = 0; p < NS; ++p)                                                     \
    _Pragma("unroll") for (int q = 0; q < NS; ++q)                                                     \
      if (p + q < NG) acc[p + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(FAC[p], FBC[q], acc[p + q], 0, 0, 0); \
    __builtin_amdgcn_sched_barrier(0);                                                                 \
  }

  for (int64_t st = 0; st < nstage; st += 4) {
    AJ_RG_STEP(st, R1, R3, fa0, fb0, fa1, fb1)
    AJ_RG_STEP(st + 1, R2, R0, fa1, fb1, fa0, fb0)
    AJ_RG_STEP(st + 2, R3, R1, fa0, fb0, fa1, fb1)
    AJ_RG_STEP(st + 3, R0, R2, fa1, fb1, fa0, fb0)
    if ((st + 3) % SPC == SPC - 1) {  // end of a chunk: flush the exact digit-order sums into W
      const double qm = qmax_of(NS);
      const int c = (int)((st + 3) / SPC);
      const double s2 = chunk_sc[c] * chunk_sc[c] / (qm * qm);
      int64_t Mo = M;
      asm volatile("" : "+s"(Mo));
      // accumulator register r of lane l holds row 8(r/4) + 4(l/32) + r%4, column l%32
      double* Wt = W + (I0 + wr + 4 * lh) * Mo + (J0 + wc + lr32);
      const int64_t rlim = M - (I0 + wr + 4 * lh), clim = M - (J0 + wc + lr32);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int ro = 8 * (r >> 2) + (r & 3);
        double v = 0.0;
#pragma unroll
        for (int g = 0; g < NG; ++g) v = v * 256.0 + (double)acc[g][r];
#pragma unroll
        for (int g = NG; g < NFULL; ++g) v *= 256.0;
        if (ro < rlim && 0 < clim) Wt[ro * Mo] += v * s2;
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  // the loads issued past the last stage must land before the workgroup's LDS is released
  __builtin_amdgcn_s_waitcnt(0);
#undef AJ_RG_STEP
#undef AJ_RG_BARRIER
#undef AJ_RG_READ
#undef AJ_RG_ISSUE
}

// Copy the lower triangle to the upper one (32x32 tiles through LDS).
__global__ void mirror_kernel(double* __restrict__ W, int64_t M) {
  __shared__ double t[32][33];
  const int64_t bi = blockIdx.y, bj = blockIdx.x;
  if (bj > bi) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int yy = ty; yy < 32; yy += 8) {
    int64_t r = bi * 32 + yy, c = bj * 32 + tx;
    t[yy][tx] = (r < M && c < M) ? W[r * M + c] : 0.0;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    int64_t r = bj * 32 + yy, c = bi * 32 + tx;  // upper element (r < c)
    if (r < M && c < M && c > r) W[r * M + c] = t[tx][yy];
  }
}

// LDS layout variant of the integer SYRK (aj_w_tilde_set_variant; chosen by the autotuner).
// Integer SYRK kernel variant; -1 selects the default per digit count: the 8-wave kernel (7) for 5
// digits, whose 4-wave forms need more than 512 registers per lane, else variant 0.
static std::atomic<int> g_i8_layout{-1};  // process-wide tuning, read once per build
constexpr int default_i8_variant(int ns) { return ns == 5 ? 7 : 0; }

struct Plan {
  int mode, ns, es;
  int64_t Mpad, ldx;
  int group;  // chunks per launch
  size_t feat_bytes, bytes;
};

// Upper bound on the feature matrix of one SYRK launch (bytes).  Every tile row and column streams
// the launch's features, so this is the re-read working set: at 256 MiB it stays within the 256 MB
// memory-side (infinity) cache and the re-reads do not go to HBM (1 GiB groups were 11% slower for
// the 4-digit integer SYRK at M = 5000).
constexpr size_t GROUP_BYTES_DEFAULT = size_t(256) << 20;
static std::atomic<size_t> g_group_bytes{GROUP_BYTES_DEFAULT};

inline Plan make_plan(int64_t K, int64_t M, int mode) {
  Plan p{};
  p.mode = mode;
  p.ns = slices_of(mode);
  p.es = esize_of(mode);
  const int64_t pad = (is_i8(mode) || mode == AJ_PREC_FP64) ? 2 * TILE : TILE;
  p.Mpad = ceil_div(M, pad) * pad;
  const int64_t nch = std::max<int64_t>(1, ceil_div(K, KC));
  const double per_chunk = (double)p.ns * p.Mpad * FC * p.es;
  int64_t g = (int64_t)((double)g_group_bytes.load(std::memory_order_relaxed) / per_chunk);
  g = std::max<int64_t>(1, std::min<int64_t>(g, nch));
  p.group = (int)g;
  p.ldx = (int64_t)p.group * FC;
  p.feat_bytes = (size_t)p.ns * p.Mpad * p.ldx * p.es;
  p.bytes = align_up(p.feat_bytes, 256) + 256 /*centre*/ + align_up(sizeof(double) * p.group, 256) + 256;
  p.bytes += align_up(sizeof(int) * 2 * (size_t)i8_tiles(ceil_div(M, 2 * TILE) * 2 * TILE), 256) + 256;
  if (mode == AJ_PREC_FP64) {
    const int64_t nt = p.Mpad / F64_TILE;
    p.bytes += align_up(sizeof(double) * (F64_SPLIT - 1) * (size_t)(nt * (nt + 1) / 2) * F64_TILE * F64_TILE, 256) + 256;
  }
  return p;
}

// With ds (FP64 only) the feature launches also accumulate the data-side sums into ds->part.
template <int MODE>
int run(const double* noise, const double* uv, const double* grid, int64_t K, int64_t M, double* W,
        const Plan& p, Arena& ar, hipStream_t st, const DataSums* ds = nullptr) {
  if (ds && MODE != AJ_PREC_FP64) {
    set_error("data-side sums fuse with the FP64 features only");
    return AJ_EINVAL;
  }
  void* X = ar.take<char>(p.feat_bytes);
  double* centre = ar.take<double>(2);
  double* sc = ar.take<double>(p.group);
  double* partials = nullptr;
  if constexpr (MODE == AJ_PREC_FP64) {
    const int64_t n128 = p.Mpad / F64_TILE;
    partials = ar.take<double>((F64_SPLIT - 1) * (size_t)(n128 * (n128 + 1) / 2) * F64_TILE * F64_TILE);
  }
  int* tiles = nullptr;
  int64_t nt8 = 0;
  [[maybe_unused]] const int layout = g_i8_layout.load(std::memory_order_relaxed);
  if (is_i8(MODE)) {
    static thread_local std::vector<int> tl;
    build_i8_tiles(p.Mpad, tl);
    nt8 = (int64_t)tl.size() / 2;
    tiles = ar.take<int>(tl.size());
    if (!ar.ok()) {
      set_error("workspace too small: need %zu bytes, got %zu", p.bytes, ar.cap);
      return AJ_ESPACE;
    }
    AJ_CHECK_HIP(hipMemcpyAsync(tiles, tl.data(), sizeof(int) * tl.size(), hipMemcpyHostToDevice, st));
  }
  if (!ar.ok()) {
    set_error("workspace too small: need %zu bytes, got %zu", p.bytes, ar.cap);
    return AJ_ESPACE;
  }
  hipLaunchKernelGGL(grid_mean_kernel, dim3(1), dim3(256), 0, st, grid, M, centre);
  AJ_LAUNCH_CHECK();
  if (ds) {
    const int rc = data_sums_coef(*ds, noise, uv, centre, K, st);
    if (rc != AJ_OK) return rc;
  }
  const int64_t nch = std::max<int64_t>(1, ceil_div(K, KC));
  const int64_t nt = p.Mpad / TILE;
  const int64_t ntiles = nt * (nt + 1) / 2;
  int first = 1;
  if (is_i8(MODE)) AJ_CHECK_HIP(hipMemsetAsync(W, 0, sizeof(double) * M * M, st));
  for (int64_t c0 = 0; c0 < nch; c0 += p.group) {
    const int ng = (int)std::min<int64_t>(p.group, nch - c0);
    const int64_t k_begin = c0 * KC;
    if (is_i8(MODE)) {
      hipLaunchKernelGGL(chunk_scale_kernel, dim3(ng), dim3(256), 0, st, noise, K, k_begin, ng, sc);
      AJ_LAUNCH_CHECK();
    }
    dim3 fg((unsigned)ceil_div((int64_t)ng * KC, 256), (unsigned)ceil_div(p.Mpad, FEAT_ROWS));
    timer_begin(AJ_TIMER_WT_FEATURES, st);
    if constexpr (MODE == AJ_PREC_FP64) {
      if (ds)
        hipLaunchKernelGGL((features_kernel<MODE, true>), fg, dim3(256), 0, st, noise, uv, grid, centre, K, M, p.Mpad,
                           k_begin, ng, sc, X, p.ldx, *ds, k_begin / 256);
      else
        hipLaunchKernelGGL((features_kernel<MODE, false>), fg, dim3(256), 0, st, noise, uv, grid, centre, K, M, p.Mpad,
                           k_begin, ng, sc, X, p.ldx, DataSums{}, (int64_t)0);
    } else {
      hipLaunchKernelGGL((features_kernel<MODE, false>), fg, dim3(256), 0, st, noise, uv, grid, centre, K, M, p.Mpad,
                         k_begin, ng, sc, X, p.ldx, DataSums{}, (int64_t)0);
    }
    timer_end(AJ_TIMER_WT_FEATURES, st);
    AJ_LAUNCH_CHECK();
    timer_begin(AJ_TIMER_WT_SYRK, st);
    if constexpr (is_i8(MODE)) {
      constexpr int NSL = Mode<MODE>::NS, NGL = Mode<MODE>::NG;
      const dim3 g8((unsigned)(8 * ((nt8 + 7) / 8)));
      const int8_t* X8 = reinterpret_cast<const int8_t*>(X);
      switch (layout >= 0 ? layout : default_i8_variant(NSL)) {
        case 1: hipLaunchKernelGGL((syrk_i8_kernel<NSL, NGL, 1>), g8, dim3(256), 0, st, X8, M, p.Mpad, p.ldx, ng, sc, W, nt8, tiles); break;
        case 2: hipLaunchKernelGGL((syrk_i8_kernel<NSL, NGL, 2>), g8, dim3(256), 0, st, X8, M, p.Mpad, p.ldx, ng, sc, W, nt8, tiles); break;
        case 3: hipLaunchKernelGGL((syrk_i8_kernel<NSL, NGL, 3>), g8, dim3(256), 0, st, X8, M, p.Mpad, p.ldx, ng, sc, W, nt8, tiles); break;
        case 4:
          if constexpr (2 * NSL * (I8_TM + I8_TN) * I8Layout<4>::pitch <= 160 * 1024) {
            hipLaunchKernelGGL((syrk_i8_kernel<NSL, NGL, 4>), g8, dim3(256), 0, st, X8, M, p.Mpad, p.ldx, ng, sc, W, nt8, tiles);
          } else {
            set_error("SYRK variant 4 does not fit in LDS with %d digit slices", NSL);
            return AJ_EINVAL;
          }
          break;
        case 8: hipLaunchKernelGGL((syrk_i8_ring_kernel<NSL, NGL>), g8, dim3(512), 0, st, X8, M, p.Mpad, p.ldx, ng, sc, W, nt8, tiles); break;
        case 7: hipLaunchKernelGGL((syrk_i8_w8_kernel<NSL, NGL>), g8, dim3(512), 0, st, X8, M, p.Mpad, p.ldx, ng, sc, W, nt8, tiles); break;
        case 6: hipLaunchKernelGGL((syrk_i8_k32_kernel<NSL, NGL>), g8, dim3(256), 0, st, X8, M, p.Mpad, p.ldx, ng, sc, W, nt8, tiles); break;
        case 5: hipLaunchKernelGGL((syrk_i8_dma_kernel<NSL, NGL>), g8, dim3(256), 0, st, X8, M, p.Mpad, p.ldx, ng, sc, W, nt8, tiles); break;
        default: hipLaunchKernelGGL((syrk_i8_kernel<NSL, NGL, 0>), g8, dim3(256), 0, st, X8, M, p.Mpad, p.ldx, ng, sc, W, nt8, tiles); break;
      }
    } else if constexpr (MODE == AJ_PREC_FP64) {
      const int64_t n128 = p.Mpad / F64_TILE, t128 = n128 * (n128 + 1) / 2;
      hipLaunchKernelGGL(syrk_f64_kernel, dim3((unsigned)(F64_SPLIT * t128)), dim3(512), 0, st,
                         static_cast<const double*>(X), M, p.ldx, (int64_t)ng * FC, (int)t128, W, partials, first);
    } else {
      hipLaunchKernelGGL(syrk_kernel<MODE>, dim3((unsigned)ntiles), dim3(256), 0, st, X, M, p.Mpad, p.ldx, ng, sc, W,
                         first);
    }
    timer_end(AJ_TIMER_WT_SYRK, st);
    AJ_LAUNCH_CHECK();
    first = 0;
  }
  if constexpr (MODE == AJ_PREC_FP64) {
    if (!first) {
      const int64_t n128 = p.Mpad / F64_TILE, t128 = n128 * (n128 + 1) / 2;
      hipLaunchKernelGGL(syrk_f64_partials_kernel, dim3(16, (unsigned)t128), dim3(256), 0, st, partials, (int)t128, M,
                         W);
      AJ_LAUNCH_CHECK();
    }
  }
  dim3 mg((unsigned)ceil_div(M, 32), (unsigned)ceil_div(M, 32));
  hipLaunchKernelGGL(mirror_kernel, mg, dim3(256), 0, st, W, M);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

// The integer-residue (CRT) mode lives in wtilde_crt.hip: it accumulates the lower triangle of a
// zeroed W; centring and the mirror are shared with the modes above.
size_t crt_workspace_bytes(int64_t M, int64_t K, int nm);
int crt_accumulate(const double* noise, const double* uv, const double* grid, const double* centre, int64_t K,
                   int64_t M, int nm, double* W, void* workspace, size_t workspace_bytes, hipStream_t st,
                   const DataSums* ds);

inline int crt_moduli(int precision) { return precision == AJ_PREC_INT8CRT11 ? 11 : 12; }
inline size_t crt_mode_bytes(int64_t M, int64_t K, int nm) { return 512 + crt_workspace_bytes(M, K, nm); }

int run_crt_mode(const double* noise, const double* uv, const double* grid, int64_t K, int64_t M, int nm, double* W,
                 void* workspace, size_t workspace_bytes, hipStream_t st, const DataSums* ds = nullptr) {
  if (workspace_bytes < crt_mode_bytes(M, K, nm)) {
    set_error("workspace too small: need %zu bytes, got %zu", crt_mode_bytes(M, K, nm), workspace_bytes);
    return AJ_ESPACE;
  }
  double* centre = static_cast<double*>(workspace);
  hipLaunchKernelGGL(grid_mean_kernel, dim3(1), dim3(256), 0, st, grid, M, centre);
  AJ_LAUNCH_CHECK();
  if (ds) {
    const int rc = data_sums_coef(*ds, noise, uv, centre, K, st);
    if (rc != AJ_OK) return rc;
  }
  AJ_CHECK_HIP(hipMemsetAsync(W, 0, sizeof(double) * M * M, st));
  const int rc = crt_accumulate(noise, uv, grid, centre, K, M, nm, W, static_cast<char*>(workspace) + 512,
                                workspace_bytes - 512, st, ds);
  if (rc != AJ_OK) return rc;
  dim3 mg((unsigned)ceil_div(M, 32), (unsigned)ceil_div(M, 32));
  hipLaunchKernelGGL(mirror_kernel, mg, dim3(256), 0, st, W, M);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

// Coefficients of the data-side sums per visibility, rotated by the centre's phase phi_k (data_sums.h):
//   dirty image   a = Re d / s^2, b = Im d / s^2:  alpha = a cos + b sin,  beta = b cos - a sin
//   w_tilde_data  w = (s^2 / v)^2 (numba.py:149):   gamma = w cos,         delta = -w sin
// with s = noise_k and cos, sin of 2 pi phi_k.  Absent inputs give zero coefficients.
__global__ void data_coef_kernel(const double* __restrict__ data, const double* __restrict__ vis,
                                 const double* __restrict__ noise, const double* __restrict__ uv,
                                 const double* __restrict__ centre, int64_t K, double* __restrict__ coef) {
  const double cy = centre[0], cx = centre[1];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < K; k += (int64_t)gridDim.x * blockDim.x) {
    double sp, cp;
    sincospi(2.0 * reduced_revs(cx, cy, uv[2 * k], uv[2 * k + 1]), &sp, &cp);
    const double n = noise[k];
    double al = 0.0, be = 0.0, ga = 0.0, de = 0.0;
    if (data) {
      const double a = data[2 * k] / (n * n), b = data[2 * k + 1] / (n * n);
      al = fma(a, cp, b * sp);
      be = fma(b, cp, -(a * sp));
    }
    if (vis) {
      const double q = n * n / vis[k];
      const double w = q * q;
      ga = w * cp;
      de = -(w * sp);
    }
    coef[4 * k] = al;
    coef[4 * k + 1] = be;
    coef[4 * k + 2] = ga;
    coef[4 * k + 3] = de;
  }
}

// out[o][s][i] = sum over the blocks b of segment s, in order, of in[o][b][i] (i < M); blockIdx.z = o.
__global__ void seg_sum_kernel(const double* __restrict__ in, int64_t nb, int64_t ld, int64_t seg, int64_t nseg,
                               int64_t M, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  const int64_t s = blockIdx.y, o = blockIdx.z;
  const int64_t b0 = s * seg, b1 = std::min<int64_t>(nb, b0 + seg);
  const double* src = in + o * nb * ld + i;
  double acc = 0.0;
  for (int64_t b = b0; b < b1; ++b) acc += src[b * ld];
  out[(o * nseg + s) * M + i] = acc;
}

// Final pass: the segment sums, in order, into the requested outputs (null ones skipped).
__global__ void seg_final_kernel(const double* __restrict__ part2, int64_t nseg, int64_t M, double* __restrict__ y,
                                 double* __restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  double a = 0.0, b = 0.0;
  for (int64_t s = 0; s < nseg; ++s) {
    a += part2[s * M + i];
    b += part2[(nseg + s) * M + i];
  }
  if (y) y[i] = a;
  if (z) z[i] = b;
}

// Workgroup blocks of the visibility axis whose row sums a fused build writes: FP64 features cover
// whole KC-chunks in blocks of 256 visibilities; the residue features cover the visibilities in blocks
// of 512.
inline int64_t data_sum_blocks(int64_t K, int precision) {
  if (precision == AJ_PREC_FP64) return std::max<int64_t>(1, ceil_div(K, KC)) * (KC / 256);
  return ceil_div(K, 512);
}
inline bool fuses_data_sums(int precision) { return precision == AJ_PREC_FP64 || precision >= AJ_PREC_INT8CRT; }
constexpr int64_t DS_SEGMENTS = 64;  // segments of the first reduction pass

inline DataSums data_sums_layout(int64_t K, int64_t M, int precision, Arena& ar) {
  DataSums d{};
  d.nblk = data_sum_blocks(K, precision);
  d.ld = ceil_div(M, 256) * 256;  // at least every feature kernel's padded row count
  d.nseg = std::max<int64_t>(1, std::min<int64_t>(d.nblk, DS_SEGMENTS));
  d.coef = ar.take<double>(4 * (size_t)std::max<int64_t>(K, 1));
  d.part = ar.take<double>(2 * (size_t)std::max<int64_t>(d.nblk, 1) * d.ld);
  d.part2 = ar.take<double>(2 * (size_t)d.nseg * M);
  return d;
}

inline size_t data_sums_bytes(int64_t K, int64_t M, int precision) {
  const int64_t nblk = data_sum_blocks(K, precision), ld = ceil_div(M, 256) * 256;
  const int64_t nseg = std::max<int64_t>(1, std::min<int64_t>(nblk, DS_SEGMENTS));
  return align_up(sizeof(double) * 4 * (size_t)std::max<int64_t>(K, 1), 256) +
         align_up(sizeof(double) * 2 * (size_t)std::max<int64_t>(nblk, 1) * ld, 256) +
         align_up(sizeof(double) * 2 * (size_t)nseg * M, 256) + 256;
}

}  // namespace wt

int data_sums_coef(const DataSums& ds, const double* noise, const double* uv, const double* centre, int64_t K,
                   hipStream_t st) {
  if (K == 0) return AJ_OK;
  hipLaunchKernelGGL(wt::data_coef_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(K, 256), 4096)), dim3(256), 0,
                     st, ds.data, ds.vis, noise, uv, centre, K, ds.coef);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

int data_sums_finish(const DataSums& ds, int64_t M, double* dirty_image, double* w_tilde_data, hipStream_t st) {
  const int64_t seg = ceil_div(ds.nblk, ds.nseg);
  hipLaunchKernelGGL(wt::seg_sum_kernel, dim3((unsigned)ceil_div(M, 256), (unsigned)ds.nseg, 2), dim3(256), 0, st,
                     ds.part, ds.nblk, ds.ld, seg, ds.nseg, M, ds.part2);
  AJ_LAUNCH_CHECK();
  hipLaunchKernelGGL(wt::seg_final_kernel, dim3((unsigned)ceil_div(M, 256)), dim3(256), 0, st, ds.part2, ds.nseg, M,
                     dirty_image, w_tilde_data);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

}  // namespace aj

using namespace aj;

extern "C" size_t aj_w_tilde_workspace_bytes(int64_t K, int64_t M, int precision) {
  if (K < 0 || M <= 0 || precision < AJ_PREC_FP64 || precision > AJ_PREC_INT8CRT11) return 0;
  if (precision >= AJ_PREC_INT8CRT) return wt::crt_mode_bytes(M, K, wt::crt_moduli(precision));
  return wt::make_plan(K, M, precision).bytes;
}

extern "C" int aj_w_tilde_curvature_interferometer_from(const double* noise_map_real, const double* uv_wavelengths,
                                                        const double* grid_radians_slim, int64_t K, int64_t M,
                                                        int precision, double* w_tilde, void* workspace,
                                                        size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && K >= 0, "bad sizes M=%lld K=%lld", (long long)M, (long long)K);
  AJ_REQUIRE(precision >= AJ_PREC_FP64 && precision <= AJ_PREC_INT8CRT11, "bad precision %d", precision);
  AJ_REQUIRE(w_tilde && grid_radians_slim && (K == 0 || (noise_map_real && uv_wavelengths)), "null pointer");
  if (precision >= AJ_PREC_INT8CRT)
    return wt::run_crt_mode(noise_map_real, uv_wavelengths, grid_radians_slim, K, M, wt::crt_moduli(precision),
                            w_tilde, workspace, workspace_bytes, as_stream(stream));
  wt::Plan p = wt::make_plan(K, M, precision);
  AJ_REQUIRE(workspace_bytes >= p.bytes && workspace, "workspace too small: need %zu bytes, got %zu", p.bytes,
             workspace_bytes);
  Arena ar(workspace, workspace_bytes);
  hipStream_t st = as_stream(stream);
  switch (precision) {
    case AJ_PREC_FP64: return wt::run<AJ_PREC_FP64>(noise_map_real, uv_wavelengths, grid_radians_slim, K, M, w_tilde, p, ar, st);
    case AJ_PREC_FP32: return wt::run<AJ_PREC_FP32>(noise_map_real, uv_wavelengths, grid_radians_slim, K, M, w_tilde, p, ar, st);
    case AJ_PREC_BF16X3: return wt::run<AJ_PREC_BF16X3>(noise_map_real, uv_wavelengths, grid_radians_slim, K, M, w_tilde, p, ar, st);
    case AJ_PREC_INT8X3: return wt::run<AJ_PREC_INT8X3>(noise_map_real, uv_wavelengths, grid_radians_slim, K, M, w_tilde, p, ar, st);
    case AJ_PREC_INT8X4: return wt::run<AJ_PREC_INT8X4>(noise_map_real, uv_wavelengths, grid_radians_slim, K, M, w_tilde, p, ar, st);
    default: return wt::run<AJ_PREC_INT8X5>(noise_map_real, uv_wavelengths, grid_radians_slim, K, M, w_tilde, p, ar, st);
  }
}

extern "C" size_t aj_w_tilde_with_data_workspace_bytes(int64_t K, int64_t M, int precision) {
  const size_t w = aj_w_tilde_workspace_bytes(K, M, precision);
  if (!w) return 0;
  const size_t d = wt::fuses_data_sums(precision) ? wt::data_sums_bytes(K, M, precision)
                                                   : aj_visibility_sum_workspace_bytes(K, M);
  return align_up(w, 256) + d;
}

extern "C" int aj_w_tilde_with_data_interferometer_from(const double* noise_map_real, const double* uv_wavelengths,
                                                        const double* grid_radians_slim, int64_t K, int64_t M,
                                                        int precision, const double* data,
                                                        const double* visibilities_real, double* w_tilde,
                                                        double* dirty_image, double* w_tilde_data, void* workspace,
                                                        size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && K >= 0, "bad sizes M=%lld K=%lld", (long long)M, (long long)K);
  AJ_REQUIRE(precision >= AJ_PREC_FP64 && precision <= AJ_PREC_INT8CRT11, "bad precision %d", precision);
  AJ_REQUIRE(w_tilde && grid_radians_slim && (K == 0 || (noise_map_real && uv_wavelengths)), "null pointer");
  AJ_REQUIRE(K == 0 || !dirty_image || data, "dirty_image requested without data");
  AJ_REQUIRE(K == 0 || !w_tilde_data || visibilities_real, "w_tilde_data requested without visibilities_real");
  const size_t need = aj_w_tilde_with_data_workspace_bytes(K, M, precision);
  AJ_REQUIRE(workspace && workspace_bytes >= need, "workspace too small: need %zu bytes, got %zu", need,
             workspace_bytes);
  const size_t wb = align_up(aj_w_tilde_workspace_bytes(K, M, precision), 256);
  void* dws = static_cast<char*>(workspace) + wb;
  const size_t dwb = workspace_bytes - wb;
  hipStream_t st = as_stream(stream);
  if (!wt::fuses_data_sums(precision)) {
    // single-precision features: the data-side sums evaluate their own FP64 phases
    int rc = aj_w_tilde_curvature_interferometer_from(noise_map_real, uv_wavelengths, grid_radians_slim, K, M,
                                                      precision, w_tilde, workspace, wb, stream);
    if (rc == AJ_OK && dirty_image)
      rc = vis_sum(data, 2, data ? data + 1 : nullptr, 2, noise_map_real, 1, 0, uv_wavelengths, grid_radians_slim,
                   K, M, dirty_image, dws, dwb, st);
    if (rc == AJ_OK && w_tilde_data)
      rc = vis_sum(visibilities_real, 1, nullptr, 0, noise_map_real, 1, 1, uv_wavelengths, grid_radians_slim, K, M,
                   w_tilde_data, dws, dwb, st);
    return rc;
  }
  Arena da(dws, dwb);
  DataSums ds = wt::data_sums_layout(K, M, precision, da);
  AJ_REQUIRE(da.ok(), "workspace too small for the data-side sums");
  ds.data = dirty_image ? data : nullptr;
  ds.vis = w_tilde_data ? visibilities_real : nullptr;
  int rc;
  if (precision >= AJ_PREC_INT8CRT) {
    rc = wt::run_crt_mode(noise_map_real, uv_wavelengths, grid_radians_slim, K, M, wt::crt_moduli(precision), w_tilde,
                          workspace, wb, st, &ds);
  } else {
    const wt::Plan p = wt::make_plan(K, M, AJ_PREC_FP64);
    Arena ar(workspace, wb);
    rc = wt::run<AJ_PREC_FP64>(noise_map_real, uv_wavelengths, grid_radians_slim, K, M, w_tilde, p, ar, st, &ds);
  }
  if (rc != AJ_OK) return rc;
  return data_sums_finish(ds, M, dirty_image, w_tilde_data, st);
}

extern "C" int aj_w_tilde_set_variant(int variant) {
  AJ_REQUIRE(variant >= -1 && variant < wt::I8_NLAYOUT, "variant %d outside [-1, %d)", variant, wt::I8_NLAYOUT);
  wt::g_i8_layout.store(variant, std::memory_order_relaxed);
  return AJ_OK;
}

extern "C" int aj_w_tilde_variant_count(void) { return wt::I8_NLAYOUT; }

extern "C" int aj_w_tilde_set_group_bytes(size_t bytes) {
  AJ_REQUIRE(bytes == 0 || bytes >= (size_t(1) << 20), "group bytes %zu below 1 MiB", bytes);
  wt::g_group_bytes.store(bytes ? bytes : wt::GROUP_BYTES_DEFAULT, std::memory_order_relaxed);
  return AJ_OK;
}
