// Data-side sums over visibilities accumulated by the W-tilde feature kernels: the dirty image
//   y_i = sum_k (Re d_k cos 2pi r_ik + Im d_k sin 2pi r_ik) / sigma_k^2
// and the reference's w_tilde_data (numba.py:89-152)
//   z_i = sum_k (sigma_k^2 / v_k)^2 cos 2pi r_ik,
// r_ik = g_i1 u_k0 + g_i0 u_k1 the absolute phase in revolutions.  The feature kernels evaluate the
// cosine and sine of the centre-relative phase r'_ik = r_ik - phi_k (phi_k = c_1 u_k0 + c_0 u_k1, c the
// grid mean) for W-tilde; with per-visibility coefficients rotated by phi_k,
//   a cos 2pi r + b sin 2pi r = cos 2pi r' (a cos 2pi phi + b sin 2pi phi) + sin 2pi r' (b cos 2pi phi - a sin 2pi phi),
// the same cos/sin values give both sums.  Each feature workgroup (16 image rows x a block of
// visibilities) writes its 16 row sums to part[o][block][row]; two fixed-order passes then sum the
// blocks (deterministic, no atomics).
#pragma once

#include "aj_common.h"

namespace aj {

struct DataSums {
  const double* data;     // K complex visibilities, (re, im) interleaved, or null
  const double* vis;      // K real visibilities of w_tilde_data, or null
  double* coef;           // [K][4]: (alpha, beta) of the dirty image, (gamma, delta) of w_tilde_data
  double* part;           // [2][nblk][ld] per-block row sums (dirty image, then w_tilde_data)
  double* part2;          // [2][nseg][M] sums over segments of blocks
  int64_t nblk, ld, nseg;
};

// Wave sums of 16 per-row values pushed one row at a time (row index r wave-uniform, 0..15, in order):
// a streaming butterfly whose exchanges at lane distance 32, 16, 8 and 4 each halve the rows a lane
// holds, then a sum over the 4 lanes that differ in bits 0-1.  After row 15 every lane holds the wave
// sum of row_of(lane).
struct RowSum16 {
  double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
  __device__ static int row_of(int lane) {
    return 8 * ((lane >> 2) & 1) + 4 * ((lane >> 3) & 1) + 2 * ((lane >> 4) & 1) + ((lane >> 5) & 1);
  }
  // lo holds row 2j of the pair, hi row 2j + 1: lanes with the mask bit set keep hi
  __device__ static double pair(double lo, double hi, int mask, int lane) {
    const bool up = (lane & mask) != 0;
    return (up ? hi : lo) + __shfl_xor(up ? lo : hi, mask, 64);
  }
  __device__ void push(int r, double v, int lane, double& out) {
    if (!(r & 1)) {
      p0 = v;
      return;
    }
    const double w = pair(p0, v, 32, lane);
    if (!(r & 2)) {
      p1 = w;
      return;
    }
    const double x = pair(p1, w, 16, lane);
    if (!(r & 4)) {
      p2 = x;
      return;
    }
    const double y = pair(p2, x, 8, lane);
    if (!(r & 8)) {
      p3 = y;
      return;
    }
    double z = pair(p3, y, 4, lane);
    z += __shfl_xor(z, 1, 64);
    z += __shfl_xor(z, 2, 64);
    out = z;
  }
};

// The workgroup's (256 threads, 4 waves) row sums s0 (dirty image) and s1 (w_tilde_data) of rows
// row0 .. row0 + 15, summed over the waves in order, to part[o][blk][row0 + row].  Every thread of the
// workgroup calls it.
__device__ inline void block_row_sums_out(double s0, double s1, const DataSums& ds, int64_t blk, int64_t row0) {
  __shared__ double red[2][4][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if ((lane & 3) == 0) {
    const int row = RowSum16::row_of(lane);
    red[0][wave][row] = s0;
    red[1][wave][row] = s1;
  }
  __syncthreads();
  if (threadIdx.x < 32) {
    const int o = threadIdx.x >> 4, row = threadIdx.x & 15;
    const double s = ((red[o][0][row] + red[o][1][row]) + red[o][2][row]) + red[o][3][row];
    ds.part[((int64_t)o * ds.nblk + blk) * ds.ld + row0 + row] = s;
  }
}

// Host side (wtilde.hip): per-visibility coefficients after the centre is known, and the reduction
// of the per-block sums into the outputs.
int data_sums_coef(const DataSums& ds, const double* noise, const double* uv, const double* centre, int64_t K,
                   hipStream_t st);
int data_sums_finish(const DataSums& ds, int64_t M, double* dirty_image, double* w_tilde_data, hipStream_t st);

// Stand-alone data-side sum (likelihood.hip), one output per call, phases evaluated on their own:
//   out_i = sum_k (a_k cos 2pi r_ik + b_k sin 2pi r_ik) scaled per mode by the noise wdiv[wstride k]
// (mode 0: the dirty image, a = Re d, b = Im d, divided by sigma^2; mode 1: w_tilde_data, a = v, b unused).
// The workspace holds aj_visibility_sum_workspace_bytes(K, M) bytes.
int vis_sum(const double* a, int64_t as, const double* b, int64_t bs, const double* wdiv, int64_t wstride, int mode,
            const double* uv, const double* grid, int64_t K, int64_t M, double* out, void* ws, size_t wsb,
            hipStream_t st);

}  // namespace aj
