// Likelihood path after W-tilde: sparse mapping, data vector, curvature,
// regularisation, FP64 Cholesky, solves and the log-evidence, for one sample
// or a batch of samples sharing one W-tilde.
//
// Reference: src/autojax/numba.py
//   mapping_matrix_from                   :293-376  (rows stop at the first -1)
//   data_vector_from                      :379-406  (D = dirty @ T)
//   curvature via sparse mapping          :552-614  (F = T^T W T)
//   constant_regularization_matrix_from   :656-695
//   reconstruction_positive_negative_from :698-731  (here: Cholesky, as jax.py:1032-1044)
//   noise_normalization_complex_from      :734-752
//   log_likelihood_function_via_w_tilde_from :755-857
//
// Design: the mapping is inverted once per sample into source-pixel buckets
// (CSC, sorted by image pixel), so every gather below is a deterministic
// ordered sum with no floating-point atomics.
#include "aj_common.h"
#include "data_sums.h"

#include <climits>
#include <mutex>
#include <vector>

#include <hip/hip_cooperative_groups.h>

namespace aj {
namespace lk {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) double lds_f64;  // a double in LDS

// A wave-uniform value read as such (from lane 0): for values the compiler cannot prove uniform, such as
// a device function's arguments or what it loads, so that branches, loops and buffer resources on them
// stay scalar.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t uni(int64_t v) {
  const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(lo | (hi << 32));
}
template <typename T>
__device__ inline T* uni(T* p) {  // keeps the pointer's address space
  return (T*)uni((int64_t)(uint64_t)p);
}
// v_mfma_f64_16x16x4: A lane l holds A[l%16][l/16], B lane l holds B[l/16][l%16], and the
// accumulator register r of lane l holds D[4r + l/16][l%16].

// Per-sample table, device resident.
struct SampleDesc {
  int64_t S;          // source pixels
  int64_t s_off;      // offset of this sample's rows in neighbor / D / reconstruction arrays
  int64_t map_off;    // offset of this sample's [M,B] mapping block
  int64_t csc_ptr;    // offset of the S+1 bucket pointers
  int64_t csc_ent;    // offset of the M*B bucket entries
  int64_t omega_off;  // offset of the M*S omega block (doubles)
  int64_t mat_off;    // offset of the S*S matrices A and H (doubles)
  double coef2;       // regularisation coefficient squared
  double beta;        // rank-one deflation shift for log|H| (0 = none)
};

// The persistent Cholesky's synchronisation words (ints): [q] head of XCD q's task queue (q < 8), [8] error
// flag, [9 + j] completed tasks of matrix j; cleared by the batch's first launch.
constexpr int FLOW_QUEUES = 8;  // one per XCD of an MI355X
constexpr int FLOW_ERR = FLOW_QUEUES;
constexpr int FLOW_SYNC_HEAD = FLOW_QUEUES + 1;
// s_getreg operand of the XCC_ID hardware register (register 20, bits 3:0): the XCD a wave runs on
constexpr int FLOW_XCC_ID_REG = (3 << 11) | (0 << 6) | 20;

// ---------------------------------------------------------------------------
// Mapping inversion (CSC): bucket s holds the entries (m, t) with idx[m,b] = s
// for b before the first -1, ordered by (m, b).
// ---------------------------------------------------------------------------
__global__ void csc_count_kernel(const SampleDesc* __restrict__ sd, const int64_t* __restrict__ idx, int64_t M,
                                 int64_t B, int* __restrict__ ptrs) {
  const int smp = blockIdx.y;
  const SampleDesc d = sd[smp];
  int* cnt = ptrs + d.csc_ptr;
  const int64_t* ix = idx + d.map_off;
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (int64_t)gridDim.x * blockDim.x) {
    for (int64_t b = 0; b < B; ++b) {
      int64_t s = ix[m * B + b];
      if (s == -1) break;
      if (s >= 0 && s < d.S) atomicAdd(&cnt[s + 1], 1);
    }
  }
}

// Inclusive scan of counts in place (ptr[0] = 0): one block per sample.
__global__ void csc_scan_kernel(const SampleDesc* __restrict__ sd, int* __restrict__ ptrs, int* __restrict__ cursor) {
  const SampleDesc d = sd[blockIdx.x];
  int* p = ptrs + d.csc_ptr;
  int* cur = cursor + d.csc_ptr;
  __shared__ int carry;
  __shared__ int buf[1024];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 1; base <= d.S; base += 1024) {
    int64_t i = base + threadIdx.x;
    int v = (i <= d.S) ? p[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int t = (threadIdx.x >= o) ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i <= d.S) p[i] = buf[threadIdx.x] + carry;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  for (int64_t i = threadIdx.x; i <= d.S; i += 1024) cur[i] = p[i];
}

__global__ void csc_fill_kernel(const SampleDesc* __restrict__ sd, const int64_t* __restrict__ idx,
                                const double* __restrict__ wts, int64_t M, int64_t B, int* __restrict__ cursor,
                                int* __restrict__ ent_key, double* __restrict__ ent_t) {
  const int smp = blockIdx.y;
  const SampleDesc d = sd[smp];
  int* cur = cursor + d.csc_ptr;
  const int64_t* ix = idx + d.map_off;
  const double* wt = wts + d.map_off;
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (int64_t)gridDim.x * blockDim.x) {
    for (int64_t b = 0; b < B; ++b) {
      int64_t s = ix[m * B + b];
      if (s == -1) break;
      if (s >= 0 && s < d.S) {
        int pos = atomicAdd(&cur[s], 1);
        ent_key[d.csc_ent + pos] = (int)(m * B + b);
        ent_t[d.csc_ent + pos] = wt[m * B + b];
      }
    }
  }
}

// Sort each bucket by key (m*B + b): insertion sort, buckets are short.
// One wave per bucket: buckets of at most 64 entries are ordered by rank (each lane counts the keys
// below its own over the wave; keys m*B+b are distinct); larger buckets by insertion sort in lane 0.
// Keys become m for the gathers.  256 threads = 4 buckets per workgroup.
__global__ void __launch_bounds__(256) csc_sort_kernel(const SampleDesc* __restrict__ sd,
                                                       const int* __restrict__ ptrs, int64_t B,
                                                       int* __restrict__ ent_key, double* __restrict__ ent_t) {
  const SampleDesc d = sd[blockIdx.y];
  const int* p = ptrs + d.csc_ptr;
  const int lane = threadIdx.x & 63;
  for (int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); s < d.S; s += (int64_t)gridDim.x * 4) {
    const int lo = p[s], hi = p[s + 1], len = hi - lo;
    int* k = ent_key + d.csc_ent;
    double* t = ent_t + d.csc_ent;
    if (len <= 64) {
      const int key = lane < len ? k[lo + lane] : INT_MAX;
      const double tv = lane < len ? t[lo + lane] : 0.0;
      int rank = 0;
      for (int j = 0; j < 64; ++j) rank += __shfl(key, j, 64) < key ? 1 : 0;
      if (lane < len) {
        k[lo + rank] = (int)(key / B);
        t[lo + rank] = tv;
      }
    } else if (lane == 0) {
      for (int i = lo + 1; i < hi; ++i) {
        int kk = k[i];
        double tt = t[i];
        int j = i - 1;
        while (j >= lo && k[j] > kk) {
          k[j + 1] = k[j];
          t[j + 1] = t[j];
          --j;
        }
        k[j + 1] = kk;
        t[j + 1] = tt;
      }
      for (int i = lo; i < hi; ++i) k[i] = (int)(k[i] / B);
    }
  }
}

// ---------------------------------------------------------------------------
// The same mapping inversion and data vector for one sample in one workgroup (the batch's first
// launch): bucket counts in LDS, their prefix sum (the pointers), the entries placed by LDS cursors, each
// bucket ordered by (m, b) by one wave (a rank sort up to 64 entries), then D of the sample's buckets.
// ---------------------------------------------------------------------------
constexpr int PREP_THREADS = 1024;
constexpr int PREP_SLICES_MAX = 256;  // slices of 64 columns per sample (S <= 16,384)

__global__ void __launch_bounds__(PREP_THREADS) prep_kernel(const SampleDesc* __restrict__ sd,
                                                            const int64_t* __restrict__ idx,
                                                            const double* __restrict__ wts, int64_t M, int64_t B,
                                                            const double* __restrict__ dirty,
                                                            int* __restrict__ ptrs, int* __restrict__ ent_key,
                                                            double* __restrict__ ent_t, double* __restrict__ D,
                                                            int* __restrict__ sl_key, double* __restrict__ sl_t,
                                                            int* __restrict__ sl_col, int* __restrict__ sl_pos,
                                                            int* __restrict__ flow_sync) {
  extern __shared__ int cnt[];  // S + 1 counters, then cursors
  __shared__ int part[PREP_THREADS];
  const SampleDesc d = sd[blockIdx.x];
  if (flow_sync != nullptr && threadIdx.x < FLOW_SYNC_HEAD) {
    if (threadIdx.x < 2) flow_sync[FLOW_SYNC_HEAD + 2 * blockIdx.x + threadIdx.x] = 0;
    if (blockIdx.x == 0) flow_sync[threadIdx.x] = 0;
  }
  const int64_t S = d.S;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t* ix = idx + d.map_off;
  const double* wt = wts + d.map_off;
  int* p = ptrs + d.csc_ptr;
  int* k = ent_key + d.csc_ent;
  double* t = ent_t + d.csc_ent;
  for (int64_t i = tid; i <= S; i += PREP_THREADS) cnt[i] = 0;
  __syncthreads();
  for (int64_t m = tid; m < M; m += PREP_THREADS)
    for (int64_t b = 0; b < B; ++b) {
      const int64_t s = ix[m * B + b];
      if (s == -1) break;
      if (s >= 0 && s < S) atomicAdd(&cnt[s + 1], 1);
    }
  __syncthreads();
  // prefix sum of cnt[0..S] (cnt[0] = 0): thread tid owns a contiguous run of `per` entries
  const int64_t per = (S + 1 + PREP_THREADS - 1) / PREP_THREADS;
  const int64_t a0 = tid * per, a1 = std::min<int64_t>(S + 1, a0 + per);
  int run = 0;
  for (int64_t i = a0; i < a1; ++i) run += cnt[i];
  part[tid] = run;
  __syncthreads();
  for (int o = 1; o < PREP_THREADS; o <<= 1) {
    const int v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int acc = part[tid] - run;  // exclusive prefix of the runs before this one
  for (int64_t i = a0; i < a1; ++i) {
    acc += cnt[i];
    cnt[i] = acc;  // inclusive: pointer to the end of bucket i - 1, the start of bucket i
    p[i] = acc;
  }
  __syncthreads();
  for (int64_t m = tid; m < M; m += PREP_THREADS)
    for (int64_t b = 0; b < B; ++b) {
      const int64_t s = ix[m * B + b];
      if (s == -1) break;
      if (s >= 0 && s < S) {
        const int pos = atomicAdd(&cnt[s], 1);
        k[pos] = (int)(m * B + b);
        t[pos] = wt[m * B + b];
      }
    }
  __syncthreads();
  // order each bucket by key m*B + b (keys are distinct), then keep m
  for (int64_t s = wave; s < S; s += PREP_THREADS / 64) {
    const int lo = p[s], hi = p[s + 1], len = hi - lo;
    if (len <= 64) {
      const int key = lane < len ? k[lo + lane] : INT_MAX;
      const double tv = lane < len ? t[lo + lane] : 0.0;
      int rank = 0;
      for (int j = 0; j < 64; ++j) rank += __shfl(key, j, 64) < key ? 1 : 0;
      if (lane < len) {
        k[lo + rank] = (int)(key / B);
        t[lo + rank] = tv;
      }
    } else if (lane == 0) {
      for (int i = lo + 1; i < hi; ++i) {
        const int kk = k[i];
        const double tt = t[i];
        int j = i - 1;
        while (j >= lo && k[j] > kk) {
          k[j + 1] = k[j];
          t[j + 1] = t[j];
          --j;
        }
        k[j + 1] = kk;
        t[j + 1] = tt;
      }
      for (int i = lo; i < hi; ++i) k[i] = (int)(k[i] / B);
    }
  }
  __syncthreads();
  for (int64_t s = tid; s < S; s += PREP_THREADS) {
    double v = 0.0;
    for (int e = p[s]; e < p[s + 1]; ++e) v = fma(t[e], dirty[k[e]], v);
    D[d.s_off + s] = v;
  }
  // The buckets again, ordered by slice for the Omega kernel.  Within each window of 256 columns the columns
  // are ranked by bucket length (longest first, ties by column) and cut into slices of 64, so that the lanes
  // of a slice run for about the same number of entries while a slice's Omega columns stay in the window's
  // 2 KB of each Omega row (the window's slices complete those lines together in the L2); within a slice the
  // lanes are ranked by decreasing length (sl_col[64 q + rank] = the column), so the columns that have an
  // entry j are ranks 0 .. a_j - 1.  Entry j of rank r sits at sl_pos[q] + (the slice's entries before j) + r:
  // a wave reading entry j of its 64 columns reads consecutive addresses and needs only a_j.  Which lane sums
  // a column does not change the column's sum.
  constexpr int SLW = 256;  // columns per window
  static_assert(SLW <= PREP_THREADS, "a window's lengths fit part[]");
  __shared__ int ltot[PREP_SLICES_MAX];  // entries per slice, then the slices' first entries
  int* order = cnt;  // the columns in window order (the bucket counters are free)
  for (int64_t w0 = 0; w0 < S; w0 += SLW) {
    const int nw = (int)std::min<int64_t>(SLW, S - w0);
    __syncthreads();
    if (tid < nw) part[tid] = p[w0 + tid + 1] - p[w0 + tid];
    __syncthreads();
    if (tid < nw) {
      const int li = part[tid];
      int rank = 0;
      for (int j = 0; j < nw; ++j) {
        const int lj = part[j];
        rank += lj > li || (lj == li && j < tid) ? 1 : 0;
      }
      order[w0 + rank] = (int)(w0 + tid);
    }
  }
  __syncthreads();
  const int nsl = (int)((S + 63) / 64);
  for (int q = wave; q < nsl; q += PREP_THREADS / 64) {
    const int64_t i = (int64_t)q * 64 + lane;
    int tot = i < S ? p[order[i] + 1] - p[order[i]] : 0;
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 64);
    if (lane == 0) ltot[q] = tot;
  }
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int q = 0; q < nsl; ++q) {
      const int v = ltot[q];
      ltot[q] = run;
      run += v;
    }
  }
  __syncthreads();
  int* sk = sl_key + d.csc_ent;
  double* st = sl_t + d.csc_ent;
  int* sc = sl_col + d.csc_ptr;
  for (int q = wave; q < nsl; q += PREP_THREADS / 64) {
    const int64_t i = (int64_t)q * 64 + lane;
    const int s = i < S ? order[i] : -1;
    const int lo = s >= 0 ? p[s] : 0;
    const int len = s >= 0 ? p[s + 1] - lo : 0;
    int rank = 0;  // the lanes past S (length 0, after every column of the slice) rank last
    for (int l = 0; l < 64; ++l) {
      const int ll = __shfl(len, l, 64);
      rank += ll > len || (ll == len && l < lane) ? 1 : 0;
    }
    if (s >= 0) sc[q * 64 + rank] = s;
    int pos = ltot[q];
    if (lane == 0) sl_pos[d.csc_ptr + q] = pos;
    for (int j = 0;; ++j) {
      const uint64_t mask = __ballot(len > j);
      if (mask == 0) break;
      if (len > j) {
        sk[pos + rank] = k[lo + j];
        st[pos + rank] = t[lo + j];
      }
      pos += __popcll(mask);
    }
  }
}

// ---------------------------------------------------------------------------
// D[s] = sum_{(m,t) in bucket s} t * dirty[m]     (numba.py:406)
// ---------------------------------------------------------------------------
__global__ void data_vector_kernel(const SampleDesc* __restrict__ sd, const int* __restrict__ ptrs,
                                   const int* __restrict__ ent_m, const double* __restrict__ ent_t,
                                   const double* __restrict__ dirty, double* __restrict__ D) {
  const SampleDesc d = sd[blockIdx.y];
  const int* p = ptrs + d.csc_ptr;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < d.S; s += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int e = p[s]; e < p[s + 1]; ++e) acc = fma(ent_t[d.csc_ent + e], dirty[ent_m[d.csc_ent + e]], acc);
    D[d.s_off + s] = acc;
  }
}

// ---------------------------------------------------------------------------
// Omega = W T for every sample: a workgroup stages R rows of W in LDS (read
// once from HBM for the whole batch), then each thread forms Omega[m, s] for
// one source pixel s and the R rows by gathering its bucket from LDS.
// ---------------------------------------------------------------------------
constexpr int OMEGA_THREADS = 1024;
constexpr size_t OMEGA_LDS_BUDGET = 160 * 1024;  // the whole LDS of a gfx950 CU: 4 rows of W at M = 5,000

// Where W-tilde comes from: the dense M x M matrix, or the lattice table w_compact [N, 2N-1] with
// the (row, column) grid index of every image pixel (numba.py:237-255 gives the element rule).
struct WSrc {
  const double* W;       // dense M x M W-tilde, or
  const double* C;       // the lattice table w_compact (N x (2N - 1)), or
  const int64_t* native;  // (pixel grid positions, for C and Pre)
  int64_t N;
  const double* Pre;  // the preload table (Py x Px)
  int64_t Py, Px;
};

// W-tilde[i, j] from the preload table as original.py:431-449 reads it: the entry at the offset of the
// later pixel in slim order from the earlier one, negative offsets wrapping as numpy indexes.
__device__ inline double preload_at(const double* P, int64_t Py, int64_t Px, const int64_t* native, int64_t i,
                                    int64_t j) {
  const int64_t sg = i <= j ? 1 : -1;
  int64_t dy = sg * (native[2 * j] - native[2 * i]), dx = sg * (native[2 * j + 1] - native[2 * i + 1]);
  dy += dy < 0 ? Py : 0;
  dx += dx < 0 ? Px : 0;
  return P[dy * Px + dx];
}

__device__ inline double wsrc_at(const WSrc& s, int64_t M, int64_t i, int64_t j) {
  if (s.Pre != nullptr) return preload_at(s.Pre, s.Py, s.Px, s.native, i, j);
  if (s.C == nullptr) return s.W[i * M + j];
  const int64_t m = s.native[2 * i] - s.native[2 * j];
  const int64_t n = (m >= 0 ? s.native[2 * i + 1] - s.native[2 * j + 1] : s.native[2 * j + 1] - s.native[2 * i + 1]) +
                    s.N - 1;
  return s.C[(m >= 0 ? m : -m) * (2 * s.N - 1) + n];
}

__global__ void __launch_bounds__(OMEGA_THREADS) omega_kernel(const WSrc W, int64_t M, int R,
                                                               int nsmp, const SampleDesc* __restrict__ sd,
                                                               const int* __restrict__ ptrs,
                                                               const int* __restrict__ ent_m,
                                                               const double* __restrict__ ent_t,
                                                               double* __restrict__ omega) {
  extern __shared__ double wrow[];  // R x M
  const int64_t m0 = (int64_t)blockIdx.x * R;
  const int rows = (int)std::min<int64_t>(R, M - m0);
  if (W.C == nullptr && W.Pre == nullptr) {
    for (int64_t i = threadIdx.x; i < (int64_t)rows * M; i += OMEGA_THREADS) wrow[i] = W.W[m0 * M + i];
  } else {
    for (int64_t i = threadIdx.x; i < (int64_t)rows * M; i += OMEGA_THREADS) wrow[i] = wsrc_at(W, M, m0 + i / M, i % M);
  }
  __syncthreads();
  for (int smp = 0; smp < nsmp; ++smp) {
    const SampleDesc d = sd[smp];
    const int* p = ptrs + d.csc_ptr;
    const int* em = ent_m + d.csc_ent;
    const double* et = ent_t + d.csc_ent;
    double* om = omega + d.omega_off;
    for (int64_t s = threadIdx.x; s < d.S; s += OMEGA_THREADS) {
      double acc[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) acc[r] = 0.0;
      int e = p[s];
      const int end = p[s + 1];
      // four bucket entries loaded together (independent loads in flight), summed in entry order
      for (; e + 4 <= end; e += 4) {
        const int m0 = em[e], m1 = em[e + 1], m2 = em[e + 2], m3 = em[e + 3];
        const double t0 = et[e], t1 = et[e + 1], t2 = et[e + 2], t3 = et[e + 3];
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (r < rows) {
            const double* wr = wrow + r * M;
            acc[r] = fma(t3, wr[m3], fma(t2, wr[m2], fma(t1, wr[m1], fma(t0, wr[m0], acc[r]))));
          }
      }
      for (; e < end; ++e) {
        const int m = em[e];
        const double t = et[e];
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (r < rows) acc[r] = fma(t, wrow[r * M + m], acc[r]);
      }
#pragma unroll
      for (int r = 0; r < 8; ++r)
        if (r < rows) om[(m0 + r) * d.S + s] = acc[r];
    }
  }
}

// Omega = W T over the slice-ordered buckets: R (1, 2, 4 or 8) rows of W in LDS, interleaved (wl[m R + r]
// = W[m0 + r, m]) so that the R values one bucket entry needs are adjacent (one 8- or 16-byte read, or R / 2
// 16-byte reads).  Each wave takes whole slices of 64 columns (slice g: sample slices[g].x, first column
// slices[g].y), lane r the slice's column of rank r (by decreasing bucket length; the slices group columns
// of about the same length), so the lanes that have an entry j are lanes 0 .. a_j - 1 and read entries
// pos_j + lane: the loop has no per-lane branch or index arithmetic, only one ballot per entry, and runs to
// the slice's longest bucket.  The loads of OMEGA_SLICE_UNROLL entries (then 4) are issued ahead of their
// sums; a lane past its bucket's end reads a neighbouring entry (the slice buffers end with 64 spare entries)
// and adds it with weight +0.0 at column 0, which leaves its sum unchanged bit for bit (the sum is never
// -0.0: it starts at +0.0 and a rounded sum is -0.0 only if both terms are).  Each column's sum runs over
// its bucket in entry order (the bits of omega_kernel).
constexpr int OMEGA_SLICE_UNROLL = 8;
constexpr int OMEGA_SLICE_PAD = 64;  // spare entries after the slice buffers

template <int R>
__device__ inline void lds_rows(const double* wl, int m, double (&w)[R]) {
  if constexpr (R == 1) {
    w[0] = wl[m];
  } else {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2* src = reinterpret_cast<const d2*>(wl + (int64_t)m * R);
#pragma unroll
    for (int h = 0; h < R / 2; ++h) {
      const d2 v = src[h];
      w[2 * h] = v.x;
      w[2 * h + 1] = v.y;
    }
  }
}

// A wave's slice: its sample's descriptor fields, the lane's column and bucket length, the slice's first
// entry.
struct OmegaSlice {
  int64_t S, csc_ent, omega_off;
  int64_t s;  // the lane's column (S for a lane past the sample's columns)
  int len;    // its bucket's length (0 past the sample's last column)
  int pos;    // the slice's first entry in its sample's buckets
};

__device__ inline OmegaSlice omega_slice(int g, const int2* __restrict__ slices, const SampleDesc* __restrict__ sd,
                                         const int* __restrict__ ptrs, const int* __restrict__ sl_col,
                                         const int* __restrict__ sl_pos, int lane) {
  const int2 sq = slices[g];
  const SampleDesc d = sd[sq.x];
  const int* p = ptrs + d.csc_ptr;
  OmegaSlice o;
  o.S = d.S;
  o.csc_ent = d.csc_ent;
  o.omega_off = d.omega_off;
  const int64_t r = sq.y + lane;
  o.s = r < d.S ? sl_col[d.csc_ptr + r] : d.S;
  o.len = o.s < d.S ? p[o.s + 1] - p[o.s] : 0;
  o.pos = sl_pos[d.csc_ptr + sq.y / 64];
  return o;
}

// A group of V entries of a wave's slice: their loads (entries j .. j + V - 1, from pos on; pos is advanced
// past them), then their sums into acc.
template <int V>
struct OmegaGroup {
  int m[V];
  double t[V];
};
template <int V>
__device__ __attribute__((always_inline)) inline void omega_load(OmegaGroup<V>& g, int j, int len, int& pos, i32x4 rk,
                                                                 i32x4 rt, int vk, int vt) {
#pragma unroll
  for (int u = 0; u < V; ++u) {
    g.m[u] = raw_buffer_load_i32(rk, vk, pos * 4, 0);
    g.t[u] = raw_buffer_load_f64(rt, vt, pos * 8, 0);
    pos += __popcll(__ballot(len > j + u));
  }
}
template <int R, int V>
__device__ __attribute__((always_inline)) inline void omega_sum(const OmegaGroup<V>& g, int j, int len,
                                                                const double* wl, double (&acc)[R]) {
#pragma unroll
  for (int u = 0; u < V; ++u) {
    const bool act = len > j + u;
    double w[R];
    lds_rows<R>(wl, act ? g.m[u] : 0, w);
    const double tu = act ? g.t[u] : 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = fma(tu, w[r], acc[r]);
  }
}

template <int R, int U>
__global__ void __launch_bounds__(OMEGA_THREADS) omega_slices_kernel(const WSrc W, int64_t M, int nsl,
                                                                      const int2* __restrict__ slices,
                                                                      const SampleDesc* __restrict__ sd,
                                                                      const int* __restrict__ ptrs,
                                                                      const int* __restrict__ sl_col,
                                                                      const int* __restrict__ sl_pos,
                                                                      const int* __restrict__ sl_key,
                                                                      const double* __restrict__ sl_t,
                                                                      int64_t sl_elems, double* __restrict__ omega) {
  // M x R, interleaved; 16-byte aligned behind the static LDS (the slice counter) so that an entry's R
  // values are whole 16-byte reads
  extern __shared__ __attribute__((aligned(16))) double wl[];
  const int64_t m0 = (int64_t)blockIdx.x * R;
  const int rows = (int)std::min<int64_t>(R, M - m0);
  const bool dense = W.C == nullptr && W.Pre == nullptr;
  for (int r = 0; r < R; ++r)
    for (int64_t m = threadIdx.x; m < M; m += OMEGA_THREADS)
      wl[m * R + r] = r >= rows ? 0.0 : dense ? W.W[(m0 + r) * M + m] : wsrc_at(W, M, m0 + r, m);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  constexpr int WAVES = OMEGA_THREADS / 64;
  // slices are taken from a counter in LDS (their lengths differ: a fixed share per wave would wait for
  // the wave with the longest ones); the next slice's descriptor is loaded while the current one is summed
  __shared__ int next_slice;
  if (threadIdx.x == 0) next_slice = WAVES;
  __syncthreads();
  auto take = [&]() {
    int v = 0;
    if (lane == 0) v = atomicAdd(&next_slice, 1);
    return __builtin_amdgcn_readfirstlane(v);
  };
  OmegaSlice nx{};
  if (wave < nsl) nx = omega_slice(wave, slices, sd, ptrs, sl_col, sl_pos, lane);
  int gn = take();
  for (int g = wave; g < nsl;) {
    const OmegaSlice cu = nx;
    if (gn < nsl) nx = omega_slice(gn, slices, sd, ptrs, sl_col, sl_pos, lane);
    // raw buffer loads over the slice buffers from the sample's first entry: the lane's byte offset in a
    // vector register, the entry position as the scalar offset (no vector address arithmetic per entry)
    const int64_t ent = uni(cu.csc_ent);
    const int64_t rem = sl_elems - ent;  // entries from there to the buffers' end
    const i32x4 rk = make_rsrc(sl_key + ent, (uint32_t)std::min<int64_t>(rem * 4, 0xFFFFFFFFll));
    const i32x4 rt = make_rsrc(sl_t + ent, (uint32_t)std::min<int64_t>(rem * 8, 0xFFFFFFFFll));
    const int vk = lane * 4, vt = lane * 8;
    const int len = cu.len;
    int pos = __builtin_amdgcn_readfirstlane(cu.pos);
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    // lane 0 has the slice's longest bucket: U entries at a time, each group's loads issued before the
    // previous group's sums, then four at a time
    const int L = uni(len);
    int j = 0;
    if (U <= L) {
      OmegaGroup<U> ga, gb;
      omega_load<U>(ga, 0, len, pos, rk, rt, vk, vt);
      for (;;) {
        const bool more = j + 2 * U <= L;
        if (more) omega_load<U>(gb, j + U, len, pos, rk, rt, vk, vt);
        omega_sum<R, U>(ga, j, len, wl, acc);
        j += U;
        if (!more) break;
        const bool more2 = j + 2 * U <= L;
        if (more2) omega_load<U>(ga, j + U, len, pos, rk, rt, vk, vt);
        omega_sum<R, U>(gb, j, len, wl, acc);
        j += U;
        if (!more2) break;
      }
    }
    for (; j < L; j += 4) {
      OmegaGroup<4> gt;
      omega_load<4>(gt, j, len, pos, rk, rt, vk, vt);
      omega_sum<R, 4>(gt, j, len, wl, acc);
    }
    if (cu.s < cu.S) {
      double* om = omega + cu.omega_off + m0 * cu.S + cu.s;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r < rows) om[r * cu.S] = acc[r];
    }
    g = gn;
    if (g < nsl) gn = take();
  }
}

// Same product when a row of W does not fit in LDS: gather from global memory.
__global__ void omega_global_kernel(const WSrc W, int64_t M, int nsmp,
                                    const SampleDesc* __restrict__ sd, const int* __restrict__ ptrs,
                                    const int* __restrict__ ent_m, const double* __restrict__ ent_t,
                                    double* __restrict__ omega) {
  const int64_t m = blockIdx.x;
  for (int smp = 0; smp < nsmp; ++smp) {
    const SampleDesc d = sd[smp];
    const int* p = ptrs + d.csc_ptr;
    for (int64_t s = threadIdx.x; s < d.S; s += blockDim.x) {
      double acc = 0.0;
      for (int e = p[s]; e < p[s + 1]; ++e) acc = fma(ent_t[d.csc_ent + e], wsrc_at(W, M, m, ent_m[d.csc_ent + e]), acc);
      omega[d.omega_off + m * d.S + s] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// F[s1, s2] = sum_{(m,t) in bucket s1} t * Omega[m, s2], written into the
// matrix block (lower triangle only when lower_only).  One workgroup per s1.
// ---------------------------------------------------------------------------
__global__ void fgather_kernel(const SampleDesc* __restrict__ sd, const int* __restrict__ ptrs,
                               const int* __restrict__ ent_m, const double* __restrict__ ent_t,
                               const double* __restrict__ omega, double* __restrict__ mats, int lower_only) {
  const SampleDesc d = sd[blockIdx.y];
  const int64_t s1 = blockIdx.x;
  if (s1 >= d.S) return;
  const int* p = ptrs + d.csc_ptr;
  const int lo = p[s1], hi = p[s1 + 1];
  const int64_t ncol = lower_only ? s1 + 1 : d.S;
  double* out = mats + d.mat_off + s1 * d.S;
  const double* om = omega + d.omega_off;
  for (int64_t s2 = threadIdx.x; s2 < ncol; s2 += blockDim.x) {
    double acc = 0.0;
    for (int e = lo; e < hi; ++e) acc = fma(ent_t[d.csc_ent + e], om[(int64_t)ent_m[d.csc_ent + e] * d.S + s2], acc);
    out[s2] = acc;
  }
}

// ---------------------------------------------------------------------------
// Row s1 of A = F + H and of H, lower triangle, in one pass (the batch's third launch):
//   F[s1, s2] = sum over bucket s1 of t Omega[m, s2],
//   H[s1, s2] = (beta / S, the rank-one deflation of log|H|, or 0) + [s1 = s2] (1e-8 + c^2 size_s1)
//               - c^2 (the number of times s2 appears in s1's neighbour list before -1)   (numba.py:686-694).
// ---------------------------------------------------------------------------
__global__ void assemble_kernel(const SampleDesc* __restrict__ sd, const int* __restrict__ ptrs,
                                const int* __restrict__ ent_m, const double* __restrict__ ent_t,
                                const double* __restrict__ omega, const int64_t* __restrict__ nbr,
                                const int64_t* __restrict__ nsz, int64_t P, double* __restrict__ mats) {
  const SampleDesc d = sd[blockIdx.y];
  const int64_t s1 = blockIdx.x;
  if (s1 >= d.S) return;
  const int* p = ptrs + d.csc_ptr;
  const int lo = p[s1], hi = p[s1 + 1];
  double* A = mats + d.mat_off + s1 * d.S;
  double* H = A + d.S * d.S;
  const double* om = omega + d.omega_off;
  const int64_t* row = nbr + (d.s_off + s1) * P;
  const double hc = d.beta > 0.0 ? d.beta / (double)d.S : 0.0;
  const double diag = 1e-8 + d.coef2 * (double)nsz[d.s_off + s1];
  for (int64_t s2 = threadIdx.x; s2 <= s1; s2 += blockDim.x) {
    double acc = 0.0;
    for (int e = lo; e < hi; ++e) acc = fma(ent_t[d.csc_ent + e], om[(int64_t)ent_m[d.csc_ent + e] * d.S + s2], acc);
    int times = 0;
    for (int64_t q = 0; q < P; ++q) {
      const int64_t kq = row[q];
      if (kq == -1) break;
      times += kq == s2 ? 1 : 0;
    }
    double r = s2 == s1 ? diag : 0.0;
    for (int q = 0; q < times; ++q) r -= d.coef2;
    A[s2] = acc + r;
    H[s2] = hc + r;
  }
}

// ---------------------------------------------------------------------------
// Regularisation (numba.py:686-694), added into A and written into H:
//   H = diag(1e-8 + c^2 sizes);  H[i, k] -= c^2 for k in neighbours[i] until -1.
// With beta > 0, H receives + beta/S everywhere (rank-one deflation of the
// constant null vector for log|H|).  lower_only restricts both to k <= i.
// ---------------------------------------------------------------------------
__global__ void fill_kernel(const SampleDesc* __restrict__ sd, double* __restrict__ mats, int which_h) {
  const SampleDesc d = sd[blockIdx.y];
  double* H = mats + d.mat_off + (which_h ? d.S * d.S : 0);
  const double v = d.beta > 0.0 ? d.beta / (double)d.S : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.S * d.S; i += (int64_t)gridDim.x * blockDim.x)
    H[i] = v;
}

__global__ void regularize_kernel(const SampleDesc* __restrict__ sd, const int64_t* __restrict__ nbr,
                                  const int64_t* __restrict__ nsz, int64_t P, double* __restrict__ mats,
                                  int add_to_a, int write_h, int lower_only) {
  const SampleDesc d = sd[blockIdx.y];
  double* A = mats + d.mat_off;
  double* H = A + d.S * d.S;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.S; i += (int64_t)gridDim.x * blockDim.x) {
    const double diag = 1e-8 + d.coef2 * (double)nsz[d.s_off + i];
    if (add_to_a) A[i * d.S + i] += diag;
    if (write_h) H[i * d.S + i] += diag;
    const int64_t* row = nbr + (d.s_off + i) * P;
    for (int64_t j = 0; j < P; ++j) {
      const int64_t k = row[j];
      if (k == -1) break;
      if (k < 0 || k >= d.S) continue;
      if (lower_only && k > i) continue;
      if (add_to_a) A[i * d.S + k] -= d.coef2;
      if (write_h) H[i * d.S + k] -= d.coef2;
    }
  }
}

// Column-block width of the blocked Cholesky factorisations.
constexpr int NB = 32;



// ---------------------------------------------------------------------------
// Step-synchronous blocked Cholesky over a set of matrices of different sizes
// (every A and H of a batch, or one matrix).  Step k handles columns
// k0 = 32k .. k0+31 of every matrix with n > k0, in three launches:
//   chol_diag_kernel   one workgroup per matrix: factor the 32x32 diagonal block in LDS,
//                      record sum log L_jj and the first non-positive pivot of the block;
//   chol_panel_kernel  256-row blocks of every matrix: X L_kk^T = A[r0:, k0:k0+32];
//   chol_trail_kernel  64x64 lower-triangle tiles of every matrix: A22 -= P P^T on
//                      v_mfma_f64_16x16x4 with the two panel row blocks staged in LDS.
// One grid spans all matrices, so the device is filled by the many tiles of a batch
// or of one large matrix.  chol_finish_kernel sums the per-step partials in step order
// (deterministic).  The factor is left in place in the lower triangle.
// ---------------------------------------------------------------------------
struct StepEntry {
  int mat;     // matrix j: sample j >> 1, A (j & 1 == 0) or H
  int pfx_p;   // panel workgroups before this matrix in the step's grid
  int pfx_t;   // trailing tiles before this matrix in the step's grid
};

__device__ inline int find_entry(const StepEntry* __restrict__ e, int ne, int blk, bool trail) {
  int lo = 0, hi = ne - 1;  // largest i with prefix(i) <= blk
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((trail ? e[mid].pfx_t : e[mid].pfx_p) <= blk) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ inline double* mat_ptr(const SampleDesc* __restrict__ sd, double* __restrict__ mats, int j, int64_t& n) {
  const SampleDesc d = sd[j >> 1];
  n = uni(d.S);
  return mats + uni(d.mat_off + ((j & 1) ? d.S * d.S : 0));
}

__device__ inline double readlane_d(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// 1 / sqrt(d) for d > 0: the hardware estimate (v_rsq_f64) refined by two Newton steps
// y <- y + y (1 - d y^2) / 2, each of which about squares the relative error.
__device__ inline double rsqrt_f64(double d) {
  double y = __builtin_amdgcn_rsq(d);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-d * y, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

// Cholesky of an nb x nb block held by one wave, lane l < nb holding row l in r[0..l]; for each column the
// pivot and the column entries are broadcast with readlane (no barriers, no LDS).  ld receives
// sum_c log L_cc, taken as one log of the product of the pivots' mantissas plus their exponents;
// fail the first non-positive pivot (1-based, offset by k0) or 0.  A column is scaled by 1 / sqrt(pivot)
// (rsqrt_f64: a short dependent chain, where a correctly rounded square root and a division per column
// made the block's factorisation a chain of some 27 dependent FP64 operations per column).  Each lane's
// updates are selects, not branches (a lane keeps r[q] where it is not updated).
__device__ inline void chol_factor_regs(double (&r)[NB], int nb, int64_t k0, int lane, double& ld, int64_t& fail) {
  double mant = 1.0;
  int ex = 0;
  fail = 0;
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    if (c < nb) {
      double d = readlane_d(r[c], c);
      if (!(d > 0.0)) {
        if (!fail) fail = k0 + c + 1;
        d = 1.0;  // keep the arithmetic finite; the result is flagged
      }
      const double y = rsqrt_f64(d);  // the column's entries are r[c] y, its pivot d y
      const double lcc = d * y;
      int e;
      mant *= frexp(lcc, &e);  // in [0.5, 1): 32 factors stay above 2^-32
      ex += e;
      r[c] = lane == c ? lcc : lane > c ? r[c] * y : r[c];
#pragma unroll
      for (int q = c + 1; q < NB; ++q) {
        const double lqc = readlane_d(r[c], q);  // L[q][c]
        r[q] = lane >= q ? fma(-r[c], lqc, r[q]) : r[q];
      }
    }
  }
  ld = log(mant) + (double)ex * 0.69314718055994530942;
}

// The same factorisation with each column broadcast through LDS (col: NB doubles of LDS owned by the wave)
// instead of one lane read per element, the column's entries read together: the same operations on the
// same values, so the same bits.
template <typename ColPtr>
__device__ inline void chol_factor_lds(double (&r)[NB], int nb, int64_t k0, int lane, double& ld, int64_t& fail,
                                       ColPtr col) {
  double mant = 1.0;
  int ex = 0;
  fail = 0;
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    if (c < nb) {
      double d = readlane_d(r[c], c);
      if (!(d > 0.0)) {
        if (!fail) fail = k0 + c + 1;
        d = 1.0;  // keep the arithmetic finite; the result is flagged
      }
      const double y = rsqrt_f64(d);
      const double lcc = d * y;
      int e;
      mant *= frexp(lcc, &e);
      ex += e;
      r[c] = lane == c ? lcc : lane > c ? r[c] * y : r[c];
      // column c: L[q][c] in lane q (one wave: its LDS accesses complete in order, and the compiler keeps
      // this write after the previous column's reads and before these)
      if (lane < NB) col[lane] = r[c];
      double lq[NB];
#pragma unroll
      for (int q = c + 2; q < NB; ++q) lq[q] = col[q];
      // the next pivot's column first, its entry L[c+1][c] by a lane read (the same value as in LDS), so
      // that the next column does not wait for the LDS round trip
      if (c + 1 < NB) r[c + 1] = lane >= c + 1 ? fma(-r[c], readlane_d(r[c], c + 1), r[c + 1]) : r[c + 1];
#pragma unroll
      for (int q = c + 2; q < NB; ++q) r[q] = lane >= q ? fma(-r[c], lq[q], r[q]) : r[q];
    }
  }
  ld = log(mant) + (double)ex * 0.69314718055994530942;
}

// Diagonal block k0 of matrix j, one wave: load, factor, store, record the block's log-determinant part and
// first failing pivot for step kstep.
__device__ inline void chol_diag_dev(const SampleDesc* __restrict__ sd, int j, int64_t k0, int kstep, int maxsteps,
                                     double* __restrict__ mats, double* __restrict__ ldp,
                                     int64_t* __restrict__ failp, int lane) {
  int64_t n;
  double* A = mat_ptr(sd, mats, j, n);
  const int nb = (int)std::min<int64_t>(NB, n - k0);
  double r[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) r[c] = (lane < nb && c <= lane) ? A[(k0 + lane) * n + k0 + c] : 0.0;
  double ld;
  int64_t fail;
  chol_factor_regs(r, nb, k0, lane, ld, fail);
#pragma unroll
  for (int c = 0; c < NB; ++c)
    if (lane < nb && c <= lane) A[(k0 + lane) * n + k0 + c] = r[c];
  if (lane == 0) {
    ldp[(int64_t)j * maxsteps + kstep] = ld;
    failp[(int64_t)j * maxsteps + kstep] = fail;
  }
}

// Panel: rows [k0 + nb + 256 rb, +256) of matrix j solve X L_kk^T = A[rows, k0:k0+nb], one thread per row
// kept in its own LDS row (no cross-thread traffic, no barriers after the L_kk load).
__device__ inline void chol_panel_dev(const SampleDesc* __restrict__ sd, int j, int rb, int64_t k0,
                                      double* __restrict__ mats, double (*Lkk)[NB + 1], double (*Xs)[NB + 1]) {
  int64_t n;
  double* A = mat_ptr(sd, mats, j, n);
  const int nb = (int)std::min<int64_t>(NB, n - k0);
  const int tid = threadIdx.x;
  for (int x = tid; x < NB * NB; x += 256) {
    const int i = x / NB, c = x % NB;
    const double v = (i < nb && c <= i) ? A[(k0 + i) * n + k0 + c] : 0.0;
    Lkk[i][c] = v;
    if (c == i) Lkk[i][NB] = i < nb ? 1.0 / v : 0.0;  // the pivot's reciprocal (the row's padding column)
  }
  __syncthreads();
  const int64_t i = k0 + nb + (int64_t)rb * 256 + tid;
  if (i < n) {
    double* row = A + i * n + k0;
    for (int c = 0; c < nb; ++c) Xs[tid][c] = row[c];
    for (int c = 0; c < nb; ++c) {
      double v = Xs[tid][c];
      for (int l = 0; l < c; ++l) v -= Xs[tid][l] * Lkk[c][l];
      Xs[tid][c] = v * Lkk[c][NB];
    }
    for (int c = 0; c < nb; ++c) row[c] = Xs[tid][c];
  }
}

constexpr int TT = 64;  // trailing tile
static_assert(TT >= NB, "the first trailing tile must hold the next diagonal block");

// Trailing tile t (lower-triangle order, tiles from row and column r0) of matrix j: A22 -= P P^T with P the
// panel of step k0 on v_mfma_f64_16x16x4, the two panel row blocks staged in LDS; wave w owns the 32 x 32 quarter (w >> 1, w & 1) of the 64 x 64 tile, so wave 0
// holds the tile's first 32 x 32 diagonal block (the next step's diagonal block when t = 0).
__device__ inline void chol_trail_dev(const SampleDesc* __restrict__ sd, int j, int64_t t, int64_t k0, int64_t r0,
                                      double* __restrict__ mats, double (*PI)[NB + 1], double (*PJ)[NB + 1]) {
  int64_t n;
  double* A = mat_ptr(sd, mats, j, n);
  const int nb = (int)std::min<int64_t>(NB, n - k0);
  int64_t bi = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (bi * (bi + 1) / 2 > t) --bi;
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  const int64_t bj = t - bi * (bi + 1) / 2;
  const int64_t R0 = r0 + bi * TT, C0 = r0 + bj * TT;
  const int tid = threadIdx.x;
  for (int x = tid; x < TT * NB; x += 256) {
    const int i = x / NB, p = x % NB;
    PI[i][p] = (R0 + i < n && p < nb) ? -A[(R0 + i) * n + k0 + p] : 0.0;
    PJ[i][p] = (C0 + i < n && p < nb) ? A[(C0 + i) * n + k0 + p] : 0.0;
  }
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  if (bi == bj && wj > wi) return;  // above the diagonal of a diagonal tile
  const int lr = lane & 15, lk = lane >> 4;
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = R0 + wi * 32 + a * 16 + lk + 4 * r, gj = C0 + wj * 32 + b * 16 + lr;
        acc[a][b][r] = (gi < n && gj <= gi) ? A[gi * n + gj] : 0.0;
      }
#pragma unroll
  for (int ks = 0; ks < NB / 4; ++ks) {
    const int p = ks * 4 + lk;
    const double fa0 = PI[wi * 32 + lr][p], fa1 = PI[wi * 32 + 16 + lr][p];
    const double fb0 = PJ[wj * 32 + lr][p], fb1 = PJ[wj * 32 + 16 + lr][p];
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
        const int64_t gi = R0 + wi * 32 + a * 16 + lk + 4 * r, gj = C0 + wj * 32 + b * 16 + lr;
        if (gi < n && gj <= gi) A[gi * n + gj] = acc[a][b][r];
      }
}

__global__ void __launch_bounds__(64) chol_diag_kernel(const SampleDesc* __restrict__ sd,
                                                       const StepEntry* __restrict__ e, int64_t k0, int kstep,
                                                       int maxsteps, double* __restrict__ mats,
                                                       double* __restrict__ ldp, int64_t* __restrict__ failp) {
  chol_diag_dev(sd, e[blockIdx.x].mat, k0, kstep, maxsteps, mats, ldp, failp, threadIdx.x);
}

__global__ void __launch_bounds__(256) chol_panel_kernel(const SampleDesc* __restrict__ sd,
                                                         const StepEntry* __restrict__ e, int ne, int64_t k0,
                                                         double* __restrict__ mats) {
  __shared__ double Lkk[NB][NB + 1];
  __shared__ double Xs[256][NB + 1];
  const int ei = find_entry(e, ne, (int)blockIdx.x, false);
  chol_panel_dev(sd, e[ei].mat, (int)blockIdx.x - e[ei].pfx_p, k0, mats, Lkk, Xs);
}

__global__ void __launch_bounds__(256) chol_trail_kernel(const SampleDesc* __restrict__ sd,
                                                         const StepEntry* __restrict__ e, int ne, int64_t k0,
                                                         double* __restrict__ mats) {
  __shared__ double PI[TT][NB + 1];
  __shared__ double PJ[TT][NB + 1];
  const int ei = find_entry(e, ne, (int)blockIdx.x, true);
  int64_t n;
  mat_ptr(sd, mats, e[ei].mat, n);
  chol_trail_dev(sd, e[ei].mat, (int64_t)blockIdx.x - e[ei].pfx_t, k0, k0 + std::min<int64_t>(NB, n - k0), mats, PI,
                 PJ);
}

// One launch per step with lookahead (mode 3).  Launch k assumes the panel of step k (columns k0 .. k0 + 31,
// every row below) final and everything right of it updated by the steps before k; it runs two task kinds
// side by side:
//   next   (one per STRIP rows below the next diagonal block, at least one per matrix with a step k + 1):
//          the step-k update of the diagonal block k + 1 and of its STRIP rows' columns k0 + 32 .. k0 + 63,
//          the block's factorisation (every task of the matrix factors it; the first stores it), and the
//          panel solve of its rows: the panel of step k + 1, final when the launch ends;
//   trail  the step-k update of the rest, rows and columns from k0 + 64, in 64 x 64 tiles.
// Each element sees the same operations in the same order as in the step-synchronous form (same MFMA
// operands and accumulation order, same substitutions), so the factors are bitwise the same, in
// maxsteps + 2 launches instead of 3 maxsteps + 1.
constexpr int STRIP = 128;

__device__ inline void chol_next_dev(const SampleDesc* __restrict__ sd, int j, int rb, int64_t k0, int kstep,
                                     int maxsteps, double* __restrict__ mats, double* __restrict__ ldp,
                                     int64_t* __restrict__ failp, double* __restrict__ stage, double (*Pd)[NB + 1],
                                     double (*Pr)[NB + 1], double (*Lkk)[NB + 1]) {
  int64_t n;
  double* A = mat_ptr(sd, mats, j, n);
  const int nb = (int)std::min<int64_t>(NB, n - k0);
  const int64_t k1 = k0 + nb;  // the next diagonal block
  const int nb1 = (int)std::min<int64_t>(NB, n - k1);
  const int64_t s0 = k1 + nb1 + (int64_t)rb * STRIP;  // first strip row
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  double* stg = stage + (int64_t)j * NB * NB;
  // The factored block k + 1 cannot go into A during this launch: the matrix's other tasks read the block's
  // values from before the update.  The first task keeps it in the matrix's stage slot and, here, moves
  // the block staged by the previous launch (block k) into place; no task of this launch reads block k.
  if (rb == 0 && kstep > 0)
    for (int x = tid; x < NB * NB; x += 256) {
      const int i = x / NB, c = x % NB;
      if (i < nb && c <= i) A[(k0 + i) * n + k0 + c] = stg[x];
    }
  for (int x = tid; x < NB * NB; x += 256) {
    const int i = x / NB, p = x % NB;
    Pd[i][p] = (i < nb1 && p < nb) ? A[(k1 + i) * n + k0 + p] : 0.0;
  }
  for (int x = tid; x < STRIP * NB; x += 256) {
    const int i = x / NB, p = x % NB;
    Pr[i][p] = (s0 + i < n && p < nb) ? A[(s0 + i) * n + k0 + p] : 0.0;
  }
  __syncthreads();
  if (wave == 0) {
    // the diagonal block's update, as wave 0 of the step-k trailing tile that holds it
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t gi = k1 + a * 16 + lk + 4 * r, gj = k1 + b * 16 + lr;
          acc[a][b][r] = (gi < n && gj <= gi) ? A[gi * n + gj] : 0.0;
        }
#pragma unroll
    for (int ks = 0; ks < NB / 4; ++ks) {
      const int p = ks * 4 + lk;
      const double fa0 = -Pd[lr][p], fa1 = -Pd[16 + lr][p];
      const double fb0 = Pd[lr][p], fb1 = Pd[16 + lr][p];
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
          const int i = a * 16 + lk + 4 * r, c = b * 16 + lr;
          Lkk[i][c] = (c <= i && k1 + i < n) ? acc[a][b][r] : 0.0;
        }
  }
  __syncthreads();
  if (wave == 0) {
    double r[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) r[c] = (lane < nb1 && c <= lane) ? Lkk[lane][c] : 0.0;
    double ld;
    int64_t fail;
    chol_factor_regs(r, nb1, k1, lane, ld, fail);
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (lane < NB) Lkk[lane][c] = (lane < nb1 && c <= lane) ? r[c] : 0.0;
    // the pivot's reciprocal (the row's padding column), as the other schedules form it from the factor
    if (lane < NB) Lkk[lane][NB] = lane < nb1 ? 1.0 / Lkk[lane][lane] : 0.0;
    if (rb == 0) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (lane < NB) stg[lane * NB + c] = r[c];
      if (lane == 0) {
        ldp[(int64_t)j * maxsteps + kstep + 1] = ld;
        failp[(int64_t)j * maxsteps + kstep + 1] = fail;
      }
    }
  }
  // the strip's step-k update, wave w on rows 32 w .. 32 w + 31 (as the trailing tiles' quarters that hold
  // them compute it), into Pr's rows (each wave overwrites only the rows it has read)
  {
    const int w0 = wave * 32;
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t gi = s0 + w0 + a * 16 + lk + 4 * r, gj = k1 + b * 16 + lr;
          acc[a][b][r] = (gi < n && b * 16 + lr < nb1) ? A[gi * n + gj] : 0.0;
        }
#pragma unroll
    for (int ks = 0; ks < NB / 4; ++ks) {
      const int p = ks * 4 + lk;
      const double fa0 = -Pr[w0 + lr][p], fa1 = -Pr[w0 + 16 + lr][p];
      const double fb0 = Pd[lr][p], fb1 = Pd[16 + lr][p];
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
        for (int r = 0; r < 4; ++r) Pr[w0 + a * 16 + lk + 4 * r][b * 16 + lr] = acc[a][b][r];
  }
  __syncthreads();
  // the panel solve of step k + 1 for the strip's rows: X L_k+1^T = strip, one thread per row
  if (tid < STRIP && s0 + tid < n) {
    for (int c = 0; c < nb1; ++c) {
      double v = Pr[tid][c];
      for (int l = 0; l < c; ++l) v -= Pr[tid][l] * Lkk[c][l];
      Pr[tid][c] = v * Lkk[c][NB];
    }
    double* row = A + (s0 + tid) * n + k1;
    for (int c = 0; c < nb1; ++c) row[c] = Pr[tid][c];
  }
}

__global__ void __launch_bounds__(256) chol_step_kernel(const SampleDesc* __restrict__ sd,
                                                        const StepEntry* __restrict__ e, int ne, int n_next,
                                                        int64_t k0, int kstep, int maxsteps,
                                                        double* __restrict__ mats, double* __restrict__ ldp,
                                                        int64_t* __restrict__ failp, double* __restrict__ stage) {
  __shared__ double smem[2 * NB * (NB + 1) + STRIP * (NB + 1)];
  double(*Pd)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(smem);
  double(*Lkk)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(smem + NB * (NB + 1));
  double(*Pr)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(smem + 2 * NB * (NB + 1));
  const int t = (int)blockIdx.x;
  if (t < n_next) {
    const int ei = find_entry(e, ne, t, false);
    chol_next_dev(sd, e[ei].mat, t - e[ei].pfx_p, k0, kstep, maxsteps, mats, ldp, failp, stage, Pd, Pr, Lkk);
  } else {
    const int ei = find_entry(e, ne, t - n_next, true);
    const int j = e[ei].mat;
    int64_t n;
    mat_ptr(sd, mats, j, n);
    chol_trail_dev(sd, j, (int64_t)(t - n_next) - e[ei].pfx_t, k0, k0 + 2 * NB, mats,
                   reinterpret_cast<double(*)[NB + 1]>(smem), reinterpret_cast<double(*)[NB + 1]>(smem + TT * (NB + 1)));
  }
}
static_assert(2 * NB * (NB + 1) + STRIP * (NB + 1) >= 2 * TT * (NB + 1), "trailing panels must fit the step kernel");
static_assert(STRIP <= 256 && STRIP == 4 * 32, "four waves of 32 strip rows");

// logdet[j] = 2 sum_k ldp[j][k] (+ the deflation correction of H), info[j] = first failing column.
__device__ inline void chol_finish_dev(const SampleDesc* __restrict__ sd, int j, int maxsteps,
                                       const double* __restrict__ ldp, const int64_t* __restrict__ failp,
                                       double* __restrict__ logdet, int64_t* __restrict__ info) {
  const SampleDesc d = sd[j >> 1];
  const int steps = (int)((d.S + NB - 1) / NB);
  double s = 0.0;
  int64_t f = 0;
  for (int k = 0; k < steps; ++k) {
    s += ldp[(int64_t)j * maxsteps + k];
    if (!f) f = failp[(int64_t)j * maxsteps + k];
  }
  s *= 2.0;
  if ((j & 1) && d.beta > 0.0) s += log(1e-8) - log(1e-8 + d.beta);
  logdet[j] = s;
  info[j] = f;
}

__global__ void chol_finish_kernel(const SampleDesc* __restrict__ sd, int nmat, int maxsteps,
                                   const double* __restrict__ ldp, const int64_t* __restrict__ failp,
                                   double* __restrict__ logdet, int64_t* __restrict__ info) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nmat) chol_finish_dev(sd, j, maxsteps, ldp, failp, logdet, info);
}

// After the lookahead (or left-looking) launches: every matrix's last diagonal block (from block
// first_staged on, each is staged by the launch that factors it) moved from its stage slot into place.
// One workgroup per matrix.
__global__ void chol_unstage_kernel(const SampleDesc* __restrict__ sd, double* __restrict__ mats,
                                    const double* __restrict__ stage, int first_staged) {
  const int j = blockIdx.x;
  int64_t n;
  double* A = mat_ptr(sd, mats, j, n);
  const int64_t steps = (n + NB - 1) / NB;
  if (steps < 1 + first_staged) return;
  const int64_t kl = (steps - 1) * NB;
  const int nb = (int)(n - kl);
  const double* stg = stage + (int64_t)j * NB * NB;
  for (int x = threadIdx.x; x < NB * NB; x += blockDim.x) {
    const int i = x / NB, c = x % NB;
    if (i < nb && c <= i) A[(kl + i) * n + kl + c] = stg[x];
  }
}

// Left-looking form (mode 4): one launch per 32-column block k after a launch that factors every
// diagonal block 0.  A task takes strip rows of block k below its diagonal block and applies to them the
// updates of every earlier block p < k (A[rows, k] -= L[rows, p] L[kblock, p]^T, accumulated in
// registers over p; the panels of each p staged in LDS while the next p's loads are in flight), then
// solves them against L_kk, one row per lane with the row in registers.  The trailing matrix is never
// written between steps: each element of the factor is written once.  The first task of a matrix
// (96 strip rows; the others take 128) also forms the next diagonal block k + 1 (its updates by the blocks
// p < k from the same staged panels, the one by block k from its own solved rows, which are that block's
// rows) and factors it, for the next launch: it goes to a stage slot (parity k + 1), the next launch's
// tasks read it from there and that launch's first task moves it into place (no task of that launch
// reads the block in place).  Each element sees the same MFMA operands, in the same order, and the same
// substitutions as in the step-synchronous form, so the factor is bitwise the same.
constexpr int LL_STRIP = 128;            // strip rows of a task (one launch per step): four waves of 32 rows
constexpr int LL_PITCH = NB + 2;         // LDS row pitch (doubles): the fragment reads are free of conflicts
constexpr int LL_THREADS = 256;

// A task's shape for STRIP strip rows: a wave per 32 rows; the first task of a matrix takes STRIP - 32
// rows and its last wave forms the next diagonal block.
template <int STRIP>
struct LeftShape {
  static constexpr int STRIP0 = STRIP - NB;          // strip rows of the first task
  static constexpr int ROWS = STRIP + NB;            // staged rows: the strip and the diagonal block's rows
  static constexpr int THREADS = 2 * STRIP;
  static constexpr int WAVES = THREADS / 64;
  static constexpr int PASS = THREADS / NB;          // staged rows per load pass
  static constexpr int LOADS = ROWS * NB / THREADS;  // doubles per thread per staged block
  static_assert(LOADS * THREADS == ROWS * NB && LOADS <= 32 && STRIP % 32 == 0, "whole load passes");
};
static_assert(LeftShape<LL_STRIP>::THREADS == LL_THREADS, "one-launch-per-step tasks: four waves");

// x L^T = b for one row held in registers (x = b on entry), L the nb x nb lower-triangular block in LDS:
// x_c = (b_c - sum_{l<c} x_l L_cl) / L_cc with the sum taken in l order, column by column.
__device__ inline void solve_row_lkk(double (&x)[NB], int nb, const lds_f64 (*__restrict__ Lkk)[NB + 1]) {
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    if (c < nb) {
      double v = x[c];
#pragma unroll
      for (int l = 0; l < c; ++l) v -= x[l] * Lkk[c][l];
      x[c] = v * Lkk[c][NB];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Index (doubles) of element (r, c) in a staged block of NB-double rows whose 16-byte segments are stored
// XOR (r mod 16), so that the 16 rows of a fragment read fall in 16 different segments.
__device__ inline int ll_at(int r, int c) { return r * NB + ((((c >> 1) ^ (r & 15)) << 1) | (c & 1)); }

constexpr int LL_OOB = 0x7ffff000;  // past any matrix's bytes (n <= SOLVE_S_MAX), also 47 blocks on

// The persistent schedule's second wait of a task (lookahead): before its last block (p = k - 1) and the
// diagonal block L_kk, both from block column k - 1, a task waits until matrix mat has `need` completed
// tasks (all of columns <= k - 1); it starts once the columns <= k - 2 are complete.  flow_wait is defined
// with the persistent kernel below.
struct FlowWait {
  int* sync;
  int mat, need, spin_limit;
  long long* tr;  // the task's trace record (phase times at tr[8 ..]), or null
};
// wave 0 records a phase time of the traced task (every lane the same value)
#define AJ_FLOW_MARK(FW, SLOT)                                                          \
  do {                                                                                  \
    if ((FW).tr != nullptr && uni((int)(threadIdx.x >> 6)) == 0) (FW).tr[SLOT] = wall_clock64(); \
  } while (0)
// the calling wave records a phase time of the traced task
#define AJ_FLOW_MARK_WAVE(FW, SLOT)                                   \
  do {                                                                \
    if ((FW).tr != nullptr) (FW).tr[SLOT] = wall_clock64();           \
  } while (0)
__device__ void flow_wait(int* __restrict__ sync, int mat, int need, int spin_limit);

// Task rb of block column k (k0 = 32 k) of matrix j; Ls (ROWS x LL_PITCH doubles), Lkk and colb (NB doubles)
// in the caller's LDS; threadIdx.x < THREADS.  LA: the persistent schedule's lookahead (fw).
template <int STRIP, bool LA = false>
__device__ __attribute__((always_inline)) inline void chol_left_task(const SampleDesc* __restrict__ sd, int j, int rb, int64_t k0, int kstep,
                                      int maxsteps, double* __restrict__ mats, double* __restrict__ ldp,
                                      int64_t* __restrict__ failp, double* __restrict__ stage,
                                      lds_f64* __restrict__ Ls, lds_f64 (*__restrict__ Lkk)[NB + 1],
                                      lds_f64* __restrict__ colb, FlowWait fw = FlowWait{}) {
  using SH = LeftShape<STRIP>;
  int64_t n;
  double* A = mat_ptr(sd, mats, j, n);
  const int nb = (int)std::min<int64_t>(NB, n - k0);
  const bool next = rb == 0 && k0 + NB < n;  // this task forms diagonal block k + 1
  const int nb1 = next ? (int)std::min<int64_t>(NB, n - k0 - NB) : 0;
  const int64_t s0 = k0 + nb + (rb == 0 ? 0 : SH::STRIP0 + (int64_t)(rb - 1) * STRIP);  // first strip row
  const int rows = (int)std::max<int64_t>(0, std::min<int64_t>(rb == 0 ? SH::STRIP0 : STRIP, n - s0));
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const double* stg_in = stage + ((int64_t)2 * j + (kstep & 1)) * NB * NB;
  double* stg_out = stage + ((int64_t)2 * j + ((kstep + 1) & 1)) * NB * NB;
  // L_kk: factored in place by the first launch (step 0) or staged by the previous launch (with lookahead:
  // read after the second wait)
  auto load_lkk = [&]() {
    for (int x = tid; x < NB * NB; x += SH::THREADS) {
      const int i = x / NB, c = x % NB;
      double v = 0.0;
      if (i < nb && c <= i) v = kstep == 0 ? A[(k0 + i) * n + k0 + c] : stg_in[x];
      Lkk[i][c] = v;
      if (c == i) Lkk[i][NB] = i < nb ? 1.0 / v : 0.0;  // the pivot's reciprocal (the row's padding column)
      if (rb == 0 && kstep > 0 && i < nb && c <= i) A[(k0 + i) * n + k0 + c] = v;
    }
  };
  if (!LA || kstep == 0) load_lkk();
  // wave w < WAVES - 1 (and the last wave when the task does not form the next block) owns strip rows
  // 32 w .. 32 w + 31; the last wave of a task that does forms block k + 1 from strip rows 0 .. 31 (its rows
  // k0 + 32 ..)
  const bool dwave = next && wave == SH::WAVES - 1;
  const int w0 = wave * 32;
  const bool active = dwave || w0 < rows;
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = a * 16 + lk + 4 * r, c = b * 16 + lr;
        double v = 0.0;
        if (dwave) {
          if (i < nb1 && c <= i) v = A[(k0 + NB + i) * n + k0 + NB + c];
        } else if (w0 + i < rows && c < nb) {
          v = A[(s0 + w0 + i) * n + k0 + c];
        }
        acc[a][b][r] = v;
      }
  if (kstep > 0) {
    // staging: thread tid loads column lc of LDS rows lr0 + PASS q: rows < STRIP are strip rows s0 + r (zero
    // past the task's rows), the last NB the diagonal block's rows k0 + i (zero past nb); block p + 1's
    // loads are in flight during block p's MFMAs.
    // (raw buffer loads: rows outside take an offset past the matrix and read as zeros)
    const int lc = tid & 31, lr0 = tid >> 5;
    const i32x4 rsrc = make_rsrc(A, (uint32_t)(n * n * (int64_t)sizeof(double)));
    int voff[SH::LOADS];
#pragma unroll
    for (int q = 0; q < SH::LOADS; ++q) {
      const int r = lr0 + SH::PASS * q;
      const bool ok = r < STRIP ? r < rows || (next && r < nb1) : r - STRIP < nb;
      const int64_t grow = r < STRIP ? s0 + r : k0 + (r - STRIP);
      voff[q] = ok ? (int)((grow * n + lc) * (int64_t)sizeof(double)) : LL_OOB;
    }
    double pre[SH::LOADS];
    const int last = kstep - 1;
    if (!LA || last > 0) {
#pragma unroll
      for (int q = 0; q < SH::LOADS; ++q) pre[q] = raw_buffer_load_f64(rsrc, voff[q], 0, 0);
    }
    if (LA) AJ_FLOW_MARK(fw, 8);
    for (int p = 0; p < kstep; ++p) {
      if (LA && p == last) {
        // lookahead: block k - 1 and L_kk once block column k - 1 is complete
        AJ_FLOW_MARK(fw, 9);
        flow_wait(fw.sync, fw.mat, fw.need, fw.spin_limit);
        AJ_FLOW_MARK(fw, 10);
        load_lkk();
        const int off = p * NB * (int)sizeof(double);
#pragma unroll
        for (int q = 0; q < SH::LOADS; ++q) pre[q] = raw_buffer_load_f64(rsrc, voff[q], off, 0);
      }
      __syncthreads();  // every wave is past its fragment reads of block p - 1
#pragma unroll
      for (int q = 0; q < SH::LOADS; ++q) Ls[(lr0 + SH::PASS * q) * LL_PITCH + lc] = pre[q];
      __syncthreads();
      if (p + 1 < kstep && !(LA && p + 1 == last)) {
        const int off = (p + 1) * NB * (int)sizeof(double);
#pragma unroll
        for (int q = 0; q < SH::LOADS; ++q) pre[q] = raw_buffer_load_f64(rsrc, voff[q], off, 0);
      }
      if (active) {
        // strip waves: rows w0 .. w0 + 31 against the diagonal block's rows; the block-(k + 1) wave: strip
        // rows 0 .. 31 against themselves (its block (0, 1), above the diagonal, is not used)
        const int ra = dwave ? 0 : w0, rbb = dwave ? 0 : STRIP;
#pragma unroll
        for (int ks = 0; ks < NB / 4; ++ks) {
          const int pc = ks * 4 + lk;
          const double fa0 = -Ls[(ra + lr) * LL_PITCH + pc], fa1 = -Ls[(ra + 16 + lr) * LL_PITCH + pc];
          const double fb0 = Ls[(rbb + lr) * LL_PITCH + pc], fb1 = Ls[(rbb + 16 + lr) * LL_PITCH + pc];
          acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa0, fb0, acc[0][0], 0, 0, 0);
          acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa0, fb1, acc[0][1], 0, 0, 0);
          acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1, fb0, acc[1][0], 0, 0, 0);
          acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1, fb1, acc[1][1], 0, 0, 0);
        }
      }
    }
  }
  // the strip rows into Xs (the staging buffer, free after this barrier)
  lds_f64* Xs = Ls;
  if (LA) AJ_FLOW_MARK(fw, 11);
  __syncthreads();
  if (!dwave) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) Xs[(w0 + a * 16 + lk + 4 * r) * LL_PITCH + b * 16 + lr] = acc[a][b][r];
  }
  __syncthreads();
  // X L_kk^T = (updated rows): lanes 0 .. 31 of wave w, row w0 + lane, the row in registers
  if (!dwave && lane < 32 && w0 + lane < rows) {
    lds_f64* xr = Xs + (w0 + lane) * LL_PITCH;
    double x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = xr[c];
    solve_row_lkk(x, nb, Lkk);
#pragma unroll
    for (int c = 0; c < NB; ++c) xr[c] = x[c];
  }
  __syncthreads();
  if (LA) AJ_FLOW_MARK(fw, 12);
  for (int x = tid; x < rows * NB; x += SH::THREADS) {
    const int i = x / NB, c = x % NB;
    if (c < nb) A[(s0 + i) * n + k0 + c] = Xs[i * LL_PITCH + c];
  }
  if (!dwave) return;
  // block k + 1: the update by block k (its rows' solved values: strip rows 0 .. nb1 - 1), then the
  // factorisation, into the stage slot of step k + 1
#pragma unroll
  for (int ks = 0; ks < NB / 4; ++ks) {
    const int pc = ks * 4 + lk;
    const double fa0 = -Xs[lr * LL_PITCH + pc], fa1 = -Xs[(16 + lr) * LL_PITCH + pc];
    const double fb0 = Xs[lr * LL_PITCH + pc], fb1 = Xs[(16 + lr) * LL_PITCH + pc];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa0, fb0, acc[0][0], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1, fb0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1, fb1, acc[1][1], 0, 0, 0);
  }
  // Lkk is free: the other waves are past the substitution (the barrier above)
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = a * 16 + lk + 4 * r, c = b * 16 + lr;
        Lkk[i][c] = (b <= a && c <= i && i < nb1) ? acc[a][b][r] : 0.0;
      }
  if (LA) AJ_FLOW_MARK_WAVE(fw, 13);
  double r[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) r[c] = (lane < nb1 && c <= lane) ? Lkk[lane][c] : 0.0;
  double ld;
  int64_t fail;
  chol_factor_lds(r, nb1, k0 + NB, lane, ld, fail, colb);
  if (LA) AJ_FLOW_MARK_WAVE(fw, 14);
#pragma unroll
  for (int c = 0; c < NB; ++c)
    if (lane < NB) stg_out[lane * NB + c] = r[c];
  if (LA && fw.tr != nullptr) {
    __builtin_amdgcn_s_waitcnt(0);  // the stage slot's stores have landed
    AJ_FLOW_MARK_WAVE(fw, 15);
  }
  if (lane == 0) {
    ldp[(int64_t)j * maxsteps + kstep + 1] = ld;
    failp[(int64_t)j * maxsteps + kstep + 1] = fail;
  }
}

__global__ void __launch_bounds__(LL_THREADS) chol_left_kernel(const SampleDesc* __restrict__ sd,
                                                              const StepEntry* __restrict__ e, int ne, int64_t k0,
                                                              int kstep, int maxsteps, double* __restrict__ mats,
                                                              double* __restrict__ ldp, int64_t* __restrict__ failp,
                                                              double* __restrict__ stage) {
  __shared__ double Ls[LeftShape<LL_STRIP>::ROWS * LL_PITCH];
  __shared__ double Lkk[NB][NB + 1];
  __shared__ double colb[NB];
  const int ei = find_entry(e, ne, (int)blockIdx.x, false);
  chol_left_task<LL_STRIP>(sd, e[ei].mat, (int)blockIdx.x - e[ei].pfx_p, k0, kstep, maxsteps, mats, ldp, failp,
                           stage, (lds_f64*)Ls, (lds_f64(*)[NB + 1])Lkk, (lds_f64*)colb);
}
static_assert(sizeof(double) * (LeftShape<LL_STRIP>::ROWS * LL_PITCH + NB * (NB + 1) + NB) <= 54 * 1024,
              "three workgroups per CU");

// The same step for the last steps, when a matrix has few rows left below its diagonal block and a task's
// time is its chain of blocks p < k: 32-row tasks, each 16 x 16 block of the strip (and of the next
// diagonal block, in the first task) on its own wave, so that a block p costs 8 MFMAs per wave; the
// staged rows (32 strip rows, then the diagonal block's 32) arrive by direct global -> LDS loads five blocks
// ahead of their use.  Same operations per element as chol_left_kernel, so the same bits.
constexpr int LS_ROWS = 2 * NB;          // staged rows per block: the strip's 32, the diagonal block's 32
constexpr int LS_STAGES = 6;             // ring of staged blocks: the one read and five in flight
constexpr int LS_THREADS = 512;          // waves 0-3: the strip's four 16 x 16 blocks; 4-7: the next diagonal block's
constexpr int LS_LOADS = LS_ROWS / 4 / (LS_THREADS / 64);  // 1 KB loads per wave per block
static_assert(LS_LOADS == 2 && LS_STAGES * LS_ROWS * NB * 8 + NB * (NB + 1) * 8 * 2 <= 160 * 1024, "small-task LDS");

// Task rb of block column k (k0 = 32 k) of matrix j; the ring Ls (LS_STAGES blocks of LS_ROWS x NB doubles,
// 1 KB aligned), Lkk, Dn and colb (NB doubles) in the caller's LDS; threadIdx.x < LS_THREADS.  LA: the
// persistent schedule's lookahead (fw).
template <bool LA = false>
__device__ __attribute__((always_inline)) inline void chol_small_task(const SampleDesc* __restrict__ sd, int j, int rb, int64_t k0, int kstep,
                                       int maxsteps, double* __restrict__ mats, double* __restrict__ ldp,
                                       int64_t* __restrict__ failp, double* __restrict__ stage,
                                       lds_f64* __restrict__ Ls, lds_f64 (*__restrict__ Lkk)[NB + 1],
                                       lds_f64 (*__restrict__ Dn)[NB + 1], lds_f64* __restrict__ colb,
                                       FlowWait fw = FlowWait{}) {
  int64_t n;
  double* A = mat_ptr(sd, mats, j, n);
  const int nb = (int)std::min<int64_t>(NB, n - k0);
  const bool next = rb == 0 && k0 + NB < n;
  const int nb1 = next ? (int)std::min<int64_t>(NB, n - k0 - NB) : 0;
  const int64_t s0 = k0 + nb + (int64_t)rb * NB;  // first strip row
  const int rows = (int)std::max<int64_t>(0, std::min<int64_t>(NB, n - s0));
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const double* stg_in = stage + ((int64_t)2 * j + (kstep & 1)) * NB * NB;
  double* stg_out = stage + ((int64_t)2 * j + ((kstep + 1) & 1)) * NB * NB;
  auto load_lkk = [&]() {
    for (int x = tid; x < NB * NB; x += LS_THREADS) {
      const int i = x / NB, c = x % NB;
      double v = 0.0;
      if (i < nb && c <= i) v = kstep == 0 ? A[(k0 + i) * n + k0 + c] : stg_in[x];
      Lkk[i][c] = v;
      if (c == i) Lkk[i][NB] = i < nb ? 1.0 / v : 0.0;  // the pivot's reciprocal (the row's padding column)
      if (rb == 0 && kstep > 0 && i < nb && c <= i) A[(k0 + i) * n + k0 + c] = v;
    }
  };
  if (!LA || kstep == 0) load_lkk();  // with lookahead: after the second wait
  // wave w < 4: strip block (w >> 1, w & 1); wave w >= 4: block ((w - 4) >> 1, (w - 4) & 1) of the next
  // diagonal block (its rows k0 + 32 .. are the strip's rows 0 .. 31; block (0, 1) is above the diagonal
  // and not used)
  const bool dwave = wave >= 4;
  const int ba = (wave >> 1) & 1, bb = wave & 1;
  d4 acc;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = ba * 16 + lk + 4 * r, c = bb * 16 + lr;
    double v = 0.0;
    if (dwave) {
      if (i < nb1 && c <= i) v = A[(k0 + NB + i) * n + k0 + NB + c];
    } else if (i < rows && c < nb) {
      v = A[(s0 + i) * n + k0 + c];
    }
    acc[r] = v;
  }
  if (kstep > 0) {
    // wave w issues LDS rows 4 (2 w + jj) .. + 3 of each block (rows < 32: strip rows s0 + r, the others the
    // diagonal block's rows k0 + r - 32); rows outside read as zeros
    const i32x4 rsrc = make_rsrc(A, (uint32_t)(n * n * (int64_t)sizeof(double)));
    int voff[LS_LOADS];
#pragma unroll
    for (int jj = 0; jj < LS_LOADS; ++jj) {
      const int r = 4 * (LS_LOADS * wave + jj) + (lane >> 4);
      const bool ok = r < NB ? r < rows || (next && r < nb1) : r - NB < nb;
      const int64_t grow = r < NB ? s0 + r : k0 + (r - NB);
      const int g = (lane & 15) ^ (r & 15);
      voff[jj] = ok ? (int)((grow * n + 2 * g) * (int64_t)sizeof(double)) : -1;
    }
    constexpr int LS_OOB = 0x7ffff000;  // past any matrix's bytes (n <= SOLVE_S_MAX)
    // (with lookahead block k - 1 is issued as zeros here and loaded after the second wait)
#define AJ_LS_ISSUE_IF(P, COND)                                                                            \
  {                                                                                                        \
    lds_f64* lb_ = Ls + ((P) % LS_STAGES) * (LS_ROWS * NB);                                                \
    _Pragma("unroll") for (int jj = 0; jj < LS_LOADS; ++jj) raw_buffer_load_lds(                          \
        rsrc, (lds_u32*)(lb_ + (LS_LOADS * wave + jj) * 128), 16,                                          \
        (COND) && voff[jj] >= 0 ? voff[jj] + (P) * NB * (int)sizeof(double) : LS_OOB, 0, 0, 0);           \
  }
#define AJ_LS_ISSUE(P) AJ_LS_ISSUE_IF(P, (P) < kstep && !(LA && (P) == kstep - 1))
    // branch-free: blocks past the last are issued as zeros, so each wait counts the same loads
#pragma unroll
    for (int q = 0; q < LS_STAGES - 1; ++q) AJ_LS_ISSUE(q)
    const int ra = 16 * ba, rbb = dwave ? 16 * bb : NB + 16 * bb;
    for (int p = 0; p < kstep; ++p) {
      if (LA && p == kstep - 1) {
        // lookahead: every earlier load (with this block's zeros) landed; once block column k - 1 is
        // complete, L_kk and this block
        __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
        flow_wait(fw.sync, fw.mat, fw.need, fw.spin_limit);
        load_lkk();
        AJ_LS_ISSUE_IF(p, true)
        __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
      }
      // block p has landed (blocks p + 1 .. p + 4 may be in flight); after the barrier every wave is past
      // its reads of block p - 1, whose buffer takes block p + 5
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(LS_LOADS * (LS_STAGES - 2)));
      asm volatile("s_barrier" ::: "memory");
      AJ_LS_ISSUE(p + LS_STAGES - 1)
      // the block's fragments as LDS loads in assembly, one group waiting for itself (the compiler would
      // wait for every direct-to-LDS load in flight before any LDS read)
      const uint32_t lb = (uint32_t)(uintptr_t)(lds_u32*)(Ls + (p % LS_STAGES) * (LS_ROWS * NB));
      uint32_t aa[NB / 4], ab[NB / 4];
#pragma unroll
      for (int ks = 0; ks < NB / 4; ++ks) {
        aa[ks] = lb + 8u * (uint32_t)ll_at(ra + lr, ks * 4 + lk);
        ab[ks] = lb + 8u * (uint32_t)ll_at(rbb + lr, ks * 4 + lk);
      }
      double fa[NB / 4], fb[NB / 4];
      asm volatile(
          "ds_read_b64 %0, %16\n\tds_read_b64 %1, %17\n\tds_read_b64 %2, %18\n\tds_read_b64 %3, %19\n\t"
          "ds_read_b64 %4, %20\n\tds_read_b64 %5, %21\n\tds_read_b64 %6, %22\n\tds_read_b64 %7, %23\n\t"
          "ds_read_b64 %8, %24\n\tds_read_b64 %9, %25\n\tds_read_b64 %10, %26\n\tds_read_b64 %11, %27\n\t"
          "ds_read_b64 %12, %28\n\tds_read_b64 %13, %29\n\tds_read_b64 %14, %30\n\tds_read_b64 %15, %31\n\t"
          "s_waitcnt lgkmcnt(0)"
          : "=&v"(fa[0]), "=&v"(fa[1]), "=&v"(fa[2]), "=&v"(fa[3]), "=&v"(fa[4]), "=&v"(fa[5]), "=&v"(fa[6]),
            "=&v"(fa[7]), "=&v"(fb[0]), "=&v"(fb[1]), "=&v"(fb[2]), "=&v"(fb[3]), "=&v"(fb[4]), "=&v"(fb[5]),
            "=&v"(fb[6]), "=&v"(fb[7])
          : "v"(aa[0]), "v"(aa[1]), "v"(aa[2]), "v"(aa[3]), "v"(aa[4]), "v"(aa[5]), "v"(aa[6]), "v"(aa[7]),
            "v"(ab[0]), "v"(ab[1]), "v"(ab[2]), "v"(ab[3]), "v"(ab[4]), "v"(ab[5]), "v"(ab[6]), "v"(ab[7])
          : "memory");
#pragma unroll
      for (int ks = 0; ks < NB / 4; ++ks) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-fa[ks], fb[ks], acc, 0, 0, 0);
    }
#undef AJ_LS_ISSUE
#undef AJ_LS_ISSUE_IF
    // the loads issued past the last block land before the buffers are reused
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
  }
  // the strip's rows into Xs, the next diagonal block into Dn
  lds_f64* Xs = Ls;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = ba * 16 + lk + 4 * r, c = bb * 16 + lr;
    if (dwave)
      Dn[i][c] = acc[r];
    else
      Xs[i * LL_PITCH + c] = acc[r];
  }
  __syncthreads();
  // X L_kk^T = (updated rows): lanes 0 .. 31 of wave 0, one row each, the row in registers
  if (wave == 0 && lane < rows) {
    lds_f64* xr = Xs + lane * LL_PITCH;
    double x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = xr[c];
    solve_row_lkk(x, nb, Lkk);
#pragma unroll
    for (int c = 0; c < NB; ++c) xr[c] = x[c];
  }
  __syncthreads();
  for (int x = tid; x < rows * NB; x += LS_THREADS) {
    const int i = x / NB, c = x % NB;
    if (c < nb) A[(s0 + i) * n + k0 + c] = Xs[i * LL_PITCH + c];
  }
  if (!next) return;
  // the next diagonal block: the update by block k (the strip's solved rows), then its factorisation
  if (dwave) {
#pragma unroll
    for (int ks = 0; ks < NB / 4; ++ks) {
      const int pc = ks * 4 + lk;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-Xs[(16 * ba + lr) * LL_PITCH + pc], Xs[(16 * bb + lr) * LL_PITCH + pc],
                                                 acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = ba * 16 + lk + 4 * r, c = bb * 16 + lr;
      Dn[i][c] = (bb <= ba && c <= i && i < nb1) ? acc[r] : 0.0;
    }
  }
  __syncthreads();
  if (wave == 4) {
    double r[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) r[c] = (lane < nb1 && c <= lane) ? Dn[lane][c] : 0.0;
    double ld;
    int64_t fail;
    chol_factor_lds(r, nb1, k0 + NB, lane, ld, fail, colb);
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (lane < NB) stg_out[lane * NB + c] = r[c];
    if (lane == 0) {
      ldp[(int64_t)j * maxsteps + kstep + 1] = ld;
      failp[(int64_t)j * maxsteps + kstep + 1] = fail;
    }
  }
}

__global__ void __launch_bounds__(LS_THREADS) chol_left_small_kernel(const SampleDesc* __restrict__ sd,
                                                                    const StepEntry* __restrict__ e, int ne,
                                                                    int64_t k0, int kstep, int maxsteps,
                                                                    double* __restrict__ mats, double* __restrict__ ldp,
                                                                    int64_t* __restrict__ failp,
                                                                    double* __restrict__ stage) {
  __shared__ __attribute__((aligned(1024))) double Ls[LS_STAGES * LS_ROWS * NB];
  __shared__ double Lkk[NB][NB + 1];
  __shared__ double Dn[NB][NB + 1];
  __shared__ double colb[NB];
  const int ei = find_entry(e, ne, (int)blockIdx.x, false);
  chol_small_task<false>(sd, e[ei].mat, (int)blockIdx.x - e[ei].pfx_p, k0, kstep, maxsteps, mats, ldp, failp, stage,
                  (lds_f64*)Ls, (lds_f64(*)[NB + 1])Lkk, (lds_f64(*)[NB + 1])Dn, (lds_f64*)colb);
}

// Persistent form of the left-looking schedule (mode 5): the whole factorisation of every matrix in one
// launch of one workgroup (eight waves) per compute unit.  Each matrix belongs to one XCD (the host
// balances their S^3 over the eight); each XCD has its own task list (its matrices' first diagonal blocks,
// then block column 0's tasks of its matrices, column 1's, ...) and queue head (an atomic counter).  A
// workgroup takes the next task of the queue of the XCD it runs on, waits until the tasks of that matrix
// in block columns <= k - 2 are complete (a per-matrix count of completed tasks), updates its strip by
// blocks 0 .. k - 2, waits for column k - 1 (lookahead: a column's chain of earlier blocks overlaps the
// previous column), applies block k - 1, solves and counts the task complete.  A task waits only for tasks
// earlier in its list, which running workgroups have taken, so every wait ends; the wait is bounded all the
// same (a spin limit), past which the task runs, the error flag is set and the solve reports the batch as
// failed (also when a matrix's count is short of its total, as it would be if an XCD ran none of the
// workgroups).  The tasks are those of the one-launch-per-step form (256 strip rows on eight waves, then,
// once an XCD's 32-row tasks of a column fit its compute units, 32-row tasks), so each element sees the
// same operations in the same order: the factor is bitwise that of the other schedules.  A matrix's data
// passes only through its XCD's L2: completion is published once every wave's stores have completed into
// that L2 (before the workgroup's barrier; no write-back beyond it), and observed with an agent-scope
// acquire, which invalidates the waiting CU's L1.  The queue's single-thread steps run on the whole of
// wave 0 (lane 0's value, atomics whose other lanes add 0), so the persistent loop has no divergent branch
// next to its barriers.
struct FlowTask {
  int mat;    // matrix j
  int kind;   // 0: its first diagonal block; 1: a 256-row task of block column k; 2: a 32-row task
  int k;      // block column
  int rb;     // task within the block column
  int need;   // completed tasks of matrix mat required before the task starts: its tasks in block columns
              // <= k - 2 (its first diagonal block's task for k <= 1)
  int need2;  // ... before its last block (lookahead): its tasks in block columns <= k - 1
};
constexpr int FL_STRIP = 256;
constexpr int FLOW_THREADS = 512;
static_assert(LeftShape<FL_STRIP>::THREADS == FLOW_THREADS && LS_THREADS == FLOW_THREADS, "eight waves per task");
constexpr int FLOW_RING = LS_STAGES * LS_ROWS * NB > LeftShape<FL_STRIP>::ROWS * LL_PITCH
                              ? LS_STAGES * LS_ROWS * NB
                              : LeftShape<FL_STRIP>::ROWS * LL_PITCH;
constexpr int FLOW_LDS_DOUBLES = FLOW_RING + 2 * NB * (NB + 1) + NB;
static_assert(FLOW_LDS_DOUBLES * sizeof(double) <= 160 * 1024, "one workgroup per CU");

// The waits of a task (every thread calls it; wave 0 polls, all its lanes reading the one counter): until
// matrix mat has `need` completed tasks, at most spin_limit polls (then the error flag is set), then an
// agent-scope acquire (the producers ran on this XCD: their stores are in its L2, and this CU's L1 may hold
// lines read before); the barrier holds the other waves until then.
__device__ void flow_wait(int* __restrict__ sync, int mat, int need, int spin_limit) {
  need = uni(need);
  if (uni((int)(threadIdx.x >> 6)) == 0 && need > 0) {
    const int* cnt = &sync[FLOW_SYNC_HEAD + uni(mat)];
    const int limit = uni(spin_limit);
    int spins = 0;
    while (uni(__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < need) {
      if (++spins > limit) {
        atomicOr(&sync[FLOW_ERR], (threadIdx.x & 63) == 0 ? 1 : 0);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// The three task kinds as separate functions, so that each keeps its own register allocation (inlined
// together they spill).  Their global pointers arrive typed as global memory, so that the inlined task
// bodies load and store with global (not flat) instructions.
typedef __attribute__((address_space(1))) double g_f64;
typedef __attribute__((address_space(1))) int64_t g_i64;
typedef __attribute__((address_space(1))) const SampleDesc g_desc;

__device__ __attribute__((noinline)) void flow_diag_task(g_desc* sd, int j, int maxsteps, g_f64* mats, g_f64* ldp,
                                                         g_i64* failp) {
  if ((threadIdx.x >> 6) == 0)
    chol_diag_dev((const SampleDesc*)uni(sd), uni(j), 0, 0, uni(maxsteps), (double*)uni(mats), (double*)uni(ldp),
                  (int64_t*)uni(failp), threadIdx.x & 63);
}

__device__ __attribute__((noinline)) void flow_left_task(g_desc* sd, int j, int rb, int k, int maxsteps, g_f64* mats,
                                                         g_f64* ldp, g_i64* failp, g_f64* stage, lds_f64* lds,
                                                         int* sync, int need2, int spin_limit, long long* tr) {
  k = uni(k);
  j = uni(j);
  chol_left_task<FL_STRIP, true>((const SampleDesc*)uni(sd), j, uni(rb), (int64_t)k * NB, k, uni(maxsteps),
                                 (double*)uni(mats), (double*)uni(ldp), (int64_t*)uni(failp), (double*)uni(stage), lds,
                                 reinterpret_cast<lds_f64(*)[NB + 1]>(lds + FLOW_RING),
                                 lds + FLOW_RING + 2 * NB * (NB + 1),
                                 FlowWait{uni(sync), j, uni(need2), uni(spin_limit), uni(tr)});
}

__device__ __attribute__((noinline)) void flow_small_task(g_desc* sd, int j, int rb, int k, int maxsteps, g_f64* mats,
                                                          g_f64* ldp, g_i64* failp, g_f64* stage, lds_f64* lds,
                                                          int* sync, int need2, int spin_limit, long long* tr) {
  k = uni(k);
  j = uni(j);
  chol_small_task<true>((const SampleDesc*)uni(sd), j, uni(rb), (int64_t)k * NB, k, uni(maxsteps), (double*)uni(mats),
                        (double*)uni(ldp), (int64_t*)uni(failp), (double*)uni(stage), lds,
                        reinterpret_cast<lds_f64(*)[NB + 1]>(lds + FLOW_RING),
                        reinterpret_cast<lds_f64(*)[NB + 1]>(lds + FLOW_RING + NB * (NB + 1)),
                        lds + FLOW_RING + 2 * NB * (NB + 1),
                        FlowWait{uni(sync), j, uni(need2), uni(spin_limit), uni(tr)});
}

__global__ void __launch_bounds__(FLOW_THREADS) chol_flow_kernel(const SampleDesc* __restrict__ sd,
                                                                 const FlowTask* __restrict__ tasks,
                                                                 const int* __restrict__ qoff,
                                                                 int maxsteps, double* __restrict__ mats,
                                                                 double* __restrict__ ldp,
                                                                 int64_t* __restrict__ failp,
                                                                 double* __restrict__ stage, int* __restrict__ sync,
                                                                 int spin_limit, long long* __restrict__ trace) {
  __shared__ __attribute__((aligned(1024))) double lds[FLOW_LDS_DOUBLES];
  __shared__ int cur;
  g_desc* gsd = (g_desc*)sd;
  const int wave = uni((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  // the queue of the XCD this workgroup runs on: a matrix's tasks all run on one XCD, so their data passes
  // through that XCD's L2 only
  const int q = (int)(__builtin_amdgcn_s_getreg(FLOW_XCC_ID_REG) % FLOW_QUEUES);
  const int q0 = qoff[q], qn = qoff[q + 1] - q0;
  int prev_mat = -1, prev_t = 0;
  long long taken = 0, ready = 0;
  for (;;) {
    if (wave == 0) {
      if (prev_mat >= 0) {
        // the previous task complete: every wave's stores completed before the barrier that ended it
        __hip_atomic_fetch_add(&sync[FLOW_SYNC_HEAD + prev_mat], lane == 0 ? 1 : 0, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (trace != nullptr) {
          const FlowTask tp = tasks[prev_t];
          const long long v[8] = {taken, ready, (long long)wall_clock64(), tp.mat, tp.kind, tp.k, tp.rb,
                                  (long long)__smid()};
          if (lane < 8) trace[16 * (int64_t)prev_t + lane] = v[lane & 7];
        }
      }
      cur = uni(atomicAdd(&sync[q], lane == 0 ? 1 : 0));
    }
    __syncthreads();
    const int i = uni(cur);
    if (i >= qn) break;
    const int t = q0 + i;
    if (trace != nullptr) taken = wall_clock64();
    const FlowTask tk = tasks[t];
    flow_wait(sync, tk.mat, tk.need, spin_limit);
    if (trace != nullptr) ready = wall_clock64();
    if (tk.kind == 0)
      flow_diag_task(gsd, tk.mat, maxsteps, (g_f64*)mats, (g_f64*)ldp, (g_i64*)failp);
    else if (tk.kind == 1)
      flow_left_task(gsd, tk.mat, tk.rb, tk.k, maxsteps, (g_f64*)mats, (g_f64*)ldp, (g_i64*)failp, (g_f64*)stage,
                     (lds_f64*)lds, sync, tk.need2, spin_limit, trace ? trace + 16 * (int64_t)t : nullptr);
    else
      flow_small_task(gsd, tk.mat, tk.rb, tk.k, maxsteps, (g_f64*)mats, (g_f64*)ldp, (g_i64*)failp, (g_f64*)stage,
                      (lds_f64*)lds, sync, tk.need2, spin_limit, trace ? trace + 16 * (int64_t)t : nullptr);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    prev_mat = tk.mat;
    prev_t = t;
  }
}

// The whole blocked Cholesky of every matrix of the batch in one cooperative launch: the workgroups stay
// resident and walk the steps together, a grid-wide barrier between the phases of a step (the launch
// boundaries of the step-synchronous form).  Within a phase, task t goes to workgroup t mod the grid, so
// each task's arithmetic, and the factor, are those of the step-synchronous form.  With lookahead, the
// workgroup that updates a matrix's first trailing tile of step k factors the diagonal block of step k + 1
// right after it (the block lies in wave 0's quarter of that tile, which wave 0 itself wrote), which saves
// the diagonal phase and its barrier at every step but the first.
constexpr int COOP_LDS_DOUBLES = NB * (NB + 1) + 256 * (NB + 1);  // the panel phase's two arrays (the largest)
static_assert(COOP_LDS_DOUBLES >= 2 * TT * (NB + 1), "the trailing phase's panels must fit the shared buffer");

__global__ void __launch_bounds__(256) chol_coop_kernel(const SampleDesc* __restrict__ sd,
                                                        const StepEntry* __restrict__ steps,
                                                        const int* __restrict__ step_tab, int maxsteps, int nmat,
                                                        int lookahead, double* __restrict__ mats,
                                                        double* __restrict__ ldp, int64_t* __restrict__ failp,
                                                        double* __restrict__ logdet, int64_t* __restrict__ info) {
  __shared__ double smem[COOP_LDS_DOUBLES];
  auto grid = cooperative_groups::this_grid();
  double(*Lkk)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(smem);
  double(*Xs)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(smem + NB * (NB + 1));
  double(*PI)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(smem);
  double(*PJ)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(smem + TT * (NB + 1));
  const int G = (int)gridDim.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // step_tab: per step k, {entry offset, entries, panel tasks, trailing tasks}
  for (int k = 0; k < maxsteps; ++k) {
    const int4 tk = reinterpret_cast<const int4*>(step_tab)[k];
    const StepEntry* e = steps + tk.x;
    const int ne = tk.y;
    const int64_t k0 = (int64_t)k * NB;
    if (k == 0 || !lookahead) {
      if (wave == 0)
        for (int t = blockIdx.x; t < ne; t += G) chol_diag_dev(sd, e[t].mat, k0, k, maxsteps, mats, ldp, failp, lane);
      grid.sync();
    }
    for (int t = blockIdx.x; t < tk.z; t += G) {
      const int ei = find_entry(e, ne, t, false);
      chol_panel_dev(sd, e[ei].mat, t - e[ei].pfx_p, k0, mats, Lkk, Xs);
      __syncthreads();  // the next task reloads the shared arrays
    }
    grid.sync();
    for (int t = blockIdx.x; t < tk.w; t += G) {
      const int ei = find_entry(e, ne, t, true);
      const int j = e[ei].mat;
      const int64_t tt = (int64_t)t - e[ei].pfx_t;
      int64_t n;
      mat_ptr(sd, mats, j, n);
      chol_trail_dev(sd, j, tt, k0, k0 + std::min<int64_t>(NB, n - k0), mats, PI, PJ);
      if (lookahead && tt == 0 && wave == 0) chol_diag_dev(sd, j, k0 + NB, k + 1, maxsteps, mats, ldp, failp, lane);
      __syncthreads();
    }
    grid.sync();
  }
  for (int j = blockIdx.x * 256 + threadIdx.x; j < nmat; j += G * 256)
    chol_finish_dev(sd, j, maxsteps, ldp, failp, logdet, info);
}

constexpr int SB = 64;  // row block of the triangular solves
constexpr int SOLVE_THREADS = 1024;  // sixteen waves per sample
// largest sample the solve kernel takes: its solution vector (FP64, dynamic LDS) beside the static
// arrays (reduction, 64 x 65 diagonal block, 16 x 64 partial sums) in the 160 KB of a CU
constexpr int SOLVE_S_MAX =
    (int)((160 * 1024 - sizeof(double) * (SOLVE_THREADS + SB * (SB + 1) + (SOLVE_THREADS / 64) * SB)) / sizeof(double));
static_assert(SOLVE_S_MAX <= 64 * PREP_SLICES_MAX, "the prep kernel's slice table holds every sample's slices");

// ---------------------------------------------------------------------------
// Solves L y = D, L^T s = y and the log-evidence, one workgroup per sample.
//   logL = -1/2 (chi.chi - s.D + log|A| - log|H| + noise_norm)   (numba.py:852-857)
// ---------------------------------------------------------------------------
// With ldp non-null the kernel also finishes the factorisation's log-determinants and failure codes of the
// sample's two matrices (chol_finish_dev), and marks a matrix -1 when the persistent schedule's error flag
// is set or the matrix's count of completed tasks is not its task total (flow_sync / flow_total).
__global__ void __launch_bounds__(SOLVE_THREADS) solve_loglike_kernel(const SampleDesc* __restrict__ sd,
                                                                      const double* __restrict__ mats,
                                                                      const double* __restrict__ D,
                                                                      const double* __restrict__ chi2_nn,
                                                                      double* __restrict__ logdet,
                                                                      double* __restrict__ recon,
                                                                      double* __restrict__ logl,
                                                                      const double* __restrict__ ldp,
                                                                      const int64_t* __restrict__ failp,
                                                                      int maxsteps, int64_t* __restrict__ info,
                                                                      const int* __restrict__ flow_sync,
                                                                      const int* __restrict__ flow_total) {
  constexpr int NW = SOLVE_THREADS / 64;  // waves
  extern __shared__ double y[];
  __shared__ double red[SOLVE_THREADS];
  __shared__ double Ld[SB][SB + 1];
  __shared__ double part[NW][SB];
  __shared__ double ldv[2];
  const SampleDesc d = sd[blockIdx.x];
  const double* L = mats + d.mat_off;
  const int64_t n = d.S;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int64_t i = tid; i < n; i += SOLVE_THREADS) y[i] = D[d.s_off + i];
  if (tid < 2) {
    const int j = 2 * (int)blockIdx.x + tid;
    if (ldp != nullptr) {
      chol_finish_dev(sd, j, maxsteps, ldp, failp, logdet, info);
      if (flow_sync != nullptr &&
          (flow_sync[FLOW_ERR] != 0 || flow_sync[FLOW_SYNC_HEAD + j] != flow_total[j]))
        info[j] = -1;
    }
    ldv[tid] = logdet[j];
  }
  __syncthreads();
  // Blocked substitution over 64-row blocks: the off-diagonal part of each block is a GEMV on
  // contiguous row segments of L (all sixteen waves, eight 512-byte row loads in flight per wave, so
  // that one workgroup keeps enough of L in flight to stream it), the 64 x 64 diagonal block is solved
  // from LDS by one wave with the solved value broadcast by a lane shuffle.
  // forward, L y = D: y_blk -= L[blk, 0:k0] y[0:k0]; then the diagonal block
  constexpr int LDQ = SB * SB / SOLVE_THREADS;  // diagonal-block entries per thread
  static_assert(LDQ * SOLVE_THREADS == SB * SB, "whole passes");
  for (int64_t k0 = 0; k0 < n; k0 += SB) {
    const int nb = (int)std::min<int64_t>(SB, n - k0);
    // the diagonal block's loads are issued before the GEMV and stored to LDS after it
    double ldq[LDQ];
#pragma unroll
    for (int q = 0; q < LDQ; ++q) {
      const int e = tid + q * SOLVE_THREADS, r = e / SB, c = e % SB;
      ldq[q] = (r < nb && c <= r) ? L[(k0 + r) * n + k0 + c] : 0.0;
    }
    // wave w: rows w, w + 16, w + 32, w + 48; per row two partial sums (j and j + 64 of each 128-column
    // step), added in that order
    {
      double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
      const double* Li[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) Li[q] = L + (k0 + std::min(wave + NW * q, nb - 1)) * n;
      int64_t j = lane;
      for (; j + 64 < k0; j += 128) {
        const double yj = y[j], yk = y[j + 64];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          a[q] = fma(Li[q][j], yj, a[q]);
          b[q] = fma(Li[q][j + 64], yk, b[q]);
        }
      }
      if (j < k0) {
        const double yj = y[j];
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] = fma(Li[q][j], yj, a[q]);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double t = wave_sum(a[q] + b[q]);
        if (lane == 0 && wave + NW * q < nb) part[0][wave + NW * q] = t;
      }
    }
#pragma unroll
    for (int q = 0; q < LDQ; ++q) {
      const int e = tid + q * SOLVE_THREADS;
      Ld[e / SB][e % SB] = ldq[q];
    }
    __syncthreads();
    if (wave == 0) {
      // lane j's value is scaled by the reciprocal of its pivot (computed by every lane at once) and read
      // by the others with a lane read: one multiply, one lane read and one FMA per row on the chain
      double v = lane < nb ? y[k0 + lane] - part[0][lane] : 0.0;
      const double inv = lane < nb ? 1.0 / Ld[lane][lane] : 0.0;
      for (int j = 0; j < nb; ++j) {
        v = lane == j ? v * inv : v;
        const double yj = readlane_d(v, j);
        v = lane > j ? fma(-Ld[lane][j], yj, v) : v;
      }
      if (lane < nb) y[k0 + lane] = v;
    }
    __syncthreads();
  }
  // backward, L^T s = y: y_blk -= L[k0+nb:n, blk]^T s[k0+nb:n] (per wave four row-interleaved partial
  // sums, the sixteen waves' sums added in a fixed order); then the transposed diagonal block
  for (int64_t k0 = ((n - 1) / SB) * SB; k0 >= 0; k0 -= SB) {
    const int nb = (int)std::min<int64_t>(SB, n - k0);
    double ldq[LDQ];
#pragma unroll
    for (int q = 0; q < LDQ; ++q) {
      const int e = tid + q * SOLVE_THREADS, r = e / SB, c = e % SB;
      ldq[q] = (r < nb && c <= r) ? L[(k0 + r) * n + k0 + c] : 0.0;
    }
    // thread (wave w, lane c): column k0 + c over rows j = k0 + nb + w, + 16, ..., in four interleaved
    // accumulators
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (lane < nb) {
      const double* col = L + k0 + lane;
      int64_t j = k0 + nb + wave;
      for (; j + 3 * NW < n; j += 4 * NW) {
        a0 = fma(col[j * n], y[j], a0);
        a1 = fma(col[(j + NW) * n], y[j + NW], a1);
        a2 = fma(col[(j + 2 * NW) * n], y[j + 2 * NW], a2);
        a3 = fma(col[(j + 3 * NW) * n], y[j + 3 * NW], a3);
      }
      for (; j < n; j += NW) a0 = fma(col[j * n], y[j], a0);
    }
    part[wave][lane] = (a0 + a1) + (a2 + a3);
#pragma unroll
    for (int q = 0; q < LDQ; ++q) {
      const int e = tid + q * SOLVE_THREADS;
      Ld[e / SB][e % SB] = ldq[q];
    }
    __syncthreads();
    if (wave == 0) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) t += part[w][lane];
      double v = lane < nb ? y[k0 + lane] - t : 0.0;
      const double inv = lane < nb ? 1.0 / Ld[lane][lane] : 0.0;
      for (int j = nb - 1; j >= 0; --j) {
        v = lane == j ? v * inv : v;
        const double sj = readlane_d(v, j);
        v = lane < j ? fma(-Ld[j][lane], sj, v) : v;
      }
      if (lane < nb) y[k0 + lane] = v;
    }
    __syncthreads();
  }
  double acc = 0.0;
  for (int64_t i = tid; i < n; i += SOLVE_THREADS) {
    acc += y[i] * D[d.s_off + i];
    if (recon) recon[d.s_off + i] = y[i];
  }
  red[tid] = acc;
  __syncthreads();
  for (int o = SOLVE_THREADS / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  if (tid == 0) {
    const double sD = red[0];
    logl[blockIdx.x] = -0.5 * (chi2_nn[0] - sD + ldv[0] - ldv[1] + chi2_nn[1]);
  }
}

// chi.chi over the 2K real entries and the noise normalisation (numba.py:752, :846-847).
__global__ void chi2_noise_kernel(const double* __restrict__ data, const double* __restrict__ noise, int64_t K2,
                                  double* __restrict__ out) {
  __shared__ double r0[256], r1[256];
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < K2; i += 256) {
    const double c = data[i] / noise[i];
    a = fma(c, c, a);
    b += log(fabs(noise[i]));
  }
  r0[threadIdx.x] = a;
  r1[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      r0[threadIdx.x] += r0[threadIdx.x + o];
      r1[threadIdx.x] += r1[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = r0[0];
    out[1] = 2.0 * ((double)(K2 / 2) * log(2.0 * M_PI) + r1[0]);
  }
}

__global__ void noise_norm_kernel(const double* __restrict__ noise, int64_t K2, double* __restrict__ out) {
  __shared__ double r1[256];
  double b = 0.0;
  for (int64_t i = threadIdx.x; i < K2; i += 256) b += log(fabs(noise[i]));
  r1[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) r1[threadIdx.x] += r1[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = 2.0 * ((double)(K2 / 2) * log(2.0 * M_PI) + r1[0]);
}

// ---------------------------------------------------------------------------
// Host-side planning of the batch workspace.
// ---------------------------------------------------------------------------
struct BatchPlan {
  int64_t n, M, B, sumS, sumS2, omega_elems, csc_ptr_elems, csc_ent_elems, maxsteps, n_entries, nsl, max_flow;
  size_t off_desc, off_ptr, off_cur, off_key, off_t, off_omega, off_mats, off_D, off_logdet, off_info,
      off_steps, off_tab, off_ldp, off_failp, off_stage, off_skey, off_st, off_scol, off_spos, off_slices, off_fsync, off_fqoff, off_ftasks, bytes;
};

inline BatchPlan plan_batch(int64_t M, int64_t B, int64_t n, const int64_t* S_host) {
  BatchPlan p{};
  p.n = n;
  p.M = M;
  p.B = B;
  for (int64_t b = 0; b < n; ++b) {
    p.sumS += S_host[b];
    p.sumS2 += S_host[b] * S_host[b];
    p.maxsteps = std::max<int64_t>(p.maxsteps, ceil_div(S_host[b], (int64_t)NB));
    p.n_entries += 2 * ceil_div(S_host[b], (int64_t)NB);
    p.nsl += ceil_div(S_host[b], (int64_t)64);
    // the persistent schedule's tasks of the sample's two matrices: at most the 32-row tasks of every step
    for (int64_t k0 = 0; k0 < S_host[b]; k0 += NB)
      p.max_flow += 2 * std::max<int64_t>(1, ceil_div(S_host[b] - std::min<int64_t>(S_host[b], k0 + NB), (int64_t)NB));
  }
  p.max_flow += 2 * n;  // the first diagonal blocks, one task per matrix
  p.omega_elems = M * p.sumS;
  p.csc_ptr_elems = p.sumS + n;
  p.csc_ent_elems = n * M * B;
  size_t o = 0;
  auto put = [&](size_t bytes) {
    o = align_up(o, 256);
    size_t r = o;
    o += bytes;
    return r;
  };
  // the tables the host writes, adjacent so that one copy uploads them: sample descriptors, Omega slices,
  // the persistent schedule's queue offsets and task totals, its task lists
  p.off_desc = put(sizeof(SampleDesc) * n);
  p.off_slices = put(sizeof(int2) * p.nsl);
  p.off_fqoff = put(sizeof(int) * (FLOW_QUEUES + 1 + 2 * n));  // queue offsets, then task totals per matrix
  p.off_ftasks = put(sizeof(FlowTask) * p.max_flow);
  p.off_fsync = put(sizeof(int) * (FLOW_SYNC_HEAD + 2 * n));  // the persistent schedule's queue heads and counts
  p.off_ptr = put(sizeof(int) * p.csc_ptr_elems);
  p.off_cur = put(sizeof(int) * p.csc_ptr_elems);
  p.off_key = put(sizeof(int) * p.csc_ent_elems);
  p.off_t = put(sizeof(double) * p.csc_ent_elems);
  p.off_omega = put(sizeof(double) * p.omega_elems);
  p.off_mats = put(sizeof(double) * 2 * p.sumS2);
  p.off_D = put(sizeof(double) * p.sumS);
  p.off_logdet = put(sizeof(double) * 2 * n);
  p.off_info = put(sizeof(int64_t) * 2 * n);
  p.off_steps = put(sizeof(StepEntry) * 2 * p.n_entries);  // the step tables of two schedules
  p.off_tab = put(sizeof(int) * 4 * p.maxsteps);
  p.off_ldp = put(sizeof(double) * 2 * n * p.maxsteps);
  p.off_failp = put(sizeof(int64_t) * 2 * n * p.maxsteps);
  p.off_stage = put(sizeof(double) * 4 * n * NB * NB);  // two slots per matrix
  p.off_skey = put(sizeof(int) * (p.csc_ent_elems + OMEGA_SLICE_PAD));  // the buckets in slice order
  p.off_st = put(sizeof(double) * (p.csc_ent_elems + OMEGA_SLICE_PAD));
  p.off_scol = put(sizeof(int) * p.csc_ptr_elems);  // the slices' columns by rank
  p.off_spos = put(sizeof(int) * p.csc_ptr_elems);  // the slices' first entries
  p.bytes = align_up(o, 256);
  return p;
}

}  // namespace lk
}  // namespace aj

using namespace aj;
using namespace aj::lk;

namespace {

// How the batched Cholesky runs (process-wide, aj_log_evidence_configure): three launches per 32-column
// step; one cooperative launch (without or with lookahead); one launch per step with lookahead; one
// left-looking launch per step; the left-looking tasks in one persistent launch with lookahead (the
// default).  All give the same factor bit for bit.
constexpr int CHOL_STEPS = 0, CHOL_COOP = 1, CHOL_COOP_LOOKAHEAD = 2, CHOL_STEPS_LOOKAHEAD = 3, CHOL_LEFT = 4,
              CHOL_FLOW = 5;
constexpr int CHOL_MODES = 6;
std::atomic<int> g_chol_mode{CHOL_FLOW};
int chol_mode() { return g_chol_mode.load(std::memory_order_relaxed); }

// Workgroups of the cooperative Cholesky: every one resident at once (occupancy x compute units), per
// device; 0 when the device cannot launch cooperatively.
int coop_grid() {
  static std::mutex mu;
  static int cache[64];
  static bool have[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  std::lock_guard<std::mutex> lock(mu);
  if (!have[dev]) {
    int coop = 0, cus = 0, per = 0;
    int g = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess && coop &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)chol_coop_kernel, 256, 0) == hipSuccess)
      g = per * cus;
    cache[dev] = g;
    have[dev] = true;
  }
  return cache[dev];
}

// Compute units of the current device (256 when the query fails).
int cu_count() {
  int dev = 0, c = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      c > 0)
    return c;
  return 256;
}

// Resident workgroups of the persistent Cholesky per compute unit (at least one).
int flow_occupancy() {
  static std::atomic<int> cached{0};
  int v = cached.load(std::memory_order_relaxed);
  if (v == 0) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)chol_flow_kernel, FLOW_THREADS, 0) != hipSuccess ||
        per < 1)
      per = 1;
    cached.store(per, std::memory_order_relaxed);
    v = per;
  }
  return v;
}

// The persistent schedule's dependency wait gives up after this many polls (each a short sleep and an
// L2 read, ~1 us): a quarter of a second, ~100 times the factorisation of a whole batch.
constexpr int FLOW_SPIN_LIMIT = 1 << 18;
std::atomic<int> g_flow_spin_limit{FLOW_SPIN_LIMIT};
// A block column runs as 32-row tasks once its 32-row tasks on an XCD number at most this many per compute
// unit of the XCD.
constexpr int FLOW_SMALL_PER_CU = 1;
std::atomic<int> g_flow_small_per_cu{FLOW_SMALL_PER_CU};
// Optional task trace of the persistent schedule (sixteen int64 per task, aj_log_evidence_flow_trace).
std::atomic<long long*> g_flow_trace{nullptr};
std::atomic<int64_t> g_flow_trace_cap{0};

// Runs the whole per-sample pipeline for a batch.  Returns the workspace
// views through `p`; outputs per-sample log-evidence into logl.
int batch_pipeline(const WSrc w_tilde, int64_t M, const double* dirty_image, const double* chi2_nn, int64_t n,
                   int64_t B, int64_t P, const int64_t* S_host, const int64_t* pix_indexes, const double* pix_weights,
                   const int64_t* neighbors, const int64_t* neighbors_sizes, const double* coef_host,
                   const double* beta_host, double* logl, double* recon, double* logdets_out, int64_t* info_out,
                   void* workspace, size_t workspace_bytes, hipStream_t st) {
  BatchPlan p = plan_batch(M, B, n, S_host);
  AJ_REQUIRE(workspace && workspace_bytes >= p.bytes, "workspace too small: need %zu bytes, got %zu", p.bytes,
             workspace_bytes);
  char* ws = static_cast<char*>(workspace);
  SampleDesc* desc = reinterpret_cast<SampleDesc*>(ws + p.off_desc);
  int* ptrs = reinterpret_cast<int*>(ws + p.off_ptr);
  int* cur = reinterpret_cast<int*>(ws + p.off_cur);
  int* ekey = reinterpret_cast<int*>(ws + p.off_key);
  double* et = reinterpret_cast<double*>(ws + p.off_t);
  double* omega = reinterpret_cast<double*>(ws + p.off_omega);
  double* mats = reinterpret_cast<double*>(ws + p.off_mats);
  double* D = reinterpret_cast<double*>(ws + p.off_D);
  double* logdet = logdets_out ? logdets_out : reinterpret_cast<double*>(ws + p.off_logdet);
  int64_t* info = info_out ? info_out : reinterpret_cast<int64_t*>(ws + p.off_info);
  int* skey = reinterpret_cast<int*>(ws + p.off_skey);
  double* st_t = reinterpret_cast<double*>(ws + p.off_st);
  int* scol = reinterpret_cast<int*>(ws + p.off_scol);
  int* spos = reinterpret_cast<int*>(ws + p.off_spos);
  int2* slices = reinterpret_cast<int2*>(ws + p.off_slices);

  // Per-sample table (host -> device).
  static thread_local std::vector<SampleDesc> hd;
  hd.resize(n);
  int64_t s_off = 0, mat_off = 0, om_off = 0, ptr_off = 0, maxS = 0;
  for (int64_t b = 0; b < n; ++b) {
    const int64_t S = S_host[b];
    AJ_REQUIRE(S > 0 && S < (1ll << 31), "sample %lld: bad pixel count %lld", (long long)b, (long long)S);
    hd[b] = SampleDesc{S, s_off, b * M * B, ptr_off, b * M * B, om_off, mat_off, coef_host[b] * coef_host[b],
                       beta_host ? beta_host[b] : 0.0};
    s_off += S;
    mat_off += 2 * S * S;
    om_off += M * S;
    ptr_off += S + 1;
    maxS = std::max(maxS, S);
  }
  AJ_REQUIRE(M * B * (int64_t)sizeof(double) + OMEGA_SLICE_PAD * 8 < (1ll << 31),
             "%lld mapping entries per sample exceed the Omega kernel's 32-bit offsets", (long long)(M * B));
  AJ_REQUIRE(maxS <= SOLVE_S_MAX, "a sample of %lld source pixels exceeds the solve kernel's LDS (%d at most)",
             (long long)maxS, SOLVE_S_MAX);
  // slices of 64 columns for the Omega kernel, samples in order
  static thread_local std::vector<int2> hs;
  hs.clear();
  for (int64_t b = 0; b < n; ++b)
    for (int64_t c = 0; c < S_host[b]; c += 64) hs.push_back(int2{(int)b, (int)c});
  AJ_REQUIRE((int64_t)hs.size() == p.nsl && p.nsl < (1ll << 31), "slice table of %zu entries", hs.size());
  static thread_local std::vector<int> ho;
  ho.resize(n);
  for (int64_t b = 0; b < n; ++b) ho[b] = (int)b;
  std::stable_sort(ho.begin(), ho.end(), [&](int x, int y) { return S_host[x] > S_host[y]; });
  (void)cur;
  const int mode = chol_mode();
  // the persistent schedule's task lists (built here, so that they go up with the sample table)
  int flow_ntasks = 0;
  static thread_local std::vector<FlowTask> hf;
  static thread_local std::vector<int> hq;
  hf.clear();
  hq.clear();
  if (mode == CHOL_FLOW) {
    // matrices to XCDs: in decreasing size, each to the XCD with the least work so far (S^3 each), so that
    // the largest matrices, whose chains of block columns are the longest, land on different XCDs
    const int cus = cu_count();
    const int64_t per_xcd = std::max(1, cus / FLOW_QUEUES);
    static thread_local std::vector<int> xcd, need, need_prev;
    static thread_local std::vector<FlowTask> lists[FLOW_QUEUES];
    xcd.assign(2 * n, 0);
    need.assign(2 * n, 1);       // tasks of the matrix in block columns <= k - 1 (and its first diagonal block)
    need_prev.assign(2 * n, 1);  // ... in block columns <= k - 2
    double work[FLOW_QUEUES] = {};
    for (int64_t q = 0; q < 2 * n; ++q) {
      const int j = 2 * ho[q >> 1] + (int)(q & 1);
      int x = 0;
      for (int y = 1; y < FLOW_QUEUES; ++y)
        if (work[y] < work[x]) x = y;
      xcd[j] = x;
      const double S = (double)S_host[j >> 1];
      work[x] += S * S * S;
    }
    // each XCD's list: its matrices' first diagonal blocks, then per block column k their tasks (matrices in
    // decreasing size): 256-row tasks, or 32-row tasks once the XCD's 32-row tasks of the column number at
    // most small_per_cu per compute unit; a task starts once its matrix's tasks in block columns <= k - 2 are
    // complete and takes its last block once those of column k - 1 are
    for (auto& l : lists) l.clear();
    for (int64_t q = 0; q < 2 * n; ++q) {
      const int j = 2 * ho[q >> 1] + (int)(q & 1);
      lists[xcd[j]].push_back(FlowTask{j, 0, 0, 0, 0, 0});
    }
    const int64_t small_per = per_xcd * g_flow_small_per_cu.load(std::memory_order_relaxed);
    for (int64_t k = 0; k < p.maxsteps; ++k) {
      const int64_t k0 = k * NB;
      int64_t nsmall[FLOW_QUEUES] = {};
      for (int64_t j = 0; j < 2 * n; ++j) {
        const int64_t S = S_host[j >> 1];
        if (S > k0) nsmall[xcd[j]] += std::max<int64_t>(1, ceil_div(S - std::min<int64_t>(S, k0 + NB), (int64_t)NB));
      }
      for (int64_t q = 0; q < 2 * n; ++q) {
        const int j = 2 * ho[q >> 1] + (int)(q & 1);
        const int64_t S = S_host[j >> 1];
        if (S <= k0) continue;
        const bool small = nsmall[xcd[j]] <= small_per;
        const int64_t below = S - std::min<int64_t>(S, k0 + NB);
        const int64_t nt = small ? std::max<int64_t>(1, ceil_div(below, (int64_t)NB))
                                 : 1 + ceil_div(std::max<int64_t>(0, below - LeftShape<FL_STRIP>::STRIP0),
                                                (int64_t)FL_STRIP);
        for (int64_t rb = 0; rb < nt; ++rb)
          lists[xcd[j]].push_back(FlowTask{j, small ? 2 : 1, (int)k, (int)rb, need_prev[j], need[j]});
        need_prev[j] = need[j];
        need[j] += (int)nt;
      }
    }
    hq.assign(FLOW_QUEUES + 1 + 2 * n, 0);
    for (int x = 0; x < FLOW_QUEUES; ++x) {
      hq[x] = (int)hf.size();
      hf.insert(hf.end(), lists[x].begin(), lists[x].end());
    }
    hq[FLOW_QUEUES] = (int)hf.size();
    for (int64_t j = 0; j < 2 * n; ++j) hq[FLOW_QUEUES + 1 + j] = need[j];  // each matrix's task total
    AJ_REQUIRE((int64_t)hf.size() <= p.max_flow && hf.size() < (1u << 30), "internal: persistent task list %zu > %lld",
               hf.size(), (long long)p.max_flow);
    flow_ntasks = (int)hf.size();
  }
  // one copy of the host tables
  timer_begin(AJ_TIMER_BATCH, st);
  {
    static thread_local std::vector<char> up;
    up.resize(p.off_ftasks + sizeof(FlowTask) * p.max_flow - p.off_desc);
    auto put_up = [&](size_t off, const void* src, size_t bytes) {
      if (bytes) std::memcpy(up.data() + (off - p.off_desc), src, bytes);
      return off - p.off_desc + bytes;
    };
    size_t used = put_up(p.off_desc, hd.data(), sizeof(SampleDesc) * n);
    used = std::max(used, put_up(p.off_slices, hs.data(), sizeof(int2) * hs.size()));
    if (mode == CHOL_FLOW) {
      used = std::max(used, put_up(p.off_fqoff, hq.data(), sizeof(int) * hq.size()));
      used = std::max(used, put_up(p.off_ftasks, hf.data(), sizeof(FlowTask) * hf.size()));
    }
    AJ_CHECK_HIP(hipMemcpyAsync(desc, up.data(), used, hipMemcpyHostToDevice, st));
  }

  // 1: mapping inversion and data vector, one workgroup per sample
  AJ_CHECK_HIP(hipFuncSetAttribute((const void*)prep_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(int) * (maxS + 1))));
  int* fsync = reinterpret_cast<int*>(ws + p.off_fsync);
  hipLaunchKernelGGL(prep_kernel, dim3((unsigned)n), dim3(PREP_THREADS), sizeof(int) * (maxS + 1), st, desc,
                     pix_indexes, pix_weights, M, B, dirty_image, ptrs, ekey, et, D, skey, st_t, scol, spos,
                     mode == CHOL_FLOW ? fsync : nullptr);
  AJ_LAUNCH_CHECK();
  // 2: Omega = W T, W read once for the batch; rows of W per workgroup: the largest power of two <= 8
  // whose rows fit in LDS
  int R = 8;
  while (R >= 1 && (size_t)R * M * sizeof(double) > OMEGA_LDS_BUDGET) R >>= 1;
  timer_begin(AJ_TIMER_OMEGA, st);
  if (R >= 1) {
    constexpr int U = OMEGA_SLICE_UNROLL;
    const void* fn = R == 8 ? (const void*)omega_slices_kernel<8, U>
                   : R == 4 ? (const void*)omega_slices_kernel<4, U>
                   : R == 2 ? (const void*)omega_slices_kernel<2, U>
                            : (const void*)omega_slices_kernel<1, U>;
    AJ_CHECK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(R * M * sizeof(double))));
    const dim3 og((unsigned)ceil_div(M, R));
    const size_t lds = (size_t)R * M * sizeof(double);
    const int nsl = (int)p.nsl;
    const int64_t sl_elems = p.csc_ent_elems + OMEGA_SLICE_PAD;
    switch (R) {
      case 8: hipLaunchKernelGGL((omega_slices_kernel<8, U>), og, dim3(OMEGA_THREADS), lds, st, w_tilde, M, nsl, slices, desc, ptrs, scol, spos, skey, st_t, sl_elems, omega); break;
      case 4: hipLaunchKernelGGL((omega_slices_kernel<4, U>), og, dim3(OMEGA_THREADS), lds, st, w_tilde, M, nsl, slices, desc, ptrs, scol, spos, skey, st_t, sl_elems, omega); break;
      case 2: hipLaunchKernelGGL((omega_slices_kernel<2, U>), og, dim3(OMEGA_THREADS), lds, st, w_tilde, M, nsl, slices, desc, ptrs, scol, spos, skey, st_t, sl_elems, omega); break;
      default: hipLaunchKernelGGL((omega_slices_kernel<1, U>), og, dim3(OMEGA_THREADS), lds, st, w_tilde, M, nsl, slices, desc, ptrs, scol, spos, skey, st_t, sl_elems, omega); break;
    }
  } else {
    hipLaunchKernelGGL(omega_global_kernel, dim3((unsigned)M), dim3(256), 0, st, w_tilde, M, (int)n, desc, ptrs,
                       ekey, et, omega);
  }
  timer_end(AJ_TIMER_OMEGA, st);
  AJ_LAUNCH_CHECK();
  // 3: lower triangles of A = F + H and of H
  timer_begin(AJ_TIMER_FGATHER, st);
  hipLaunchKernelGGL(assemble_kernel, dim3((unsigned)maxS, (unsigned)n), dim3(256), 0, st, desc, ptrs, ekey, et, omega,
                     neighbors, neighbors_sizes, P, mats);
  timer_end(AJ_TIMER_FGATHER, st);
  AJ_LAUNCH_CHECK();
  // 4: blocked Cholesky of every A and H.  Step tables: matrices in decreasing size, per step the ones
  // with columns left, with prefix counts of panel workgroups and trailing tiles.
  static thread_local std::vector<StepEntry> he;
  static thread_local std::vector<int> tab;
  static thread_local std::vector<int64_t> grid_p, grid_t;
  he.clear();
  tab.assign(4 * p.maxsteps, 0);
  grid_p.assign(p.maxsteps, 0);
  grid_t.assign(p.maxsteps, 0);
  for (int64_t k = 0; k < p.maxsteps; ++k) {
    const int64_t k0 = k * NB;
    tab[4 * k] = (int)he.size();
    for (int64_t q = 0; q < 2 * n; ++q) {
      const int j = 2 * ho[q >> 1] + (int)(q & 1);
      const int64_t S = S_host[j >> 1];
      if (S <= k0) continue;
      const int64_t m = S - std::min<int64_t>(S, k0 + NB);  // rows below the diagonal block
      const int64_t tb = ceil_div(m, (int64_t)TT);
      AJ_REQUIRE(grid_p[k] + ceil_div(m, (int64_t)256) < (1ll << 31) && grid_t[k] + tb * (tb + 1) / 2 < (1ll << 31),
                 "batch too large for one Cholesky step grid");
      he.push_back(StepEntry{j, (int)grid_p[k], (int)grid_t[k]});
      grid_p[k] += ceil_div(m, (int64_t)256);
      grid_t[k] += tb * (tb + 1) / 2;
    }
    tab[4 * k + 1] = (int)he.size() - tab[4 * k];
    tab[4 * k + 2] = (int)grid_p[k];
    tab[4 * k + 3] = (int)grid_t[k];
  }
  AJ_REQUIRE((int64_t)he.size() == p.n_entries, "internal: step table size %zu != %lld", he.size(),
             (long long)p.n_entries);
  StepEntry* steps = reinterpret_cast<StepEntry*>(ws + p.off_steps);
  int* step_tab = reinterpret_cast<int*>(ws + p.off_tab);
  double* ldp = reinterpret_cast<double*>(ws + p.off_ldp);
  int64_t* failp = reinterpret_cast<int64_t*>(ws + p.off_failp);
  double* stage = reinterpret_cast<double*>(ws + p.off_stage);
  if (mode != CHOL_FLOW) {  // (the persistent schedule's tables went up with the sample table)
    AJ_CHECK_HIP(hipMemcpyAsync(steps, he.data(), sizeof(StepEntry) * he.size(), hipMemcpyHostToDevice, st));
    AJ_CHECK_HIP(hipMemcpyAsync(step_tab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice, st));
  }
  timer_begin(AJ_TIMER_CHOL, st);
  bool done = false;
  bool solve_finishes = false;
  if (mode == CHOL_COOP || mode == CHOL_COOP_LOOKAHEAD) {
    const int G = coop_grid();
    if (G > 0) {
      int maxsteps_i = (int)p.maxsteps, nmat = (int)(2 * n), look = mode == CHOL_COOP_LOOKAHEAD ? 1 : 0;
      const SampleDesc* desc_c = desc;
      const StepEntry* steps_c = steps;
      const int* tab_c = step_tab;
      void* args[] = {(void*)&desc_c, (void*)&steps_c, (void*)&tab_c, (void*)&maxsteps_i, (void*)&nmat, (void*)&look,
                      (void*)&mats, (void*)&ldp, (void*)&failp, (void*)&logdet, (void*)&info};
      AJ_CHECK_HIP(hipLaunchCooperativeKernel((const void*)chol_coop_kernel, dim3((unsigned)G), dim3(256), args, 0, st));
      done = true;
    }
  }
  if (!done && mode == CHOL_FLOW) {
    // the task lists went up with the sample table
    const int ntasks = flow_ntasks;
    const int grid = cu_count() * flow_occupancy();
    FlowTask* ftasks = reinterpret_cast<FlowTask*>(ws + p.off_ftasks);
    int* fqoff = reinterpret_cast<int*>(ws + p.off_fqoff);
    long long* trace = g_flow_trace.load(std::memory_order_relaxed);
    if (g_flow_trace_cap.load(std::memory_order_relaxed) < ntasks) trace = nullptr;
    hipLaunchKernelGGL(chol_flow_kernel, dim3((unsigned)grid), dim3(FLOW_THREADS), 0, st, desc, ftasks, fqoff,
                       (int)p.maxsteps, mats, ldp, failp, stage, fsync,
                       g_flow_spin_limit.load(std::memory_order_relaxed), trace);
    AJ_LAUNCH_CHECK();
    solve_finishes = true;
    done = true;
  }
  if (!done && mode == CHOL_LEFT) {
    // per step, the matrices with columns left, each with its strip tasks: 128-row tasks (the first
    // task 96 rows), or, once the 32-row tasks of a step fit the compute units one workgroup each,
    // 32-row tasks (a step's time is then its tasks' chains of blocks, which they run four times faster)
    static int cus = 0;
    if (cus == 0) {
      int dev = 0, c = 0;
      if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
        cus = c;
      if (cus <= 0) cus = 256;
    }
    static thread_local std::vector<StepEntry> h4;
    static thread_local std::vector<int64_t> off4, cnt4, nt4;
    static thread_local std::vector<char> small4;
    h4.clear();
    off4.assign(p.maxsteps, 0);
    cnt4.assign(p.maxsteps, 0);
    nt4.assign(p.maxsteps, 0);
    small4.assign(p.maxsteps, 0);
    for (int64_t k = 0; k < p.maxsteps; ++k) {
      const int64_t k0 = k * NB;
      int64_t nsmall = 0;
      for (int64_t b = 0; b < n; ++b) {
        const int64_t S = S_host[b];
        if (S > k0) nsmall += 2 * std::max<int64_t>(1, ceil_div(S - std::min<int64_t>(S, k0 + NB), (int64_t)NB));
      }
      const bool small = nsmall <= cus;
      small4[k] = small;
      off4[k] = (int64_t)h4.size();
      for (int64_t q = 0; q < 2 * n; ++q) {
        const int j = 2 * ho[q >> 1] + (int)(q & 1);
        const int64_t S = S_host[j >> 1];
        if (S <= k0) continue;
        const int64_t below = S - std::min<int64_t>(S, k0 + NB);
        const int64_t nt = small ? std::max<int64_t>(1, ceil_div(below, (int64_t)NB))
                                 : 1 + ceil_div(std::max<int64_t>(0, below - LeftShape<LL_STRIP>::STRIP0),
                                                (int64_t)LL_STRIP);
        AJ_REQUIRE(nt4[k] + nt < (1ll << 30), "batch too large for one Cholesky step grid");
        h4.push_back(StepEntry{j, (int)nt4[k], 0});
        nt4[k] += nt;
      }
      cnt4[k] = (int64_t)h4.size() - off4[k];
    }
    AJ_REQUIRE((int64_t)h4.size() == p.n_entries, "internal: left-looking step table size %zu", h4.size());
    StepEntry* steps4 = steps + p.n_entries;
    AJ_CHECK_HIP(hipMemcpyAsync(steps4, h4.data(), sizeof(StepEntry) * h4.size(), hipMemcpyHostToDevice, st));
    // every matrix's diagonal block 0, factored in place; then one launch per block column
    hipLaunchKernelGGL(chol_diag_kernel, dim3((unsigned)tab[1]), dim3(64), 0, st, desc, steps, (int64_t)0, 0,
                       (int)p.maxsteps, mats, ldp, failp);
    AJ_LAUNCH_CHECK();
    for (int64_t k = 0; k < p.maxsteps; ++k) {
      if (small4[k])
        hipLaunchKernelGGL(chol_left_small_kernel, dim3((unsigned)nt4[k]), dim3(LS_THREADS), 0, st, desc,
                           steps4 + off4[k], (int)cnt4[k], k * NB, (int)k, (int)p.maxsteps, mats, ldp, failp, stage);
      else
        hipLaunchKernelGGL(chol_left_kernel, dim3((unsigned)nt4[k]), dim3(LL_THREADS), 0, st, desc, steps4 + off4[k],
                           (int)cnt4[k], k * NB, (int)k, (int)p.maxsteps, mats, ldp, failp, stage);
      AJ_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(chol_finish_kernel, dim3((unsigned)ceil_div(2 * n, (int64_t)256)), dim3(256), 0, st, desc,
                       (int)(2 * n), (int)p.maxsteps, ldp, failp, logdet, info);
    done = true;
  }
  if (!done && mode == CHOL_STEPS_LOOKAHEAD) {
    // prologue: diagonal blocks and panels of step 0; then one launch per step
    static thread_local std::vector<StepEntry> h3;
    static thread_local std::vector<int64_t> off3, cnt3, na3, nt3;
    h3.clear();
    off3.assign(p.maxsteps, 0);
    cnt3.assign(p.maxsteps, 0);
    na3.assign(p.maxsteps, 0);
    nt3.assign(p.maxsteps, 0);
    for (int64_t k = 0; k + 1 < p.maxsteps; ++k) {
      const int64_t k0 = k * NB;
      off3[k] = (int64_t)h3.size();
      for (int64_t q = 0; q < 2 * n; ++q) {
        const int j = 2 * ho[q >> 1] + (int)(q & 1);
        const int64_t S = S_host[j >> 1];
        if (S <= k0 + NB) continue;  // no step k + 1
        const int64_t below = S - (k0 + 2 * NB);  // rows below the next diagonal block
        const int64_t na = std::max<int64_t>(1, ceil_div(std::max<int64_t>(below, 0), (int64_t)STRIP));
        const int64_t tb = below > 0 ? ceil_div(below, (int64_t)TT) : 0;
        AJ_REQUIRE(na3[k] + na < (1ll << 30) && nt3[k] + tb * (tb + 1) / 2 < (1ll << 30),
                   "batch too large for one Cholesky step grid");
        h3.push_back(StepEntry{j, (int)na3[k], (int)nt3[k]});
        na3[k] += na;
        nt3[k] += tb * (tb + 1) / 2;
      }
      cnt3[k] = (int64_t)h3.size() - off3[k];
    }
    AJ_REQUIRE((int64_t)h3.size() <= p.n_entries, "internal: lookahead step table size %zu", h3.size());
    StepEntry* steps3 = steps + p.n_entries;
    if (!h3.empty())
      AJ_CHECK_HIP(hipMemcpyAsync(steps3, h3.data(), sizeof(StepEntry) * h3.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(chol_diag_kernel, dim3((unsigned)tab[1]), dim3(64), 0, st, desc, steps, (int64_t)0, 0,
                       (int)p.maxsteps, mats, ldp, failp);
    AJ_LAUNCH_CHECK();
    if (grid_p[0] > 0) {
      hipLaunchKernelGGL(chol_panel_kernel, dim3((unsigned)grid_p[0]), dim3(256), 0, st, desc, steps, tab[1],
                         (int64_t)0, mats);
      AJ_LAUNCH_CHECK();
    }
    for (int64_t k = 0; k + 1 < p.maxsteps; ++k) {
      if (cnt3[k] == 0) continue;
      hipLaunchKernelGGL(chol_step_kernel, dim3((unsigned)(na3[k] + nt3[k])), dim3(256), 0, st, desc,
                         steps3 + off3[k], (int)cnt3[k], (int)na3[k], k * NB, (int)k, (int)p.maxsteps, mats, ldp,
                         failp, stage);
      AJ_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(chol_unstage_kernel, dim3((unsigned)(2 * n)), dim3(256), 0, st, desc, mats, stage, 1);
    AJ_LAUNCH_CHECK();
    hipLaunchKernelGGL(chol_finish_kernel, dim3((unsigned)ceil_div(2 * n, (int64_t)256)), dim3(256), 0, st, desc,
                       (int)(2 * n), (int)p.maxsteps, ldp, failp, logdet, info);
    done = true;
  }
  if (!done) {
    for (int64_t k = 0; k < p.maxsteps; ++k) {
      const StepEntry* e = steps + tab[4 * k];
      const int ne = tab[4 * k + 1];
      hipLaunchKernelGGL(chol_diag_kernel, dim3((unsigned)ne), dim3(64), 0, st, desc, e, k * NB, (int)k,
                         (int)p.maxsteps, mats, ldp, failp);
      AJ_LAUNCH_CHECK();
      if (grid_p[k] > 0) {
        hipLaunchKernelGGL(chol_panel_kernel, dim3((unsigned)grid_p[k]), dim3(256), 0, st, desc, e, ne, k * NB, mats);
        AJ_LAUNCH_CHECK();
      }
      if (grid_t[k] > 0) {
        hipLaunchKernelGGL(chol_trail_kernel, dim3((unsigned)grid_t[k]), dim3(256), 0, st, desc, e, ne, k * NB, mats);
        AJ_LAUNCH_CHECK();
      }
    }
    hipLaunchKernelGGL(chol_finish_kernel, dim3((unsigned)ceil_div(2 * n, (int64_t)256)), dim3(256), 0, st, desc,
                       (int)(2 * n), (int)p.maxsteps, ldp, failp, logdet, info);
  }
  timer_end(AJ_TIMER_CHOL, st);
  AJ_LAUNCH_CHECK();
  timer_begin(AJ_TIMER_SOLVE, st);
  AJ_CHECK_HIP(hipFuncSetAttribute((const void*)solve_loglike_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(double) * maxS)));
  hipLaunchKernelGGL(solve_loglike_kernel, dim3((unsigned)n), dim3(SOLVE_THREADS), sizeof(double) * maxS, st, desc, mats, D,
                     chi2_nn, logdet, recon, logl, solve_finishes ? ldp : nullptr, failp, (int)p.maxsteps, info,
                     solve_finishes ? fsync : nullptr,
                     reinterpret_cast<const int*>(ws + p.off_fqoff) + FLOW_QUEUES + 1);
  timer_end(AJ_TIMER_SOLVE, st);
  timer_end(AJ_TIMER_BATCH, st);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

}  // namespace

extern "C" size_t aj_log_evidence_batched_workspace_bytes(int64_t M, int64_t B, int64_t n_samples,
                                                          const int64_t* S_host) {
  if (M <= 0 || B <= 0 || n_samples <= 0 || !S_host) return 0;
  return plan_batch(M, B, n_samples, S_host).bytes;
}

extern "C" int aj_chi_squared_noise_normalization(const double* data, const double* noise_map, int64_t K,
                                                  double* out2, void* stream) {
  AJ_REQUIRE(K >= 0 && data && noise_map && out2, "bad arguments");
  hipLaunchKernelGGL(chi2_noise_kernel, dim3(1), dim3(256), 0, as_stream(stream), data, noise_map, 2 * K, out2);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" int aj_noise_normalization_complex_from(const double* noise_map, int64_t K, double* out, void* stream) {
  AJ_REQUIRE(K >= 0 && noise_map && out, "bad arguments");
  hipLaunchKernelGGL(noise_norm_kernel, dim3(1), dim3(256), 0, as_stream(stream), noise_map, 2 * K, out);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" int aj_log_evidence_batched(const double* w_tilde, int64_t M, const double* dirty_image,
                                       const double* chi2_noise_norm, int64_t n_samples, int64_t B, int64_t P,
                                       const int64_t* S_host, const int64_t* pix_indexes, const double* pix_weights,
                                       const int64_t* neighbors, const int64_t* neighbors_sizes,
                                       const double* coefficients_host, const double* deflation_beta_host,
                                       double* log_evidence, double* reconstruction, double* logdets,
                                       int64_t* info, void* workspace, size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && B > 0 && P > 0 && n_samples > 0, "bad sizes");
  AJ_REQUIRE(w_tilde && dirty_image && chi2_noise_norm && S_host && pix_indexes && pix_weights && neighbors &&
                 neighbors_sizes && coefficients_host && log_evidence,
             "null pointer");
  return batch_pipeline(WSrc{w_tilde, nullptr, nullptr, 0}, M, dirty_image, chi2_noise_norm, n_samples, B, P, S_host,
                        pix_indexes, pix_weights,
                        neighbors, neighbors_sizes, coefficients_host, deflation_beta_host, log_evidence,
                        reconstruction, logdets, info, workspace, workspace_bytes, as_stream(stream));
}

extern "C" size_t aj_log_likelihood_workspace_bytes(int64_t M, int64_t K, int64_t B, int64_t S, int64_t P) {
  (void)K;
  (void)P;
  if (M <= 0 || B <= 0 || S <= 0) return 0;
  return align_up(plan_batch(M, B, 1, &S).bytes, 256) + 256;
}

extern "C" int aj_log_likelihood_function_via_w_tilde_from(
    const double* dirty_image, const double* data, const double* noise_map, const double* w_tilde,
    const int64_t* pix_indexes_for_sub_slim_index, const double* pix_weights_for_sub_slim_index,
    const int64_t* neighbors, const int64_t* neighbors_sizes, int64_t M, int64_t K, int64_t B, int64_t S, int64_t P,
    double* log_likelihood, int64_t* info, void* workspace, size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && K >= 0 && B > 0 && S > 0 && P > 0, "bad sizes");
  AJ_REQUIRE(dirty_image && data && noise_map && w_tilde && pix_indexes_for_sub_slim_index &&
                 pix_weights_for_sub_slim_index && neighbors && neighbors_sizes && log_likelihood && info,
             "null pointer");
  const size_t need = aj_log_likelihood_workspace_bytes(M, K, B, S, P);
  AJ_REQUIRE(workspace && workspace_bytes >= need, "workspace too small: need %zu bytes, got %zu", need,
             workspace_bytes);
  hipStream_t st = as_stream(stream);
  double* chi2nn = reinterpret_cast<double*>(static_cast<char*>(workspace) + need - 256);
  hipLaunchKernelGGL(chi2_noise_kernel, dim3(1), dim3(256), 0, st, data, noise_map, 2 * K, chi2nn);
  AJ_LAUNCH_CHECK();
  const double coef = 1.0;  // numba.py:819
  return batch_pipeline(WSrc{w_tilde, nullptr, nullptr, 0}, M, dirty_image, chi2nn, 1, B, P, &S,
                        pix_indexes_for_sub_slim_index,
                        pix_weights_for_sub_slim_index, neighbors, neighbors_sizes, &coef, nullptr, log_likelihood,
                        nullptr, nullptr, info, workspace, need - 256, st);
}

extern "C" int aj_log_likelihood_function_via_w_compact_from(
    const double* dirty_image, const double* data, const double* noise_map, const double* w_compact, int64_t grid_size,
    const int64_t* native_index_for_slim_index, const int64_t* pix_indexes_for_sub_slim_index,
    const double* pix_weights_for_sub_slim_index, const int64_t* neighbors, const int64_t* neighbors_sizes, int64_t M,
    int64_t K, int64_t B, int64_t S, int64_t P, double* log_likelihood, int64_t* info, void* workspace,
    size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && K >= 0 && B > 0 && S > 0 && P > 0 && grid_size > 0, "bad sizes");
  AJ_REQUIRE(dirty_image && data && noise_map && w_compact && native_index_for_slim_index &&
                 pix_indexes_for_sub_slim_index && pix_weights_for_sub_slim_index && neighbors && neighbors_sizes &&
                 log_likelihood && info,
             "null pointer");
  const size_t need = aj_log_likelihood_workspace_bytes(M, K, B, S, P);
  AJ_REQUIRE(workspace && workspace_bytes >= need, "workspace too small: need %zu bytes, got %zu", need,
             workspace_bytes);
  hipStream_t st = as_stream(stream);
  double* chi2nn = reinterpret_cast<double*>(static_cast<char*>(workspace) + need - 256);
  hipLaunchKernelGGL(chi2_noise_kernel, dim3(1), dim3(256), 0, st, data, noise_map, 2 * K, chi2nn);
  AJ_LAUNCH_CHECK();
  const double coef = 1.0;  // numba.py:923
  return batch_pipeline(WSrc{nullptr, w_compact, native_index_for_slim_index, grid_size}, M, dirty_image, chi2nn, 1, B,
                        P, &S, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, neighbors,
                        neighbors_sizes, &coef, nullptr, log_likelihood, nullptr, nullptr, info, workspace, need - 256,
                        st);
}

extern "C" int aj_log_evidence_batched_compact(const double* w_compact, int64_t grid_size,
                                               const int64_t* native_index_for_slim_index, int64_t M,
                                               const double* dirty_image, const double* chi2_noise_norm,
                                               int64_t n_samples, int64_t B, int64_t P, const int64_t* S_host,
                                               const int64_t* pix_indexes, const double* pix_weights,
                                               const int64_t* neighbors, const int64_t* neighbors_sizes,
                                               const double* coefficients_host, const double* deflation_beta_host,
                                               double* log_evidence, double* reconstruction, double* logdets,
                                               int64_t* info, void* workspace, size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && B > 0 && P > 0 && n_samples > 0 && grid_size > 0, "bad sizes");
  AJ_REQUIRE(w_compact && native_index_for_slim_index && dirty_image && chi2_noise_norm && S_host && pix_indexes &&
                 pix_weights && neighbors && neighbors_sizes && coefficients_host && log_evidence,
             "null pointer");
  return batch_pipeline(WSrc{nullptr, w_compact, native_index_for_slim_index, grid_size}, M, dirty_image,
                        chi2_noise_norm, n_samples, B, P, S_host, pix_indexes, pix_weights, neighbors,
                        neighbors_sizes, coefficients_host, deflation_beta_host, log_evidence, reconstruction, logdets,
                        info, workspace, workspace_bytes, as_stream(stream));
}

// ===========================================================================
// Single-function entry points mirroring the individual reference functions.
// ===========================================================================
namespace aj {
namespace lk {

__global__ void mapping_dense_kernel(const int64_t* __restrict__ idx, const double* __restrict__ w, int64_t M,
                                     int64_t B, int64_t S, double* __restrict__ T) {
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (int64_t)gridDim.x * blockDim.x) {
    for (int64_t b = 0; b < B; ++b) {
      const int64_t s = idx[m * B + b];
      if (s == -1) break;
      if (s >= 0 && s < S) T[m * S + s] = w[m * B + b];  // assignment, as numba.py:375
    }
  }
}

__global__ void dense_data_vector_kernel(const double* __restrict__ T, const double* __restrict__ dirty, int64_t M,
                                         int64_t S, double* __restrict__ D) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < S; s += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int64_t m = 0; m < M; ++m) acc = fma(dirty[m], T[m * S + s], acc);
    D[s] = acc;
  }
}

// C[i][j] = sum_k opA(i,k) B[k][j]; opA(i,k) = A[i*lda+k] (trans=0) or A[k*lda+i] (trans=1).
// B row-major with ldb, C row-major with ldc.  64x64 tiles, FP64 MFMA.
__global__ void __launch_bounds__(256) gemm_f64_kernel(const double* __restrict__ A, int64_t lda, int transA,
                                                       const double* __restrict__ Bm, int64_t ldb,
                                                       double* __restrict__ C, int64_t ldc, int64_t Mr, int64_t N,
                                                       int64_t Kd) {
  __shared__ double As[64][17];
  __shared__ double Bs[16][65];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t i0 = (int64_t)blockIdx.y * 64, j0 = (int64_t)blockIdx.x * 64;
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0, 0, 0, 0};
  for (int64_t k0 = 0; k0 < Kd; k0 += 16) {
    for (int e = tid; e < 64 * 16; e += 256) {
      int r, c;
      if (transA) {
        c = e / 64;
        r = e % 64;
      } else {
        r = e / 16;
        c = e % 16;
      }
      const int64_t gi = i0 + r, gk = k0 + c;
      double v = 0.0;
      if (gi < Mr && gk < Kd) v = transA ? A[gk * lda + gi] : A[gi * lda + gk];
      As[r][c] = v;
      const int kr = e / 64, jc = e % 64;
      const int64_t bk = k0 + kr, bj = j0 + jc;
      Bs[kr][jc] = (bk < Kd && bj < N) ? Bm[bk * ldb + bj] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      double fa[2], fb[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        fa[a] = As[wr + a * 16 + lr][ks * 4 + lk];
        fb[a] = Bs[ks * 4 + lk][wc + a * 16 + lr];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = i0 + wr + a * 16 + lk + 4 * r, gj = j0 + wc + b * 16 + lr;
        if (gi < Mr && gj < N) C[gi * ldc + gj] = acc[a][b][r];
      }
}

// out[s1, :] = sum_{(m,t) in bucket s1} t * X[m, :]   (numba.py:578-586), X is [M, W].
__global__ void tgather_kernel(const SampleDesc* __restrict__ sd, const int* __restrict__ ptrs,
                               const int* __restrict__ ent_m, const double* __restrict__ ent_t,
                               const double* __restrict__ X, int64_t Wd, double* __restrict__ out) {
  const SampleDesc d = sd[0];
  const int64_t s1 = blockIdx.x;
  const int* p = ptrs + d.csc_ptr;
  const int lo = p[s1], hi = p[s1 + 1];
  for (int64_t c = threadIdx.x; c < Wd; c += blockDim.x) {
    double acc = 0.0;
    for (int e = lo; e < hi; ++e) acc = fma(ent_t[d.csc_ent + e], X[(int64_t)ent_m[d.csc_ent + e] * Wd + c], acc);
    out[s1 * Wd + c] = acc;
  }
}

__global__ void reg_standalone_kernel(double coef2, const int64_t* __restrict__ nbr, const int64_t* __restrict__ nsz,
                                      int64_t S, int64_t P, double* __restrict__ H) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S; i += (int64_t)gridDim.x * blockDim.x) {
    H[i * S + i] += 1e-8 + coef2 * (double)nsz[i];
    for (int64_t j = 0; j < P; ++j) {
      const int64_t k = nbr[i * P + j];
      if (k == -1) break;
      if (k >= 0 && k < S) H[i * S + k] -= coef2;
    }
  }
}

// Data-side sums over visibilities, per image pixel, absolute phases:
//   out_i = sum_k (a_k cos 2pi r_ik + b_k sin 2pi r_ik),  r_ik = g_i1 u_k0 + g_i0 u_k1.
// Partial sums per K-split, reduced in a fixed order (deterministic).
constexpr int DSPLIT = 8192;
__global__ void __launch_bounds__(256) vis_sum_partial_kernel(const double* __restrict__ a, int64_t a_stride,
                                                              const double* __restrict__ b, int64_t b_stride,
                                                              const double* __restrict__ wdiv, int64_t wstride,
                                                              int mode,
                                                              const double* __restrict__ uv,
                                                              const double* __restrict__ grid, int64_t K, int64_t M,
                                                              double* __restrict__ partial) {
  __shared__ double su0[256], su1[256], sa[256], sb[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t kb = (int64_t)blockIdx.y * DSPLIT;
  const int64_t ke = std::min<int64_t>(K, kb + DSPLIT);
  double gx = 0.0, gy = 0.0;
  if (i < M) {
    gy = grid[2 * i];
    gx = grid[2 * i + 1];
  }
  double acc = 0.0;
  for (int64_t k0 = kb; k0 < ke; k0 += 256) {
    const int64_t k = k0 + threadIdx.x;
    __syncthreads();
    if (k < ke) {
      su0[threadIdx.x] = uv[2 * k];
      su1[threadIdx.x] = uv[2 * k + 1];
      double av, bv;
      if (mode == 0) {
        // dirty image: (Re d, Im d) / sigma^2 with sigma = Re noise
        const double n = wdiv[wstride * k];
        av = a[a_stride * k] / (n * n);
        bv = b[b_stride * k] / (n * n);
      } else {
        // numba.py:149: w_k = (n_k^2 / v_k)^2 on the cos term only
        const double n = wdiv[wstride * k];
        const double q = n * n / a[k];
        av = q * q;
        bv = 0.0;
      }
      sa[threadIdx.x] = av;
      sb[threadIdx.x] = bv;
    } else {
      su0[threadIdx.x] = su1[threadIdx.x] = sa[threadIdx.x] = sb[threadIdx.x] = 0.0;
    }
    __syncthreads();
    const int nk = (int)std::min<int64_t>(256, ke - k0);
    if (i < M) {
      for (int j = 0; j < nk; ++j) {
        double r = fma(gx, su0[j], gy * su1[j]);
        r -= rint(r);
        double sv, cv;
        sincospi(2.0 * r, &sv, &cv);
        acc = fma(sa[j], cv, acc);
        if (mode == 0) acc = fma(sb[j], sv, acc);
      }
    }
  }
  if (i < M) partial[(int64_t)blockIdx.y * M + i] = acc;
}

__global__ void vis_sum_reduce_kernel(const double* __restrict__ partial, int64_t nsplit, int64_t M,
                                      double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int64_t s = 0; s < nsplit; ++s) acc += partial[s * M + i];
    out[i] = acc;
  }
}

}  // namespace lk
}  // namespace aj

namespace {

// CSC for one sample (mapping [M,B]) in a workspace; returns the views.
struct OneCsc {
  SampleDesc* desc;
  int* ptrs;
  int* ekey;
  double* et;
};

int build_one_csc(const int64_t* idx, const double* w, int64_t M, int64_t B, int64_t S, Arena& ar, OneCsc& o,
                  hipStream_t st) {
  o.desc = ar.take<SampleDesc>(1);
  o.ptrs = ar.take<int>(S + 1);
  int* cur = ar.take<int>(S + 1);
  o.ekey = ar.take<int>(M * B);
  o.et = ar.take<double>(M * B);
  if (!ar.ok()) {
    set_error("workspace too small");
    return AJ_ESPACE;
  }
  SampleDesc h{S, 0, 0, 0, 0, 0, 0, 0.0, 0.0};
  AJ_CHECK_HIP(hipMemcpyAsync(o.desc, &h, sizeof(h), hipMemcpyHostToDevice, st));
  AJ_CHECK_HIP(hipMemsetAsync(o.ptrs, 0, sizeof(int) * (S + 1), st));
  const unsigned gm = (unsigned)std::min<int64_t>(ceil_div(M, 256), 1024);
  hipLaunchKernelGGL(csc_count_kernel, dim3(gm, 1), dim3(256), 0, st, o.desc, idx, M, B, o.ptrs);
  hipLaunchKernelGGL(csc_scan_kernel, dim3(1), dim3(1024), 0, st, o.desc, o.ptrs, cur);
  hipLaunchKernelGGL(csc_fill_kernel, dim3(gm, 1), dim3(256), 0, st, o.desc, idx, w, M, B, cur, o.ekey, o.et);
  hipLaunchKernelGGL(csc_sort_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(S, 4), 65535), 1), dim3(256), 0, st,
                     o.desc, o.ptrs, B, o.ekey, o.et);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

size_t csc_bytes(int64_t M, int64_t B, int64_t S) {
  return 4 * 256 + sizeof(SampleDesc) + 2 * sizeof(int) * (S + 1) + (sizeof(int) + sizeof(double)) * M * B + 1024;
}

}  // namespace

extern "C" int aj_mapping_matrix_from(const int64_t* pix_indexes_for_sub_slim_index,
                                      const double* pix_weights_for_sub_slim_index, int64_t M, int64_t B,
                                      int64_t pixels, double* mapping_matrix, void* stream) {
  AJ_REQUIRE(M > 0 && B > 0 && pixels > 0, "bad sizes");
  AJ_REQUIRE(pix_indexes_for_sub_slim_index && pix_weights_for_sub_slim_index && mapping_matrix, "null pointer");
  hipStream_t st = as_stream(stream);
  AJ_CHECK_HIP(hipMemsetAsync(mapping_matrix, 0, sizeof(double) * M * pixels, st));
  hipLaunchKernelGGL(mapping_dense_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(M, 256), 1024)), dim3(256), 0, st,
                     pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, M, B, pixels, mapping_matrix);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" int aj_data_vector_from(const double* mapping_matrix, const double* dirty_image, int64_t M, int64_t S,
                                   double* data_vector, void* stream) {
  AJ_REQUIRE(M > 0 && S > 0 && mapping_matrix && dirty_image && data_vector, "bad arguments");
  hipLaunchKernelGGL(dense_data_vector_kernel, dim3((unsigned)ceil_div(S, 256)), dim3(256), 0, as_stream(stream),
                     mapping_matrix, dirty_image, M, S, data_vector);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" size_t aj_curvature_dense_workspace_bytes(int64_t M, int64_t S) {
  return (M > 0 && S > 0) ? sizeof(double) * M * S : 0;
}

extern "C" int aj_curvature_matrix_via_w_tilde_from(const double* w_tilde, const double* mapping_matrix, int64_t M,
                                                    int64_t S, double* curvature_matrix, void* workspace,
                                                    size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && S > 0 && w_tilde && mapping_matrix && curvature_matrix, "bad arguments");
  AJ_REQUIRE(workspace && workspace_bytes >= aj_curvature_dense_workspace_bytes(M, S), "workspace too small");
  hipStream_t st = as_stream(stream);
  double* omega = static_cast<double*>(workspace);
  // Omega = W T  (M x S), then F = T^T Omega (S x S): the two products of numba.py:435.
  hipLaunchKernelGGL(gemm_f64_kernel, dim3((unsigned)ceil_div(S, 64), (unsigned)ceil_div(M, 64)), dim3(256), 0, st,
                     w_tilde, M, 0, mapping_matrix, S, omega, S, M, S, M);
  hipLaunchKernelGGL(gemm_f64_kernel, dim3((unsigned)ceil_div(S, 64), (unsigned)ceil_div(S, 64)), dim3(256), 0, st,
                     mapping_matrix, S, 1, omega, S, curvature_matrix, S, S, S, M);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" size_t aj_sparse_transpose_matmul_workspace_bytes(int64_t M, int64_t B, int64_t S) {
  return (M > 0 && B > 0 && S > 0) ? csc_bytes(M, B, S) : 0;
}

extern "C" int aj_sparse_mapping_matrix_transpose_matmul(const double* matrix, const int64_t* pix_indexes_for_sub_slim_index,
                                                         const double* pix_weights_for_sub_slim_index, int64_t M,
                                                         int64_t S2, int64_t B, int64_t pixels, double* out,
                                                         void* workspace, size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && S2 > 0 && B > 0 && pixels > 0, "bad sizes");
  AJ_REQUIRE(matrix && pix_indexes_for_sub_slim_index && pix_weights_for_sub_slim_index && out && workspace,
             "null pointer");
  hipStream_t st = as_stream(stream);
  Arena ar(workspace, workspace_bytes);
  OneCsc c;
  int rc = build_one_csc(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, M, B, pixels, ar, c, st);
  if (rc) return rc;
  hipLaunchKernelGGL(tgather_kernel, dim3((unsigned)pixels), dim3(256), 0, st, c.desc, c.ptrs, c.ekey, c.et, matrix,
                     S2, out);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" size_t aj_curvature_sparse_workspace_bytes(int64_t M, int64_t B, int64_t S) {
  return (M > 0 && B > 0 && S > 0) ? csc_bytes(M, B, S) + sizeof(double) * M * S + 256 : 0;
}

static int sparse_curvature(const WSrc w_tilde, const int64_t* pix_indexes_for_sub_slim_index,
                            const double* pix_weights_for_sub_slim_index, int64_t M, int64_t B, int64_t pixels,
                            double* curvature_matrix, void* workspace, size_t workspace_bytes, hipStream_t st) {
  Arena ar(workspace, workspace_bytes);
  OneCsc c;
  int rc = build_one_csc(pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, M, B, pixels, ar, c, st);
  if (rc) return rc;
  double* omega = ar.take<double>(M * pixels);
  AJ_REQUIRE(ar.ok(), "workspace too small");
  const int R = (int)std::min<int64_t>(8, (int64_t)(OMEGA_LDS_BUDGET / (sizeof(double) * M)));
  if (R >= 1) {
    AJ_CHECK_HIP(hipFuncSetAttribute((const void*)omega_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(R * M * sizeof(double))));
    hipLaunchKernelGGL(omega_kernel, dim3((unsigned)ceil_div(M, R)), dim3(OMEGA_THREADS),
                       (size_t)R * M * sizeof(double), st, w_tilde, M, R, 1, c.desc, c.ptrs, c.ekey, c.et, omega);
  } else {
    hipLaunchKernelGGL(omega_global_kernel, dim3((unsigned)M), dim3(256), 0, st, w_tilde, M, 1, c.desc, c.ptrs,
                       c.ekey, c.et, omega);
  }
  hipLaunchKernelGGL(tgather_kernel, dim3((unsigned)pixels), dim3(256), 0, st, c.desc, c.ptrs, c.ekey, c.et, omega,
                     pixels, curvature_matrix);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" int aj_curvature_matrix_via_w_wilde_sparse_mapping_matrix_from(
    const double* w_tilde, const int64_t* pix_indexes_for_sub_slim_index, const double* pix_weights_for_sub_slim_index,
    int64_t M, int64_t B, int64_t pixels, double* curvature_matrix, void* workspace, size_t workspace_bytes,
    void* stream) {
  AJ_REQUIRE(M > 0 && B > 0 && pixels > 0, "bad sizes");
  AJ_REQUIRE(w_tilde && pix_indexes_for_sub_slim_index && pix_weights_for_sub_slim_index && curvature_matrix &&
                 workspace,
             "null pointer");
  return sparse_curvature(WSrc{w_tilde, nullptr, nullptr, 0}, pix_indexes_for_sub_slim_index,
                          pix_weights_for_sub_slim_index, M, B, pixels, curvature_matrix, workspace, workspace_bytes,
                          as_stream(stream));
}

extern "C" int aj_curvature_matrix_via_w_compact_sparse_mapping_matrix_from(
    const double* w_compact, int64_t grid_size, const int64_t* native_index_for_slim_index,
    const int64_t* pix_indexes_for_sub_slim_index, const double* pix_weights_for_sub_slim_index, int64_t M, int64_t B,
    int64_t pixels, double* curvature_matrix, void* workspace, size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && B > 0 && pixels > 0 && grid_size > 0, "bad sizes");
  AJ_REQUIRE(w_compact && native_index_for_slim_index && pix_indexes_for_sub_slim_index &&
                 pix_weights_for_sub_slim_index && curvature_matrix && workspace,
             "null pointer");
  return sparse_curvature(WSrc{nullptr, w_compact, native_index_for_slim_index, grid_size},
                          pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, M, B, pixels,
                          curvature_matrix, workspace, workspace_bytes, as_stream(stream));
}

extern "C" int aj_curvature_matrix_via_w_tilde_curvature_preload_interferometer_from(
    const double* curvature_preload, int64_t preload_rows, int64_t preload_cols,
    const int64_t* native_index_for_slim_index, const int64_t* pix_indexes_for_sub_slim_index,
    const double* pix_weights_for_sub_slim_index, int64_t M, int64_t B, int64_t pixels, double* curvature_matrix,
    void* workspace, size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && B > 0 && pixels > 0 && preload_rows > 0 && preload_cols > 0, "bad sizes");
  AJ_REQUIRE(curvature_preload && native_index_for_slim_index && pix_indexes_for_sub_slim_index &&
                 pix_weights_for_sub_slim_index && curvature_matrix && workspace,
             "null pointer");
  return sparse_curvature(WSrc{nullptr, nullptr, native_index_for_slim_index, 0, curvature_preload, preload_rows,
                               preload_cols},
                          pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index, M, B, pixels,
                          curvature_matrix, workspace, workspace_bytes, as_stream(stream));
}

extern "C" int aj_constant_regularization_matrix_from(double coefficient, const int64_t* neighbors,
                                                      const int64_t* neighbors_sizes, int64_t S, int64_t P,
                                                      double* regularization_matrix, void* stream) {
  AJ_REQUIRE(S > 0 && P > 0 && neighbors && neighbors_sizes && regularization_matrix, "bad arguments");
  hipStream_t st = as_stream(stream);
  AJ_CHECK_HIP(hipMemsetAsync(regularization_matrix, 0, sizeof(double) * S * S, st));
  hipLaunchKernelGGL(reg_standalone_kernel, dim3((unsigned)ceil_div(S, 256)), dim3(256), 0, st,
                     coefficient * coefficient, neighbors, neighbors_sizes, S, P, regularization_matrix);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" size_t aj_visibility_sum_workspace_bytes(int64_t K, int64_t M) {
  return (M > 0 && K >= 0) ? sizeof(double) * M * std::max<int64_t>(1, ceil_div(K, DSPLIT)) : 0;
}

int aj::vis_sum(const double* a, int64_t as, const double* b, int64_t bs, const double* wdiv, int64_t wstride,
                int mode, const double* uv, const double* grid, int64_t K, int64_t M, double* out, void* ws,
                size_t wsb, hipStream_t st) {
  AJ_REQUIRE(M > 0 && K >= 0 && grid && out && (K == 0 || (a && wdiv && uv)), "bad arguments");
  AJ_REQUIRE(ws && wsb >= aj_visibility_sum_workspace_bytes(K, M), "workspace too small");
  const int64_t ns = std::max<int64_t>(1, ceil_div(K, DSPLIT));
  double* part = static_cast<double*>(ws);
  hipLaunchKernelGGL(vis_sum_partial_kernel, dim3((unsigned)ceil_div(M, 256), (unsigned)ns), dim3(256), 0, st, a, as,
                     b, bs, wdiv, wstride, mode, uv, grid, K, M, part);
  hipLaunchKernelGGL(vis_sum_reduce_kernel, dim3((unsigned)ceil_div(M, 256)), dim3(256), 0, st, part, ns, M, out);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" int aj_dirty_image_from(const double* data, const double* noise_map, const double* uv_wavelengths,
                                   const double* grid_radians_slim, int64_t K, int64_t M, double* dirty_image,
                                   void* workspace, size_t workspace_bytes, void* stream) {
  return vis_sum(data, 2, data ? data + 1 : nullptr, 2, noise_map, 2, 0, uv_wavelengths, grid_radians_slim, K, M,
                 dirty_image, workspace, workspace_bytes, as_stream(stream));
}

extern "C" int aj_w_tilde_data_interferometer_from(const double* visibilities_real, const double* noise_map_real,
                                                   const double* uv_wavelengths, const double* grid_radians_slim,
                                                   int64_t K, int64_t M, double* w_tilde_data, void* workspace,
                                                   size_t workspace_bytes, void* stream) {
  return vis_sum(visibilities_real, 1, nullptr, 0, noise_map_real, 1, 1, uv_wavelengths, grid_radians_slim, K, M,
                 w_tilde_data, workspace, workspace_bytes, as_stream(stream));
}

// ===========================================================================
// Backward of the batched log-evidence (no reference counterpart; the
// reference differentiates its JAX form, jax.py:1032-1058, with autodiff).
// With A = F + H, s = A^-1 D, F = T^T W T, D = T^T y, H = c^2 L + eps I:
//   dlogL/dc       = -c [ s^T L s + tr(A^-1 L) - tr(H^-1 L) ],
//                    tr(H^-1 L) = (S - eps tr H^-1) / c^2
//   dlogL/dT[m,s]  = s_s (y_m - (W T s)_m) - (W T A^-1)[m, s]
// evaluated at the nonzeros (m, b) of the mapping.  Uses the Cholesky
// factors left in the forward workspace.
// ===========================================================================
namespace aj {
namespace lk {

// Triangular inverses X = L^-1 of every factor of the batch (A and H), in place in a copy of the factors, by
// recursive halving: the 32 x 32 diagonal blocks are inverted one per wave; then at each level the sibling
// blocks (a, a + h) of size h (the second h2 <= h rows) are joined,
//   X21 = -X22 T,   T = L21 X11,
// two MFMA products (T kept transposed in the mirror of X21's block above the diagonal, which X does not
// use).  ceil(log2(S / 32)) levels of two launches, all matrices' blocks of a level in one grid.
constexpr int IB = 32;  // diagonal block of the recursion
struct InvTask {
  int mat;  // matrix: sample mat >> 1, A (even) or H (odd)
  int a;    // first row and column of the pair
  int ti;   // 64-row tile of the h2 rows below the pair's first half
  int tj;   // 64-column tile of the h columns of the first half
};

// One wave per (matrix, diagonal block): lane l holds row l of the block's L (r) and builds row l of its
// inverse (x), row k final when lane k divides by L_kk, broadcast by readlane to the rows below.
__global__ void __launch_bounds__(256) trtri_diag_kernel(const SampleDesc* __restrict__ sd,
                                                         const int2* __restrict__ blocks, int nblocks,
                                                         double* __restrict__ X) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = (int)blockIdx.x * 4 + wave;
  if (t >= nblocks) return;
  const int j = blocks[t].x;
  const int64_t k0 = (int64_t)blocks[t].y * IB;
  int64_t n;
  double* A = mat_ptr(sd, X, j, n);
  const int nb = (int)std::min<int64_t>(IB, n - k0);
  double r[IB], x[IB];
#pragma unroll
  for (int c = 0; c < IB; ++c) {
    r[c] = (lane < nb && c <= lane) ? A[(k0 + lane) * n + k0 + c] : 0.0;
    x[c] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < IB; ++k) {
    if (k < nb) {
      if (lane == k) {
#pragma unroll
        for (int c = 0; c <= k; ++c) x[c] = ((c == k ? 1.0 : 0.0) - x[c]) / r[k];
      }
#pragma unroll
      for (int c = 0; c <= k; ++c) {
        const double xkc = readlane_d(x[c], k);
        if (lane > k) x[c] = fma(r[k], xkc, x[c]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < IB; ++c)
    if (lane < nb && c <= lane) A[(k0 + lane) * n + k0 + c] = x[c];
}

// One 64 x 64 output tile of a level's product, K in chunks of 32 staged in LDS, four waves of 32 x 32.
//   step 1: T[i][c] = sum_{k >= c} L21[i][k] X11[k][c], i < h2, c < h; stored as T^T above the diagonal.
//   step 2: X21[i][c] = -sum_{k <= i} X22[i][k] T[k][c]; stored in X21's place.
template <int STEP>
__global__ void __launch_bounds__(256) trtri_level_kernel(const SampleDesc* __restrict__ sd,
                                                          const InvTask* __restrict__ tasks, int64_t h,
                                                          double* __restrict__ X) {
  __shared__ double As[64][33];
  __shared__ double Bs[32][65];
  const InvTask tk = tasks[blockIdx.x];
  int64_t n;
  double* Xm = mat_ptr(sd, X, tk.mat, n);
  const int64_t a = tk.a, h2 = std::min<int64_t>(h, n - a - h);
  const int64_t i0 = (int64_t)tk.ti * 64, c0 = (int64_t)tk.tj * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32, lr = lane & 15, lk = lane >> 4;
  // k range: step 1 from c0 (X11 is lower triangular) to h; step 2 from 0 to the tile's last row (X22 lower)
  const int64_t kbeg = STEP == 1 ? c0 : 0;
  const int64_t kend = STEP == 1 ? h : std::min<int64_t>(h2, i0 + 64);
  d4 acc[2][2];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q) acc[p][q] = d4{0.0, 0.0, 0.0, 0.0};
  for (int64_t kc = kbeg; kc < kend; kc += 32) {
    for (int e = tid; e < 64 * 32; e += 256) {
      const int i = e >> 5, k = e & 31;  // A operand (64 rows x 32 k), k fastest: coalesced in k
      const int64_t gi = i0 + i, gk = kc + k;
      double v = 0.0;
      if (gi < h2 && gk < kend) {
        if (STEP == 1) v = Xm[(a + h + gi) * n + a + gk];                          // L21
        else if (gk <= gi) v = Xm[(a + h + gi) * n + a + h + gk];                  // X22, lower part
      }
      As[i][k] = v;
    }
    for (int e = tid; e < 32 * 64; e += 256) {
      double v = 0.0;
      if (STEP == 1) {
        const int k = e >> 6, c = e & 63;  // X11[k][c]: a row of X11, coalesced in c
        const int64_t gk = kc + k, gc = c0 + c;
        if (gk < kend && gc < h && gc <= gk) v = Xm[(a + gk) * n + a + gc];
        Bs[k][c] = v;
      } else {
        const int c = e >> 5, k = e & 31;  // T[k][c] = T^T[c][k]: a row of the stored T^T, coalesced in k
        const int64_t gk = kc + k, gc = c0 + c;
        if (gk < kend && gc < h) v = Xm[(a + gc) * n + a + h + gk];
        Bs[k][c] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const int k = ks * 4 + lk;
      const double fa0 = As[wr + lr][k], fa1 = As[wr + 16 + lr][k];
      const double fb0 = Bs[k][wc + lr], fb1 = Bs[k][wc + 16 + lr];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa0, fb0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa0, fb1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1, fb0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1, fb1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = i0 + wr + p * 16 + lk + 4 * r, gc = c0 + wc + q * 16 + lr;
        if (gi < h2 && gc < h) {
          if (STEP == 1) Xm[(a + gc) * n + a + h + gi] = acc[p][q][r];
          else Xm[(a + h + gi) * n + a + gc] = -acc[p][q][r];
        }
      }
}

// A^-1 = X^T X for the A factor of every sample (X = L_A^-1 lower triangular), both triangles, into the
// sample's S x S block of inv: tile (I, J), I >= J, sums X[k][I0 + i] X[k][J0 + j] over k >= I0.
struct SymTask {
  int smp;
  int ti;
  int tj;
};

__global__ void __launch_bounds__(256) inverse_gram_kernel(const SampleDesc* __restrict__ sd,
                                                           const SymTask* __restrict__ tasks,
                                                           const double* __restrict__ X, double* __restrict__ inv) {
  __shared__ double As[32][65];
  __shared__ double Bs[32][65];
  const SymTask tk = tasks[blockIdx.x];
  const SampleDesc d = sd[tk.smp];
  const int64_t n = d.S;
  const double* Xm = X + d.mat_off;
  double* Ai = inv + d.mat_off / 2;
  const int64_t I0 = (int64_t)tk.ti * 64, J0 = (int64_t)tk.tj * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32, lr = lane & 15, lk = lane >> 4;
  d4 acc[2][2];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q) acc[p][q] = d4{0.0, 0.0, 0.0, 0.0};
  for (int64_t kc = I0; kc < n; kc += 32) {
    for (int e = tid; e < 32 * 64; e += 256) {
      const int k = e >> 6, c = e & 63;  // rows k of X, columns of tile I (As) and J (Bs): coalesced in c
      const int64_t gk = kc + k;
      As[k][c] = (gk < n && I0 + c < n && I0 + c <= gk) ? Xm[gk * n + I0 + c] : 0.0;
      Bs[k][c] = (gk < n && J0 + c < n && J0 + c <= gk) ? Xm[gk * n + J0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const int k = ks * 4 + lk;
      const double fa0 = As[k][wr + lr], fa1 = As[k][wr + 16 + lr];
      const double fb0 = Bs[k][wc + lr], fb1 = Bs[k][wc + 16 + lr];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa0, fb0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa0, fb1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1, fb0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1, fb1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gi = I0 + wr + p * 16 + lk + 4 * r, gj = J0 + wc + q * 16 + lr;
        if (gi < n && gj < n) {
          Ai[gi * n + gj] = acc[p][q][r];
          if (tk.ti != tk.tj) Ai[gj * n + gi] = acc[p][q][r];
        }
      }
}

// hn2[row] = sum_c X_H[row][c]^2 (c <= row): squared row norms of L_H^-1, whose sum is tr(H^-1).  One wave
// per row, four rows per workgroup, blockIdx.y the sample.
__global__ void __launch_bounds__(256) inverse_rownorm_kernel(const SampleDesc* __restrict__ sd,
                                                              const double* __restrict__ X,
                                                              double* __restrict__ hn2) {
  const SampleDesc d = sd[blockIdx.y];
  const int64_t n = d.S;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const double* row = X + d.mat_off + n * n + i * n;
  double acc = 0.0;
  for (int64_t c = lane; c <= i; c += 64) acc = fma(row[c], row[c], acc);
  acc = wave_sum(acc);
  if (lane == 0) hn2[d.s_off + i] = acc;
}

__global__ void copy_doubles_kernel(const double* __restrict__ src, int64_t count, double* __restrict__ dst) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

// Per-sample scalar gradient: one workgroup per sample.
__global__ void __launch_bounds__(256) coef_grad_kernel(const SampleDesc* __restrict__ sd,
                                                        const int64_t* __restrict__ nbr,
                                                        const int64_t* __restrict__ nsz, int64_t P,
                                                        const double* __restrict__ recon,
                                                        const double* __restrict__ inv_all,
                                                        const double* __restrict__ hnorm2,
                                                        const double* __restrict__ coef, double* __restrict__ dcoef) {
  __shared__ double r0[256], r1[256], r2[256];
  const SampleDesc d = sd[blockIdx.x];
  const int64_t n = d.S;
  const double* s = recon + d.s_off;
  const double* Ai = inv_all + d.mat_off / 2;
  double sls = 0.0, tra = 0.0, trh = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const double sz = (double)nsz[d.s_off + i];
    sls += sz * s[i] * s[i];
    tra += sz * Ai[i * n + i];
    trh += hnorm2[d.s_off + i];
    const int64_t* row = nbr + (d.s_off + i) * P;
    for (int64_t p = 0; p < P; ++p) {
      const int64_t k = row[p];
      if (k == -1) break;
      if (k < 0 || k >= n) continue;
      sls -= s[i] * s[k];
      tra -= Ai[i * n + k];
    }
  }
  r0[threadIdx.x] = sls;
  r1[threadIdx.x] = tra;
  r2[threadIdx.x] = trh;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      r0[threadIdx.x] += r0[threadIdx.x + o];
      r1[threadIdx.x] += r1[threadIdx.x + o];
      r2[threadIdx.x] += r2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double eps = 1e-8;
    const double c = coef[blockIdx.x];
    double tr_hinv = r2[0];
    if (d.beta > 0.0) tr_hinv += 1.0 / eps - 1.0 / (eps + d.beta);  // undo the rank-one deflation
    const double tr_hl = ((double)n - eps * tr_hinv) / (c * c);
    dcoef[blockIdx.x] = -c * (r0[0] + r1[0] - tr_hl);
  }
}

// u = T s per sample (M-vector), u_all[b*M + m].
__global__ void map_recon_kernel(const SampleDesc* __restrict__ sd, const int64_t* __restrict__ idx,
                                 const double* __restrict__ wts, int64_t M, int64_t B,
                                 const double* __restrict__ recon, double* __restrict__ u_all) {
  const int smp = blockIdx.y;
  const SampleDesc d = sd[smp];
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int64_t b = 0; b < B; ++b) {
      const int64_t s = idx[d.map_off + m * B + b];
      if (s == -1) break;
      if (s >= 0 && s < d.S) acc = fma(wts[d.map_off + m * B + b], recon[d.s_off + s], acc);
    }
    u_all[(int64_t)smp * M + m] = acc;
  }
}

// v_b = W u_b for all samples; one workgroup per row of W (read once for the batch).
__global__ void __launch_bounds__(256) batched_gemv_kernel(const double* __restrict__ W, int64_t M, int nsmp,
                                                           const double* __restrict__ u_all,
                                                           double* __restrict__ v_all) {
  __shared__ double red[4];
  const int64_t m = blockIdx.x;
  const double* w = W + m * M;
  for (int b = 0; b < nsmp; ++b) {
    const double* u = u_all + (int64_t)b * M;
    double acc = 0.0;
    for (int64_t j = threadIdx.x; j < M; j += 256) acc = fma(w[j], u[j], acc);
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) v_all[(int64_t)b * M + m] = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
  }
}

// dT[m,b] = s_s (y_m - v_m) - sum_j Omega[m, j] Ainv[s, j]; one wave per image pixel.
__global__ void __launch_bounds__(64) weight_grad_kernel(const SampleDesc* __restrict__ sd,
                                                         const int64_t* __restrict__ idx, int64_t M, int64_t B,
                                                         const double* __restrict__ dirty,
                                                         const double* __restrict__ recon,
                                                         const double* __restrict__ v_all,
                                                         const double* __restrict__ omega,
                                                         const double* __restrict__ inv_all,
                                                         double* __restrict__ dw) {
  const int smp = blockIdx.y;
  const SampleDesc d = sd[smp];
  const int64_t m = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t n = d.S;
  const double* om = omega + d.omega_off + m * n;
  const double* Ai = inv_all + d.mat_off / 2;
  bool live = true;
  for (int64_t b = 0; b < B; ++b) {
    const int64_t s = idx[d.map_off + m * B + b];
    if (s == -1) live = false;
    double g = 0.0;
    if (live && s >= 0 && s < n) {
      double acc = 0.0;
      for (int64_t j = lane; j < n; j += 64) acc = fma(om[j], Ai[s * n + j], acc);
      acc = wave_sum(acc);
      g = recon[d.s_off + s] * (dirty[m] - v_all[(int64_t)smp * M + m]) - acc;
    }
    if (lane == 0) dw[d.map_off + m * B + b] = g;
  }
}

// ---------------------------------------------------------------------------
// Barycentric weights of fixed points p_m in triangles of moving vertices,
// and the vertex gradient of a loss through them.
// ---------------------------------------------------------------------------
__device__ inline double area2(double y0, double x0, double y1, double x1, double y2, double x2) {
  return (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
}

__global__ void barycentric_kernel(const double* __restrict__ V, const double* __restrict__ Ppos,
                                   const int64_t* __restrict__ tri, int64_t M, double* __restrict__ t) {
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = tri[3 * m], b = tri[3 * m + 1], c = tri[3 * m + 2];
    const double py = Ppos[2 * m], px = Ppos[2 * m + 1];
    const double ya = V[2 * a], xa = V[2 * a + 1], yb = V[2 * b], xb = V[2 * b + 1], yc = V[2 * c], xc = V[2 * c + 1];
    const double A = area2(ya, xa, yb, xb, yc, xc);
    t[3 * m] = area2(py, px, yb, xb, yc, xc) / A;
    t[3 * m + 1] = area2(ya, xa, py, px, yc, xc) / A;
    t[3 * m + 2] = area2(ya, xa, yb, xb, py, px) / A;
  }
}

// contrib[m, k, 2] = sum_i dt_i d t_i / d V_{tri[m,k]} for the three corners k.
__global__ void barycentric_vjp_kernel(const double* __restrict__ V, const double* __restrict__ Ppos,
                                       const int64_t* __restrict__ tri, const double* __restrict__ dt, int64_t M,
                                       double* __restrict__ contrib) {
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (int64_t)gridDim.x * blockDim.x) {
    double y[3], x[3];
    for (int k = 0; k < 3; ++k) {
      const int64_t v = tri[3 * m + k];
      y[k] = V[2 * v];
      x[k] = V[2 * v + 1];
    }
    const double py = Ppos[2 * m], px = Ppos[2 * m + 1];
    const double A = area2(y[0], x[0], y[1], x[1], y[2], x[2]);
    double Ai[3], t[3];
    Ai[0] = area2(py, px, y[1], x[1], y[2], x[2]);
    Ai[1] = area2(y[0], x[0], py, px, y[2], x[2]);
    Ai[2] = area2(y[0], x[0], y[1], x[1], py, px);
    double st = 0.0;
    for (int k = 0; k < 3; ++k) {
      t[k] = Ai[k] / A;
      st += dt[3 * m + k] * t[k];
    }
    // d area2(P0,P1,P2)/d(y_k, x_k): k=0: (x2-x1, y1-y2), k=1: (x0-x2, y2-y0), k=2: (x1-x0, y0-y1)
    for (int k = 0; k < 3; ++k) {
      const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
      // full triangle
      const double dAy = x[k2] - x[k1], dAx = y[k1] - y[k2];
      // sub-triangles i != k contain corner k with p replacing corner i
      double gy = 0.0, gx = 0.0;
      for (int i = 0; i < 3; ++i) {
        if (i == k) continue;
        double yy[3] = {y[0], y[1], y[2]}, xx[3] = {x[0], x[1], x[2]};
        yy[i] = py;
        xx[i] = px;
        gy += dt[3 * m + i] * (xx[k2] - xx[k1]);
        gx += dt[3 * m + i] * (yy[k1] - yy[k2]);
      }
      contrib[6 * m + 2 * k] = (gy - st * dAy) / A;
      contrib[6 * m + 2 * k + 1] = (gx - st * dAx) / A;
    }
  }
}

// dV[s] = sum over the bucket of s (entries (m, b) with tri[m, b] = s) of contrib[m, b].
__global__ void vertex_reduce_kernel(const int* __restrict__ ptrs, const int* __restrict__ ekey,
                                     const double* __restrict__ contrib, int64_t S, double* __restrict__ dV) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < S; s += (int64_t)gridDim.x * blockDim.x) {
    double gy = 0.0, gx = 0.0;
    for (int e = ptrs[s]; e < ptrs[s + 1]; ++e) {
      const int key = ekey[e];  // m*3 + b before conversion; here raw keys are kept
      gy += contrib[2 * key];
      gx += contrib[2 * key + 1];
    }
    dV[2 * s] = gy;
    dV[2 * s + 1] = gx;
  }
}

__global__ void csc_keys_kernel(const int64_t* __restrict__ tri, int64_t M, int64_t S, int* __restrict__ cnt) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 3 * M; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = tri[e];
    if (s >= 0 && s < S) atomicAdd(&cnt[s + 1], 1);
  }
}

__global__ void csc_keys_fill_kernel(const int64_t* __restrict__ tri, int64_t M, int64_t S, int* __restrict__ cur,
                                     int* __restrict__ ekey) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 3 * M; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = tri[e];
    if (s >= 0 && s < S) ekey[atomicAdd(&cur[s], 1)] = (int)e;
  }
}

__global__ void csc_keys_sort_kernel(const int* __restrict__ ptrs, int64_t S, int* __restrict__ ekey) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < S; s += (int64_t)gridDim.x * blockDim.x) {
    for (int i = ptrs[s] + 1; i < ptrs[s + 1]; ++i) {
      int kk = ekey[i], j = i - 1;
      while (j >= ptrs[s] && ekey[j] > kk) {
        ekey[j + 1] = ekey[j];
        --j;
      }
      ekey[j + 1] = kk;
    }
  }
}

__global__ void scan_small_kernel(int* __restrict__ p, int64_t S, int* __restrict__ cur) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int64_t i = 1; i <= S; ++i) p[i] += p[i - 1];
    for (int64_t i = 0; i <= S; ++i) cur[i] = p[i];
  }
}

}  // namespace lk
}  // namespace aj

namespace {

// Task tables of the backward's inverses: diagonal blocks of every factor, the tiles of every recursion level
// (one list per level, the same tiles for both products of the level), the tiles of every A^-1 = X^T X.
struct BackwardTasks {
  std::vector<int2> diag;
  std::vector<InvTask> level;
  std::vector<int64_t> level_off, level_h;  // level l: tasks [level_off[l], level_off[l + 1]), block size h
  std::vector<SymTask> gram;
};

void backward_tasks(int64_t n, const int64_t* S_host, BackwardTasks& t) {
  t.diag.clear();
  t.level.clear();
  t.level_off.assign(1, 0);
  t.level_h.clear();
  t.gram.clear();
  int64_t maxS = 0;
  for (int64_t b = 0; b < n; ++b) maxS = std::max(maxS, S_host[b]);
  for (int64_t j = 0; j < 2 * n; ++j)
    for (int64_t k = 0; k < ceil_div(S_host[j >> 1], (int64_t)IB); ++k) t.diag.push_back(int2{(int)j, (int)k});
  for (int64_t h = IB; h < maxS; h *= 2) {
    for (int64_t j = 0; j < 2 * n; ++j) {
      const int64_t S = S_host[j >> 1];
      for (int64_t a = 0; a + h < S; a += 2 * h) {
        const int64_t h2 = std::min<int64_t>(h, S - a - h);
        for (int64_t ti = 0; ti < ceil_div(h2, (int64_t)64); ++ti)
          for (int64_t tj = 0; tj < ceil_div(h, (int64_t)64); ++tj)
            t.level.push_back(InvTask{(int)j, (int)a, (int)ti, (int)tj});
      }
    }
    t.level_off.push_back((int64_t)t.level.size());
    t.level_h.push_back(h);
  }
  for (int64_t b = 0; b < n; ++b) {
    const int64_t nt = ceil_div(S_host[b], (int64_t)64);
    for (int64_t ti = 0; ti < nt; ++ti)
      for (int64_t tj = 0; tj <= ti; ++tj) t.gram.push_back(SymTask{(int)b, (int)ti, (int)tj});
  }
}

struct BackwardLayout {
  size_t off_x, off_inv, off_hn2, off_u, off_v, off_diag, off_level, off_gram, bytes;
};

BackwardLayout backward_layout(int64_t M, int64_t n, int64_t sumS, int64_t sumS2, const BackwardTasks& t) {
  BackwardLayout L{};
  size_t o = 0;
  auto put = [&](size_t bytes) {
    o = align_up(o, 256);
    const size_t r = o;
    o += bytes;
    return r;
  };
  L.off_x = put(sizeof(double) * 2 * sumS2);
  L.off_inv = put(sizeof(double) * sumS2);
  L.off_hn2 = put(sizeof(double) * sumS);
  L.off_u = put(sizeof(double) * M * n);
  L.off_v = put(sizeof(double) * M * n);
  L.off_diag = put(sizeof(int2) * t.diag.size());
  L.off_level = put(sizeof(InvTask) * t.level.size());
  L.off_gram = put(sizeof(SymTask) * t.gram.size());
  L.bytes = align_up(o, 256);
  return L;
}

}  // namespace

extern "C" size_t aj_log_evidence_backward_workspace_bytes(int64_t M, int64_t n_samples, const int64_t* S_host) {
  if (M <= 0 || n_samples <= 0 || !S_host) return 0;
  int64_t sumS = 0, sumS2 = 0;
  for (int64_t b = 0; b < n_samples; ++b) {
    if (S_host[b] <= 0) return 0;
    sumS += S_host[b];
    sumS2 += S_host[b] * S_host[b];
  }
  static thread_local BackwardTasks t;
  backward_tasks(n_samples, S_host, t);
  return backward_layout(M, n_samples, sumS, sumS2, t).bytes;
}

extern "C" int aj_log_evidence_batched_backward(const double* w_tilde, int64_t M, const double* dirty_image,
                                                int64_t n_samples, int64_t B, int64_t P, const int64_t* S_host,
                                                const int64_t* pix_indexes, const double* pix_weights,
                                                const int64_t* neighbors, const int64_t* neighbors_sizes,
                                                const double* coefficients_dev, const double* reconstruction,
                                                const void* forward_workspace, size_t forward_workspace_bytes,
                                                double* dcoefficients, double* dweights, void* workspace,
                                                size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && B > 0 && P > 0 && n_samples > 0 && S_host, "bad sizes");
  AJ_REQUIRE(w_tilde && dirty_image && pix_indexes && pix_weights && neighbors && neighbors_sizes &&
                 coefficients_dev && reconstruction && forward_workspace && dcoefficients && dweights,
             "null pointer");
  for (int64_t b = 0; b < n_samples; ++b) AJ_REQUIRE(S_host[b] > 0, "sample %lld: no source pixels", (long long)b);
  BatchPlan p = plan_batch(M, B, n_samples, S_host);
  AJ_REQUIRE(forward_workspace_bytes >= p.bytes, "forward workspace does not match this batch");
  static thread_local BackwardTasks t;
  backward_tasks(n_samples, S_host, t);
  const BackwardLayout L = backward_layout(M, n_samples, p.sumS, p.sumS2, t);
  AJ_REQUIRE(workspace && workspace_bytes >= L.bytes, "workspace too small: need %zu bytes, got %zu", L.bytes,
             workspace_bytes);
  AJ_REQUIRE(t.diag.size() < (1u << 31) && t.level.size() < (1u << 31) && t.gram.size() < (1u << 31),
             "batch too large for the backward's task grids");
  hipStream_t st = as_stream(stream);
  const char* fw = static_cast<const char*>(forward_workspace);
  const SampleDesc* desc = reinterpret_cast<const SampleDesc*>(fw + p.off_desc);
  const double* omega = reinterpret_cast<const double*>(fw + p.off_omega);
  const double* mats = reinterpret_cast<const double*>(fw + p.off_mats);
  char* ws = static_cast<char*>(workspace);
  double* X = reinterpret_cast<double*>(ws + L.off_x);
  double* inv = reinterpret_cast<double*>(ws + L.off_inv);
  double* hn2 = reinterpret_cast<double*>(ws + L.off_hn2);
  double* u = reinterpret_cast<double*>(ws + L.off_u);
  double* v = reinterpret_cast<double*>(ws + L.off_v);
  int2* diag = reinterpret_cast<int2*>(ws + L.off_diag);
  InvTask* level = reinterpret_cast<InvTask*>(ws + L.off_level);
  SymTask* gram = reinterpret_cast<SymTask*>(ws + L.off_gram);
  int64_t maxS = 0;
  for (int64_t b = 0; b < n_samples; ++b) maxS = std::max(maxS, S_host[b]);
  AJ_CHECK_HIP(hipMemcpyAsync(diag, t.diag.data(), sizeof(int2) * t.diag.size(), hipMemcpyHostToDevice, st));
  if (!t.level.empty())
    AJ_CHECK_HIP(hipMemcpyAsync(level, t.level.data(), sizeof(InvTask) * t.level.size(), hipMemcpyHostToDevice, st));
  AJ_CHECK_HIP(hipMemcpyAsync(gram, t.gram.data(), sizeof(SymTask) * t.gram.size(), hipMemcpyHostToDevice, st));
  // X = L^-1 of every factor (the forward's factors are left as they are)
  hipLaunchKernelGGL(copy_doubles_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(2 * p.sumS2, 256), 8192)),
                     dim3(256), 0, st, mats, 2 * p.sumS2, X);
  hipLaunchKernelGGL(trtri_diag_kernel, dim3((unsigned)ceil_div((int64_t)t.diag.size(), (int64_t)4)), dim3(256), 0,
                     st, desc, diag, (int)t.diag.size(), X);
  AJ_LAUNCH_CHECK();
  for (size_t l = 0; l < t.level_h.size(); ++l) {
    const int64_t cnt = t.level_off[l + 1] - t.level_off[l];
    if (cnt == 0) continue;
    hipLaunchKernelGGL(trtri_level_kernel<1>, dim3((unsigned)cnt), dim3(256), 0, st, desc, level + t.level_off[l],
                       t.level_h[l], X);
    hipLaunchKernelGGL(trtri_level_kernel<2>, dim3((unsigned)cnt), dim3(256), 0, st, desc, level + t.level_off[l],
                       t.level_h[l], X);
    AJ_LAUNCH_CHECK();
  }
  // A^-1 = X_A^T X_A (both triangles) and tr(H^-1) by rows of X_H
  hipLaunchKernelGGL(inverse_gram_kernel, dim3((unsigned)t.gram.size()), dim3(256), 0, st, desc, gram, X, inv);
  hipLaunchKernelGGL(inverse_rownorm_kernel, dim3((unsigned)ceil_div(maxS, (int64_t)4), (unsigned)n_samples),
                     dim3(256), 0, st, desc, X, hn2);
  AJ_LAUNCH_CHECK();
  hipLaunchKernelGGL(coef_grad_kernel, dim3((unsigned)n_samples), dim3(256), 0, st, desc, neighbors, neighbors_sizes,
                     P, reconstruction, inv, hn2, coefficients_dev, dcoefficients);
  const unsigned gm = (unsigned)std::min<int64_t>(ceil_div(M, 256), 1024);
  hipLaunchKernelGGL(map_recon_kernel, dim3(gm, (unsigned)n_samples), dim3(256), 0, st, desc, pix_indexes,
                     pix_weights, M, B, reconstruction, u);
  hipLaunchKernelGGL(batched_gemv_kernel, dim3((unsigned)M), dim3(256), 0, st, w_tilde, M, (int)n_samples, u, v);
  hipLaunchKernelGGL(weight_grad_kernel, dim3((unsigned)M, (unsigned)n_samples), dim3(64), 0, st, desc, pix_indexes,
                     M, B, dirty_image, reconstruction, v, omega, inv, dweights);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" int aj_barycentric_weights(const double* vertices, const double* positions, const int64_t* triangles,
                                      int64_t M, double* weights, void* stream) {
  AJ_REQUIRE(M > 0 && vertices && positions && triangles && weights, "bad arguments");
  hipLaunchKernelGGL(barycentric_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(M, 256), 1024)), dim3(256), 0,
                     as_stream(stream), vertices, positions, triangles, M, weights);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" size_t aj_barycentric_vjp_workspace_bytes(int64_t M, int64_t S) {
  return (M > 0 && S > 0) ? sizeof(double) * 6 * M + sizeof(int) * (2 * (S + 1) + 3 * M) + 1024 : 0;
}

extern "C" int aj_barycentric_vjp(const double* vertices, const double* positions, const int64_t* triangles,
                                  int64_t M, int64_t S, const double* dweights, double* dvertices, void* workspace,
                                  size_t workspace_bytes, void* stream) {
  AJ_REQUIRE(M > 0 && S > 0 && vertices && positions && triangles && dweights && dvertices, "bad arguments");
  AJ_REQUIRE(workspace && workspace_bytes >= aj_barycentric_vjp_workspace_bytes(M, S), "workspace too small");
  hipStream_t st = as_stream(stream);
  Arena ar(workspace, workspace_bytes);
  double* contrib = ar.take<double>(6 * M);
  int* ptrs = ar.take<int>(S + 1);
  int* cur = ar.take<int>(S + 1);
  int* ekey = ar.take<int>(3 * M);
  AJ_CHECK_HIP(hipMemsetAsync(ptrs, 0, sizeof(int) * (S + 1), st));
  const unsigned g = (unsigned)std::min<int64_t>(ceil_div(3 * M, 256), 1024);
  hipLaunchKernelGGL(barycentric_vjp_kernel, dim3(g), dim3(256), 0, st, vertices, positions, triangles, dweights, M,
                     contrib);
  hipLaunchKernelGGL(csc_keys_kernel, dim3(g), dim3(256), 0, st, triangles, M, S, ptrs);
  hipLaunchKernelGGL(scan_small_kernel, dim3(1), dim3(64), 0, st, ptrs, S, cur);
  hipLaunchKernelGGL(csc_keys_fill_kernel, dim3(g), dim3(256), 0, st, triangles, M, S, cur, ekey);
  hipLaunchKernelGGL(csc_keys_sort_kernel, dim3((unsigned)ceil_div(S, 256)), dim3(256), 0, st, ptrs, S, ekey);
  hipLaunchKernelGGL(vertex_reduce_kernel, dim3((unsigned)ceil_div(S, 256)), dim3(256), 0, st, ptrs, ekey, contrib, S,
                     dvertices);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" int aj_log_evidence_flow_configure(int small_per_cu, int spin_limit) {
  AJ_REQUIRE(small_per_cu >= -1 && small_per_cu <= (1 << 20), "small_per_cu %d: -1 (default) or 0 .. 2^20",
             small_per_cu);
  AJ_REQUIRE(spin_limit >= -1, "spin_limit %d: -1 (default) or >= 0", spin_limit);
  g_flow_small_per_cu.store(small_per_cu < 0 ? FLOW_SMALL_PER_CU : small_per_cu, std::memory_order_relaxed);
  g_flow_spin_limit.store(spin_limit < 0 ? FLOW_SPIN_LIMIT : spin_limit, std::memory_order_relaxed);
  return AJ_OK;
}

extern "C" int aj_log_evidence_flow_trace(void* device_buffer, int64_t capacity_tasks) {
  AJ_REQUIRE(capacity_tasks >= 0 && (capacity_tasks == 0 || device_buffer != nullptr),
             "trace buffer: a device buffer of capacity_tasks x 16 int64, or null and 0");
  g_flow_trace.store(capacity_tasks > 0 ? static_cast<long long*>(device_buffer) : nullptr, std::memory_order_relaxed);
  g_flow_trace_cap.store(capacity_tasks, std::memory_order_relaxed);
  return AJ_OK;
}

extern "C" int aj_log_evidence_configure(int chol_mode_in) {
  AJ_REQUIRE(chol_mode_in >= -1 && chol_mode_in < CHOL_MODES, "chol_mode %d: -1 (default) or 0 .. %d",
             chol_mode_in, CHOL_MODES - 1);
  g_chol_mode.store(chol_mode_in < 0 ? CHOL_FLOW : chol_mode_in, std::memory_order_relaxed);
  return AJ_OK;
}
