// W-tilde as an exact integer Gram matrix assembled from its residues modulo 12 (or 11) small moduli,
// for gfx950.
//
// Reference: src/autojax/numba.py:155-207 (w_tilde_curvature_interferometer_from)
//   W_ij = sum_k n_k^-2 cos(2 pi [(g_i1 - g_j1) u_k0 + (g_i0 - g_j0) u_k1])
//
// As in the other modes, W = [C S][C S]^T over cos/sin features of the FP64-reduced phases.  Per chunk
// of kc visibilities every feature sqrt(w_k) cos(.) (and sin) is divided by the chunk's largest
// sqrt(w_k) and rounded to an integer q with |q| <= Q: Q = 127 * 256^4 (the five-digit mode's grid)
// with all 12 moduli (chunks of 16384 by default, at most 32768), Q = 60 * 2^30 with the first 11
// (chunks of at most 16384).  The chunk's integer Gram matrix G_ij = sum_f q_if q_jf is then bounded
// by kc (Q^2 + 2Q) < P/2, P = prod_i m_i over the moduli in use, so G is determined by its residues
// G mod m_i (Chinese remainder theorem).  Each residue is one int8 Gram product of the features'
// balanced residues (|r| <= 128) on v_mfma_i32_16x16x64_i8, exact in int32 because
// 2 kc * 128^2 < 2^31.  That is 12 (or 11) int8 products per FP64-equivalent product, where
// splitting q into five base-256 digits needs 22.  The residues are recombined in FP64 with the CRT
// basis split into an exact high part (a multiple of 2^53) and a low part below 2^53, the
// integer-modular emulation of Ozaki, Uchino and Imamura ("Ozaki Scheme II: a GEMM-oriented emulation
// of floating-point matrix multiplication using an integer modular technique", 2025).  The result is
// the Gram matrix of the quantised features to FP64 rounding, so it stays positive semi-definite.
//
// Launches per chunk: scale (max 1/|n_k|, once for all chunks), features (one residue plane per
// modulus), the SYRK (by default persistent: one workgroup per CU working through (256 x 256
// lower-triangle tile, modulus) units and pieces of units, writing int8 residues of G; a fix-up sums the
// pieces' residues), combine (CRT + FP64 accumulation into the lower triangle of W).

#include "aj_common.h"
#include "data_sums.h"

#include <vector>

namespace aj {
namespace crt {

typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int NMOD = 12;           // moduli of the default mode; the 11-moduli mode uses the first 11
constexpr int KC_DEFAULT = 16384;  // visibilities per exact integer Gram matrix (a chunk)
constexpr int KC_MAX = 32768;      // chunk bound: KC_MAX (Q^2 + 2Q) < P/2 and 2 KC_MAX 128^2 < 2^31
constexpr int CT = 256;            // output tile edge (rows of A and of B per workgroup)
constexpr int KB = 64;             // features per 1 KB block (16 rows x 64 bytes) and per LDS stage
// A chunk of kc visibilities has 2 kc features (cos block, then sin block) = nkb = kc / 32 blocks
// per row-block.
constexpr int UNIT_BYTES = CT * CT;  // int8 residues of one tile

// Pairwise coprime moduli, all <= 256 so that balanced residues fit in int8.
constexpr int MODULI[NMOD] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217};

// Largest feature integer: 127 * 256^4 (about 2^39, the five-digit grid) with all 12 moduli
// (P about 2^94.8, chunks up to 58,091 visibilities); 60 * 2^30 (about 2^35.9) with the first 11
// (P about 2^87.0, chunks up to 23,300 visibilities).
constexpr double QMAX12 = 127.0 * 256.0 * 256.0 * 256.0 * 256.0;
constexpr double QMAX11 = 60.0 * 1073741824.0;
__host__ __device__ constexpr double qmax_for(int nm) { return nm == 12 ? QMAX12 : QMAX11; }
// largest chunk (visibilities) whose integer Gram stays inside (-P/2, P/2)
constexpr int kc_max_for(int nm) { return nm == 12 ? 32768 : 16384; }

struct Consts {
  double m[NMOD];      // moduli
  double inv_m[NMOD];  // 1 / m_i (FP64)
  double y_hi[NMOD];   // CRT basis y_i = (P/m_i) * ((P/m_i)^-1 mod m_i) mod P, high part: multiple of 2^53
  double y_lo[NMOD];   // y_i - y_hi, an integer below 2^53 (exact)
  double p_hi, p_lo;   // P split the same way
  double inv_p;        // 1 / P
  int nm;              // moduli in use (the first nm of MODULI)
};

typedef unsigned __int128 u128;

inline int64_t inv_mod(int64_t a, int64_t m) {
  // extended Euclid; a and m coprime
  int64_t t = 0, nt = 1, r = m, nr = a % m;
  while (nr != 0) {
    const int64_t q = r / nr;
    int64_t tmp = t - q * nt;
    t = nt;
    nt = tmp;
    tmp = r - q * nr;
    r = nr;
    nr = tmp;
  }
  return t < 0 ? t + m : t;
}

inline double u128_to_double_exact_hi(u128 v) {  // v is a multiple of 2^53 below 2^106
  return std::ldexp((double)(uint64_t)(v >> 53), 53);
}

// CRT constants of the first nm moduli, computed once in exact 128-bit integer arithmetic (P < 2^95).
inline Consts make_consts(int nm) {
  Consts c{};
  c.nm = nm;
  u128 P = 1;
  for (int i = 0; i < nm; ++i) P *= (u128)MODULI[i];
  const u128 mask53 = (((u128)1) << 53) - 1;
  for (int i = 0; i < nm; ++i) {
    const int64_t m = MODULI[i];
    const u128 Pi = P / (u128)m;
    const int64_t inv = inv_mod((int64_t)(Pi % (u128)m), m);
    const u128 y = (Pi * (u128)inv) % P;
    c.m[i] = (double)m;
    c.inv_m[i] = 1.0 / (double)m;
    c.y_hi[i] = u128_to_double_exact_hi(y & ~mask53);
    c.y_lo[i] = (double)(uint64_t)(y & mask53);
  }
  c.p_hi = u128_to_double_exact_hi(P & ~mask53);
  c.p_lo = (double)(uint64_t)(P & mask53);
  c.inv_p = 1.0 / ((double)(uint64_t)(P >> 64) * 18446744073709551616.0 + (double)(uint64_t)P);
  return c;
}

// ---------------------------------------------------------------------------
// Chunk scales: sc[c] = max_k 1/|n_k| over chunk c's visibilities (1 if the chunk is empty); one
// workgroup per chunk.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) scale_kernel(const double* __restrict__ noise, int64_t K, int kc,
                                                    double* __restrict__ sc) {
  __shared__ double red[256];
  const int64_t k_begin = (int64_t)blockIdx.x * kc;
  double m = 0.0;
  for (int k = threadIdx.x; k < kc; k += 256) {
    const int64_t kg = k_begin + k;
    if (kg < K) m = fmax(m, 1.0 / fabs(noise[kg]));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) sc[blockIdx.x] = red[0] > 0.0 ? red[0] : 1.0;
}

// sin(pi t) and cos(pi t) for |t| <= 1: t = x + n/2 with n = rint(2t), |x| <= 1/4, then the Taylor
// series of sin(pi x) and cos(pi x) to degree 15 and 16 (truncation below 5e-17 at |pi x| <= pi/4,
// under FP64 rounding; absolute error about 1e-16 overall) and the quadrant n mod 4.  Coefficients
// (-1)^k pi^(2k+1)/(2k+1)! and (-1)^k pi^(2k)/(2k)! rounded to FP64.
__device__ inline void sincospi_small(double t, double& sn, double& cs) {
  const double n = rint(2.0 * t);
  const double x = fma(-0.5, n, t);
  const double y = x * x;
  double ps = -2.1915353447830217e-05;
  ps = fma(ps, y, 0.00046630280576761255);
  ps = fma(ps, y, -0.0073704309457143504);
  ps = fma(ps, y, 0.08214588661112823);
  ps = fma(ps, y, -0.5992645293207921);
  ps = fma(ps, y, 2.5501640398773455);
  ps = fma(ps, y, -5.16771278004997);
  ps = fma(ps, y, 3.141592653589793);
  double pc = 4.303069587032947e-06;
  pc = fma(pc, y, -0.0001046381049248457);
  pc = fma(pc, y, 0.0019295743094039231);
  pc = fma(pc, y, -0.02580689139001406);
  pc = fma(pc, y, 0.2353306303588932);
  pc = fma(pc, y, -1.3352627688545895);
  pc = fma(pc, y, 4.0587121264167685);
  pc = fma(pc, y, -4.934802200544679);
  pc = fma(pc, y, 1.0);
  const double sx = x * ps;
  const int q = (int)n & 3;  // quadrant: pi t = pi x + q pi/2 (mod 2 pi)
  const double a = (q & 1) ? pc : sx;   // sin for q = 0, 2 (negated for 2); cos for q = 1, 3
  const double b = (q & 1) ? sx : pc;
  sn = (q & 2) ? -a : a;
  cs = ((q + 1) & 2) ? -b : b;
}

// Residues of an integer-valued double q (|q| < 2^40) as int8 bytes.  qm = q + 1.5 * 2^52 (exact) holds q
// in the low bits of its significand: the low 32 bits of qm's bit pattern are q mod 2^32 in two's
// complement, so its low byte is the balanced residue modulo 256, and for an odd modulus m
// fma(-n, m, qm) with n = rint(q / m) is exactly qm - n m = 1.5 * 2^52 + r (|r| <= (m - 1)/2; q/m is
// never a half-integer and the product's rounding is far below the 1/(2m) gap), whose low byte is r.
constexpr double RES_SHIFT = 6755399441055744.0;  // 1.5 * 2^52

__device__ inline uint32_t low_word(double d) { return (uint32_t)__double_as_longlong(d); }

template <int i>
__device__ inline uint32_t residue_word(double q, double qm) {
  if constexpr (i == 0) return low_word(qm);
  constexpr double m = (double)MODULI[i], inv_m = 1.0 / (double)MODULI[i];
  return low_word(fma(-rint(q * inv_m), m, qm));
}

// Packed residues of (cos, sin) of two visibilities: bytes cos0, sin0, cos1, sin1.
template <int i>
__device__ inline uint32_t packed_residues(const double (&qc)[2], const double (&qn)[2], const double (&qcm)[2],
                                           const double (&qnm)[2]) {
  const uint32_t a = residue_word<i>(qc[0], qcm[0]), b = residue_word<i>(qn[0], qnm[0]);
  const uint32_t c = residue_word<i>(qc[1], qcm[1]), d = residue_word<i>(qn[1], qnm[1]);
  // v_perm_b32 selects bytes of {src0, src1} (src1 = bytes 0-3, src0 = bytes 4-7; selector 12 = zero)
  const uint32_t ab = __builtin_amdgcn_perm(b, a, 0x0c0c0400u);  // a0 b0 0 0
  const uint32_t cd = __builtin_amdgcn_perm(d, c, 0x0c0c0400u);  // c0 d0 0 0
  return __builtin_amdgcn_perm(cd, ab, 0x05040100u);             // a0 b0 c0 d0
}

template <int i, int NM>
__device__ inline void store_residues4(int8_t* __restrict__ X, int64_t plane, uint32_t off, const double (&qc)[2],
                                       const double (&qn)[2], const double (&qcm)[2], const double (&qnm)[2]) {
  *reinterpret_cast<uint32_t*>(X + i * plane + off) = packed_residues<i>(qc, qn, qcm, qnm);
  if constexpr (i + 1 < NM) store_residues4<i + 1, NM>(X, plane, off, qc, qn, qcm, qnm);
}

// ---------------------------------------------------------------------------
// Features: residue planes X[m][rb][kb][16][64] (1 KB blocks: 16 image rows x 64 features); feature
// 2k is the cos and 2k + 1 the sin feature of the chunk's visibility k.  Thread = 2 consecutive
// visibilities (4 bytes per row and plane) x the 16 rows of row-block blockIdx.y.  Padded rows and
// visibilities past K hold zeros.
// ---------------------------------------------------------------------------
// With DS the kernel also accumulates the data-side sums (data_sums.h) of its 512 visibilities and 16
// rows from the FP64 cos/sin values before they are quantised; workgroup (x, y) writes block
// blk0 + x of the row sums.
template <int NM, bool DS>
__global__ void __launch_bounds__(256) features_kernel(const double* __restrict__ noise,
                                                       const double* __restrict__ uv,
                                                       const double* __restrict__ grid,
                                                       const double* __restrict__ centre, int64_t K, int64_t M,
                                                       int64_t Mpad, int64_t k_begin, int kc,
                                                       const double* __restrict__ sc, int8_t* __restrict__ X,
                                                       DataSums dsum, int64_t blk0) {
  const int nkb = kc / 32;
  const int k0 = (blockIdx.x * 256 + threadIdx.x) * 2;  // first of 2 visibilities (chunk-local)
  const int rb = blockIdx.y;
  double u0[2], u1[2], qs[2];
  const double inv_sc = qmax_for(NM) / sc[0];
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    const int64_t kg = k_begin + k0 + v;
    const bool ok = kg < K;
    u0[v] = ok ? uv[2 * kg] : 0.0;
    u1[v] = ok ? uv[2 * kg + 1] : 0.0;
    qs[v] = ok ? inv_sc / fabs(noise[kg]) : 0.0;  // zero scale: features of absent visibilities are 0
  }
  double cf[2][4] = {};  // data-side coefficients of the two visibilities (zero when absent)
  if constexpr (DS) {
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int64_t kg = k_begin + k0 + v;
      if (kg < K) {
#pragma unroll
        for (int j = 0; j < 4; ++j) cf[v][j] = dsum.coef[4 * kg + j];
      }
    }
  }
  RowSum16 rs0, rs1;
  double s0 = 0.0, s1 = 0.0;
  const double cy = centre[0], cx = centre[1];
  const int64_t plane = Mpad * 2 * (int64_t)kc;
  // byte offset within a plane (a plane holds fewer than 2^32 bytes: checked on the host)
  uint32_t off = ((uint32_t)rb * (uint32_t)nkb + (uint32_t)(k0 / 32)) * 1024u + (uint32_t)((2 * k0) % 64);
#pragma unroll 1
  for (int r = 0; r < 16; ++r, off += 64) {
    const int64_t i = (int64_t)rb * 16 + r;
    double qc[2], qn[2];
    double d0 = 0.0, d1 = 0.0;
    if (i < M) {
      const double gx = grid[2 * i + 1] - cx, gy = grid[2 * i] - cy;
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        double rv = fma(gx, u0[v], gy * u1[v]);
        rv -= rint(rv);
        double s, co;
        sincospi_small(2.0 * rv, s, co);
        qc[v] = rint(co * qs[v]);
        qn[v] = rint(s * qs[v]);
        if constexpr (DS) {
          d0 = fma(cf[v][0], co, fma(cf[v][1], s, d0));
          d1 = fma(cf[v][2], co, fma(cf[v][3], s, d1));
        }
      }
    } else {
#pragma unroll
      for (int v = 0; v < 2; ++v) qc[v] = qn[v] = 0.0;
    }
    if constexpr (DS) {
      rs0.push(r, d0, threadIdx.x & 63, s0);
      rs1.push(r, d1, threadIdx.x & 63, s1);
    }
    double qcm[2], qnm[2];
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      qcm[v] = qc[v] + RES_SHIFT;
      qnm[v] = qn[v] + RES_SHIFT;
    }
    store_residues4<0, NM>(X, plane, off, qc, qn, qcm, qnm);
  }
  if constexpr (DS) block_row_sums_out(s0, s1, dsum, blk0 + blockIdx.x, (int64_t)rb * 16);
}

// Residues of a wave's exact int32 Gram block (|acc| < 2^30), balanced, 4 per word.  Layout of a
// unit: R[unit][rb][cb][lane] for the 16 x 16 block (rb, cb) of the 256 x 256 tile, each block's 64
// words in MFMA output order (lane holds rows 4*(lane/16) + 0..3, column lane%16).  The wave's
// block (a, b) is tile block (rb0 + a, cb0 + b).
// Balanced residue of v mod m_mi (modulus 0 is 256: the low byte as int8; the others odd, in
// [-(m - 1)/2, (m - 1)/2]); exact for |v| < 2^30.
__device__ inline int balanced_residue(int v, int mi, double m, double inv_m) {
  if (mi == 0) return (int)(int8_t)(v & 0xFF);
  const double d = (double)v;
  return (int)fma(-rint(d * inv_m), m, d);
}

template <int NA, int NB>
__device__ inline void store_residues_to(const i4 (&acc)[NA][NB], const Consts& c, int mi, int8_t* __restrict__ unit,
                                         int rb0, int cb0, int lane) {
  const double m = c.m[mi], inv_m = c.inv_m[mi];
  uint32_t* out = reinterpret_cast<uint32_t*>(unit) + lane;
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      uint32_t w = 0;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        w |= (uint32_t)(uint8_t)(int8_t)balanced_residue(acc[a][b][r], mi, m, inv_m) << (8 * r);
      out[((rb0 + a) * 16 + cb0 + b) * 64] = w;
    }
}

template <int NA, int NB>
__device__ inline void store_residues(const i4 (&acc)[NA][NB], const Consts& c, int mi, int ntiles, int t, int rb0,
                                      int cb0, int lane, int8_t* __restrict__ R) {
  store_residues_to(acc, c, mi, R + ((int64_t)mi * ntiles + t) * UNIT_BYTES, rb0, cb0, lane);
}

// Direct global->LDS load of 16 bytes per lane: lane L of the wave writes bytes 16 L .. 16 L + 15 of
// the wave-uniform destination block.
__device__ inline void lds_dma16(const int8_t* g, int8_t* lds_block) {
  __builtin_amdgcn_global_load_lds((const void*)g,
                                   (void __attribute__((address_space(3)))*)(
                                       (__attribute__((address_space(3))) int8_t*)lds_block),
                                   16, 0, 0);
}

// ---------------------------------------------------------------------------
// SYRK of one (tile, modulus) unit: the int32 Gram block of A = rows [I*256, +256) and B = rows
// [J*256, +256) over the first nst stages (64 features each; all 2 kc / 64 = nkb of them except in a
// short last chunk) of residue plane m, reduced to balanced int8 residues
// mod m_m.  Eight waves (two per SIMD), each a 128 x 64 block of the tile (8 x 4 MFMA 16x16x64
// blocks, 128 int32 accumulators per lane).  The feature blocks (1 KB: 16 rows x 64 bytes; 32 per
// stage, 16 of A and 16 of B) move by direct global->LDS loads into a ring of four 32 KB stage
// buffers, the XOR swizzle (16-byte unit u of row r stored at u ^ (r/2)%4, free of bank conflicts
// for the fragment reads) applied through each lane's source address.  One barrier per interval of
// two stages, the loads of the next interval issued right after it.
// Units: blockIdx % 8 selects the XCD; XCD x works through its run of `per` tiles (spatially compact
// in the tile list) for modulus 0, then modulus 1, ..., so that all XCDs read the same residue plane
// at a time and the tiles an XCD holds share row panels in its L2.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(512, 1) syrk_ring8_kernel(const int8_t* __restrict__ X, int64_t plane, int nkb, int nst,
                                                            int ntiles, int per, const int* __restrict__ tiles,
                                                            Consts c, int8_t* __restrict__ R) {
  __shared__ __attribute__((aligned(1024))) int8_t lds[4][32][1024];
  const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const int mi = slot / per, loc = slot % per;
  if (mi >= c.nm) return;
  const int t = xcd * per + loc;
  if (t >= ntiles) return;
  const int64_t I = tiles[2 * t], J = tiles[2 * t + 1];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;  // 8 waves
  const int lr = lane & 15, lk = lane >> 4;
  const int srow = lane >> 2;
  const int soff = srow * 64 + 16 * ((lane & 3) ^ ((srow >> 1) & 3));
  // wave w moves blocks w and w + 8 of A (w < 8 covers 16 blocks) and of B
  const int8_t* gA = X + mi * plane + (I * 16 + wave) * (int64_t)nkb * 1024 + soff;
  const int8_t* gB = X + mi * plane + (J * 16 + wave) * (int64_t)nkb * 1024 + soff;
  const int64_t jstride = 8 * (int64_t)nkb * 1024;
  const int roff = lr * 64 + 16 * (lk ^ ((lr >> 1) & 3));
  const int ablk = (wave >> 2) * 8, bblk = 16 + (wave & 3) * 4;

#define AJ_RING8_ISSUE(ST)                                                                   \
  {                                                                                          \
    const int64_t ko_ = (int64_t)(ST) * 1024;                                                \
    int8_t* lb_ = &lds[(ST) & 3][0][0];                                                      \
    _Pragma("unroll") for (int j = 0; j < 2; ++j) {                                          \
      lds_dma16(gA + j * jstride + ko_, lb_ + (wave + 8 * j) * 1024);                        \
      lds_dma16(gB + j * jstride + ko_, lb_ + (16 + wave + 8 * j) * 1024);                   \
    }                                                                                        \
  }

  i4 acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = i4{0, 0, 0, 0};

  AJ_RING8_ISSUE(0)
  AJ_RING8_ISSUE(1)
  const int NI = nst / 2;
  for (int j = 0; j < NI; ++j) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (j + 1 < NI) {
      AJ_RING8_ISSUE(2 * j + 2)
      AJ_RING8_ISSUE(2 * j + 3)
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      __builtin_amdgcn_sched_barrier(0);
      const int8_t* l0 = &lds[(2 * j + h) & 3][0][0];
      i4 fa[8], fb[4];
      fa[0] = *reinterpret_cast<const i4*>(l0 + ablk * 1024 + roff);
#pragma unroll
      for (int b = 0; b < 4; ++b) fb[b] = *reinterpret_cast<const i4*>(l0 + (bblk + b) * 1024 + roff);
#pragma unroll
      for (int a = 1; a < 8; ++a) fa[a] = *reinterpret_cast<const i4*>(l0 + (ablk + a) * 1024 + roff);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
#undef AJ_RING8_ISSUE
  store_residues(acc, c, mi, ntiles, t, (wave >> 2) * 8, (wave & 3) * 4, lane, R);
}

// ---------------------------------------------------------------------------
// Eight waves with the fragment reads software-pipelined across stages: after the barrier of
// stage s every wave issues the loads of stage s + 3 (raw buffer loads straight to LDS, into the
// buffer of stage s - 1) and the fragment reads of stage s + 1, then runs the MFMAs of stage s on
// the fragments it read during stage s - 1.  The barrier of stage s follows each wave's wait for
// its loads of stage s + 1 (the loads of stage s + 2 may still be in flight).
// ---------------------------------------------------------------------------

// Pipeline steps shared by the residue SYRK kernels.  They use the kernel's locals: lds (the ring of
// four 32 KB stage buffers), rsrc (buffer resource of the residue plane), oA / oB (this wave's byte
// offsets of its A and B blocks at stage 0), jstr, wave, ablk, bblk, roff, acc and the SPLIT parameter.
#define AJ_PIPE_ISSUE(ST)                                                                                \
  {                                                                                                      \
    const uint32_t ko_ = (uint32_t)(ST) * 1024u;                                                         \
    int8_t* lb_ = &lds[(ST) & 3][0][0];                                                                  \
    _Pragma("unroll") for (int j = 0; j < 2; ++j) {                                                      \
      raw_buffer_load_lds(rsrc, (lds_u32*)(lb_ + (wave + 8 * j) * 1024), 16, oA + j * jstr + ko_, 0, 0, 0);      \
      raw_buffer_load_lds(rsrc, (lds_u32*)(lb_ + (16 + wave + 8 * j) * 1024), 16, oB + j * jstr + ko_, 0, 0, 0); \
    }                                                                                                    \
  }
#define AJ_PIPE_READ(FA, FB, ST)                                                                         \
  {                                                                                                      \
    const int8_t* l_ = &lds[(ST) & 3][0][0];                                                             \
    FA[0] = *reinterpret_cast<const i4*>(l_ + ablk * 1024 + roff);                                       \
    _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                        \
      FB[b] = *reinterpret_cast<const i4*>(l_ + (bblk + b) * 1024 + roff);                               \
    _Pragma("unroll") for (int a = 1; a < 8; ++a)                                                        \
      FA[a] = *reinterpret_cast<const i4*>(l_ + (ablk + a) * 1024 + roff);                               \
  }
#define AJ_PIPE_MFMA(FA, FB, A0, A1)                                                                     \
  _Pragma("unroll") for (int a = (A0); a < (A1); ++a)                                                    \
  _Pragma("unroll") for (int b = 0; b < 4; ++b)                                                          \
    acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(FA[a], FB[b], acc[a][b], 0, 0, 0);
  // wait until at most N of this wave's buffer loads are in flight, then the workgroup barrier (no
  // release fence: it would wait for every load); the memory clobber keeps LDS accesses in place
#define AJ_PIPE_SYNC(N)                                                                                  \
  {                                                                                                      \
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(N));                                                            \
    asm volatile("s_barrier" ::: "memory");                                                              \
  }
// Branch-free, so that the waits before the MFMAs count only the reads issued one step earlier.  Past
// the chunk's last stage the loads fetch the next row-block's features, or past the end of the plane
// (the buffer range) zeros, into buffers that are no longer read, and the reads of stage nst are
// never used.
#define AJ_PIPE_STEP(FA, FB, GA, GB, ST)                                                                 \
  {                                                                                                      \
    const int s_ = (ST);                                                                                 \
    if constexpr (SPLIT) {                                                                               \
      AJ_PIPE_MFMA(FA, FB, 0, 4)                                                                         \
      __builtin_amdgcn_sched_barrier(0);                                                                 \
    }                                                                                                    \
    AJ_PIPE_SYNC(4)                                                                                      \
    AJ_PIPE_ISSUE(s_ + 3)                                                                                \
    AJ_PIPE_READ(GA, GB, s_ + 1)                                                                         \
    __builtin_amdgcn_sched_barrier(0);                                                                   \
    AJ_PIPE_MFMA(FA, FB, SPLIT ? 4 : 0, 8)                                                               \
    __builtin_amdgcn_sched_barrier(0);                                                                   \
  }


// SPLIT = 1 runs the first half of a stage's MFMAs before the barrier (their fragments are already in
// registers), so a SIMD has MFMA work while its waves wait at the barrier and issue the next loads.
template <int SPLIT>
__global__ void __launch_bounds__(512, 1) syrk_pipe8_kernel(const int8_t* __restrict__ X, int64_t plane, int nkb, int nst,
                                                            int ntiles, int per, const int* __restrict__ tiles,
                                                            Consts c, int8_t* __restrict__ R) {
  __shared__ __attribute__((aligned(1024))) int8_t lds[4][32][1024];
  const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const int mi = slot / per, loc = slot % per;
  if (mi >= c.nm) return;
  const int t = xcd * per + loc;
  if (t >= ntiles) return;
  const int I = tiles[2 * t], J = tiles[2 * t + 1];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int srow = lane >> 2;
  const int soff = srow * 64 + 16 * ((lane & 3) ^ ((srow >> 1) & 3));
  const i32x4 rsrc = make_rsrc(X + mi * plane, (uint32_t)plane);
  const uint32_t rbytes = (uint32_t)nkb * 1024u;  // bytes per row-block
  const uint32_t oA = (uint32_t)(I * 16 + wave) * rbytes + soff;
  const uint32_t oB = (uint32_t)(J * 16 + wave) * rbytes + soff;
  const uint32_t jstr = 8u * rbytes;
  const int roff = lr * 64 + 16 * (lk ^ ((lr >> 1) & 3));
  const int ablk = (wave >> 2) * 8, bblk = 16 + (wave & 3) * 4;

  i4 acc[8][4];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = i4{0, 0, 0, 0};

  // nst even (>= 2); nkb (the row-block stride in blocks) >= nst
  i4 fa0[8], fb0[4], fa1[8], fb1[4];
  AJ_PIPE_ISSUE(0)
  AJ_PIPE_ISSUE(1)
  AJ_PIPE_ISSUE(2)
  // stage 0 landed everywhere; waiting for every counter here also leaves nothing but this wave's
  // LDS reads pending at the loop head, so the waits before the MFMAs can count them
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  AJ_PIPE_READ(fa0, fb0, 0)
  for (int st = 0; st < nst; st += 2) {
    AJ_PIPE_STEP(fa0, fb0, fa1, fb1, st)
    AJ_PIPE_STEP(fa1, fb1, fa0, fb0, st + 1)
  }
  // the loads issued past the last stage must land before the workgroup's LDS is released
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
  store_residues(acc, c, mi, ntiles, t, (wave >> 2) * 8, (wave & 3) * 4, lane, R);
}

// ---------------------------------------------------------------------------
// Persistent form of syrk_pipe8_kernel<SPLIT>: one workgroup per CU (blockIdx % 8 selects the XCD),
// each running a host-built list of segments beg[blockIdx.x] .. beg[blockIdx.x + 1] - 1.  A segment
// is a run of stages [k0, k0 + ns) of one (tile, modulus) unit; its residues go to the unit's place
// in R, or, for a piece of a unit split between workgroups, to a partial slot of Rp that
// fixup_kernel sums.  The host deals every CU the same number of stages (whole units in rounds of
// consecutive units per XCD, as a dispatch of one workgroup per unit would, then the remainder split
// into pieces), so that no CU idles through a last, partial round.
// ---------------------------------------------------------------------------
struct __attribute__((aligned(16))) SkSeg {
  int t, mi, k0, ns;  // tile, modulus, first stage, stages (even)
  int slot;           // -1: whole unit, residues to R; else partial slot in Rp
  int pad[3];
};

struct __attribute__((aligned(16))) SkFix {
  int t, mi, slot0, n;  // unit (tile, modulus) whose residues are the sum of Rp slots slot0 .. slot0 + n - 1
};

template <int SPLIT>
__global__ void __launch_bounds__(512, 1) syrk_sk_kernel(const int8_t* __restrict__ X, int64_t plane, int nkb,
                                                         const int* __restrict__ tiles, const SkSeg* __restrict__ segs,
                                                         const int* __restrict__ beg, Consts c, int ntiles,
                                                         int8_t* __restrict__ R, int8_t* __restrict__ Rp) {
  __shared__ __attribute__((aligned(1024))) int8_t lds[4][32][1024];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int srow = lane >> 2;
  const int soff = srow * 64 + 16 * ((lane & 3) ^ ((srow >> 1) & 3));
  const uint32_t rbytes = (uint32_t)nkb * 1024u;  // bytes per row-block
  const uint32_t jstr = 8u * rbytes;
  const int roff = lr * 64 + 16 * (lk ^ ((lr >> 1) & 3));
  const int ablk = (wave >> 2) * 8, bblk = 16 + (wave & 3) * 4;
  const int s1 = beg[blockIdx.x + 1];
  for (int s = beg[blockIdx.x]; s < s1; ++s) {
    const SkSeg g = segs[s];
    const int I = tiles[2 * g.t], J = tiles[2 * g.t + 1];
    const i32x4 rsrc = make_rsrc(X + g.mi * plane, (uint32_t)plane);
    const uint32_t oA = (uint32_t)(I * 16 + wave) * rbytes + (uint32_t)g.k0 * 1024u + soff;
    const uint32_t oB = (uint32_t)(J * 16 + wave) * rbytes + (uint32_t)g.k0 * 1024u + soff;

    i4 acc[8][4];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = i4{0, 0, 0, 0};

    i4 fa0[8], fb0[4], fa1[8], fb1[4];
    AJ_PIPE_ISSUE(0)
    AJ_PIPE_ISSUE(1)
    AJ_PIPE_ISSUE(2)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    AJ_PIPE_READ(fa0, fb0, 0)
    for (int st = 0; st < g.ns; st += 2) {
      AJ_PIPE_STEP(fa0, fb0, fa1, fb1, st)
      AJ_PIPE_STEP(fa1, fb1, fa0, fb0, st + 1)
    }
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
    int8_t* out = g.slot < 0 ? R + ((int64_t)g.mi * ntiles + g.t) * UNIT_BYTES : Rp + (int64_t)g.slot * UNIT_BYTES;
    store_residues_to(acc, c, g.mi, out, (wave >> 2) * 8, (wave & 3) * 4, lane);
    // every wave is past its reads of this segment's stages before the next segment's loads
    __syncthreads();
  }
}
#undef AJ_PIPE_ISSUE
#undef AJ_PIPE_READ
#undef AJ_PIPE_MFMA
#undef AJ_PIPE_SYNC
#undef AJ_PIPE_STEP

// Residues of each split unit: the balanced residue of the sum of its pieces' balanced residues (the
// same value as the residue of the unit's whole int32 Gram block).  Grid (64, split units); a thread
// is one packed word (4 residues) of the unit.
__global__ void __launch_bounds__(256) fixup_kernel(const SkFix* __restrict__ fix, const int8_t* __restrict__ Rp,
                                                    Consts c, int ntiles, int8_t* __restrict__ R) {
  const SkFix f = fix[blockIdx.y];
  const int wd = blockIdx.x * 256 + threadIdx.x;
  int v[4] = {0, 0, 0, 0};
  for (int p = 0; p < f.n; ++p) {
    const uint32_t w = reinterpret_cast<const uint32_t*>(Rp + (int64_t)(f.slot0 + p) * UNIT_BYTES)[wd];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] += (int)(int8_t)(w >> (8 * r));
  }
  const double m = c.m[f.mi], inv_m = c.inv_m[f.mi];
  uint32_t w = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) w |= (uint32_t)(uint8_t)(int8_t)balanced_residue(v[r], f.mi, m, inv_m) << (8 * r);
  reinterpret_cast<uint32_t*>(R + ((int64_t)f.mi * ntiles + f.t) * UNIT_BYTES)[wd] = w;
}

// ---------------------------------------------------------------------------
// Combine: per element of the lower triangle, G from its NM residues, W += G * scale.
//   S_hi = sum_i r_i y_hi_i  (exact: every term and partial sum is a multiple of 2^53 below 2^106)
//   S_lo = sum_i r_i y_lo_i  (|.| < 2^64, rounding error below 2^16)
//   n    = rint((S_hi + S_lo) / P)
//   G    = (S_hi - n p_hi) + (S_lo - n p_lo)   (first bracket exact)
// One launch recombines the ng chunks whose residues sit in sets 0 .. ng - 1 (set g at R + g set_bytes, scale
// sc[g]), adding them to W in chunk order with one FMA each, as ng launches of one chunk would: W is read and
// written once per group instead of once per chunk.
// Thread = one packed word of one unit: 4 elements (one column, 4 rows) of a 16 x 16 block.
// ---------------------------------------------------------------------------
constexpr int COMBINE_GROUP = 8;  // chunks per combine launch
template <int NM>
__global__ void __launch_bounds__(256) combine_kernel(const int8_t* __restrict__ R, int64_t set_bytes, int ng,
                                                      int ntiles, const int* __restrict__ tiles, Consts c,
                                                      const double* __restrict__ sc, double* __restrict__ W,
                                                      int64_t M) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int t = (int)(g >> 14);
  if (t >= ntiles) return;
  const int rem = (int)(g & 16383);
  const int rb = rem >> 10, cb = (rem >> 6) & 15, lane = rem & 63;
  const int64_t i0 = (int64_t)tiles[2 * t] * CT + rb * 16 + 4 * (lane >> 4);
  const int64_t j = (int64_t)tiles[2 * t + 1] * CT + cb * 16 + (lane & 15);
  if (j >= M || i0 >= M || i0 + 3 < j) return;
  double wv[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t i = i0 + r;
    wv[r] = (i < M && i >= j) ? W[i * M + j] : 0.0;
  }
  constexpr double Q = qmax_for(NM);
  for (int s = 0; s < ng; ++s) {
    const int8_t* Rs = R + s * set_bytes;
    uint32_t w[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m)
      w[m] = reinterpret_cast<const uint32_t*>(Rs + ((int64_t)m * ntiles + t) * UNIT_BYTES)[rem];
    const double s2 = sc[s] * sc[s] / (Q * Q);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double shi = 0.0, slo = 0.0;
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const double rv = (double)(int8_t)(w[m] >> (8 * r));
        shi = fma(rv, c.y_hi[m], shi);
        slo = fma(rv, c.y_lo[m], slo);
      }
      const double n = rint((shi + slo) * c.inv_p);
      const double G = fma(-n, c.p_hi, shi) + fma(-n, c.p_lo, slo);
      wv[r] = fma(G, s2, wv[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t i = i0 + r;
    if (i < M && i >= j) W[i * M + j] = wv[r];
  }
}

// Lower-triangle 256-tiles in 4 x 4 super-tiles, so that a run of consecutive tiles is compact.
inline void build_tiles(int64_t nI, std::vector<int>& out) {
  out.clear();
  const int64_t nS = (nI + 3) / 4;
  for (int64_t SI = 0; SI < nS; ++SI)
    for (int64_t SJ = 0; SJ <= SI; ++SJ)
      for (int64_t i = 0; i < 4; ++i)
        for (int64_t j = 0; j < 4; ++j) {
          const int64_t I = 4 * SI + i, J = 4 * SJ + j;
          if (I < nI && J <= I) {
            out.push_back((int)I);
            out.push_back((int)J);
          }
        }
}

// Schedule of syrk_sk_kernel: the ntiles * nm units of `pairs` stage pairs each dealt to G workgroups
// (G a multiple of 8; workgroup w runs on XCD w % 8).  XCD x owns the tiles [x ntiles / 8,
// (x + 1) ntiles / 8) and their units in modulus-major order; its C = G / 8 workgroups take them in
// rounds of C consecutive units (so the units in flight on an XCD share row panels in its L2, and all
// XCDs read the same residue plane at a time, which the memory-side cache holds for all of them).  The
// units left over after the last full round of every XCD are then dealt, as runs of stage pairs in
// XCD order, raising the workgroups below a common level to it (to one pair): every workgroup ends at
// the same stage count when the XCDs run the same number of full rounds; an XCD with one more full
// round than the level ends up to one unit later.  A unit covered by one run is written whole, a unit
// cut between runs is written as partial residues to consecutive slots and summed by fixup_kernel.
constexpr int SK_WG_MAX = 256;  // workgroups of the persistent SYRK (one per CU on MI355X)

struct SkSched {
  std::vector<SkSeg> segs;
  std::vector<int> beg;  // G + 1 offsets into segs
  std::vector<SkFix> fix;
  int nslots = 0;
};

inline void build_sk(int ntiles, int nm, int pairs, int G, SkSched& S) {
  const int C = G / 8;
  std::vector<std::vector<SkSeg>> per(G);
  std::vector<int64_t> base(G, 0);
  struct Unit {
    int t, mi;
  };
  std::vector<Unit> left;
  for (int x = 0; x < 8; ++x) {
    const int t0 = (int)((int64_t)x * ntiles / 8), nx = (int)((int64_t)(x + 1) * ntiles / 8) - t0;
    const int U = nm * nx, F = U / C;
    for (int j = 0; j < C; ++j) {
      const int w = j * 8 + x;
      for (int r = 0; r < F; ++r) {
        const int u = r * C + j;
        per[w].push_back(SkSeg{t0 + u % nx, u / nx, 0, 2 * pairs, -1, {0, 0, 0}});
      }
      base[w] = (int64_t)F * pairs;
    }
    for (int u = F * C; u < U; ++u) left.push_back(Unit{t0 + u % nx, u / nx});
  }
  // leftover pairs to each workgroup: fill every workgroup up to the common level H - 1, then one more
  // pair to as many workgroups below H as needed
  const int64_t LP = (int64_t)left.size() * pairs;
  auto fill = [&](int64_t H) {
    int64_t f = 0;
    for (int w = 0; w < G; ++w) f += std::max<int64_t>(0, H - base[w]);
    return f;
  };
  int64_t lo = 0, hi = LP + *std::max_element(base.begin(), base.end());
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (fill(mid) >= LP)
      hi = mid;
    else
      lo = mid + 1;
  }
  std::vector<int64_t> d(G);
  int64_t rem = LP;
  for (int w = 0; w < G; ++w) {
    d[w] = std::max<int64_t>(0, lo - 1 - base[w]);
    rem -= d[w];
  }
  for (int x = 0; x < 8; ++x)
    for (int j = 0; j < C && rem > 0; ++j)
      if (base[j * 8 + x] < lo) {
        ++d[j * 8 + x];
        --rem;
      }
  struct Piece {
    int w;
    int64_t u;
    int p0, np;
  };
  std::vector<Piece> pcs;
  int64_t pos = 0;
  for (int x = 0; x < 8; ++x)
    for (int j = 0; j < C; ++j) {
      const int w = j * 8 + x;
      for (int64_t a = pos, b = pos + d[w]; a < b;) {
        const int64_t u = a / pairs;
        const int p0 = (int)(a % pairs), np = (int)std::min<int64_t>(pairs - p0, b - a);
        pcs.push_back(Piece{w, u, p0, np});
        a += np;
      }
      pos += d[w];
    }
  S.fix.clear();
  S.nslots = 0;
  for (size_t i = 0; i < pcs.size();) {
    size_t k = i;
    while (k < pcs.size() && pcs[k].u == pcs[i].u) ++k;
    const Unit lu = left[pcs[i].u];
    if (k - i == 1) {
      per[pcs[i].w].push_back(SkSeg{lu.t, lu.mi, 0, 2 * pairs, -1, {0, 0, 0}});
    } else {
      S.fix.push_back(SkFix{lu.t, lu.mi, S.nslots, (int)(k - i)});
      for (size_t q = i; q < k; ++q)
        per[pcs[q].w].push_back(SkSeg{lu.t, lu.mi, 2 * pcs[q].p0, 2 * pcs[q].np, S.nslots++, {0, 0, 0}});
    }
    i = k;
  }
  S.segs.clear();
  S.beg.assign(G + 1, 0);
  for (int w = 0; w < G; ++w) {
    S.beg[w] = (int)S.segs.size();
    S.segs.insert(S.segs.end(), per[w].begin(), per[w].end());
  }
  S.beg[G] = (int)S.segs.size();
}

// Capacities of one schedule's tables for at most SK_WG_MAX workgroups: every unit once plus at most
// one piece per workgroup; leftover units are fewer than 8 per round slot (at most 8 (C - 1)).
inline int64_t sk_seg_cap(int64_t units) { return units + SK_WG_MAX; }
inline int64_t sk_slot_cap(int64_t units) { return std::min<int64_t>(units, 8 * (SK_WG_MAX / 8 - 1)) + SK_WG_MAX; }

struct Plan {
  int64_t Mpad, nI, ntiles, per;
  size_t feat_bytes, set_bytes, res_bytes, bytes;
  int nset;  // residue sets: the chunks recombined by one combine launch (at most COMBINE_GROUP)
  int64_t seg_cap, slot_cap;  // persistent SYRK schedule tables (two: full chunks, last chunk)
};

inline Plan make_plan(int64_t M, int64_t K, int kc, int nm) {
  Plan p{};
  p.Mpad = ceil_div(M, CT) * CT;
  p.nI = p.Mpad / CT;
  p.ntiles = p.nI * (p.nI + 1) / 2;
  p.per = ceil_div(p.ntiles, 8);
  p.feat_bytes = (size_t)nm * p.Mpad * 2 * kc;  // one chunk
  const int64_t nch = std::max<int64_t>(1, ceil_div(K, kc));
  p.set_bytes = (size_t)nm * p.ntiles * UNIT_BYTES;
  p.nset = (int)std::min<int64_t>(COMBINE_GROUP, nch);
  p.res_bytes = p.set_bytes * p.nset;
  const int64_t units = (int64_t)nm * p.ntiles;
  p.seg_cap = sk_seg_cap(units);
  p.slot_cap = sk_slot_cap(units);
  const size_t sched_bytes = align_up(sizeof(SkSeg) * p.seg_cap, 256) + align_up(sizeof(int) * (SK_WG_MAX + 1), 256) +
                             align_up(sizeof(SkFix) * units, 256);
  p.bytes = align_up(p.feat_bytes, 256) + align_up(p.res_bytes, 256) +
            align_up(sizeof(int) * 2 * p.ntiles, 256) + align_up(sizeof(double) * nch, 256) + 2 * sched_bytes +
            align_up((size_t)p.slot_cap * UNIT_BYTES, 256) + 256;
  return p;
}

}  // namespace crt

namespace wt {

// Residue SYRK kernel (0 = syrk_pipe8_kernel<1>; 1 = syrk_pipe8_kernel<0>; 2 = syrk_ring8_kernel;
// 3 = syrk_sk_kernel<1> + fixup_kernel, the default) and visibilities per chunk; process-wide tuning,
// atomic so that a configure call may race a build on another thread (each build reads them once).
std::atomic<int> g_crt_variant{3};
std::atomic<int> g_crt_kc{crt::KC_DEFAULT};

// Visibilities per chunk for nm moduli: the configured chunk, capped by the modulus product's range.
inline int crt_kc(int nm) { return std::min(g_crt_kc.load(std::memory_order_relaxed), crt::kc_max_for(nm)); }

size_t crt_workspace_bytes(int64_t M, int64_t K, int nm) { return crt::make_plan(M, K, crt_kc(nm), nm).bytes; }

// Accumulate the lower triangle of W (zeroed by the caller) over all visibilities.
int crt_accumulate(const double* noise, const double* uv, const double* grid, const double* centre, int64_t K,
                   int64_t M, int nm, double* W, void* workspace, size_t workspace_bytes, hipStream_t st,
                   const DataSums* ds) {
  using namespace crt;
  if (nm != 11 && nm != 12) {
    set_error("int8crt: %d moduli (11 or 12 supported)", nm);
    return AJ_EINVAL;
  }
  // the tuning read once: the schedule built below and the kernel launched must agree
  const int variant = g_crt_variant.load(std::memory_order_relaxed);
  const int kc = crt_kc(nm), nkb = kc / 32;
  const Plan p = make_plan(M, K, kc, nm);
  const int64_t nch = std::max<int64_t>(1, ceil_div(K, kc));
  Arena ar(workspace, workspace_bytes);
  int8_t* X = ar.take<int8_t>(p.feat_bytes);
  int8_t* R = ar.take<int8_t>(p.res_bytes);
  int* tiles = ar.take<int>(2 * p.ntiles);
  double* sc = ar.take<double>(nch);
  const int64_t units = (int64_t)nm * p.ntiles;
  SkSeg* dsegs[2];
  int* dbeg[2];
  SkFix* dfix[2];
  for (int i = 0; i < 2; ++i) {
    dsegs[i] = ar.take<SkSeg>(p.seg_cap);
    dbeg[i] = ar.take<int>(SK_WG_MAX + 1);
    dfix[i] = ar.take<SkFix>(units);
  }
  int8_t* Rp = ar.take<int8_t>((size_t)p.slot_cap * UNIT_BYTES);
  if (!ar.ok()) {
    set_error("workspace too small: need %zu bytes, got %zu", p.bytes, workspace_bytes);
    return AJ_ESPACE;
  }
  if ((uint64_t)p.Mpad * 2 * (uint64_t)kc >= (uint64_t(1) << 32)) {
    set_error("int8crt: %lld image pixels x %d visibilities per chunk exceed the 32-bit plane offsets; use "
              "a smaller chunk (aj_w_tilde_crt_configure)", (long long)M, kc);
    return AJ_EINVAL;
  }
  if (K == 0) return AJ_OK;
  static const Consts cst12 = make_consts(12), cst11 = make_consts(11);
  const Consts& cst = nm == 12 ? cst12 : cst11;
  static thread_local std::vector<int> tl;
  build_tiles(p.nI, tl);
  AJ_CHECK_HIP(hipMemcpyAsync(tiles, tl.data(), sizeof(int) * tl.size(), hipMemcpyHostToDevice, st));
  // persistent SYRK: schedules for a full chunk (i = 0) and for the last chunk (i = 1)
  int G = 0;
  SkSched* sk[2] = {nullptr, nullptr};
  static thread_local SkSched skt[2];
  if (variant == 3) {
    int dev = 0, cus = 0;
    AJ_CHECK_HIP(hipGetDevice(&dev));
    AJ_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    G = std::max(8, std::min(SK_WG_MAX, cus / 8 * 8));
    for (int i = 0; i < 2; ++i) {
      const int64_t kv = i == 0 ? std::min<int64_t>(kc, K) : K - (nch - 1) * kc;
      const int nst = std::min<int>(nkb, (int)(2 * ceil_div(kv, 64)));
      build_sk((int)p.ntiles, nm, nst / 2, G, skt[i]);
      if ((int64_t)skt[i].segs.size() > p.seg_cap || (int64_t)skt[i].nslots > p.slot_cap) {
        set_error("int8crt: SYRK schedule exceeds its tables (%zu segments, %d slots)", skt[i].segs.size(),
                  skt[i].nslots);
        return AJ_EINVAL;
      }
      AJ_CHECK_HIP(hipMemcpyAsync(dsegs[i], skt[i].segs.data(), sizeof(SkSeg) * skt[i].segs.size(),
                                  hipMemcpyHostToDevice, st));
      AJ_CHECK_HIP(hipMemcpyAsync(dbeg[i], skt[i].beg.data(), sizeof(int) * skt[i].beg.size(), hipMemcpyHostToDevice,
                                  st));
      if (!skt[i].fix.empty())
        AJ_CHECK_HIP(hipMemcpyAsync(dfix[i], skt[i].fix.data(), sizeof(SkFix) * skt[i].fix.size(),
                                    hipMemcpyHostToDevice, st));
      sk[i] = &skt[i];
    }
  }
  hipLaunchKernelGGL(scale_kernel, dim3((unsigned)nch), dim3(256), 0, st, noise, K, kc, sc);
  AJ_LAUNCH_CHECK();
  const int64_t plane = p.Mpad * 2 * (int64_t)kc;
  const dim3 fg((unsigned)(kc / 512), (unsigned)(p.Mpad / 16));
  const dim3 sg((unsigned)(8 * p.per * nm));
  const dim3 cg((unsigned)ceil_div(p.ntiles * 16384, 256));
  const int nt = (int)p.ntiles, per = (int)p.per;
  for (int64_t ch = 0; ch < nch; ++ch) {
    // a short last chunk runs only the stages that hold its visibilities (an even count; the features
    // kernel zero-fills every visibility of those stages past K)
    const int64_t kv = std::min<int64_t>(kc, K - ch * kc);
    const int nst = std::min<int>(nkb, (int)(2 * ceil_div(kv, 64)));
    const dim3 fgc((unsigned)std::min<int64_t>(fg.x, ceil_div((int64_t)nst * 32, 512)), fg.y);
    timer_begin(AJ_TIMER_WT_FEATURES, st);
    // (with data-side sums: workgroup x of chunk ch covers visibilities ch kc + 512 x ..; kc is a multiple
    // of 512, so its block index is ch kc / 512 + x)
    const int64_t blk0 = ch * (kc / 512);
    if (nm == 12) {
      if (ds)
        hipLaunchKernelGGL((features_kernel<12, true>), fgc, dim3(256), 0, st, noise, uv, grid, centre, K, M, p.Mpad,
                           ch * kc, kc, sc + ch, X, *ds, blk0);
      else
        hipLaunchKernelGGL((features_kernel<12, false>), fgc, dim3(256), 0, st, noise, uv, grid, centre, K, M, p.Mpad,
                           ch * kc, kc, sc + ch, X, DataSums{}, blk0);
    } else {
      if (ds)
        hipLaunchKernelGGL((features_kernel<11, true>), fgc, dim3(256), 0, st, noise, uv, grid, centre, K, M, p.Mpad,
                           ch * kc, kc, sc + ch, X, *ds, blk0);
      else
        hipLaunchKernelGGL((features_kernel<11, false>), fgc, dim3(256), 0, st, noise, uv, grid, centre, K, M, p.Mpad,
                           ch * kc, kc, sc + ch, X, DataSums{}, blk0);
    }
    timer_end(AJ_TIMER_WT_FEATURES, st);
    AJ_LAUNCH_CHECK();
    timer_begin(AJ_TIMER_WT_SYRK, st);
    const int set = (int)(ch % p.nset);
    int8_t* Rc = R + set * p.set_bytes;  // this chunk's residues
    switch (variant) {
      case 1: hipLaunchKernelGGL(syrk_pipe8_kernel<0>, sg, dim3(512), 0, st, X, plane, nkb, nst, nt, per, tiles, cst, Rc); break;
      case 2: hipLaunchKernelGGL(syrk_ring8_kernel, sg, dim3(512), 0, st, X, plane, nkb, nst, nt, per, tiles, cst, Rc); break;
      case 3: {
        const int i = ch + 1 < nch ? 0 : 1;
        hipLaunchKernelGGL(syrk_sk_kernel<1>, dim3((unsigned)G), dim3(512), 0, st, X, plane, nkb, tiles, dsegs[i], dbeg[i],
                           cst, nt, Rc, Rp);
        if (!sk[i]->fix.empty()) {
          AJ_LAUNCH_CHECK();
          hipLaunchKernelGGL(fixup_kernel, dim3(UNIT_BYTES / 4 / 256, (unsigned)sk[i]->fix.size()), dim3(256), 0, st,
                             dfix[i], Rp, cst, nt, Rc);
        }
        break;
      }
      default: hipLaunchKernelGGL(syrk_pipe8_kernel<1>, sg, dim3(512), 0, st, X, plane, nkb, nst, nt, per, tiles, cst, Rc); break;
    }
    timer_end(AJ_TIMER_WT_SYRK, st);
    AJ_LAUNCH_CHECK();
    // the group's chunks into W once its last chunk (or the W-tilde's last) has its residues
    if (set + 1 == p.nset || ch + 1 == nch) {
      const int ng = set + 1;
      timer_begin(AJ_TIMER_WT_COMBINE, st);
      if (nm == 12)
        hipLaunchKernelGGL(combine_kernel<12>, cg, dim3(256), 0, st, R, (int64_t)p.set_bytes, ng, nt, tiles, cst,
                           sc + ch + 1 - ng, W, M);
      else
        hipLaunchKernelGGL(combine_kernel<11>, cg, dim3(256), 0, st, R, (int64_t)p.set_bytes, ng, nt, tiles, cst,
                           sc + ch + 1 - ng, W, M);
      timer_end(AJ_TIMER_WT_COMBINE, st);
      AJ_LAUNCH_CHECK();
    }
  }
  return AJ_OK;
}

}  // namespace wt
}  // namespace aj

// Residue SYRK kernel variant (0..3, -1 = default 3) and visibilities per chunk (a multiple of 4096 up to
// 32768, 0 = default 16384); tuning, no reference counterpart.
extern "C" int aj_w_tilde_crt_configure(int variant, int64_t chunk_visibilities) {
  if (variant < -1 || variant > 3) {
    aj::set_error("residue SYRK variant %d outside [-1, 3]", variant);
    return AJ_EINVAL;
  }
  if (chunk_visibilities != 0 &&
      (chunk_visibilities < 4096 || chunk_visibilities > aj::crt::KC_MAX || chunk_visibilities % 4096 != 0)) {
    aj::set_error("chunk of %lld visibilities: need a multiple of 4096 in [4096, %d]", (long long)chunk_visibilities,
                  aj::crt::KC_MAX);
    return AJ_EINVAL;
  }
  aj::wt::g_crt_variant.store(variant < 0 ? 3 : variant, std::memory_order_relaxed);
  aj::wt::g_crt_kc.store(chunk_visibilities ? (int)chunk_visibilities : aj::crt::KC_DEFAULT, std::memory_order_relaxed);
  return AJ_OK;
}

// CRT constants as the kernels use them for nmod = 12 or 11 moduli (host query, no reference
// counterpart): fills moduli[nmod], y_hi[nmod], y_lo[nmod] and pq[4] = {p_hi, p_lo, inv_p, Q} with Q
// the largest feature integer of that mode.
extern "C" int aj_crt_constants(int nmod, double* moduli, double* y_hi, double* y_lo, double* pq) {
  if (!moduli || !y_hi || !y_lo || !pq || (nmod != 11 && nmod != 12)) {
    aj::set_error("aj_crt_constants: null pointer or nmod %d not 11 or 12", nmod);
    return AJ_EINVAL;
  }
  const aj::crt::Consts c = aj::crt::make_consts(nmod);
  for (int i = 0; i < nmod; ++i) {
    moduli[i] = c.m[i];
    y_hi[i] = c.y_hi[i];
    y_lo[i] = c.y_lo[i];
  }
  pq[0] = c.p_hi;
  pq[1] = c.p_lo;
  pq[2] = c.inv_p;
  pq[3] = aj::crt::qmax_for(nmod);
  return AJ_OK;
}

// The persistent residue SYRK's schedule for ntiles tiles, nm moduli, `pairs` stage pairs per unit and
// G workgroups, as crt_accumulate builds and uploads it (host-only query for checks).  segs: 8 ints per
// segment {tile, modulus, first stage, stages, slot (-1 whole), 0, 0, 0}; beg: G + 1 offsets; fix: 4 ints
// per split unit {tile, modulus, first slot, slots}; counts = {segments, split units, slots}.  Fails when
// a table exceeds its capacity (seg_cap segments, fix_cap split units) or the kernel's table bounds.
extern "C" int aj_crt_sk_schedule(int ntiles, int nm, int pairs, int G, int* segs, int64_t seg_cap, int* beg,
                                  int* fix, int64_t fix_cap, int64_t* counts) {
  using namespace aj::crt;
  if (!segs || !beg || !fix || !counts || ntiles < 1 || ntiles > (1 << 20) || (nm != 11 && nm != 12) || pairs < 1 ||
      pairs > (1 << 16) || G < 8 || G > SK_WG_MAX || G % 8 != 0) {
    aj::set_error("aj_crt_sk_schedule: bad arguments ntiles=%d nm=%d pairs=%d G=%d", ntiles, nm, pairs, G);
    return AJ_EINVAL;
  }
  SkSched S;
  build_sk(ntiles, nm, pairs, G, S);
  const int64_t units = (int64_t)nm * ntiles;
  if ((int64_t)S.segs.size() > seg_cap || (int64_t)S.fix.size() > fix_cap ||
      (int64_t)S.segs.size() > sk_seg_cap(units) || (int64_t)S.nslots > sk_slot_cap(units)) {
    aj::set_error("aj_crt_sk_schedule: %zu segments, %zu split units, %d slots exceed the tables", S.segs.size(),
                  S.fix.size(), S.nslots);
    return AJ_EINVAL;
  }
  static_assert(sizeof(SkSeg) == 8 * sizeof(int) && sizeof(SkFix) == 4 * sizeof(int), "table layout");
  std::memcpy(segs, S.segs.data(), sizeof(SkSeg) * S.segs.size());
  std::memcpy(beg, S.beg.data(), sizeof(int) * S.beg.size());
  if (!S.fix.empty()) std::memcpy(fix, S.fix.data(), sizeof(SkFix) * S.fix.size());
  counts[0] = (int64_t)S.segs.size();
  counts[1] = (int64_t)S.fix.size();
  counts[2] = S.nslots;
  return AJ_OK;
}
