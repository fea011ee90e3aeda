// Source-mesh mapper on the device: Delaunay triangulation of each lens-model sample's source vertices,
// the mesh's neighbour lists, and each image pixel's source-plane position located in the mesh with
// barycentric weights -- the inputs pix_indexes_for_sub_slim_index / pix_weights_for_sub_slim_index /
// neighbors / neighbors_sizes of the likelihood (numba.py:355-360, :672-677), which the reference's
// upstream builds on the host (external/get_dataset.py:38-60: a Delaunay mesh and its mapper).
//
// Triangulation: Bowyer-Watson insertion, one workgroup per sample, everything in LDS.  The triangle
// set includes "ghost" triangles (u, v, G) on the hull edges, G a vertex at infinity, so that the
// triangles form a closed surface: every directed edge u -> v of one triangle appears as v -> u in
// exactly one other.  Finite triangles are counter-clockwise in the (x, y) plane; a ghost triangle
// (u, v, G) has the finite side to the right of u -> v.  Inserting point p: every thread tests its share
// of the triangles for conflict (p strictly inside the circumcircle of a finite triangle; strictly left
// of a ghost's edge, or on the open edge), the conflicting triangles form a cavity whose boundary edges
// (those whose reverse is not in the cavity) are joined to p; the new triangles reuse the cavity's slots
// in order, then append two.  A point on an existing vertex conflicts with nothing and is left out (it
// belongs to no triangle), as Qhull leaves out coincident points.  The predicates are FP64 (relative
// to p); for points in general position the result is the unique Delaunay triangulation.  Cocircular
// points give one of the valid triangulations.
//
// Location: a pixel is in the first triangle (slot order) whose three sub-areas are all >= 0; its
// weights are the barycentric coordinates computed as aj_barycentric_weights computes them.  A pixel
// outside the hull maps to its nearest vertex with weight 1 (index -1 and weight 0 in the other slots).

#include "aj_common.h"

namespace aj {
namespace mp {

constexpr int DT_THREADS = 1024;  // sixteen waves: the only workgroup on its CU (LDS), so they hide each other's LDS latency
constexpr int DT_NMAX = 3072;          // vertices per sample
constexpr int DT_TMAX = 2 * DT_NMAX;   // triangles including ghosts: 2n - 2
constexpr int DT_CAV = 512;            // conflicting triangles per insertion
constexpr int GHOST = -1;
constexpr int LOC_CHUNK = 512;         // triangles staged per pass of the location kernel

enum Status { ST_OK = 0, ST_TOO_MANY = 1, ST_COLLINEAR = 2, ST_CAVITY_SIZE = 3, ST_CAVITY_SHAPE = 4, ST_DEGREE = 5 };

// The predicates are evaluated without fused multiply-adds: contracting x*y - z*w into fma(x, y, -z*w)
// makes the result depend on the operand order (orient(a, b, b) comes out as the rounding error of a
// product instead of 0), and the insertion then meets cavities that are not star-shaped.

// twice the signed area of (a, b, c) in the (x, y) plane, > 0 counter-clockwise; the same expression as
// the barycentric kernels use
__device__ inline double orient(double ax, double ay, double bx, double by, double cx, double cy) {
#pragma clang fp contract(off)
  return (bx - ax) * (cy - ay) - (cx - ax) * (by - ay);
}

// > 0 when p is strictly inside the circumcircle of the counter-clockwise triangle (a, b, c)
__device__ inline double incircle(double ax, double ay, double bx, double by, double cx, double cy, double px,
                                  double py) {
#pragma clang fp contract(off)
  const double adx = ax - px, ady = ay - py, bdx = bx - px, bdy = by - py, cdx = cx - px, cdy = cy - py;
  return (adx * adx + ady * ady) * (bdx * cdy - cdx * bdy) + (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy) +
         (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady);
}

__device__ inline double dot(double ax, double ay, double bx, double by) {
#pragma clang fp contract(off)
  return ax * bx + ay * by;
}

// Vertices V[voff[s] + i] = (y, x); finite triangles out to tri[toff[s] + 3 t ...] (capacity 2 n per
// sample, n = voff[s + 1] - voff[s]), their count to ntri[s]; neighbours (sorted, -1 padded, at most
// P) to nbr[(voff[s] + i) * P ...] and nsz[voff[s] + i]; status[s].
__global__ void __launch_bounds__(DT_THREADS) delaunay_kernel(const double* __restrict__ V,
                                                              const int64_t* __restrict__ voff, int P,
                                                              int64_t* __restrict__ tri, int64_t* __restrict__ ntri_out,
                                                              int64_t* __restrict__ nbr, int64_t* __restrict__ nsz,
                                                              int* __restrict__ status) {
  __shared__ double sx[DT_NMAX], sy[DT_NMAX];
  __shared__ int tv[3 * DT_TMAX];
  __shared__ int bad[DT_CAV];
  __shared__ int isb[3 * DT_CAV];
  __shared__ int eu[DT_CAV + 2], ev[DT_CAV + 2];
  __shared__ int cnt[DT_NMAX];
  __shared__ int part[DT_THREADS + 1];
  __shared__ int s_nbad, s_ntri, s_err, s_init[3];

  const int s = blockIdx.x, tid = threadIdx.x;
  const int64_t v0 = voff[s];
  const int64_t n64 = voff[s + 1] - v0;
  const int64_t t0 = 2 * v0;  // triangle output offset (in triangles)
  if (n64 > DT_NMAX || n64 < 3) {
    if (tid == 0) {
      status[s] = n64 > DT_NMAX ? ST_TOO_MANY : ST_COLLINEAR;
      ntri_out[s] = 0;
    }
    return;
  }
  const int n = (int)n64;
  for (int i = tid; i < n; i += DT_THREADS) {
    sy[i] = V[2 * (v0 + i)];
    sx[i] = V[2 * (v0 + i) + 1];
    cnt[i] = 0;
  }
  __syncthreads();
  if (tid == 0) {
    // first triangle: vertex 0, the first vertex distinct from it, the first vertex off their line
    int a = 0, b = -1, c = -1;
    for (int i = 1; i < n && b < 0; ++i)
      if (sx[i] != sx[0] || sy[i] != sy[0]) b = i;
    for (int i = 1; b >= 0 && i < n && c < 0; ++i)
      if (i != b && orient(sx[a], sy[a], sx[b], sy[b], sx[i], sy[i]) != 0.0) c = i;
    s_err = c < 0 ? ST_COLLINEAR : ST_OK;
    if (c >= 0) {
      if (orient(sx[a], sy[a], sx[b], sy[b], sx[c], sy[c]) < 0.0) {
        const int t = a;
        a = b;
        b = t;
      }
      const int init[12] = {a, b, c, b, a, GHOST, c, b, GHOST, a, c, GHOST};
      for (int i = 0; i < 12; ++i) tv[i] = init[i];
      s_ntri = 4;
      s_init[0] = a;
      s_init[1] = b;
      s_init[2] = c;
    }
  }
  __syncthreads();
  if (s_err == ST_OK) {
    const int ia = s_init[0], ib = s_init[1], ic = s_init[2];
    for (int k = 0; k < n; ++k) {
      if (k == ia || k == ib || k == ic) continue;
      const double px = sx[k], py = sy[k];
      if (tid == 0) s_nbad = 0;
      __syncthreads();
      const int nt = s_ntri;
      for (int t = tid; t < nt; t += DT_THREADS) {
        const int a = tv[3 * t], b = tv[3 * t + 1], c = tv[3 * t + 2];
        bool conflict;
        if (c == GHOST) {
          const double o = orient(sx[a], sy[a], sx[b], sy[b], px, py);
          conflict = o > 0.0 || (o == 0.0 && dot(px - sx[a], py - sy[a], sx[b] - sx[a], sy[b] - sy[a]) > 0.0 &&
                                 dot(px - sx[b], py - sy[b], sx[a] - sx[b], sy[a] - sy[b]) > 0.0);
        } else {
          conflict = incircle(sx[a], sy[a], sx[b], sy[b], sx[c], sy[c], px, py) > 0.0;
        }
        if (conflict) {
          const int i = atomicAdd(&s_nbad, 1);
          if (i < DT_CAV) bad[i] = t;
        }
      }
      __syncthreads();
      const int nb = s_nbad;
      if (nb > DT_CAV) {
        if (tid == 0) s_err = ST_CAVITY_SIZE;
        break;
      }
      if (nb == 0) continue;  // coincides with a vertex already in the mesh
      if (3 * nb <= 64) {
        // One wave: the cavity's slots in slot order (so that the result does not depend on the atomics'
        // order) by a rank sort over lane shuffles; lane i < 3 nb takes edge i % 3 of the i / 3-th
        // triangle; an edge is on the boundary when its reverse is no other cavity edge (an edge's reverse
        // is never in its own triangle); boundary edge of rank r among them makes the new triangle in
        // slot r (the cavity's slots, then two appended).
        if (tid < 64) {
          const int sv = tid < nb ? bad[tid] : 0x7fffffff;
          int rank = 0;
          for (int j = 0; j < nb; ++j) rank += __shfl(sv, j) < sv;
          if (tid < nb) bad[rank] = sv;  // every lane has read bad[] above (one wave, program order)
          const bool live = tid < 3 * nb;
          int u = -2, v = -3;  // ids that match no vertex (GHOST is -1)
          if (live) {
            const int t = bad[tid / 3], e = tid % 3;  // reads the sorted slots written above
            u = tv[3 * t + e];
            v = tv[3 * t + (e + 1) % 3];
          }
          bool inner = false;
          for (int j = 0; j < 3 * nb; ++j) inner |= (__shfl(u, j) == v) & (__shfl(v, j) == u);
          const bool isbnd = live && !inner;
          const unsigned long long mask = __ballot(isbnd);
          const int m = __popcll(mask);
          if (m != nb + 2 || s_ntri + 2 > DT_TMAX) {
            if (tid == 0) s_err = m != nb + 2 ? ST_CAVITY_SHAPE : ST_CAVITY_SIZE;
          } else {
            int slot = 0;
            if (isbnd) {
              const int r = __popcll(mask & ((1ull << tid) - 1));
              slot = r < nb ? bad[r] : s_ntri + (r - nb);
            }
            // the cavity's triangles and s_ntri are read above, before any lane writes below
            if (isbnd) {
              tv[3 * slot] = u == GHOST ? v : (v == GHOST ? k : u);
              tv[3 * slot + 1] = u == GHOST ? k : (v == GHOST ? u : v);
              tv[3 * slot + 2] = (u == GHOST || v == GHOST) ? GHOST : k;
            }
            if (tid == 0) s_ntri += 2;
          }
        }
      } else {
        // larger cavities: the same steps with workgroup barriers between them
        if (nb <= 64) {
          if (tid < 64) {
            const int sv = tid < nb ? bad[tid] : 0x7fffffff;
            int rank = 0;
            for (int j = 0; j < nb; ++j) rank += __shfl(sv, j) < sv;
            if (tid < nb) bad[rank] = sv;
          }
        } else if (tid == 0) {
          for (int i = 1; i < nb; ++i) {
            const int sv = bad[i];
            int j = i - 1;
            while (j >= 0 && bad[j] > sv) {
              bad[j + 1] = bad[j];
              --j;
            }
            bad[j + 1] = sv;
          }
        }
        __syncthreads();
        for (int i = tid; i < 3 * nb; i += DT_THREADS) {
          const int t = bad[i / 3], e = i % 3;
          const int u = tv[3 * t + e], v = tv[3 * t + (e + 1) % 3];
          bool inner = false;
          for (int j = 0; j < nb && !inner; ++j) {
            const int t2 = bad[j];
            if (t2 == t) continue;
            for (int e2 = 0; e2 < 3; ++e2)
              if (tv[3 * t2 + e2] == v && tv[3 * t2 + (e2 + 1) % 3] == u) inner = true;
          }
          isb[i] = inner ? 0 : 1;
        }
        __syncthreads();
        if (tid == 0) {
          int m = 0;
          for (int i = 0; i < 3 * nb && m <= DT_CAV + 1; ++i)
            if (isb[i]) {
              const int t = bad[i / 3], e = i % 3;
              eu[m] = tv[3 * t + e];
              ev[m] = tv[3 * t + (e + 1) % 3];
              ++m;
            }
          if (m != nb + 2 || s_ntri + 2 > DT_TMAX) {
            s_err = m != nb + 2 ? ST_CAVITY_SHAPE : ST_CAVITY_SIZE;
          } else {
            for (int j = 0; j < m; ++j) {
              const int slot = j < nb ? bad[j] : s_ntri + (j - nb);
              const int u = eu[j], v = ev[j];
              // (u, v, k), rotated so that a ghost vertex comes last
              tv[3 * slot] = u == GHOST ? v : (v == GHOST ? k : u);
              tv[3 * slot + 1] = u == GHOST ? k : (v == GHOST ? u : v);
              tv[3 * slot + 2] = (u == GHOST || v == GHOST) ? GHOST : k;
            }
            s_ntri += 2;
          }
        }
      }
      __syncthreads();
      if (s_err != ST_OK) break;
    }
  }
  __syncthreads();
  const int err = s_err;
  if (err != ST_OK) {
    if (tid == 0) {
      status[s] = err;
      ntri_out[s] = 0;
    }
    return;
  }
  // finite triangles in slot order: per-thread contiguous ranges, their counts prefix-summed
  const int nt = s_ntri;
  const int per = (nt + DT_THREADS - 1) / DT_THREADS;
  const int lo = tid * per, hi = min(nt, lo + per);
  int mine = 0;
  for (int t = lo; t < hi; ++t) mine += tv[3 * t + 2] != GHOST;
  part[tid + 1] = mine;
  __syncthreads();
  if (tid == 0) {
    part[0] = 0;
    for (int i = 1; i <= DT_THREADS; ++i) part[i] += part[i - 1];
  }
  __syncthreads();
  int w = part[tid];
  for (int t = lo; t < hi; ++t)
    if (tv[3 * t + 2] != GHOST) {
      for (int e = 0; e < 3; ++e) tri[3 * (t0 + w) + e] = tv[3 * t + e];
      ++w;
    }
  // neighbours: every directed edge u -> v between finite vertices occurs once on the closed surface
  __shared__ int s_over;
  if (tid == 0) s_over = 0;
  __syncthreads();
  for (int t = tid; t < nt; t += DT_THREADS)
    for (int e = 0; e < 3; ++e) {
      const int u = tv[3 * t + e], v = tv[3 * t + (e + 1) % 3];
      if (u == GHOST || v == GHOST) continue;
      const int pos = atomicAdd(&cnt[u], 1);
      if (pos < P)
        nbr[(v0 + u) * P + pos] = v;
      else
        s_over = 1;
    }
  __syncthreads();
  for (int i = tid; i < n; i += DT_THREADS) {
    int64_t* row = nbr + (v0 + i) * P;
    const int c = min(cnt[i], P);
    for (int a = 1; a < c; ++a) {
      const int64_t v = row[a];
      int b = a - 1;
      while (b >= 0 && row[b] > v) {
        row[b + 1] = row[b];
        --b;
      }
      row[b + 1] = v;
    }
    for (int a = c; a < P; ++a) row[a] = -1;
    nsz[v0 + i] = c;
  }
  if (tid == 0) {
    status[s] = s_over ? ST_DEGREE : ST_OK;
    ntri_out[s] = part[DT_THREADS];
  }
}

// Each pixel of sample blockIdx.y (position pos[(blockIdx.y * pstride + m) * 2 ...] = (y, x)) located in
// the sample's finite triangles; idx / wts [ns, M, 3].
__global__ void __launch_bounds__(256) locate_kernel(const double* __restrict__ V, const int64_t* __restrict__ voff,
                                                     const int64_t* __restrict__ tri,
                                                     const int64_t* __restrict__ ntri, const double* __restrict__ pos,
                                                     int64_t pstride, int64_t M, int64_t* __restrict__ idx,
                                                     double* __restrict__ wts) {
  __shared__ double cx[3][LOC_CHUNK], cy[3][LOC_CHUNK];
  const int s = blockIdx.y, tid = threadIdx.x;
  const int64_t m = (int64_t)blockIdx.x * 256 + tid;
  const bool live = m < M;
  const int64_t v0 = voff[s], n = voff[s + 1] - v0, t0 = 2 * v0, nt = ntri[s];
  double py = 0.0, px = 0.0;
  if (live) {
    py = pos[2 * ((int64_t)s * pstride + m)];
    px = pos[2 * ((int64_t)s * pstride + m) + 1];
  }
  int64_t found = -1;
  double w0 = 0.0, w1 = 0.0, w2 = 0.0;
  for (int64_t c0 = 0; c0 < nt; c0 += LOC_CHUNK) {
    const int cn = (int)min((int64_t)LOC_CHUNK, nt - c0);
    for (int i = tid; i < 3 * cn; i += 256) {
      const int t = i / 3, e = i % 3;
      const int64_t v = v0 + tri[3 * (t0 + c0 + t) + e];
      cy[e][t] = V[2 * v];
      cx[e][t] = V[2 * v + 1];
    }
    __syncthreads();
    if (live && found < 0) {
      for (int t = 0; t < cn; ++t) {
        const double a0 = orient(px, py, cx[1][t], cy[1][t], cx[2][t], cy[2][t]);
        const double a1 = orient(cx[0][t], cy[0][t], px, py, cx[2][t], cy[2][t]);
        const double a2 = orient(cx[0][t], cy[0][t], cx[1][t], cy[1][t], px, py);
        if (a0 >= 0.0 && a1 >= 0.0 && a2 >= 0.0) {
          const double A = orient(cx[0][t], cy[0][t], cx[1][t], cy[1][t], cx[2][t], cy[2][t]);
          found = c0 + t;
          w0 = a0 / A;
          w1 = a1 / A;
          w2 = a2 / A;
          break;
        }
      }
    }
    __syncthreads();
  }
  if (!live) return;
  int64_t* io = idx + ((int64_t)s * M + m) * 3;
  double* wo = wts + ((int64_t)s * M + m) * 3;
  if (found >= 0) {
    for (int e = 0; e < 3; ++e) io[e] = tri[3 * (t0 + found) + e];
    wo[0] = w0;
    wo[1] = w1;
    wo[2] = w2;
  } else {
    int64_t best = 0;
    double bd = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double dy = V[2 * (v0 + i)] - py, dx = V[2 * (v0 + i) + 1] - px;
      const double d = dx * dx + dy * dy;
      if (i == 0 || d < bd) {
        bd = d;
        best = i;
      }
    }
    io[0] = best;
    io[1] = -1;
    io[2] = -1;
    wo[0] = 1.0;
    wo[1] = 0.0;
    wo[2] = 0.0;
  }
}

}  // namespace mp
}  // namespace aj

using namespace aj;

extern "C" int aj_delaunay_mesh_batched(const double* vertices, const int64_t* vertex_offsets,
                                        const int64_t* vertex_offsets_host, int64_t n_samples, int64_t max_neighbors,
                                        int64_t* triangles, int64_t* n_triangles, int64_t* neighbors,
                                        int64_t* neighbors_sizes, int* status, void* stream) {
  AJ_REQUIRE(n_samples > 0 && n_samples < (1 << 20) && max_neighbors > 0 && max_neighbors < (1 << 20),
             "bad sizes n_samples=%lld max_neighbors=%lld", (long long)n_samples, (long long)max_neighbors);
  AJ_REQUIRE(vertices && vertex_offsets && vertex_offsets_host && triangles && n_triangles && neighbors &&
                 neighbors_sizes && status,
             "null pointer");
  for (int64_t s = 0; s < n_samples; ++s) {
    const int64_t n = vertex_offsets_host[s + 1] - vertex_offsets_host[s];
    AJ_REQUIRE(n >= 3 && n <= mp::DT_NMAX, "sample %lld has %lld vertices (3..%d supported)", (long long)s,
               (long long)n, mp::DT_NMAX);
  }
  hipLaunchKernelGGL(mp::delaunay_kernel, dim3((unsigned)n_samples), dim3(mp::DT_THREADS), 0, as_stream(stream),
                     vertices, vertex_offsets, (int)max_neighbors, triangles, n_triangles, neighbors, neighbors_sizes,
                     status);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}

extern "C" int aj_delaunay_locate_batched(const double* vertices, const int64_t* vertex_offsets, int64_t n_samples,
                                          const int64_t* triangles, const int64_t* n_triangles,
                                          const double* positions, int64_t positions_stride, int64_t M,
                                          int64_t* pix_indexes_for_sub_slim_index,
                                          double* pix_weights_for_sub_slim_index, void* stream) {
  AJ_REQUIRE(n_samples > 0 && n_samples < 65536 && M > 0 && M < (1ll << 31) && positions_stride >= 0,
             "bad sizes n_samples=%lld M=%lld", (long long)n_samples, (long long)M);
  AJ_REQUIRE(vertices && vertex_offsets && triangles && n_triangles && positions && pix_indexes_for_sub_slim_index &&
                 pix_weights_for_sub_slim_index,
             "null pointer");
  hipLaunchKernelGGL(mp::locate_kernel, dim3((unsigned)ceil_div(M, (int64_t)256), (unsigned)n_samples), dim3(256), 0,
                     as_stream(stream), vertices, vertex_offsets, triangles, n_triangles, positions, positions_stride,
                     M, pix_indexes_for_sub_slim_index, pix_weights_for_sub_slim_index);
  AJ_LAUNCH_CHECK();
  return AJ_OK;
}
