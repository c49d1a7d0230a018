// supermesh.hip -- clipping kernels, reductions and P1 x P1 mixed mass.
//
// One candidate pair per lane.  The kernels read element vertex coordinates
// from the element-major arrays staged by finder.hip (48 B per triangle,
// 96 B per tet) and, for tets, B's four planes precomputed once per element
// (get_planes_tet is a pure function of the element, so this is bitwise the
// same as recomputing it per pair).  Per-pair counts and measures are
// bit-identical to the reference's intersect_tris / intersect_tets followed
// by a plain in-order sum of triangle_area / tetrahedron_volume; the global
// sum uses double-double (TwoSum) accumulation in a fixed reduction tree, so
// it is deterministic for a given input.
#include <algorithm>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "lsm_geometry.hpp"
#include "lsm_runtime.hpp"
#include "predictor.hpp"

namespace lsm {

constexpr int kClipBlock = 256;
constexpr int kClipGrid = 2048;  // fixed: the reduction tree depends only on npairs
// The tet triage's blocks: each compacts one contiguous chunk of the pairs
constexpr int kTriageGrid = 16384;

struct Partial {
  double hi, lo;
  long long nonempty, nout;
  long long maxc, pad;
};

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void dd_add(double& hi, double& lo, double x) {
  double s, e;
  two_sum(hi, x, s, e);
  hi = s;
  lo = lo + e;
}

__device__ __forceinline__ void dd_merge(double& hi, double& lo, double hi2, double lo2) {
  double s, e;
  two_sum(hi, hi2, s, e);
  hi = s;
  lo = lo + (e + lo2);
}

// Quadrature rules of the reference's mixed-mass tests: Dunavant degree 2 on
// triangles (barycentric (4,1,1)/6, weights 1/3; test_parallel_p1_inner_product_2d.F90:59-68)
// and the 4-point Stroud rule on tets (..._3d.F90:61-80).
__constant__ double kQ2[3][3] = {{4.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0},
                                 {1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0},
                                 {1.0 / 6.0, 1.0 / 6.0, 4.0 / 6.0}};
#define LSM_QA 5.8541019662496852e-1
#define LSM_QB 1.3819660112501050e-1
__constant__ double kQ3[4][4] = {{LSM_QA, LSM_QB, LSM_QB, LSM_QB},
                                 {LSM_QB, LSM_QA, LSM_QB, LSM_QB},
                                 {LSM_QB, LSM_QB, LSM_QA, LSM_QB},
                                 {LSM_QB, LSM_QB, LSM_QB, LSM_QA}};

template <int DIM>
__device__ __forceinline__ void basis(const double* __restrict__ aff, const double* x, double* f) {
  if constexpr (DIM == 2) {
    double dx = x[0] - aff[0], dy = x[1] - aff[1];
    double f2 = aff[2] * dx;
    f2 = f2 + aff[3] * dy;
    double f3 = aff[4] * dx;
    f3 = f3 + aff[5] * dy;
    f[1] = f2;
    f[2] = f3;
    f[0] = 1.0 - f2 - f3;
  } else {
    double dx = x[0] - aff[0], dy = x[1] - aff[1], dz = x[2] - aff[2];
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < 3; r++) {
      double v = aff[3 + 3 * r] * dx;
      v = v + aff[4 + 3 * r] * dy;
      v = v + aff[5 + 3 * r] * dz;
      f[1 + r] = v;
    }
    s = 1.0 - f[1] - f[2] - f[3];
    f[0] = s;
  }
}

// Adds the element's contributions to blk[LOC][LOC].  2D: the product of two
// P1 functions is quadratic, so the 3-point rule of the reference tests
// (test_parallel_p1_inner_product_2d.F90:59-68) equals the closed form
// int_c f g = |c|/12 (sum_v f_v g_v + (sum_v f_v)(sum_v g_v)) over c's
// vertex values; this evaluates the latter (about half the operations; sums
// may use fused multiply-adds: the mass is compared within 1e-12, not
// bitwise).  3D: the reference's 4-point rule as written.
template <int DIM>
__device__ __forceinline__ void mass_element(const double* el, double meas, const double* affa,
                                             const double* affb, double* blk) {
  if constexpr (DIM == 2) {
    double fa[3][3], fb[3][3];  // [vertex][basis]
#pragma unroll
    for (int v = 0; v < 3; v++) {
      basis<2>(affa, el + 2 * v, fa[v]);
      basis<2>(affb, el + 2 * v, fb[v]);
    }
    const double w = meas * (1.0 / 12.0);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const double sa = fa[0][i] + fa[1][i] + fa[2][i];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const double sb = fb[0][j] + fb[1][j] + fb[2][j];
        double d = __builtin_fma(fa[0][i], fb[0][j], sa * sb);
        d = __builtin_fma(fa[1][i], fb[1][j], d);
        d = __builtin_fma(fa[2][i], fb[2][j], d);
        blk[i * 3 + j] = __builtin_fma(w, d, blk[i * 3 + j]);
      }
    }
    return;
  }
  constexpr int LOC = DIM + 1;
  constexpr int NQ = DIM == 2 ? 3 : 4;
  const double wq = DIM == 2 ? 1.0 / 3.0 : 0.25;
  const double wm = wq * meas;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    double x[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      double s = 0.0;
#pragma unroll
      for (int v = 0; v < LOC; v++) s = s + (DIM == 2 ? kQ2[q][v] : kQ3[q][v]) * el[v * DIM + d];
      x[d] = s;
    }
    double fa[LOC], fb[LOC];
    basis<DIM>(affa, x, fa);
    basis<DIM>(affb, x, fb);
#pragma unroll
    for (int i = 0; i < LOC; i++)
#pragma unroll
      for (int j = 0; j < LOC; j++) blk[i * LOC + j] = blk[i * LOC + j] + wm * fa[i] * fb[j];
  }
}

// 2D P1 x P1 mass of a clipped polygon from its moments, the polygon's fan
// triangles accumulated in coordinates d = x - o local to A's first vertex o:
// mom = {int 1, int dx, int dy, int dx^2, int dx dy, int dy^2}.  Per
// triangle (area ar, vertices d_0..d_2; exact for quadratics):
// int dx = ar/3 sum dx_v, int dx dy = ar/12 (sum dx_v dy_v + sum dx_v sum dy_v).
__device__ __forceinline__ void tri_moments(double x0, double y0, double x1, double y1, double x2, double y2,
                                            double ar, double ox, double oy, double* mom) {
  const double a0 = x0 - ox, b0 = y0 - oy, a1 = x1 - ox, b1 = y1 - oy, a2 = x2 - ox, b2 = y2 - oy;
  const double sa = a0 + a1 + a2, sb = b0 + b1 + b2;
  const double w3 = ar * (1.0 / 3.0), w12 = ar * (1.0 / 12.0);
  mom[0] += ar;
  mom[1] += w3 * sa;
  mom[2] += w3 * sb;
  mom[3] += w12 * (a0 * a0 + a1 * a1 + a2 * a2 + sa * sa);
  mom[4] += w12 * (a0 * b0 + a1 * b1 + a2 * b2 + sa * sb);
  mom[5] += w12 * (b0 * b0 + b1 * b1 + b2 * b2 + sb * sb);
}

// blk[i][j] = int phi^A_i phi^B_j over the polygon of moments mom (origin o =
// A's first vertex, where phi^A = (1, 0, 0)): with phi^A_i = alpha_i + g_i.d
// and phi^B_j = beta_j + h_j.d, the integral is alpha_i beta_j I + alpha_i
// h_j.Id + beta_j g_i.Id + g_i^T Idd h_j.  The same integral as summing
// mass_element<2> over the fan, in exact arithmetic.
__device__ __forceinline__ void mass_from_moments(const double* mom, const double* affa, const double* affb,
                                                  double* blk) {
  double g[3][2], h[3][2], beta[3];
  g[1][0] = affa[2];
  g[1][1] = affa[3];
  g[2][0] = affa[4];
  g[2][1] = affa[5];
  g[0][0] = -(g[1][0] + g[2][0]);
  g[0][1] = -(g[1][1] + g[2][1]);
  h[1][0] = affb[2];
  h[1][1] = affb[3];
  h[2][0] = affb[4];
  h[2][1] = affb[5];
  h[0][0] = -(h[1][0] + h[2][0]);
  h[0][1] = -(h[1][1] + h[2][1]);
  const double o[2] = {affa[0], affa[1]};
  basis<2>(affb, o, beta);
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const double hd = h[j][0] * mom[1] + h[j][1] * mom[2];            // h_j . Id
    const double tx = mom[3] * h[j][0] + mom[4] * h[j][1];            // Idd h_j
    const double ty = mom[4] * h[j][0] + mom[5] * h[j][1];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      double v = beta[j] * (g[i][0] * mom[1] + g[i][1] * mom[2]) + (g[i][0] * tx + g[i][1] * ty);
      if (i == 0) v += beta[j] * mom[0] + hd;  // alpha = (1, 0, 0)
      blk[i * 3 + j] = v;
    }
  }
}

struct ClipArgs {
  long npairs;
  const long* npairs_dev;  // when set, the pair count is read here (a recorded graph's count pass)
  const int* rows;
  const int* cols;
  const int* work;    // when non-null: process pairs work[0..nwork)
  const unsigned long long* nwork;
  const double* acoords;
  const double* bcoords;
  const double* asp;
  const double* bsp;
  const double* ascale;
  const double* bscale;
  const double* bplanes;
  const double* aaff;
  const double* baff;
  int* counts;
  double* measure;
  double* blocks;
  const long* offsets;
  double* elements;
  // single pass (predicted budget) clip: per work item budget; each pair's
  // run starts at offsets[p] (the budgets laid end to end)
  const unsigned* budget;  // packed budget levels; this call uses byte bshift / 8
  int bshift;
  // outputs beyond a budget continue in the walking lane's own arena region
  // (no atomics: a lane finishes one pair before the next), elements
  // [arena_base + lane * arena_per_lane, + arena_per_lane); aoff[w] = where
  // work item w's continue.  dropped[w]: 1 the run is split (budget, then
  // arena), 2 the lane's arena ran out (the pair is clipped again)
  long arena_base;
  long arena_per_lane;
  long* aoff;
  unsigned char* dropped;
  unsigned long long* ndropped;  // [0] split runs, [1] pairs clipped again
  // fill pass: when non-null, only items flagged 2 here, each written at
  // aoff[w] (its place in the spill reserve)
  const unsigned char* ovf;
  // tet walks: one work-item counter per XCD share (zeroed before each launch)
  unsigned long long* queue;
  // 2D: pairs whose working polygon outgrew its LDS slots, for the
  // streaming clipper
  int* olist;
  unsigned long long* nolist;
  // mass row slots (2D): the count pass also writes each pair's contributions
  // (B node, value) to the rows of A's nodes, at
  // rbase[a * LOC + li] + (p - indptr[a]) * LOC + j, j = 0..LOC-1 (B's local
  // nodes); an empty pair writes node -1 and no values.  Per-pair blocks
  // only when `blocks` is set.
  const long* rbase;
  const long* indptr;
  const int* enlb;
  int* rkey;
  double* rval;
};

__device__ __forceinline__ void load_planes(const double* __restrict__ bp, long b, Planes& P) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
#pragma unroll
    for (int k = 0; k < 3; k++) P.n[i][k] = bp[b * 16 + 4 * i + k];
    P.c[i] = bp[b * 16 + 4 * i + 3];
  }
}

// Output elements are written once and not read again by the pass that
// writes them: streaming (nontemporal) stores keep them from evicting the
// meshes' coordinates and planes that the clippers read from L2 (materialised
// tet 96^3 vs 90^3: the fill walk and the single pass each 6 ms faster).
__device__ __forceinline__ void stream_tri(double* o, double x0, double y0, double x1, double y1, double x2,
                                           double y2) {
  __builtin_nontemporal_store(x0, o);
  __builtin_nontemporal_store(y0, o + 1);
  __builtin_nontemporal_store(x1, o + 2);
  __builtin_nontemporal_store(y1, o + 3);
  __builtin_nontemporal_store(x2, o + 4);
  __builtin_nontemporal_store(y2, o + 5);
}

__device__ __forceinline__ void stream_tet(double* o, const double* t) {
#pragma unroll
  for (int i = 0; i < 12; i++) __builtin_nontemporal_store(t[i], o + i);
}

// One 16-byte word of a run copy, stored the same way.
__device__ __forceinline__ void stream_word(double2* d, const double2* s) {
  using v2 = double __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(*reinterpret_cast<const v2*>(s), reinterpret_cast<v2*>(d));
}

__device__ __forceinline__ long pair_count(const ClipArgs& g) { return g.npairs_dev ? *g.npairs_dev : g.npairs; }

// Tet pairs: pairs whose A lies provably beyond one of B's face planes get
// their (empty) result here; the rest are listed for the full clip.  Block b
// owns pairs [b*chunk, (b+1)*chunk), chunk = ceil(npairs / grid), and writes
// the survivors, in pair order, to its own region work[b*chunk ..] (no
// atomics); k_gather_work then packs the regions into one dense list.
template <bool MASS>
__global__ void __launch_bounds__(kClipBlock) k_tet_triage(ClipArgs g, int* __restrict__ work,
                                                           long* __restrict__ block_count) {
  __shared__ int wave_count[kClipBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long np = pair_count(g);
  const long chunk = (np + gridDim.x - 1) / gridDim.x;
  const long lo = (long)blockIdx.x * chunk;
  const long hi = lo + chunk < np ? lo + chunk : np;
  long kept = 0;
  for (long base = lo; base < hi; base += kClipBlock) {
    const long p = base + threadIdx.x;
    bool keep = false;
    if (p < hi) {
      const long a = g.rows[p], b = g.cols[p];
      double A[12];
#pragma unroll
      for (int k = 0; k < 12; k++) A[k] = g.acoords[a * 12 + k];
      Planes P;
      load_planes(g.bplanes, b, P);
      const double sc = fmax_f(g.ascale[a], g.bscale[b]);
      keep = !tet_separated_by_b(A, P, sc);
      if (!keep) {
        __builtin_nontemporal_store(0, g.counts + p);
        __builtin_nontemporal_store(0.0, g.measure + p);
        if (MASS) {
#pragma unroll
          for (int k = 0; k < 16; k++) g.blocks[p * 16 + k] = 0.0;
        }
      }
    }
    const unsigned long long m = __ballot(keep);
    if (lane == 0) wave_count[wave] = __popcll(m);
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kClipBlock / 64; w++) {
      before += w < wave ? wave_count[w] : 0;
      total += wave_count[w];
    }
    if (keep) work[lo + kept + before + __popcll(m & ((1ULL << lane) - 1))] = (int)p;
    kept += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) block_count[blockIdx.x] = kept;
}

__global__ void __launch_bounds__(kClipBlock)
k_gather_work(long npairs, const long* __restrict__ npairs_dev, const long* __restrict__ block_offset,
              const int* __restrict__ src, int* __restrict__ dst) {
  const long np = npairs_dev ? *npairs_dev : npairs;
  const long chunk = (np + gridDim.x - 1) / gridDim.x;
  const long b = blockIdx.x;
  const long o = block_offset[b], n = block_offset[b + 1] - o;
  for (long i = threadIdx.x; i < n; i += kClipBlock) dst[o + i] = src[b * chunk + i];
}

// Exclusive offsets of the triage's kTriageGrid block counts (one block):
// offset[0] = 0, offset[b + 1] = count[0] + .. + count[b];
// offset[kTriageGrid] is the work list's length.
__global__ void __launch_bounds__(1024) k_scan_blocks(const long* __restrict__ count, long* __restrict__ offset) {
  constexpr int K = kTriageGrid / 1024;  // consecutive counts per thread
  static_assert(kTriageGrid % 1024 == 0, "whole counts per thread");
  __shared__ long part[1024];
  const int t = threadIdx.x;
  long c[K], sum = 0;
#pragma unroll
  for (int i = 0; i < K; i++) {
    c[i] = count[K * t + i];
    sum += c[i];
  }
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive scan of the threads' sums (Hillis-Steele)
    const long v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long run = t > 0 ? part[t - 1] : 0;
  if (t == 0) offset[0] = 0;
#pragma unroll
  for (int i = 0; i < K; i++) {
    run += c[i];
    offset[K * t + i + 1] = run;
  }
}

template <int DIM, bool MASS, bool FILL>
__global__ void __launch_bounds__(kClipBlock) k_clip(ClipArgs g) {
  constexpr int LOC = DIM + 1;
  constexpr int ESZ = LOC * DIM;
  const long stride = (long)gridDim.x * blockDim.x;
  const long n = g.work ? (long)*g.nwork : pair_count(g);
  for (long w = (long)blockIdx.x * blockDim.x + threadIdx.x; w < n; w += stride) {
    const long p = g.work ? g.work[w] : w;
    const long a = g.rows[p], b = g.cols[p];
    long off = 0;
    int cap = 0;  // fill pass: the count pass's result bounds the writes
    if (FILL) {
      cap = g.counts[p];
      if (cap == 0) continue;
      off = g.offsets[p];
    }
    double A[ESZ];
#pragma unroll
    for (int k = 0; k < ESZ; k++) A[k] = g.acoords[a * ESZ + k];
    const double tol = 10.0 * fmin_f(g.asp[a], g.bsp[b]);
    double m = 0.0;
    double blk[LOC * LOC];
    double affa[DIM + DIM * DIM], affb[DIM + DIM * DIM];
    if (MASS) {
#pragma unroll
      for (int k = 0; k < LOC * LOC; k++) blk[k] = 0.0;
#pragma unroll
      for (int k = 0; k < DIM + DIM * DIM; k++) {
        affa[k] = g.aaff[a * (DIM + DIM * DIM) + k];
        affb[k] = g.baff[b * (DIM + DIM * DIM) + k];
      }
    }
    int nk;
    if constexpr (DIM == 2) {
      double Bv[ESZ];
#pragma unroll
      for (int k = 0; k < ESZ; k++) Bv[k] = g.bcoords[b * ESZ + k];
      Line L[3];
      tri_lines(Bv, L);
      int k = 0;
      nk = clip_tri(A, L, tol, [&](double x0, double y0, double x1, double y1, double x2, double y2,
                                   double ar) {
        m = m + ar;
        if (MASS) {
          double el[6] = {x0, y0, x1, y1, x2, y2};
          mass_element<2>(el, ar, affa, affb, blk);
        }
        if (FILL && k < cap) {
          stream_tri(g.elements + (off + k) * 6, x0, y0, x1, y1, x2, y2);
        }
        k++;
      });
    } else {
      Planes P;
      load_planes(g.bplanes, b, P);
      Tet TA;
#pragma unroll
      for (int kk = 0; kk < 12; kk++) TA.v[kk] = A[kk];
      int k = 0;
      nk = clip_tet(TA, P, tol, [&](const Tet& T, double) {
        double vr = tet_volume_real(T.v);
        m = m + vr;
        if (MASS) mass_element<3>(T.v, vr, affa, affb, blk);
        if (FILL && k < cap) {
          stream_tet(g.elements + (off + k) * 12, T.v);
        }
        k++;
      });
    }
    if (FILL) continue;
    if (nk == 0) m = 0.0;
    g.counts[p] = nk;
    g.measure[p] = m;
    if (MASS) {
      if (g.blocks) {
#pragma unroll
        for (int k = 0; k < LOC * LOC; k++) g.blocks[p * LOC * LOC + k] = nk == 0 ? 0.0 : blk[k];
      }
      if (g.rkey) {
        const long r = (p - g.indptr[a]) * LOC;
#pragma unroll
        for (int li = 0; li < LOC; li++) {
          const long o = g.rbase[a * LOC + li] + r;
#pragma unroll
          for (int j = 0; j < LOC; j++) g.rkey[o + j] = nk == 0 ? -1 : g.enlb[b * LOC + j];
          if (nk > 0) {
#pragma unroll
            for (int j = 0; j < LOC; j++) g.rval[o + j] = blk[li * LOC + j];
          }
        }
      }
    }
  }
}

// The polygon clip of intersect_tris / intersect_polys_lines
// (Triangle_Intersection.F90:144-185, 216-276, 282-335) with the
// intermediate polygons in LDS: pass 0 reads A from registers, passes
// 1..NL-2 read the previous pass's polygon from LDS through one load site and
// write theirs through one store site, and the last pass streams into the fan.
// The streaming clip_poly instead inlines every later pass, and the sink, at
// every emit site of every earlier pass.  Same operations in the same order as
// clip_poly.  Pass s writes buffer s & 1 (capacity NA + s + 1, the most a
// convex polygon reaches); lane t's point k of buffer b is at
// q[(off(b) + k) * 2 * LANES] (x) and the next word row (y), q = base + t.
// Returns -1 when a pass would overflow its buffer (only possible when
// rounding makes the polygon non-convex); the caller then hands the pair to
// clip_poly.
template <int NA, int NL>
struct LdsPoly {
  static constexpr int cap(int b) {
    int c = 0;
    for (int s = b; s < NL - 1; s += 2) c = NA + s + 1 > c ? NA + s + 1 : c;
    return c;
  }
  static constexpr int kCap0 = cap(0), kCap1 = cap(1);
  static constexpr int kPoints = kCap0 + kCap1;  // LDS points per lane
};

template <int NL, int NA, int LANES, class Sink>
__device__ __forceinline__ int clip_poly_lds(const double* A, const Line* L, double tol, double* q,
                                             Sink&& sink) {
  using P = LdsPoly<NA, NL>;
  auto X = [&](int b, int k) -> double& { return q[((b ? P::kCap0 : 0) + k) * 2 * LANES]; };
  auto Y = [&](int b, int k) -> double& { return q[(((b ? P::kCap0 : 0) + k) * 2 + 1) * LANES]; };
  // the fan of the last pass's points (triangle (q1, q_{k-1}, q_k) for k >= 3)
  double fx0 = 0.0, fy0 = 0.0, fx1 = 0.0, fy1 = 0.0;
  int fn = 0, kept = 0;
  auto fan_push = [&](double x, double y) {
    if (fn == 0) {
      fx0 = x;
      fy0 = y;
    } else if (fn >= 2) {
      const double ar = triangle_area6(fx0, fy0, fx1, fy1, x, y);
      if (!(ar < tol)) {
        sink(fx0, fy0, fx1, fy1, x, y, ar);
        kept++;
      }
    }
    fx1 = x;
    fy1 = y;
    fn++;
  };
  int n = NA;
#pragma unroll
  for (int s = 0; s < NL; s++) {
    const int src = (s + 1) & 1, dst = s & 1;
    const int cap = dst ? P::kCap1 : P::kCap0;
    const bool last = s == NL - 1;
    const Line l = L[s];
    const double fx = s == 0 ? A[0] : X(src, 0), fy = s == 0 ? A[1] : Y(src, 0);
    double fd = l.nx * (fx - l.px);
    fd = fd + l.ny * (fy - l.py);
    double x1 = fx, y1 = fy, d1 = fd;
    int m = 0;
    bool over = false;
    auto store = [&](double u, double v) {
      if (m < cap) {
        X(dst, m) = u;
        Y(dst, m) = v;
      } else {
        over = true;
      }
    };
    auto edge = [&](double x2, double y2, double d2) {
      if (last) clip_edge(x1, y1, d1, x2, y2, d2, m, fan_push);
      else clip_edge(x1, y1, d1, x2, y2, d2, m, store);
      x1 = x2;
      y1 = y2;
      d1 = d2;
    };
    if (s == 0) {
#pragma unroll
      for (int i = 1; i < NA; i++) {
        double d2 = l.nx * (A[2 * i] - l.px);
        d2 = d2 + l.ny * (A[2 * i + 1] - l.py);
        edge(A[2 * i], A[2 * i + 1], d2);
      }
    } else {
      for (int i = 1; i < n; i++) {
        const double x2 = X(src, i), y2 = Y(src, i);
        double d2 = l.nx * (x2 - l.px);
        d2 = d2 + l.ny * (y2 - l.py);
        edge(x2, y2, d2);
      }
    }
    edge(fx, fy, fd);
    if (over) return -1;
    if (m < 3) return 0;
    n = m;
  }
  return kept;
}

constexpr int kTriBlock = 128;

// The 2D clippers' grid: one block per kTriBlock (or `block`) pairs, every
// lane one pair.  Finished blocks make room for new ones at once, so lanes
// with short polygons never wait behind a block's slowest lane more than
// once (tri 1024^2 vs 1000^2 with the mixed mass: 1.19 ms, against 1.37 ms
// with 4096 blocks looping over the pairs).
static unsigned one_round_grid(long n, int block = kTriBlock) {
  const long g = (n + block - 1) / block;
  return (unsigned)(g > 0 ? g : 1);
}

// 2D pairs, one per lane: the count pass (counts, measures, mass blocks and
// mass row slots) or the fill pass, as k_clip<2, ...>, on clip_poly_lds.
// Pairs that outgrow the LDS polygon go to g.olist for k_clip.
template <bool MASS, bool FILL, int BLOCK = kTriBlock>
__global__ void __launch_bounds__(BLOCK) k_clip_tri(ClipArgs g) {
  __shared__ double poly[LdsPoly<3, 3>::kPoints * 2 * BLOCK];
  double* q = poly + threadIdx.x;
  const long np = pair_count(g);
  // count pass with mass: the mass row slots and the 9-value blocks of the
  // block's pairs leave through the (then idle) polygon slots, written by
  // consecutive threads to consecutive words
  constexpr bool STAGE = MASS && !FILL;
  static_assert(LdsPoly<3, 3>::kPoints * 2 >= 9 + 3 + 2, "room for a lane's rows: values, targets, keys");
  double* svals = poly;                                       // [BLOCK][9]
  long* sdst = reinterpret_cast<long*>(poly + 9 * BLOCK);     // [BLOCK][3]: row slot of (pair, li), or -1
  int* skey = reinterpret_cast<int*>(poly + 12 * BLOCK);      // [BLOCK][3]: B node, or -1 (empty pair)
  const long stride = (long)gridDim.x * blockDim.x;
  // the same trip count for every lane of the block (barriers below)
  for (long base = (long)blockIdx.x * blockDim.x; base < np; base += stride) {
    const long p = base + threadIdx.x;
    double blk[9];
    bool keep = false;  // this lane's block is written from here
    bool rows = false;  // this lane's row slots are written from here
    long rdst[3] = {0, 0, 0};
    int rkey[3] = {0, 0, 0};
    do {
      if (p >= np) break;
      const long a = g.rows[p], b = g.cols[p];
      long off = 0;
      int cap = 0;
      if (FILL) {
        cap = g.counts[p];
        if (cap == 0) break;
        off = g.offsets[p];
      }
      double A[6], Bv[6];
#pragma unroll
      for (int k = 0; k < 6; k++) A[k] = g.acoords[a * 6 + k];
#pragma unroll
      for (int k = 0; k < 6; k++) Bv[k] = g.bcoords[b * 6 + k];
      const double tol = 10.0 * fmin_f(g.asp[a], g.bsp[b]);
      double m = 0.0;
      // mass: the polygon's moments about A's first vertex (the origin of
      // its barycentric map), turned into the block once the fan is done
      double mom[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, ox = 0.0, oy = 0.0;
      if (MASS) {
        ox = g.aaff[a * 6];
        oy = g.aaff[a * 6 + 1];
      }
      Line L[3];
      tri_lines(Bv, L);
      int k = 0;
      const int nk = clip_poly_lds<3, 3, BLOCK>(
          A, L, tol, q, [&](double x0, double y0, double x1, double y1, double x2, double y2, double ar) {
            m = m + ar;
            if (MASS) tri_moments(x0, y0, x1, y1, x2, y2, ar, ox, oy, mom);
            if (FILL && k < cap) {
              stream_tri(g.elements + (off + k) * 6, x0, y0, x1, y1, x2, y2);
            }
            k++;
          });
      if (nk < 0) {  // the streaming clipper writes this pair's results later
        if (!FILL) g.olist[atomicAdd(g.nolist, 1ULL)] = (int)p;
        break;
      }
      if (FILL) break;
      g.counts[p] = nk;
      g.measure[p] = nk == 0 ? 0.0 : m;
      if (MASS) {
        keep = nk > 0;
        if (keep) {
          double affa[6], affb[6];
#pragma unroll
          for (int i = 0; i < 6; i++) {
            affa[i] = g.aaff[a * 6 + i];
            affb[i] = g.baff[b * 6 + i];
          }
          mass_from_moments(mom, affa, affb, blk);
        }
        if (g.rkey) {
          rows = true;
          const long r = (p - g.indptr[a]) * 3;
#pragma unroll
          for (int li = 0; li < 3; li++) rdst[li] = g.rbase[a * 3 + li] + r;
#pragma unroll
          for (int j = 0; j < 3; j++) rkey[j] = nk == 0 ? -1 : g.enlb[b * 3 + j];
        }
      }
    } while (false);
    if (STAGE && g.rkey) {  // the same for the whole grid: the barriers match
      __syncthreads();  // every lane is done with its polygon slots
      const int t = threadIdx.x;
#pragma unroll
      for (int i = 0; i < 9; i++) svals[t * 9 + i] = keep ? blk[i] : 0.0;
#pragma unroll
      for (int li = 0; li < 3; li++) sdst[t * 3 + li] = rows ? rdst[li] : -1;
#pragma unroll
      for (int j = 0; j < 3; j++) skey[t * 3 + j] = rkey[j];
      __syncthreads();
      // row li of every pair: slot j of pair q at sdst[q][li] + j, pairs in
      // order, so the threads' writes run along each A element's row chunk
      const int nv = np - base < BLOCK ? (int)(np - base) : BLOCK;
#pragma unroll
      for (int li = 0; li < 3; li++) {
        for (int x = t; x < 3 * nv; x += BLOCK) {
          const int q = x / 3, j = x - 3 * q;
          const long d = sdst[q * 3 + li];
          if (d >= 0) {
            const int key = skey[q * 3 + j];
            g.rkey[d + j] = key;
            if (key >= 0) g.rval[d + j] = svals[q * 9 + li * 3 + j];
          }
        }
      }
      __syncthreads();  // the slots hold polygons again
    }
    if (STAGE && g.blocks) {  // the same for the whole grid: the barriers match
      __syncthreads();  // every lane is done with its polygon slots
#pragma unroll
      for (int i = 0; i < 9; i++) poly[threadIdx.x * 9 + i] = keep ? blk[i] : 0.0;
      __syncthreads();
      const long nv = np - base < BLOCK ? np - base : BLOCK;
      for (long t = threadIdx.x; t < 9 * nv; t += BLOCK) g.blocks[base * 9 + t] = poly[t];
      __syncthreads();  // the slots hold polygons again
    }
  }
}

// Tet pairs on the compacted work list, one pair per lane at a time.  Each
// lane walks its pair's clip tree with walk_next (every step builds one child
// and classifies or measures it) and, when its pair is done, moves on to its
// next work item by itself, so lanes never wait for the slowest pair of the
// wave -- only for the slowest chain of work items, which evens out.
// The walk of k_clip_tet_walk keeps its depth-1 tet in registers and its
// depth-2 and depth-3 tets in LDS: a parent there is one LDS read per
// coordinate and a child one write.  The stack is what limits how many waves
// share a CU: 192 bytes per lane let three waves per SIMD run (two with all
// three depths in LDS), at the cost of a register select per coordinate and
// step.  Lane t's depth-j tet (j = 2, 3), coordinate i, is
// stk[((j - 2) * 12 + i) * kWalkBlock] with stk = stack + t (consecutive
// lanes, consecutive words).
constexpr int kWalkBlock = 128;
constexpr int kWalkReg = 1;

// The CDNA4 dispatcher hands consecutive workgroups to its 8 XCDs (each with
// its own L2) in turn: with a grid whose size is a multiple of 8, block b
// runs on XCD b % 8, which takes the share [x * per, (x + 1) * per) of the
// work list.
constexpr int kXcds = 8;

// The walk's lanes take work items from queues (counters handing out up to
// kQueueBatch consecutive items at a time): each XCD's share of the work list
// is cut into kSubQueues contiguous parts, each with its own counter, so the
// items in flight on an XCD stay a few compact windows of the list whose A
// and B data its L2 holds (with fixed strides the lanes drift apart over the
// whole share), and no counter takes more than a fraction of the atomics.
// A short list is handed out in smaller batches (about a quarter of the
// items per lane, at least one): with 16 at a time, a list of a few
// thousand items went to a few hundred lanes that walked 16 pairs each in
// turn while the rest of the resident grid idled.
constexpr int kQueueBatch = 16;
constexpr int kSubQueues = 4;
constexpr int kQueues = kXcds * kSubQueues;
struct WalkQueue {
  unsigned long long* counter;
  long begin, end;  // this queue's part of the work list
  long w = 0, wend = 0;
  long batch = kQueueBatch;  // items per counter increment
  __device__ __forceinline__ bool next() {  // advance to the next item; false when none is left
    if (++w < wend) return true;
    const long s = begin + (long)atomicAdd(counter, (unsigned long long)batch);
    if (s >= end) return false;
    w = s;
    wend = s + batch < end ? s + batch : end;
    return true;
  }
};

__device__ __forceinline__ WalkQueue walk_queue(unsigned long long* queues, long n) {
  const int nx = gridDim.x % kXcds == 0 ? kXcds : 1;
  // every part needs blocks to drain it: a grid with fewer blocks per share
  // than parts keeps one part per share
  const int nsub = (int)(gridDim.x / nx) >= kSubQueues ? kSubQueues : 1;
  const long xcd = blockIdx.x % nx, sub = (blockIdx.x / nx) % nsub;
  const long per = (n + nx - 1) / nx;
  const long b0 = xcd * per, e0 = b0 + per < n ? b0 + per : n;
  const long sper = (e0 - b0 + nsub - 1) / nsub;
  WalkQueue q;
  q.counter = queues + xcd * kSubQueues + sub;
  q.begin = b0 + sub * sper < e0 ? b0 + sub * sper : e0;
  q.end = q.begin + sper < e0 ? q.begin + sper : e0;
  q.w = -1;
  q.wend = 0;
  const long per_lane = n / (4L * gridDim.x * blockDim.x);
  q.batch = per_lane < 1 ? 1 : (per_lane < kQueueBatch ? per_lane : kQueueBatch);
  return q;
}

__global__ void k_zero_queues(unsigned long long* __restrict__ q) {
  if (threadIdx.x < kQueues && blockIdx.x == 0) q[threadIdx.x] = 0;
}

// The walks' work queues, zeroed by a kernel before each launch.
static unsigned long long* walk_queues(Context& C) {
  unsigned long long* q = C.queue.as<unsigned long long>(kQueues);
  hipLaunchKernelGGL(k_zero_queues, dim3(1), dim3(64), 0, C.stream, q);
  return q;
}

// MODES: count (FILL and LEARNED false): counts, measures (and mass blocks);
// fill (FILL): writes each pair's counts[p] outputs at offsets[p];
// single pass (LEARNED): the count pass that also writes up to budget[w]
// outputs of work item w at offsets[p] and the rest to the lane's arena.
// BLOCK and REG (tet depths held in registers, 0 or 1; the other depths of
// 1..3 in LDS) are launch shapes the count pass can be tuned over.
template <bool MASS, bool FILL, bool LEARNED = false, int BLOCK = kWalkBlock, int REG = kWalkReg>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(MASS || REG == 0 ? 2 : 3)))
k_clip_tet_walk(ClipArgs g) {
  static_assert(REG == 0 || REG == 1, "depth 1 in registers or not");
  __shared__ double stack[(3 - REG) * 12 * BLOCK];
  double* stk = stack + threadIdx.x;
  double R[REG > 0 ? 12 * REG : 1];
  // blocks go to the XCDs round robin: each XCD takes one contiguous share
  // of the work list, so that its L2 holds one region of A and B
  WalkQueue Q = walk_queue(g.queue, (long)*g.nwork);
  bool more = Q.next();
  long p = 0, off = 0;
  int cap = 0, k = 0;
  // single pass: this lane's arena region, its fill level, and whether the
  // current pair's output outgrew budget and arena
  const long abase =
      LEARNED ? g.arena_base + ((long)blockIdx.x * blockDim.x + threadIdx.x) * g.arena_per_lane : 0;
  long atop = 0;
  bool lost = false;
  unsigned nsplit = 0, nlost = 0;  // this lane's split runs and pairs to clip again
  double A[12], tol = 0.0, m = 0.0;
  double blk[16], affa[12], affb[12];
  // B's four planes are read from memory at each step (the pair's 128 bytes
  // stay in cache): held in registers they would cost the third wave per
  // SIMD, and even planes 2 and 3 alone (16 registers more, no fewer waves)
  // made the count pass slower (30.2 against 28.1 ms on tet 64^3 vs 60^3)
  const double* bpl = nullptr;
  auto plane = [&](int j, double* o) {
#pragma unroll
    for (int i = 0; i < 4; i++) o[i] = bpl[4 * j + i];
  };
  WalkState<4> W;
  W.level = 0;
  bool active = false;
  while (true) {
    if (!active) {
      // next work item of this lane (fill pass: skip pairs with no output)
      while (more) {
        p = g.work[Q.w];
        if (!FILL) break;
        cap = g.counts[p];
        if (cap > 0 && (!g.ovf || g.ovf[Q.w] == 2)) break;
        more = Q.next();
      }
      if (!more) break;
      const long w = Q.w;
      if (FILL) off = g.ovf ? g.aoff[w] : g.offsets[p];
      if (LEARNED) {
        cap = (int)((g.budget[w] >> g.bshift) & 0xffu);
        off = g.offsets[p];
        lost = false;
      }
      const long a = g.rows[p], b = g.cols[p];
#pragma unroll
      for (int i = 0; i < 12; i++) A[i] = g.acoords[a * 12 + i];
      bpl = g.bplanes + b * 16;
      tol = 10.0 * fmin_f(g.asp[a], g.bsp[b]);
      m = 0.0;
      k = 0;
      if (MASS) {
#pragma unroll
        for (int i = 0; i < 16; i++) blk[i] = 0.0;
#pragma unroll
        for (int i = 0; i < 12; i++) {
          affa[i] = g.aaff[a * 12 + i];
          affb[i] = g.baff[b * 12 + i];
        }
      }
      double p0[4], d0[4];
      plane(0, p0);
      walk_distances(A, p0, d0);
      walk_begin(W, pack_class(d0));
      active = true;
    }
    if (W.level > 0) {
      walk_next<4, BLOCK, REG>(W, A, R, stk, plane, tol, [&](const double* t, double, bool keep) {
        // every step hands over its child; only kept leaves count, and the
        // measure adds 0 for the others (exactly the sum of the kept ones)
        const double vr = tet_volume_real(t);
        m = m + (keep ? vr : 0.0);
        if (keep) {
          if (MASS) mass_element<3>(t, vr, affa, affb, blk);
          if ((FILL || LEARNED) && k < cap) {
            stream_tet(g.elements + (off + k) * 12, t);
          } else if (LEARNED) {
            // beyond the budget: the run continues in this lane's arena
            const long e = atop + (k - cap);
            if (e < g.arena_per_lane) {
              stream_tet(g.elements + (abase + e) * 12, t);
            } else {
              lost = true;
            }
          }
        }
        k += keep ? 1 : 0;
      });
      if (W.level > 0) continue;
    }
    if (LEARNED) {
      const bool split = k > cap && !lost;
      g.aoff[Q.w] = split ? abase + atop : -1;
      if (split) atop += k - cap;
      g.dropped[Q.w] = lost ? 2 : (split ? 1 : 0);
      nsplit += split ? 1u : 0u;
      nlost += lost ? 1u : 0u;
    }
    if (!FILL) {
      __builtin_nontemporal_store(k, g.counts + p);
      __builtin_nontemporal_store(k == 0 ? 0.0 : m, g.measure + p);
      if (MASS) {
#pragma unroll
        for (int i = 0; i < 16; i++) g.blocks[p * 16 + i] = k == 0 ? 0.0 : blk[i];
      }
    }
    more = Q.next();
    active = false;
  }
  if (LEARNED) {
    // one atomic per wave: one per pair on a single counter serialised the
    // single pass (tet 96^3 at budget level 0.9: 7.4 M of them, ~20 ms)
    unsigned long long s = nsplit, l = nlost;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      s += __shfl_xor(s, o, 64);
      l += __shfl_xor(l, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      if (s) atomicAdd(g.ndropped, s);
      if (l) atomicAdd(g.ndropped + 1, l);
    }
  }
}

// Totals over pairs in pair order: fixed grid, fixed tree, so the sum is
// reproducible for a given input.
__global__ void __launch_bounds__(kClipBlock)
k_reduce_pairs(long npairs, const long* __restrict__ npairs_dev, const int* __restrict__ counts,
               const double* __restrict__ measure, Partial* __restrict__ partials) {
  if (npairs_dev) npairs = *npairs_dev;
  double hi = 0.0, lo = 0.0;
  long long nonempty = 0, nout = 0, maxc = 0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += stride) {
    int c = counts[p];
    dd_add(hi, lo, measure[p]);
    nonempty += c > 0;
    nout += c;
    maxc = c > maxc ? c : maxc;
  }
  __shared__ Partial red[kClipBlock];
  red[threadIdx.x].hi = hi;
  red[threadIdx.x].lo = lo;
  red[threadIdx.x].nonempty = nonempty;
  red[threadIdx.x].nout = nout;
  red[threadIdx.x].maxc = maxc;
  __syncthreads();
  for (int s = kClipBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      Partial& x = red[threadIdx.x];
      const Partial& y = red[threadIdx.x + s];
      dd_merge(x.hi, x.lo, y.hi, y.lo);
      x.nonempty += y.nonempty;
      x.nout += y.nout;
      x.maxc = y.maxc > x.maxc ? y.maxc : x.maxc;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kClipBlock) k_reduce_partials(int n, Partial* __restrict__ parts) {
  __shared__ Partial red[kClipBlock];
  Partial t = {0.0, 0.0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < n; i += kClipBlock) {
    dd_merge(t.hi, t.lo, parts[i].hi, parts[i].lo);
    t.nonempty += parts[i].nonempty;
    t.nout += parts[i].nout;
    t.maxc = parts[i].maxc > t.maxc ? parts[i].maxc : t.maxc;
  }
  red[threadIdx.x] = t;
  __syncthreads();
  for (int s = kClipBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      Partial& x = red[threadIdx.x];
      const Partial& y = red[threadIdx.x + s];
      dd_merge(x.hi, x.lo, y.hi, y.lo);
      x.nonempty += y.nonempty;
      x.nout += y.nout;
      x.maxc = y.maxc > x.maxc ? y.maxc : x.maxc;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) parts[0] = red[0];
}

// ---------------------------------------------------------------------------
// Mixed element types in bulk: the dispatch of intersect_elements
// (Supermesh.F90:162-209) for one pair per lane.  2D: tri x quad clips the
// triangle by the quad's lines (tol from the triangle), quad x tri clips B's
// triangle by A's lines, quad x quad clips A's quad by B's lines (tol 10*tiny);
// 3D: tet x hex clips the tet by the hex's 6 planes, hex x tet the reverse,
// hex x hex clips A's five tets (Supermesh.F90:303-334) in turn.  Counts,
// measures (triangle_area / tetrahedron_volume_real of each output) and, in
// the fill pass, the output elements.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256)
k_count_to_long(long n, const int* __restrict__ counts, long* __restrict__ out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = counts[i];
}

struct MixedArgs {
  long npairs;
  const int* rows;
  const int* cols;
  const double* acoords;
  const double* bcoords;
  int* counts;
  double* measure;
  const long* offsets;
  double* elements;
  // pairs whose polygon outgrew its LDS slots (STREAM pass: the list to do)
  int* olist;
  unsigned long long* nolist;
};

// STREAM false: every pair on clip_poly_lds, handing overflows to olist;
// STREAM true: the pairs of olist on the streaming clip_poly.
template <int LA, int LB, bool FILL, bool STREAM>
__global__ void __launch_bounds__(kTriBlock) k_clip_mixed(MixedArgs g) {
  constexpr int NA = LA == 4 && LB == 4 ? 4 : 3;
  __shared__ double poly[STREAM ? 1 : LdsPoly<NA, 4>::kPoints * 2 * kTriBlock];
  double* q = poly + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  const long total = STREAM ? (long)*g.nolist : g.npairs;
  for (long w = (long)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += stride) {
    const long p = STREAM ? g.olist[w] : w;
    long off = 0;
    int cap = 0;
    if (FILL) {
      cap = g.counts[p];
      if (cap == 0) continue;
      off = g.offsets[p];
    }
    const long a = g.rows[p], b = g.cols[p];
    double A[LA * 2], B[LB * 2];
#pragma unroll
    for (int k = 0; k < LA * 2; k++) A[k] = g.acoords[a * LA * 2 + k];
#pragma unroll
    for (int k = 0; k < LB * 2; k++) B[k] = g.bcoords[b * LB * 2 + k];
    double m = 0.0;
    int k = 0;
    auto sink = [&](double x0, double y0, double x1, double y1, double x2, double y2, double ar) {
      m = m + ar;
      if (FILL && k < cap) {
        stream_tri(g.elements + (off + k) * 6, x0, y0, x1, y1, x2, y2);
      }
      k++;
    };
    Line L[4];
    const double* poly_in = A;
    double tol = 10.0 * 2.2250738585072014e-308;
    if constexpr (LA == 3 && LB == 4) {
      quad_lines(B, L);
      tol = 10.0 * fspacing(triangle_area(A));
    } else if constexpr (LA == 4 && LB == 3) {
      quad_lines(A, L);
      poly_in = B;
      tol = 10.0 * fspacing(triangle_area(B));
    } else {
      quad_lines(B, L);
    }
    int n;
    if constexpr (STREAM) n = clip_poly<4, NA>(poly_in, L, tol, sink);
    else n = clip_poly_lds<4, NA, kTriBlock>(poly_in, L, tol, q, sink);
    if (n < 0) {
      if (!FILL) g.olist[atomicAdd(g.nolist, 1ULL)] = (int)p;
      continue;
    }
    if (FILL) continue;
    // a pass left with fewer than 3 points: no output, whatever was fanned
    g.counts[p] = n == 0 ? 0 : k;
    g.measure[p] = n == 0 ? 0.0 : m;
  }
}

// 3D: one pair per lane at a time over the six planes of the hexahedron
// (walk_next<6>, depths 1..2 in registers, 3..5 in LDS); a lane moves on to its next pair
// as soon as its walk ends.  MODE 0: tet A x hex B; 1: hex A x tet B (the
// tet is B's); 2: hex A x hex B, A's five tets in turn, outputs and measure
// accumulated over them in order.  hplanes: planes of the hexahedra,
// [n][6][4] (hex_planes).
struct HexWalkArgs {
  long npairs;
  const int* rows;
  const int* cols;
  const double* acoords;
  const double* bcoords;
  const double* hplanes;
  int* counts;
  double* measure;
  const long* offsets;
  double* elements;
  unsigned long long* queue;  // the walk's work-item counters (zeroed before each launch)
};

// One wave per block; the walk keeps its depth-1 and depth-2 tets in
// registers and depths 3..5 in LDS (18 KB per block, eight blocks per CU).
constexpr int kHexBlock = 64;
constexpr int kHexReg = 2;

template <int MODE, bool FILL>
__global__ void __launch_bounds__(kHexBlock) __attribute__((amdgpu_waves_per_eu(2)))
k_clip_hex_walk(HexWalkArgs g) {
  __shared__ double stack[(5 - kHexReg) * 12 * kHexBlock];
  double* stk = stack + threadIdx.x;
  double R[12 * kHexReg];
  // pairs from the walk's queues, as k_clip_tet_walk takes its work items
  WalkQueue Q = walk_queue(g.queue, g.npairs);
  bool more = Q.next();
  long p = 0, off = 0, a = 0;
  int cap = 0, k = 0, sub = 0;
  double T[12], tol = 0.0, m = 0.0;
  const double* planes = nullptr;
  auto plane = [&](int j, double* o) {
#pragma unroll
    for (int i = 0; i < 4; i++) o[i] = planes[4 * j + i];
  };
  WalkState<6> W;
  W.level = 0;
  bool active = false;
  auto start = [&]() {
    tol = 10.0 * fspacing(tet_volume_tet(T));
    double d0[4];
    walk_distances(T, planes, d0);
    walk_begin(W, pack_class(d0));
  };
  while (true) {
    if (!active) {
      while (more) {
        p = Q.w;
        if (!FILL) break;
        cap = g.counts[p];
        if (cap > 0) break;
        more = Q.next();
      }
      if (!more) break;
      if (FILL) off = g.offsets[p];
      a = g.rows[p];
      const long b = g.cols[p];
      if (MODE == 0) {
#pragma unroll
        for (int i = 0; i < 12; i++) T[i] = g.acoords[a * 12 + i];
        planes = g.hplanes + b * 24;
      } else if (MODE == 1) {
#pragma unroll
        for (int i = 0; i < 12; i++) T[i] = g.bcoords[b * 12 + i];
        planes = g.hplanes + a * 24;
      } else {
        sub = 0;
        hex_split_tet(g.acoords + a * 24, 0, T);
        planes = g.hplanes + b * 24;
      }
      m = 0.0;
      k = 0;
      start();
      active = true;
    }
    if (W.level > 0) {
      walk_next<6, kHexBlock, kHexReg>(W, T, R, stk, plane, tol, [&](const double* t, double, bool keep) {
        const double vr = tet_volume_real(t);
        m = m + (keep ? vr : 0.0);
        if (FILL && keep && k < cap) {
          stream_tet(g.elements + (off + k) * 12, t);
        }
        k += keep ? 1 : 0;
      });
      if (W.level > 0) continue;
    }
    if (MODE == 2 && sub < 4) {
      sub++;
      hex_split_tet(g.acoords + a * 24, sub, T);
      start();
      continue;
    }
    if (!FILL) {
      __builtin_nontemporal_store(k, g.counts + p);
      __builtin_nontemporal_store(k == 0 ? 0.0 : m, g.measure + p);
    }
    more = Q.next();
    active = false;
  }
}

__global__ void __launch_bounds__(256) k_hex_planes(long n, const double* __restrict__ coords,
                                                    double* __restrict__ planes) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double h[24];
#pragma unroll
  for (int i = 0; i < 24; i++) h[i] = coords[e * 24 + i];
  double pl[24];
  hex_planes(h, pl);
#pragma unroll
  for (int i = 0; i < 24; i++) planes[e * 24 + i] = pl[i];
}

template <bool FILL, bool STREAM>
static void launch_mixed_as(Context& C, int la, int lb, const MixedArgs& g) {
  const dim3 grid(STREAM ? 64 : one_round_grid(g.npairs)), block(kTriBlock);
  if (la == 3 && lb == 4) hipLaunchKernelGGL((k_clip_mixed<3, 4, FILL, STREAM>), grid, block, 0, C.stream, g);
  else if (la == 4 && lb == 3) hipLaunchKernelGGL((k_clip_mixed<4, 3, FILL, STREAM>), grid, block, 0, C.stream, g);
  else hipLaunchKernelGGL((k_clip_mixed<4, 4, FILL, STREAM>), grid, block, 0, C.stream, g);
  LSM_HIP(hipGetLastError());
}

// LDS clipper over all pairs, then the streaming one over the pairs it handed
// back (counted on the device; usually none).
template <bool FILL>
static void launch_mixed(Context& C, int la, int lb, MixedArgs g) {
  g.olist = C.work.as<int>((size_t)(g.npairs > 0 ? g.npairs : 1));
  g.nolist = C.scalars.as<unsigned long long>(16) + 2;
  if (!FILL) LSM_HIP(hipMemsetAsync(g.nolist, 0, sizeof(unsigned long long), C.stream));
  launch_mixed_as<FILL, false>(C, la, lb, g);
  launch_mixed_as<FILL, true>(C, la, lb, g);
}

// Persistent lanes for the hexahedron walk: one full wave of resident blocks.
template <int MODE, bool FILL>
static void launch_hex_walk_as(Context& C, const HexWalkArgs& g) {
  static int blocks_per_cu = 0;
  if (blocks_per_cu == 0) {
    int nb = 0;
    LSM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, reinterpret_cast<const void*>(k_clip_hex_walk<MODE, FILL>), kHexBlock, 0));
    blocks_per_cu = nb > 0 ? nb : 1;
  }
  int cus = 0;
  LSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, C.device));
  const long lanes = (long)blocks_per_cu * (cus > 0 ? cus : 256) * kHexBlock;
  const long need = (g.npairs + kHexBlock - 1) / kHexBlock;
  unsigned blocks = (unsigned)std::max(1L, std::min(lanes / kHexBlock, need));
  if (blocks >= kXcds) blocks -= blocks % kXcds;  // whole XCD shares (walk_queue)
  HexWalkArgs h = g;
  h.queue = walk_queues(C);
  hipLaunchKernelGGL((k_clip_hex_walk<MODE, FILL>), dim3(blocks), dim3(kHexBlock), 0, C.stream, h);
  LSM_HIP(hipGetLastError());
}

template <bool FILL>
static void launch_hex_walk(Context& C, int la, int lb, const HexWalkArgs& g) {
  if (la == 4 && lb == 8) launch_hex_walk_as<0, FILL>(C, g);
  else if (la == 8 && lb == 4) launch_hex_walk_as<1, FILL>(C, g);
  else launch_hex_walk_as<2, FILL>(C, g);
}

void clip_pairs_mixed(Context& C, const DevMesh& A, const DevMesh& B, long npairs, int flags,
                      libsupermesh_gpu_result* out) {
  LSM_REQUIRE(npairs < (1L << 31), LIBSUPERMESH_GPU_ERROR_OVERFLOW,
              "more than 2^31 - 1 candidate pairs to clip in one call");
  const int dim = A.dim;
  MixedArgs g = {};
  g.npairs = npairs;
  g.rows = out->rows;
  g.cols = out->indices;
  g.acoords = static_cast<const double*>(A.coords.ptr);
  g.bcoords = static_cast<const double*>(B.coords.ptr);
  g.counts = C.counts.as<int>(npairs);
  g.measure = C.measure.as<double>(npairs);
  HexWalkArgs h = {};
  if (dim == 3) {
    // planes of the hexahedra (B's, or A's when only A is one)
    DevMesh& H = const_cast<DevMesh&>(A.loc == 8 && B.loc == 4 ? A : B);
    h.npairs = npairs;
    h.rows = g.rows;
    h.cols = g.cols;
    h.acoords = g.acoords;
    h.bcoords = g.bcoords;
    h.counts = g.counts;
    h.measure = g.measure;
    double* hp = H.planes.as<double>((size_t)(H.nelements > 0 ? H.nelements : 1) * 24);
    h.hplanes = hp;
    if (H.nelements > 0)
      hipLaunchKernelGGL(k_hex_planes, dim3(grid_for(H.nelements, 256)), dim3(256), 0, C.stream, H.nelements,
                         static_cast<const double*>(H.coords.ptr), hp);
    LSM_HIP(hipGetLastError());
  }
  if (npairs > 0) {
    if (dim == 2) launch_mixed<false>(C, A.loc, B.loc, g);
    else launch_hex_walk<false>(C, A.loc, B.loc, h);
  }
  Partial* parts = C.partials.as<Partial>(kClipGrid);
  hipLaunchKernelGGL(k_reduce_pairs, dim3(kClipGrid), dim3(kClipBlock), 0, C.stream, npairs, nullptr, g.counts, g.measure,
                     parts);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kClipBlock), 0, C.stream, kClipGrid, parts);
  LSM_HIP(hipGetLastError());
  Partial tot;
  LSM_HIP(hipMemcpyAsync(&tot, parts, sizeof(Partial), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  out->counts = g.counts;
  out->measure = g.measure;
  out->blocks = nullptr;
  out->sum = npairs > 0 ? tot.hi + tot.lo : 0.0;
  out->nonempty = npairs > 0 ? tot.nonempty : 0;
  out->nout = npairs > 0 ? tot.nout : 0;
  out->max_count = npairs > 0 ? (int)tot.maxc : 0;
  out->offsets = nullptr;
  out->elements = nullptr;
  if (flags & LIBSUPERMESH_GPU_ELEMENTS) {
    const int esz = (dim + 1) * dim;
    long* offs = C.offsets.as<long>(npairs + 1);
    LSM_HIP(hipMemsetAsync(offs, 0, sizeof(long), C.stream));
    if (npairs > 0) {
      long* wide = C.row_count.as<long>(npairs);
      hipLaunchKernelGGL(k_count_to_long, dim3(grid_for(npairs, 256)), dim3(256), 0, C.stream, npairs, g.counts,
                         wide);
      size_t bytes = 0;
      LSM_HIP(rocprim::inclusive_scan(nullptr, bytes, wide, offs + 1, (size_t)npairs, rocprim::plus<long>(),
                                      C.stream));
      void* tmp = C.scan_tmp.get(bytes);
      LSM_HIP(rocprim::inclusive_scan(tmp, bytes, wide, offs + 1, (size_t)npairs, rocprim::plus<long>(),
                                      C.stream));
      g.offsets = offs;
      g.elements = C.elements.as<double>((size_t)(out->nout > 0 ? out->nout : 1) * esz);
      if (dim == 2) {
        launch_mixed<true>(C, A.loc, B.loc, g);
      } else {
        h.offsets = offs;
        h.elements = g.elements;
        launch_hex_walk<true>(C, A.loc, B.loc, h);
      }
    }
    LSM_HIP(hipStreamSynchronize(C.stream));
    out->offsets = offs;
    out->elements = static_cast<const double*>(C.elements.get(8));
    out->elements_size = out->nout;
  }
}

// Batches: candidate pair range of each mesh (its A elements are the rows
// a_off[g] .. a_off[g+1]) ...
__global__ void __launch_bounds__(kClipBlock)
k_batch_pairs(long ngroups, const long* __restrict__ a_off, const long* __restrict__ indptr,
              long* __restrict__ gpairs) {
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g <= ngroups) gpairs[g] = indptr[a_off[g]];
}

// ... and its measure sum: one block per mesh, double-double in a fixed order.
__global__ void __launch_bounds__(kClipBlock)
k_batch_sum(const long* __restrict__ gpairs, const double* __restrict__ measure,
            double* __restrict__ gsum) {
  const long g = blockIdx.x;
  double hi = 0.0, lo = 0.0;
  for (long p = gpairs[g] + threadIdx.x; p < gpairs[g + 1]; p += kClipBlock) dd_add(hi, lo, measure[p]);
  __shared__ double rh[kClipBlock], rl[kClipBlock];
  rh[threadIdx.x] = hi;
  rl[threadIdx.x] = lo;
  __syncthreads();
  for (int s = kClipBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double h = rh[threadIdx.x], l = rl[threadIdx.x];
      dd_merge(h, l, rh[threadIdx.x + s], rl[threadIdx.x + s]);
      rh[threadIdx.x] = h;
      rl[threadIdx.x] = l;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) gsum[g] = rh[0] + rl[0];
}

// Order-sensitive 64-bit checksum of a device buffer: sum over its 32-bit
// words w_i of w_i * (2 i + 1), modulo 2^64 (wrapping adds commute, so the
// value does not depend on the reduction order).
__global__ void __launch_bounds__(256)
k_checksum(long nwords, const unsigned* __restrict__ words, unsigned long long* __restrict__ out) {
  unsigned long long h = 0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride)
    h += (unsigned long long)words[i] * (unsigned long long)(2 * i + 1);
  __shared__ unsigned long long red[256];
  red[threadIdx.x] = h;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

unsigned long long checksum(Context& C, const void* ptr, size_t bytes) {
  unsigned long long* d = C.scalars.as<unsigned long long>(8) + 2;
  LSM_HIP(hipMemsetAsync(d, 0, sizeof(*d), C.stream));
  const long n = (long)(bytes / 4);
  if (n > 0)
    hipLaunchKernelGGL(k_checksum, dim3((unsigned)std::min<long>(4096, (n + 255) / 256)), dim3(256), 0,
                       C.stream, n, static_cast<const unsigned*>(ptr), d);
  LSM_HIP(hipGetLastError());
  unsigned long long h = 0;
  LSM_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  return h;
}

long last_work_list(Context& C, const int** work) {
  long n = 0;
  const long* bcount = static_cast<const long*>(C.nwork.ptr);
  LSM_HIP(hipMemcpyAsync(&n, bcount + 2 * kTriageGrid + 1, sizeof(long), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  *work = static_cast<const int*>(C.work.ptr);
  return n;
}

void batch_sums(Context& C, long ngroups, const long* a_offsets, const long* indptr,
                const double* measure) {
  long* gp = C.gpairs.as<long>((size_t)ngroups + 1);
  double* gs = C.gsum.as<double>((size_t)ngroups);
  hipLaunchKernelGGL(k_batch_pairs, dim3(grid_for(ngroups + 1, kClipBlock)), dim3(kClipBlock), 0,
                     C.stream, ngroups, a_offsets, indptr, gp);
  if (ngroups > 0)
    hipLaunchKernelGGL(k_batch_sum, dim3((unsigned)ngroups), dim3(kClipBlock), 0, C.stream, gp,
                       measure, gs);
  LSM_HIP(hipGetLastError());
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

template <int DIM>
static void launch_clip(Context& C, const ClipArgs& args, bool mass, bool fill) {
  dim3 grid(kClipGrid), block(kClipBlock);
  if (DIM == 2 && args.work == nullptr) {
    // LDS clipper over all pairs, then the streaming clipper over the pairs
    // it handed back (a device-side count: usually none, no host sync)
    ClipArgs g = args;
    g.olist = C.work.as<int>((size_t)(args.npairs > 0 ? args.npairs : 1));
    g.nolist = C.scalars.as<unsigned long long>(16) + 2;
    if (!fill) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, C.stream, g.nolist);
    const dim3 tgrid(one_round_grid(args.npairs)), tblock(kTriBlock);
    if (fill) hipLaunchKernelGGL((k_clip_tri<false, true>), tgrid, tblock, 0, C.stream, g);
    else if (mass) hipLaunchKernelGGL((k_clip_tri<true, false>), tgrid, tblock, 0, C.stream, g);
    else hipLaunchKernelGGL((k_clip_tri<false, false>), tgrid, tblock, 0, C.stream, g);
    LSM_HIP(hipGetLastError());
    g.work = g.olist;
    g.nwork = g.nolist;
    const dim3 fgrid(64);
    if (fill) hipLaunchKernelGGL((k_clip<DIM, false, true>), fgrid, block, 0, C.stream, g);
    else if (mass) hipLaunchKernelGGL((k_clip<DIM, true, false>), fgrid, block, 0, C.stream, g);
    else hipLaunchKernelGGL((k_clip<DIM, false, false>), fgrid, block, 0, C.stream, g);
    LSM_HIP(hipGetLastError());
    return;
  }
  if (fill) {
    hipLaunchKernelGGL((k_clip<DIM, false, true>), grid, block, 0, C.stream, args);
  } else if (mass) {
    hipLaunchKernelGGL((k_clip<DIM, true, false>), grid, block, 0, C.stream, args);
  } else {
    hipLaunchKernelGGL((k_clip<DIM, false, false>), grid, block, 0, C.stream, args);
  }
  LSM_HIP(hipGetLastError());
}

// which: 0 count, 1 count + mass, 2 fill, 3 single pass, 4 single pass + mass.
// Persistent lanes: one full wave of resident blocks.
static unsigned tet_walk_blocks(Context& C, int which) {
  static int blocks_per_cu[5] = {0, 0, 0, 0, 0};
  static const void* kernels[5] = {
      reinterpret_cast<const void*>(k_clip_tet_walk<false, false>),
      reinterpret_cast<const void*>(k_clip_tet_walk<true, false>),
      reinterpret_cast<const void*>(k_clip_tet_walk<false, true>),
      reinterpret_cast<const void*>(k_clip_tet_walk<false, false, true>),
      reinterpret_cast<const void*>(k_clip_tet_walk<true, false, true>)};
  if (blocks_per_cu[which] == 0) {
    int nb = 0;
    LSM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernels[which], kWalkBlock, 0));
    blocks_per_cu[which] = nb > 0 ? nb : 1;
  }
  int cus = 0;
  LSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, C.device));
  return (unsigned)(blocks_per_cu[which] * (cus > 0 ? cus : 256));
}

static long tet_walk_lanes(Context& C, int which) { return (long)tet_walk_blocks(C, which) * kWalkBlock; }


static void launch_tet_walk(Context& C, const ClipArgs& a, int which) {
  ClipArgs args = a;
  args.queue = walk_queues(C);
  dim3 grid(tet_walk_blocks(C, which)), block(kWalkBlock);
  switch (which) {
    case 0: hipLaunchKernelGGL((k_clip_tet_walk<false, false>), grid, block, 0, C.stream, args); break;
    case 1: hipLaunchKernelGGL((k_clip_tet_walk<true, false>), grid, block, 0, C.stream, args); break;
    case 2: hipLaunchKernelGGL((k_clip_tet_walk<false, true>), grid, block, 0, C.stream, args); break;
    case 3: hipLaunchKernelGGL((k_clip_tet_walk<false, false, true>), grid, block, 0, C.stream, args); break;
    default: hipLaunchKernelGGL((k_clip_tet_walk<true, false, true>), grid, block, 0, C.stream, args); break;
  }
  LSM_HIP(hipGetLastError());
}

static void launch_tet_walk(Context& C, const ClipArgs& args, bool mass, bool fill) {
  launch_tet_walk(C, args, fill ? 2 : (mass ? 1 : 0));
}

__global__ void __launch_bounds__(256) k_fill_long(long n, long v, long* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}

// Single pass, work items flagged `from` or above in dropped (1: split runs
// and dropped pairs, 2: dropped pairs only): their output lengths (0 for the
// others), scanned to place their runs in the spill reserve; then the
// dropped ones' runs placed from `base` (aoff[w], and offsets[p] when
// set_offsets).
struct DroppedLength {
  const int* work;
  const unsigned char* dropped;
  unsigned char from;
  const int* counts;
  __host__ __device__ long operator()(long w) const { return dropped[w] >= from ? (long)counts[work[w]] : 0L; }
};

__global__ void __launch_bounds__(256)
k_dropped_offsets(long nwork, const int* __restrict__ work, const unsigned char* __restrict__ dropped,
                  const long* __restrict__ incl, const int* __restrict__ counts, long base,
                  long* __restrict__ aoff, long* __restrict__ offsets, bool set_offsets) {
  const long w = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nwork && dropped[w] == 2) {
    const long p = work[w];
    const long o = base + incl[w] - counts[p];
    aoff[w] = o;
    if (set_offsets) offsets[p] = o;
  }
}

// Single pass, split runs: the budget's part and the arena's part of each
// joined end to end at base + incl[w] - counts[p], which becomes the pair's
// offset.  A wave takes 64 consecutive work items and copies the split ones
// one after another, its lanes on consecutive 16-byte words.
__global__ void __launch_bounds__(kClipBlock)
k_join_runs(long nwork, const int* __restrict__ work, const unsigned char* __restrict__ dropped,
            const int* __restrict__ counts, const unsigned* __restrict__ budget, int shift,
            const long* __restrict__ aoff, const long* __restrict__ incl, long base, long* __restrict__ offsets,
            double* __restrict__ el) {
  const int lane = threadIdx.x & 63;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long nwaves = ((long)gridDim.x * blockDim.x) >> 6;
  for (long b0 = wave * 64; b0 < nwork; b0 += nwaves * 64) {
    const long w = b0 + lane;
    long p = 0, s1 = 0, s2 = 0, dst = 0;
    int n1 = 0, n2 = 0;
    if (w < nwork && dropped[w] == 1) {
      p = work[w];
      const int c = counts[p];
      n1 = (int)((budget[w] >> shift) & 0xffu);
      n2 = c - n1;
      s1 = offsets[p];
      s2 = aoff[w];
      dst = base + incl[w] - c;
    }
    for (int k = 0; k < 64; k++) {
      const int m1 = __shfl(n1, k, 64), m2 = __shfl(n2, k, 64);
      if (m1 + m2 == 0) continue;
      double2* d2 = reinterpret_cast<double2*>(el + __shfl(dst, k, 64) * 12);
      const double2* a2 = reinterpret_cast<const double2*>(el + __shfl(s1, k, 64) * 12);
      const double2* b2 = reinterpret_cast<const double2*>(el + __shfl(s2, k, 64) * 12);
      for (int i = lane; i < 6 * m1; i += 64) stream_word(d2 + i, a2 + i);
      for (int i = lane; i < 6 * m2; i += 64) stream_word(d2 + 6 * m1 + i, b2 + i);
    }
    if (n1 + n2 > 0) offsets[p] = dst;
  }
}

// Dense layout of a single pass's output: pair p's run (counts[p] elements
// of EW doubles at src_off[p]) to dst_off[p], the prefix sums of the counts.
// A wave takes 64 consecutive pairs and copies their runs one after
// another, its lanes on consecutive 16-byte words.
template <int EW>
__global__ void __launch_bounds__(kClipBlock)
k_compact_runs(long npairs, const int* __restrict__ counts, const long* __restrict__ src_off,
               const double* __restrict__ src, const long* __restrict__ dst_off, double* __restrict__ dst) {
  static_assert(EW % 2 == 0, "whole 16-byte words per element");
  const int lane = threadIdx.x & 63;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long nwaves = ((long)gridDim.x * blockDim.x) >> 6;
  for (long base = wave * 64; base < npairs; base += nwaves * 64) {
    const long p = base + lane;
    long so = 0, dof = 0;
    int c = 0;
    if (p < npairs) {
      c = counts[p];
      if (c > 0) {
        so = src_off[p];
        dof = dst_off[p];
      }
    }
    for (int k = 0; k < 64; k++) {
      const int n = __shfl(c, k, 64);
      if (n == 0) continue;
      const double2* s2 = reinterpret_cast<const double2*>(src + __shfl(so, k, 64) * EW);
      double2* d2 = reinterpret_cast<double2*>(dst + __shfl(dof, k, 64) * EW);
      for (int i = lane; i < (EW / 2) * n; i += 64) stream_word(d2 + i, s2 + i);
    }
  }
}

void compact_runs(Context& C, int dim, long npairs, const int* counts, const long* src_off, const double* src,
                  const long* dst_off, double* dst) {
  if (npairs <= 0) return;
  if (dim == 2)
    hipLaunchKernelGGL(k_compact_runs<6>, dim3(kClipGrid), dim3(kClipBlock), 0, C.stream, npairs, counts, src_off,
                       src, dst_off, dst);
  else
    hipLaunchKernelGGL(k_compact_runs<12>, dim3(kClipGrid), dim3(kClipBlock), 0, C.stream, npairs, counts,
                       src_off, src, dst_off, dst);
  LSM_HIP(hipGetLastError());
}

// The dense layout straight from a single pass (LIBSUPERMESH_GPU_DENSE),
// moved inside its own buffer.  Pair p's run starts at roff[p] of `src` (the
// budgets laid end to end, roff[npairs] = the slots; src lies `gap` elements
// into the buffer) and goes to doff[p] of `dst` (the buffer's start; doff =
// the prefix sums of the counts).  When doff[p] <= gap + roff[p] for every p
// in [0, npairs], no run lands beyond its own start, and the runs move in
// steps from the first pair p0 not yet moved: every run up to p1, the last
// pair with doff[p1] <= gap + roff[p0], lands below every run not yet read
// (runs lie in pair order, the arena and spill reserve above them all), so
// those runs move at once; when not even p0 + 1 qualifies, p0's run moves
// alone, one wave copying it forward (a word is written only after the
// source words below it have been read).  A run that outgrew its budget
// moves its budget's part here and the rest afterwards (k_dense_overflow).

// bad = 1 when some p has doff[p] > gap + roff[p].
__global__ void __launch_bounds__(256)
k_dense_fits(long n, const long* __restrict__ roff, long gap, const long* __restrict__ doff,
             unsigned long long* __restrict__ bad) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n && doff[p] > gap + roff[p]) atomicOr(bad, 1ULL);
}

// One step (cursor[step & 1] = p0; block 0 leaves p1 in cursor[(step+1) & 1]),
// or every pair at once into a separate buffer (cursor null).  src and dst
// may share memory: no __restrict__ on them.
template <int EW>
__global__ void __launch_bounds__(kClipBlock)
k_dense_step(long npairs, const int* __restrict__ counts, const long* __restrict__ roff, long gap,
             const long* __restrict__ doff, long* __restrict__ cursor, int step, const double* src, double* dst) {
  static_assert(EW % 2 == 0, "whole 16-byte words per element");
  __shared__ long range[2];
  if (threadIdx.x == 0) {
    long p0 = 0, p1 = npairs;
    if (cursor) {
      p0 = cursor[step & 1];
      p1 = p0;
      if (p0 < npairs) {
        const long bound = gap + roff[p0];
        p1 = p0 + 1;
        if (doff[p1] <= bound) {
          long hi = npairs;  // the last q in [p0 + 1, npairs] with doff[q] <= bound
          while (p1 < hi) {
            const long mid = (p1 + hi + 1) >> 1;
            if (doff[mid] <= bound) p1 = mid;
            else hi = mid - 1;
          }
        }
      }
      if (blockIdx.x == 0) cursor[(step + 1) & 1] = p1;
    }
    range[0] = p0;
    range[1] = p1;
  }
  __syncthreads();
  const long p0 = range[0], p1 = range[1];
  const int lane = threadIdx.x & 63;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long nwaves = ((long)gridDim.x * blockDim.x) >> 6;
  for (long b0 = p0 + wave * 64; b0 < p1; b0 += nwaves * 64) {
    const long p = b0 + lane;
    long so = 0, dof = 0;
    int n = 0;
    if (p < p1) {
      const int c = counts[p];
      if (c > 0) {
        so = roff[p];
        const long b = roff[p + 1] - so;
        n = c < b ? c : (int)b;
        dof = doff[p];
      }
    }
    for (int k = 0; k < 64; k++) {
      const int m = __shfl(n, k, 64);
      if (m == 0) continue;
      const double2* s2 = reinterpret_cast<const double2*>(src + __shfl(so, k, 64) * EW);
      double2* d2 = reinterpret_cast<double2*>(dst + __shfl(dof, k, 64) * EW);
      // plain stores: streaming ones made the in-place steps slower (33.1 -> 34.9 ms at tet 96^3)
      for (int i = lane; i < (EW / 2) * m; i += 64) d2[i] = s2[i];
    }
  }
}

// Then the runs that outgrew their budget: a split run (dropped 1) its arena
// part, behind its budget's part; a pair clipped again (dropped 2) its whole
// run from the spill reserve; both at aoff[w] of src.
template <int EW>
__global__ void __launch_bounds__(kClipBlock)
k_dense_overflow(long nwork, const int* __restrict__ work, const unsigned char* __restrict__ dropped,
                 const int* __restrict__ counts, const long* __restrict__ roff, const long* __restrict__ aoff,
                 const long* __restrict__ doff, const double* __restrict__ src, double* __restrict__ dst) {
  const int lane = threadIdx.x & 63;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long nwaves = ((long)gridDim.x * blockDim.x) >> 6;
  for (long b0 = wave * 64; b0 < nwork; b0 += nwaves * 64) {
    const long w = b0 + lane;
    long so = 0, dof = 0;
    int n = 0;
    if (w < nwork && dropped[w]) {
      const long p = work[w];
      const long b = dropped[w] == 1 ? roff[p + 1] - roff[p] : 0;
      n = counts[p] - (int)b;
      so = aoff[w];
      dof = doff[p] + b;
    }
    for (int k = 0; k < 64; k++) {
      const int m = __shfl(n, k, 64);
      if (m == 0) continue;
      const double2* s2 = reinterpret_cast<const double2*>(src + __shfl(so, k, 64) * EW);
      double2* d2 = reinterpret_cast<double2*>(dst + __shfl(dof, k, 64) * EW);
      for (int i = lane; i < (EW / 2) * m; i += 64) stream_word(d2 + i, s2 + i);
    }
  }
}

template <class T>
static void inclusive_scan_long(Context& C, T in, long* out, long n) {
  size_t bytes = 0;
  LSM_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out, (size_t)n, rocprim::plus<long>(), C.stream));
  void* tmp = C.scan_tmp.get(bytes);
  LSM_HIP(rocprim::inclusive_scan(tmp, bytes, in, out, (size_t)n, rocprim::plus<long>(), C.stream));
}

// Single pass: work item w's budget at the chosen level (byte shift / 8 of
// the packed levels), as the run length the placement scans.
struct BudgetAt {
  const unsigned* budget;
  int shift;
  __host__ __device__ long operator()(long w) const { return (long)((budget[w] >> shift) & 0xffu); }
};

// Single pass: every pair's run start from the work items' inclusive budget
// sums (pairs off the work list, all empty, start where the next work
// item's run does; offsets[npairs] = the slots).  Item w owns the pairs from
// just after item w - 1's pair up to its own (w = nwork: the tail up to
// npairs).  A wave takes 64 consecutive items, whose pairs are one
// contiguous range, and writes it with its lanes on consecutive offsets,
// each finding its pair's item by a binary search over the lanes' ranges.
__global__ void __launch_bounds__(256)
k_place_runs(long npairs, long nwork, const int* __restrict__ work, const long* __restrict__ incl,
             long* __restrict__ offsets) {
  const int lane = threadIdx.x & 63;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long nwaves = ((long)gridDim.x * blockDim.x) >> 6;
  for (long w0 = wave * 64; w0 <= nwork; w0 += nwaves * 64) {
    const long w = w0 + lane;
    const int last = nwork - w0 < 63 ? (int)(nwork - w0) : 63;  // the wave's last item (lane)
    long hi = 0x7fffffffffffffffL, base = 0;
    if (w <= nwork) {
      hi = w < nwork ? (long)work[w] : npairs;
      base = w == 0 ? 0 : incl[w - 1];
    }
    const long first = w0 == 0 ? 0 : (long)work[w0 - 1] + 1;
    const long end = __shfl(hi, last, 64);
    for (long b = first; b <= end; b += 64) {
      const long p = b + lane <= end ? b + lane : end;
      int j = 0;  // the first lane whose range ends at or after p
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (__shfl(hi, j + step - 1, 64) < p) j += step;
      const long v = __shfl(base, j, 64);
      if (b + lane <= end) offsets[p] = v;
    }
  }
}

void compact_result(Context& C, int dim, libsupermesh_gpu_result* out);
// Tet triage and the work list of the pairs that need the full clip
// (g.work, g.nwork on the device); buffers sized for `cap` pairs.  Only
// kernels, so that a graph can record it.
static void triage_tets(Context& C, ClipArgs& g, bool mass, long cap) {
  int* region = C.mflag.as<int>(cap);
  long* bcount = C.nwork.as<long>(2 * kTriageGrid + 2);
  long* boff = bcount + kTriageGrid + 1;
  if (mass)
    hipLaunchKernelGGL(k_tet_triage<true>, dim3(kTriageGrid), dim3(kClipBlock), 0, C.stream, g, region, bcount);
  else
    hipLaunchKernelGGL(k_tet_triage<false>, dim3(kTriageGrid), dim3(kClipBlock), 0, C.stream, g, region, bcount);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, C.stream, bcount, boff);
  int* work = C.work.as<int>(cap);
  hipLaunchKernelGGL(k_gather_work, dim3(kTriageGrid), dim3(kClipBlock), 0, C.stream, g.npairs, g.npairs_dev, boff,
                     region, work);
  LSM_HIP(hipGetLastError());
  g.work = work;
  g.nwork = reinterpret_cast<const unsigned long long*>(boff + kTriageGrid);
}

bool mass_prepare_rows(Context& C, const DevMesh& A, const DevMesh& B, const long* indptr, ClipArgs& g);

// LIBSUPERMESH_GPU_DENSE after a single pass whose runs are placed: runs at
// roff[p] (split runs' arena parts and the runs clipped again at aoff[w]) of
// g.elements, which lies `gap` elements into C.elements, moved to the prefix
// sums of the counts at C.elements' start (k_dense_step), or, when some pair
// lies ahead of its run, copied to the second buffer if device memory holds
// it.  False when neither: the caller fills every pair again.
static bool dense_from_runs(Context& C, const ClipArgs& g, long gap, long nwork, const long* roff,
                            libsupermesh_gpu_result* out) {
  constexpr int kMaxSteps = 512;  // in-place steps before giving up (a few dozen in practice)
  const long npairs = out->npairs, nout = out->nout;
  long* dofs = C.offsets_dense.as<long>(npairs + 1);
  long* tmp_counts = C.row_count.as<long>(npairs);
  hipLaunchKernelGGL(k_fill_long, dim3(1), dim3(64), 0, C.stream, 1L, 0L, dofs);
  hipLaunchKernelGGL(k_count_to_long, dim3(grid_for(npairs, 256)), dim3(256), 0, C.stream, npairs, g.counts,
                     tmp_counts);
  LSM_HIP(hipGetLastError());
  inclusive_scan_long(C, tmp_counts, dofs + 1, npairs);
  long* ctl = C.dense_ctl.as<long>(4);  // [0] pairs ahead of their runs, [2..3] the step cursor
  hipLaunchKernelGGL(k_fill_long, dim3(1), dim3(64), 0, C.stream, 4L, 0L, ctl);
  hipLaunchKernelGGL(k_dense_fits, dim3(grid_for(npairs + 1, 256)), dim3(256), 0, C.stream, npairs + 1, roff, gap,
                     dofs, reinterpret_cast<unsigned long long*>(ctl));
  LSM_HIP(hipGetLastError());
  long ahead = 0;
  LSM_HIP(hipMemcpyAsync(&ahead, ctl, sizeof(long), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  const double* src = g.elements;
  double* dst = static_cast<double*>(C.elements.ptr);
  if (ahead == 0) {
    long* cursor = ctl + 2;
    long p0 = 0;
    int step = 0;
    while (p0 < npairs && step < kMaxSteps) {
      for (int j = 0; j < 8; j++, step++)
        hipLaunchKernelGGL(k_dense_step<12>, dim3(kClipGrid), dim3(kClipBlock), 0, C.stream, npairs, g.counts,
                           roff, gap, dofs, cursor, step, src, dst);
      LSM_HIP(hipGetLastError());
      LSM_HIP(hipMemcpyAsync(&p0, cursor + (step & 1), sizeof(long), hipMemcpyDeviceToHost, C.stream));
      LSM_HIP(hipStreamSynchronize(C.stream));
    }
    if (p0 < npairs) return false;
  } else {
    size_t free_b = 0, total_b = 0;
    LSM_HIP(hipMemGetInfo(&free_b, &total_b));
    if ((double)nout * 96.0 + 1e9 > (double)free_b) return false;
    dst = C.elements_dense.as<double>((size_t)(nout > 0 ? nout : 1) * 12);
    hipLaunchKernelGGL(k_dense_step<12>, dim3(kClipGrid), dim3(kClipBlock), 0, C.stream, npairs, g.counts, roff,
                       gap, dofs, nullptr, 0, src, dst);
  }
  if (nwork > 0)
    hipLaunchKernelGGL(k_dense_overflow<12>, dim3(kClipGrid), dim3(kClipBlock), 0, C.stream, nwork, g.work,
                       g.dropped, g.counts, roff, g.aoff, dofs, src, dst);
  LSM_HIP(hipGetLastError());
  LSM_HIP(hipStreamSynchronize(C.stream));
  out->offsets = dofs;
  out->elements = dst;
  out->elements_size = nout;
  out->dense_route = ahead == 0 ? 1 : 2;
  return true;
}

void clip_pairs(Context& C, const DevMesh& A, const DevMesh& B, long npairs, int flags,
                libsupermesh_gpu_result* out) {
  // work lists, overflow lists and the single pass's items are 32-bit pair numbers
  LSM_REQUIRE(npairs < (1L << 31), LIBSUPERMESH_GPU_ERROR_OVERFLOW,
              "more than 2^31 - 1 candidate pairs to clip in one call");
  const int dim = A.dim, loc = dim + 1;
  const bool mass = flags & LIBSUPERMESH_GPU_MASS;
  const bool elems = flags & LIBSUPERMESH_GPU_ELEMENTS;
  ClipArgs g = {};
  g.npairs = npairs;
  g.rows = out->rows;
  g.cols = out->indices;
  g.acoords = static_cast<const double*>(A.coords.ptr);
  g.bcoords = static_cast<const double*>(B.coords.ptr);
  g.asp = static_cast<const double*>(A.meas_sp.ptr);
  g.bsp = static_cast<const double*>(B.meas_sp.ptr);
  g.ascale = static_cast<const double*>(A.scale.ptr);
  g.bscale = static_cast<const double*>(B.scale.ptr);
  g.bplanes = static_cast<const double*>(B.planes.ptr);
  g.aaff = static_cast<const double*>(A.affine.ptr);
  g.baff = static_cast<const double*>(B.affine.ptr);
  g.counts = C.counts.as<int>(npairs);
  g.measure = C.measure.as<double>(npairs);
  // 2D: the mass goes to the rows of A's nodes laid out before the clip;
  // per-pair blocks only on request (or without rows, as the only output)
  C.mass_rows_ready = false;
  if (mass && dim == 2 && out->indptr) mass_prepare_rows(C, A, B, out->indptr, g);
  const bool blocks = mass && (dim == 3 || !C.mass_rows_ready || (flags & LIBSUPERMESH_GPU_BLOCKS));
  g.blocks = blocks ? C.blocks.as<double>((size_t)npairs * loc * loc) : nullptr;
  Partial* parts = C.partials.as<Partial>(kClipGrid);
  if (dim == 3 && npairs > 0) triage_tets(C, g, mass, npairs);
  bool learned = elems && dim == 3 && (flags & LIBSUPERMESH_GPU_LEARNED) && npairs > 0;
  const bool dense = flags & LIBSUPERMESH_GPU_DENSE;
  long nwork = 0, nslots = 0, nspill = 0, capacity = 0, arena_per_lane = 0, gap = 0;
  long* offs = nullptr;
  C.runs_layout = false;
  out->dense_route = 0;
  if (learned) {
    // single pass: predicted budgets -> runs placed from them -> one walk
    // that counts and writes
    Predictor& P = C.predictor;
    LSM_REQUIRE(P.ready, LIBSUPERMESH_GPU_ERROR_STATE,
                "LIBSUPERMESH_GPU_LEARNED needs predictor weights (libsupermesh_gpu_predictor_set)");
    LSM_HIP(hipMemcpyAsync(&nwork, g.nwork, sizeof(long), hipMemcpyDeviceToHost, C.stream));
    LSM_HIP(hipStreamSynchronize(C.stream));
    unsigned* budget = P.budget.as<unsigned>((size_t)(nwork > 0 ? nwork : 1));
    unsigned long long* sums = C.scalars.as<unsigned long long>(16) + 8;
    LSM_HIP(hipMemsetAsync(sums, 0, 5 * sizeof(unsigned long long), C.stream));
    PredictArgs pa = {};
    pa.nwork = nwork;
    pa.work = g.work;
    pa.rows = g.rows;
    pa.cols = g.cols;
    pa.acoords = g.acoords;
    pa.bcoords = g.bcoords;
    pa.weights = static_cast<const __bf16*>(P.weights.ptr);
    pa.bias = static_cast<const float*>(P.bias.ptr);
    budget_levels(P.quantile, pa.levels);
    pa.budget = budget;
    pa.sums = sums;
    pa.spare = reinterpret_cast<unsigned*>(sums + 5);
    if (C.timing) LSM_HIP(hipEventRecord(C.ev[12], C.stream));
    predict_budgets(C, pa);
    if (C.timing) LSM_HIP(hipEventRecord(C.ev[13], C.stream));
    unsigned long long hs[5] = {0, 0, 0, 0, 0};
    LSM_HIP(hipMemcpyAsync(hs, sums, sizeof(hs), hipMemcpyDeviceToHost, C.stream));
    LSM_HIP(hipStreamSynchronize(C.stream));
    // the highest quantile level whose runs, with the arena and the spill
    // reserve behind them, fit in the device memory left.  The arena takes
    // the outputs beyond a budget (a fraction of the expected output, spread
    // over the walk's lanes), the spill reserve the joined runs of those
    // pairs (run layout only) and the runs of the pairs whose lane's arena
    // ran out (clipped again).  The dense layout is moved into place in the
    // same buffer afterwards; its runs start a gap into the buffer, so that
    // the pairs ahead of their budgets do not reach runs not yet moved.
    C.elements_dense.release();
    if (dense) C.offsets_dense.as<long>(npairs + 1);  // the dense layout's offsets
    else C.offsets_dense.release();
    size_t free_b = 0, total_b = 0;
    LSM_HIP(hipMemGetInfo(&free_b, &total_b));
    const double avail = (double)free_b + (double)C.elements.bytes - 4e9;
    const double expect = (double)hs[4] / 1024.0;
    const long lanes = tet_walk_lanes(C, mass ? 4 : 3);  // the grid of the single-pass launch
    arena_per_lane = ((long)(P.arena_fraction * expect) + P.arena_min) / lanes;
    nspill = (long)(P.spill_fraction * expect) + P.spill_min;
    gap = dense ? (long)(expect / 256.0) + 4096 : 0;
    int level = -1;
    for (int k = 0; k < 4 && level < 0; k++) {
      const double need =
          (double)gap + (double)hs[k] + (double)(arena_per_lane * lanes) + (double)nspill + 1.0;
      if (need * 96.0 <= avail) level = k;
    }
    if (level < 0) {
      learned = false;  // no room: the plain count-then-fill path (same output)
    } else {
      g.bshift = 8 * level;
      P.level = level;
      out->learned_quantile = pa.levels[level];
      // runs: budgets laid end to end in pair order (pairs outside the work
      // list have none)
      offs = C.offsets.as<long>(npairs + 1);
      long* incl = C.row_count.as<long>((size_t)(nwork > 0 ? nwork : 1));
      if (nwork > 0)
        inclusive_scan_long(C, rocprim::make_transform_iterator(rocprim::counting_iterator<long>(0),
                                                                BudgetAt{budget, g.bshift}),
                            incl, nwork);
      hipLaunchKernelGGL(k_place_runs, dim3(grid_for(nwork + 1, 256)), dim3(256), 0, C.stream, npairs, nwork,
                         g.work, incl, offs);
      LSM_HIP(hipGetLastError());
      LSM_HIP(hipMemcpyAsync(&nslots, offs + npairs, sizeof(long), hipMemcpyDeviceToHost, C.stream));
      LSM_HIP(hipStreamSynchronize(C.stream));
      capacity = nslots + arena_per_lane * lanes + nspill;
      g.offsets = offs;
      g.elements = C.elements.as<double>((size_t)(gap + capacity) * 12) + gap * 12;
      g.budget = budget;
      g.arena_base = nslots;
      g.arena_per_lane = arena_per_lane;
      g.aoff = P.aoff.as<long>((size_t)(nwork > 0 ? nwork : 1));
      g.dropped = P.overflow.as<unsigned char>((size_t)(nwork > 0 ? nwork : 1));
      g.ndropped = C.scalars.as<unsigned long long>(16) + 4;
      LSM_HIP(hipMemsetAsync(g.ndropped, 0, 2 * sizeof(unsigned long long), C.stream));
    }
  }
  if (C.timing) LSM_HIP(hipEventRecord(C.ev[14], C.stream));
  if (dim == 2) launch_clip<2>(C, g, mass, false);
  else if (learned) launch_tet_walk(C, g, mass ? 4 : 3);
  else launch_tet_walk(C, g, mass, false);
  if (C.timing) LSM_HIP(hipEventRecord(C.ev[15], C.stream));
  hipLaunchKernelGGL(k_reduce_pairs, dim3(kClipGrid), dim3(kClipBlock), 0, C.stream, npairs, nullptr, g.counts,
                     g.measure, parts);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kClipBlock), 0, C.stream, kClipGrid, parts);
  LSM_HIP(hipGetLastError());
  Partial tot;
  LSM_HIP(hipMemcpyAsync(&tot, parts, sizeof(Partial), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  if (C.timing) {
    float ms = 0.f;
    LSM_HIP(hipEventElapsedTime(&ms, C.ev[14], C.ev[15]));
    out->ms[7] = ms;
    if (learned) {
      LSM_HIP(hipEventElapsedTime(&ms, C.ev[12], C.ev[13]));
      out->ms[5] = ms;
    }
  }
  out->counts = g.counts;
  out->measure = g.measure;
  out->blocks = g.blocks;
  out->sum = npairs > 0 ? tot.hi + tot.lo : 0.0;
  out->nonempty = npairs > 0 ? tot.nonempty : 0;
  out->nout = npairs > 0 ? tot.nout : 0;
  out->max_count = npairs > 0 ? (int)tot.maxc : 0;
  out->offsets = nullptr;
  out->elements = nullptr;
  out->elements_size = 0;
  if (!elems) return;
  if (learned) {
    unsigned long long h[2] = {0, 0};
    LSM_HIP(hipMemcpyAsync(h, g.ndropped, sizeof(h), hipMemcpyDeviceToHost, C.stream));
    LSM_HIP(hipStreamSynchronize(C.stream));
    out->learned_slots = nslots;
    out->learned_overflow = (long)(h[0] + h[1]);
    bool placed = true;
    // the spill reserve takes the runs of the pairs clipped again and, for
    // the run layout, the split runs joined (the dense move takes a split
    // run's two parts where they lie)
    if (h[1] > 0 || (!dense && h[0] > 0)) {
      long* incl = C.predictor.spill.as<long>((size_t)nwork);
      inclusive_scan_long(C, rocprim::make_transform_iterator(
                                 rocprim::counting_iterator<long>(0),
                                 DroppedLength{g.work, g.dropped, (unsigned char)(dense ? 2 : 1), g.counts}),
                          incl, nwork);
      long need = 0;
      LSM_HIP(hipMemcpyAsync(&need, incl + nwork - 1, sizeof(long), hipMemcpyDeviceToHost, C.stream));
      LSM_HIP(hipStreamSynchronize(C.stream));
      placed = need <= nspill;
      if (placed) {
        const long base = capacity - nspill;
        if (!dense && h[0] > 0)
          hipLaunchKernelGGL(k_join_runs, dim3(kClipGrid), dim3(kClipBlock), 0, C.stream, nwork, g.work, g.dropped,
                             g.counts, g.budget, g.bshift, g.aoff, incl, base, offs, g.elements);
        if (h[1] > 0) {
          hipLaunchKernelGGL(k_dropped_offsets, dim3(grid_for(nwork, 256)), dim3(256), 0, C.stream, nwork,
                             g.work, g.dropped, incl, g.counts, base, g.aoff, offs, !dense);
          LSM_HIP(hipGetLastError());
          g.ovf = g.dropped;
          launch_tet_walk(C, g, 2);
          g.ovf = nullptr;
        }
        LSM_HIP(hipGetLastError());
      }
    }
    if (placed && !dense) {
      out->offsets = offs;
      out->elements = static_cast<const double*>(C.elements.get(8));
      out->elements_size = capacity;
      C.runs_layout = true;
      return;
    }
    if (placed && dense_from_runs(C, g, gap, nwork, offs, out)) return;
    out->dense_route = dense ? 3 : 0;
    // the spill reserve is too small, or the dense layout could not be moved
    // into place nor copied: the plain fill pass over every pair (the counts
    // are known), into the dense layout
  }
  long* dofs = C.offsets.as<long>(npairs + 1);
  hipLaunchKernelGGL(k_fill_long, dim3(1), dim3(64), 0, C.stream, 1L, 0L, dofs);
  LSM_HIP(hipGetLastError());
  if (npairs > 0) {
    long* tmp_counts = C.row_count.as<long>(npairs);
    hipLaunchKernelGGL(k_count_to_long, dim3(grid_for(npairs, 256)), dim3(256), 0, C.stream, npairs, g.counts,
                       tmp_counts);
    inclusive_scan_long(C, tmp_counts, dofs + 1, npairs);
    g.offsets = dofs;
    g.elements = C.elements.as<double>((size_t)(out->nout > 0 ? out->nout : 1) * loc * dim);
    if (dim == 2) launch_clip<2>(C, g, false, true);
    else launch_tet_walk(C, g, false, true);
  }
  out->offsets = dofs;
  out->elements = static_cast<const double*>(C.elements.get(8));
  out->elements_size = out->nout;
}

// The dense layout of a result whose runs are placed apart (the learned
// single pass): offsets become the prefix sums of the counts, elements the
// runs end to end, in the context's second pair of output buffers.
void compact_result(Context& C, int dim, libsupermesh_gpu_result* out) {
  const long npairs = out->npairs;
  long* dofs = C.offsets_dense.as<long>(npairs + 1);
  hipLaunchKernelGGL(k_fill_long, dim3(1), dim3(64), 0, C.stream, 1L, 0L, dofs);
  LSM_HIP(hipGetLastError());
  const int ew = (dim + 1) * dim;
  double* del = C.elements_dense.as<double>((size_t)(out->nout > 0 ? out->nout : 1) * ew);
  if (npairs > 0) {
    long* tmp_counts = C.row_count.as<long>(npairs);
    hipLaunchKernelGGL(k_count_to_long, dim3(grid_for(npairs, 256)), dim3(256), 0, C.stream, npairs, out->counts,
                       tmp_counts);
    inclusive_scan_long(C, tmp_counts, dofs + 1, npairs);
    compact_runs(C, dim, npairs, out->counts, out->offsets, out->elements, dofs, del);
  }
  LSM_HIP(hipStreamSynchronize(C.stream));
  out->offsets = dofs;
  out->elements = del;
  out->elements_size = out->nout;
}

// ---------------------------------------------------------------------------
// The count pass (counts, measures, sums; no elements, no mass) in a launch
// shape picked by measurement, recorded as one hipGraph
// (LIBSUPERMESH_GPU_GRAPH).  Every kernel of the recorded pass reads the pair
// count from the device (indptr[na]) and everything it clears is cleared by
// a kernel, so a replay serves every call whose arguments (buffers, staged
// meshes, launch shape) are those it was recorded with, whatever its pairs.
// ---------------------------------------------------------------------------

__global__ void k_min_u64(const unsigned long long* __restrict__ n, unsigned long long cap,
                          unsigned long long* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = *n < cap ? *n : cap;
}

template <int BLOCK, int REG>
static int walk_blocks_per_cu() {
  static int nb = 0;
  if (nb == 0) {
    LSM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, reinterpret_cast<const void*>(k_clip_tet_walk<false, false, false, BLOCK, REG>), BLOCK, 0));
    nb = nb > 0 ? nb : 1;
  }
  return nb;
}

static int device_cus(Context& C) {
  int cus = 0;
  LSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, C.device));
  return cus > 0 ? cus : 256;
}

template <int BLOCK, int REG>
static void walk_count_as(Context& C, const ClipArgs& g, int grid) {
  hipLaunchKernelGGL((k_clip_tet_walk<false, false, false, BLOCK, REG>), dim3(grid), dim3(BLOCK), 0, C.stream, g);
}

static void launch_count_walk(Context& C, const ClipArgs& a, const CountShape& s) {
  ClipArgs g = a;
  g.queue = walk_queues(C);
  switch (s.block * 2 + s.reg) {
    case 128: walk_count_as<64, 0>(C, g, s.grid); break;
    case 129: walk_count_as<64, 1>(C, g, s.grid); break;
    case 256: walk_count_as<128, 0>(C, g, s.grid); break;
    case 257: walk_count_as<128, 1>(C, g, s.grid); break;
    case 512: walk_count_as<256, 0>(C, g, s.grid); break;
    default: walk_count_as<256, 1>(C, g, s.grid); break;
  }
  LSM_HIP(hipGetLastError());
}

// 2D: the LDS clipper in shape s, then the streaming clipper over the pairs
// it handed back (g.olist, g.nolist set by the caller).
// s.grid 0: one round over `cap` pairs (the most the call, or a replay of
// its recording, can have).
static void launch_count_tri(Context& C, const ClipArgs& g, const CountShape& s, long cap) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, C.stream, g.nolist);
  const dim3 grid(s.grid > 0 ? (unsigned)s.grid : one_round_grid(cap, s.block));
  if (s.block == 64) hipLaunchKernelGGL((k_clip_tri<false, false, 64>), grid, dim3(64), 0, C.stream, g);
  else if (s.block == 128) hipLaunchKernelGGL((k_clip_tri<false, false, 128>), grid, dim3(128), 0, C.stream, g);
  else hipLaunchKernelGGL((k_clip_tri<false, false, 256>), grid, dim3(256), 0, C.stream, g);
  ClipArgs f = g;
  f.work = g.olist;
  f.nwork = g.nolist;
  hipLaunchKernelGGL((k_clip<2, false, false>), dim3(64), dim3(kClipBlock), 0, C.stream, f);
  LSM_HIP(hipGetLastError());
}

// Times every candidate shape on a prefix of the pairs (3D: of the work list
// the triage left in g) and keeps the fastest in C.shape.  Candidates:
// blocks of 64, 128 or 256 lanes; 3D: the depth-1 tet in registers or in
// LDS, and every resident block (the walk's lanes are persistent) or half
// of them; 2D: grids of 1024 to 8192 blocks (pairs per lane of the
// grid-stride loop) or one block per `block` pairs.  Each candidate runs once
// to warm up, then once timed.
static void tune_count_pass(Context& C, ClipArgs g, int dim) {
  constexpr unsigned long long kSample = 1ULL << 23;
  std::vector<CountShape> cand;
  const int cus = device_cus(C);
  if (dim == 3) {
    const int bpc[6] = {walk_blocks_per_cu<64, 0>(), walk_blocks_per_cu<64, 1>(), walk_blocks_per_cu<128, 0>(),
                        walk_blocks_per_cu<128, 1>(), walk_blocks_per_cu<256, 0>(), walk_blocks_per_cu<256, 1>()};
    const int blocks[3] = {64, 128, 256};
    for (int b = 0; b < 3; b++)
      for (int reg = 0; reg < 2; reg++)
        for (int half = 0; half < 2; half++) {
          CountShape c;
          c.block = blocks[b];
          c.reg = reg;
          c.grid = std::max(kXcds, (bpc[2 * b + reg] * cus) >> half);
          cand.push_back(c);
        }
    unsigned long long* ns = C.scalars.as<unsigned long long>(16) + 3;
    hipLaunchKernelGGL(k_min_u64, dim3(1), dim3(64), 0, C.stream, g.nwork, kSample, ns);
    g.nwork = ns;
  } else {
    for (int block : {64, 128, 256})
      for (int grid : {0, 1024, 2048, 4096, 8192}) {
        CountShape c;
        c.block = block;
        c.grid = grid;
        cand.push_back(c);
      }
    g.npairs = std::min<long>(g.npairs, (long)kSample);
    g.npairs_dev = nullptr;
  }
  hipEvent_t e0, e1;
  LSM_HIP(hipEventCreate(&e0));
  LSM_HIP(hipEventCreate(&e1));
  CountShape best;
  for (CountShape& c : cand) {
    for (int rep = 0; rep < 2; rep++) {
      if (rep == 1) LSM_HIP(hipEventRecord(e0, C.stream));
      if (dim == 3) launch_count_walk(C, g, c);
      else launch_count_tri(C, g, c, g.npairs);
      if (rep == 1) LSM_HIP(hipEventRecord(e1, C.stream));
    }
    LSM_HIP(hipEventSynchronize(e1));
    LSM_HIP(hipEventElapsedTime(&c.ms, e0, e1));
    if (best.block == 0 || c.ms < best.ms) best = c;
  }
  LSM_HIP(hipEventDestroy(e0));
  LSM_HIP(hipEventDestroy(e1));
  C.shape[dim - 2] = best;
}

// Everything the count pass launches, in shape s: only kernels.
static void enqueue_count_pass(Context& C, ClipArgs& g, int dim, long cap, const CountShape& s, Partial* parts,
                               long ngroups, const long* a_off, const long* indptr) {
  if (dim == 3) {
    triage_tets(C, g, false, cap);
    launch_count_walk(C, g, s);
  } else {
    launch_count_tri(C, g, s, cap);
  }
  hipLaunchKernelGGL(k_reduce_pairs, dim3(kClipGrid), dim3(kClipBlock), 0, C.stream, g.npairs, g.npairs_dev,
                     g.counts, g.measure, parts);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kClipBlock), 0, C.stream, kClipGrid, parts);
  LSM_HIP(hipGetLastError());
  if (ngroups > 0) batch_sums(C, ngroups, a_off, indptr, g.measure);
}

// Records the count pass as a graph; false (and nothing kept) when the
// recording fails or holds nodes other than kernels.
static bool record_count_pass(Context& C, ClipArgs& g, int dim, long cap, const CountShape& s, Partial* parts,
                              long ngroups, const long* a_off, const long* indptr) {
  hipGraph_t graph = nullptr;
  if (hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  bool ok = true;
  try {
    enqueue_count_pass(C, g, dim, cap, s, parts, ngroups, a_off, indptr);
  } catch (...) {
    ok = false;
  }
  const hipError_t end = hipStreamEndCapture(C.stream, &graph);
  if (end != hipSuccess || !ok || graph == nullptr) {
    (void)hipGetLastError();
    if (graph) (void)hipGraphDestroy(graph);
    return false;
  }
  size_t n = 0;
  LSM_HIP(hipGraphGetNodes(graph, nullptr, &n));
  std::vector<hipGraphNode_t> nodes(n);
  LSM_HIP(hipGraphGetNodes(graph, nodes.data(), &n));
  for (hipGraphNode_t node : nodes) {
    hipGraphNodeType t;
    LSM_HIP(hipGraphNodeGetType(node, &t));
    if (t != hipGraphNodeTypeKernel) ok = false;
  }
  hipGraphExec_t exec = nullptr;
  if (ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    ok = false;
  }
  if (!ok) {
    (void)hipGraphDestroy(graph);
    C.graph_refused = true;
    return false;
  }
  C.graph[dim - 2] = graph;
  C.graph_exec[dim - 2] = exec;
  return true;
}

static void release_graph(Context& C, int d) {
  if (C.graph_exec[d]) (void)hipGraphExecDestroy(C.graph_exec[d]);
  if (C.graph[d]) (void)hipGraphDestroy(C.graph[d]);
  C.graph_exec[d] = nullptr;
  C.graph[d] = nullptr;
  C.graph_key[d].clear();
}

void release_graph(Context& C) {
  release_graph(C, 0);
  release_graph(C, 1);
}

void count_pass_graph(Context& C, const DevMesh& A, const DevMesh& B, long npairs, libsupermesh_gpu_result* out,
                      long ngroups, const long* a_off) {
  LSM_REQUIRE(npairs < (1L << 31), LIBSUPERMESH_GPU_ERROR_OVERFLOW,
              "more than 2^31 - 1 candidate pairs to clip in one call");
  const int dim = A.dim;
  // buffers with headroom, so that later calls with a few more pairs replay
  // the same recording
  if (npairs > C.graph_cap) C.graph_cap = npairs + npairs / 8 + 1024;
  const long cap = C.graph_cap;
  ClipArgs g = {};
  g.npairs = npairs;
  g.npairs_dev = out->indptr + out->nelements_a;
  g.rows = out->rows;
  g.cols = out->indices;
  g.acoords = static_cast<const double*>(A.coords.ptr);
  g.bcoords = static_cast<const double*>(B.coords.ptr);
  g.asp = static_cast<const double*>(A.meas_sp.ptr);
  g.bsp = static_cast<const double*>(B.meas_sp.ptr);
  g.ascale = static_cast<const double*>(A.scale.ptr);
  g.bscale = static_cast<const double*>(B.scale.ptr);
  g.bplanes = static_cast<const double*>(B.planes.ptr);
  g.counts = C.counts.as<int>(cap);
  g.measure = C.measure.as<double>(cap);
  Partial* parts = C.partials.as<Partial>(kClipGrid);
  // every buffer the pass touches, allocated before any recording
  C.mflag.as<int>(cap);
  C.work.as<int>(cap);
  C.nwork.as<long>(2 * kTriageGrid + 2);
  C.queue.as<unsigned long long>(kQueues);
  unsigned long long* scal = C.scalars.as<unsigned long long>(16);
  if (ngroups > 0) {
    C.gpairs.as<long>((size_t)ngroups + 1);
    C.gsum.as<double>((size_t)ngroups);
  }
  if (dim == 2) {
    g.olist = static_cast<int*>(C.work.ptr);
    g.nolist = scal + 2;
  }
  if (C.shape[dim - 2].block == 0) {
    if (dim == 3) triage_tets(C, g, false, cap);  // the work list the walk's candidates run on
    tune_count_pass(C, g, dim);
  }
  const CountShape s = C.shape[dim - 2];
  const std::vector<const void*> key = {
      reinterpret_cast<const void*>((long)dim), reinterpret_cast<const void*>(cap),
      reinterpret_cast<const void*>(ngroups), reinterpret_cast<const void*>((long)(s.block * 4 + s.reg)),
      reinterpret_cast<const void*>((long)s.grid), g.npairs_dev, g.rows, g.cols, g.acoords, g.bcoords, g.asp,
      g.bsp, g.ascale, g.bscale, g.bplanes, g.counts, g.measure, parts, C.mflag.ptr, C.work.ptr, C.nwork.ptr,
      C.queue.ptr, scal, a_off, out->indptr, C.gpairs.ptr, C.gsum.ptr};
  const int d = dim - 2;
  if (C.graph_exec[d] && key != C.graph_key[d]) release_graph(C, d);
  if (!C.graph_exec[d] && !C.graph_refused &&
      record_count_pass(C, g, dim, cap, s, parts, ngroups, a_off, out->indptr))
    C.graph_key[d] = key;
  if (C.timing) LSM_HIP(hipEventRecord(C.ev[14], C.stream));
  if (C.graph_exec[d]) {
    LSM_HIP(hipGraphLaunch(C.graph_exec[d], C.stream));
  } else {
    ClipArgs e = g;
    enqueue_count_pass(C, e, dim, cap, s, parts, ngroups, a_off, out->indptr);
  }
  if (C.timing) LSM_HIP(hipEventRecord(C.ev[15], C.stream));
  Partial tot;
  LSM_HIP(hipMemcpyAsync(&tot, parts, sizeof(Partial), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  if (C.timing) {
    float ms = 0.f;
    LSM_HIP(hipEventElapsedTime(&ms, C.ev[14], C.ev[15]));
    out->ms[7] = ms;
  }
  out->counts = g.counts;
  out->measure = g.measure;
  out->blocks = nullptr;
  out->sum = npairs > 0 ? tot.hi + tot.lo : 0.0;
  out->nonempty = npairs > 0 ? tot.nonempty : 0;
  out->nout = npairs > 0 ? tot.nout : 0;
  out->max_count = npairs > 0 ? (int)tot.maxc : 0;
  out->count_graph = C.graph_exec[d] != nullptr;
}

// ---------------------------------------------------------------------------
// Clipping an explicit pair list with no allocation, for callers that cannot
// allocate device memory (code under stream capture, serving loops).  The
// kernels read the meshes' positions and element lists directly (no staged
// copies), each lane gathering its pair's two elements; the pass's only device
// memory beyond the caller's inputs and outputs is the file-scope `scratch`
// below: the 2D pairs handed from the LDS clipper to the streaming one, their
// count, and the partial sums of the measure reduction.  Every call on a GPU
// shares that GPU's copy, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------

constexpr long kNoAllocMaxPairs = 1L << 22;
constexpr size_t kScratchOlist = 0;                                         // int [kNoAllocMaxPairs]
constexpr size_t kScratchCount = kScratchOlist + sizeof(int) * kNoAllocMaxPairs;  // unsigned long long
constexpr size_t kScratchParts = kScratchCount + 256;                        // Partial [kClipGrid]
constexpr size_t kScratchBytes = kScratchParts + sizeof(Partial) * kClipGrid;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DirectArgs {
  long npairs;
  const int* pair_a;
  const int* pair_b;
  long na, nb, nnodes_a, nnodes_b;
  const double* pos_a;
  const int* enl_a;
  const double* pos_b;
  const int* enl_b;
  int* counts;
  double* measure;
  int* olist;
  unsigned long long* nolist;
};

// The coordinates of element e (LOC = DIM + 1 nodes); false when a node id is
// outside [0, nnodes) (node 0 is read instead).
template <int DIM>
__device__ __forceinline__ bool gather_element(const double* __restrict__ pos, const int* __restrict__ enl,
                                               long nnodes, long e, double* c) {
  bool ok = true;
#pragma unroll
  for (int j = 0; j < DIM + 1; j++) {
    long v = enl[e * (DIM + 1) + j];
    if (v < 0 || v >= nnodes) {
      ok = false;
      v = 0;
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) c[j * DIM + d] = pos[v * DIM + d];
  }
  return ok;
}

// Both elements of pair p; false (and nothing read through a bad id) when an
// element or node id is out of range.
template <int DIM>
__device__ __forceinline__ bool gather_pair(const DirectArgs& g, long p, double* A, double* B) {
  const long a = g.pair_a[p], b = g.pair_b[p];
  if (a < 0 || a >= g.na || b < 0 || b >= g.nb) return false;
  const bool oa = gather_element<DIM>(g.pos_a, g.enl_a, g.nnodes_a, a, A);
  const bool ob = gather_element<DIM>(g.pos_b, g.enl_b, g.nnodes_b, b, B);
  return oa && ob;
}

__global__ void k_noalloc_zero(unsigned long long* __restrict__ n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *n = 0;
}

// 2D: clip_poly_lds as k_clip_tri; pairs that outgrow the LDS polygon go to
// the scratch list.  Invalid pairs get count -1 and measure 0.
__global__ void __launch_bounds__(kTriBlock) k_clip_direct_tri(DirectArgs g) {
  __shared__ double poly[LdsPoly<3, 3>::kPoints * 2 * kTriBlock];
  double* q = poly + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < g.npairs; p += stride) {
    double A[6], Bv[6];
    if (!gather_pair<2>(g, p, A, Bv)) {
      g.counts[p] = -1;
      g.measure[p] = 0.0;
      continue;
    }
    const double tol = tri_tol(triangle_area(A), triangle_area(Bv));
    Line L[3];
    tri_lines(Bv, L);
    double m = 0.0;
    const int nk = clip_poly_lds<3, 3, kTriBlock>(
        A, L, tol, q, [&](double, double, double, double, double, double, double ar) { m = m + ar; });
    if (nk < 0) {
      g.olist[atomicAdd(g.nolist, 1ULL)] = (int)p;
      continue;
    }
    g.counts[p] = nk;
    g.measure[p] = nk == 0 ? 0.0 : m;
  }
}

// 2D pairs of the scratch list on the streaming clipper.
__global__ void __launch_bounds__(kClipBlock) k_clip_direct_tri_stream(DirectArgs g) {
  const long n = (long)*g.nolist;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long w = (long)blockIdx.x * blockDim.x + threadIdx.x; w < n; w += stride) {
    const long p = g.olist[w];
    double A[6], Bv[6];
    gather_pair<2>(g, p, A, Bv);  // valid: checked by k_clip_direct_tri
    const double tol = tri_tol(triangle_area(A), triangle_area(Bv));
    Line L[3];
    tri_lines(Bv, L);
    double m = 0.0;
    const int nk = clip_poly<3>(A, L, tol, [&](double, double, double, double, double, double, double ar) {
      m = m + ar;
    });
    g.counts[p] = nk;
    g.measure[p] = nk == 0 ? 0.0 : m;
  }
}

// 3D: B's planes per pair (get_planes_tet, as staged for the bulk path) and
// the depth-first walk of k_clip_tet_walk.
__global__ void __launch_bounds__(kWalkBlock) k_clip_direct_tet(DirectArgs g) {
  __shared__ double stack[(3 - kWalkReg) * 12 * kWalkBlock];
  double* stk = stack + threadIdx.x;
  double R[12 * kWalkReg];
  const long stride = (long)gridDim.x * blockDim.x;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < g.npairs; p += stride) {
    double A[12], Bv[12];
    if (!gather_pair<3>(g, p, A, Bv)) {
      g.counts[p] = -1;
      g.measure[p] = 0.0;
      continue;
    }
    Planes P;
    tet_planes(Bv, P);
    auto plane = [&](int j, double* o) {
#pragma unroll
      for (int i = 0; i < 3; i++) {
        double v = in_reg(P.n[3][i]);
        v = j == 2 ? in_reg(P.n[2][i]) : v;
        v = j == 1 ? in_reg(P.n[1][i]) : v;
        v = j == 0 ? in_reg(P.n[0][i]) : v;
        o[i] = v;
      }
      double c = in_reg(P.c[3]);
      c = j == 2 ? in_reg(P.c[2]) : c;
      c = j == 1 ? in_reg(P.c[1]) : c;
      c = j == 0 ? in_reg(P.c[0]) : c;
      o[3] = c;
    };
    const double tol = tet_tol(tet_volume_tet(A), tet_volume_tet(Bv));
    double p0[4], d0[4];
    plane(0, p0);
    walk_distances(A, p0, d0);
    WalkState<4> W;
    walk_begin(W, pack_class(d0));
    double m = 0.0;
    int nk = 0;
    while (W.level > 0)
      walk_next<4, kWalkBlock, kWalkReg>(W, A, R, stk, plane, tol, [&](const double* t, double, bool keep) {
        const double vr = tet_volume_real(t);
        m = m + (keep ? vr : 0.0);
        nk += keep ? 1 : 0;
      });
    g.counts[p] = nk;
    g.measure[p] = nk == 0 ? 0.0 : m;
  }
}

// Sum of the measures: the double-double of k_reduce_partials.
__global__ void k_noalloc_sum(const Partial* __restrict__ parts, double* __restrict__ sum) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *sum = parts[0].hi + parts[0].lo;
}

// This GPU's address of `scratch`, looked up once per device with that
// device current (each GPU holds its own copy).  The array's name is a host
// stand-in: host code uses only this address.
static void* scratch_address(int device) {
  static std::mutex mu;
  static void* addr[64] = {};
  LSM_REQUIRE(device >= 0 && device < 64, LIBSUPERMESH_GPU_ERROR_INVALID, "invalid device index");
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    void* ptr = nullptr;
    LSM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    LSM_REQUIRE(ptr != nullptr, LIBSUPERMESH_GPU_ERROR_HIP, "no device address for the scratch array");
    addr[device] = ptr;
  }
  return addr[device];
}

void clip_pairs_noalloc(int device, hipStream_t stream, const libsupermesh_gpu_mesh& a,
                        const libsupermesh_gpu_mesh& b, long npairs, const int* pair_a, const int* pair_b,
                        int* counts, double* measure, double* sum) {
  LSM_REQUIRE(a.dim == b.dim && (a.dim == 2 || a.dim == 3), LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension");
  LSM_REQUIRE(npairs >= 0 && npairs <= kNoAllocMaxPairs, LIBSUPERMESH_GPU_ERROR_INVALID,
              "npairs exceeds the no-allocation limit of 4194304 pairs per call (split the list)");
  LSM_REQUIRE(sum != nullptr && (npairs == 0 || (pair_a && pair_b && counts && measure)),
              LIBSUPERMESH_GPU_ERROR_INVALID, "null argument");
  LSM_REQUIRE(a.nelements >= 0 && b.nelements >= 0 && (a.nelements == 0 || (a.nnodes > 0 && a.positions && a.enlist)) &&
                  (b.nelements == 0 || (b.nnodes > 0 && b.positions && b.enlist)),
              LIBSUPERMESH_GPU_ERROR_INVALID, "invalid mesh");
  LSM_HIP(hipSetDevice(device));
  unsigned char* base = static_cast<unsigned char*>(scratch_address(device));
  DirectArgs g;
  g.npairs = npairs;
  g.pair_a = pair_a;
  g.pair_b = pair_b;
  g.na = a.nelements;
  g.nb = b.nelements;
  g.nnodes_a = a.nnodes;
  g.nnodes_b = b.nnodes;
  g.pos_a = a.positions;
  g.enl_a = a.enlist;
  g.pos_b = b.positions;
  g.enl_b = b.enlist;
  g.counts = counts;
  g.measure = measure;
  g.olist = reinterpret_cast<int*>(base + kScratchOlist);
  g.nolist = reinterpret_cast<unsigned long long*>(base + kScratchCount);
  Partial* parts = reinterpret_cast<Partial*>(base + kScratchParts);
  if (a.dim == 2) {
    hipLaunchKernelGGL(k_noalloc_zero, dim3(1), dim3(64), 0, stream, g.nolist);
    if (npairs > 0) {
      hipLaunchKernelGGL(k_clip_direct_tri, dim3(grid_for(npairs, kTriBlock)), dim3(kTriBlock), 0, stream, g);
      hipLaunchKernelGGL(k_clip_direct_tri_stream, dim3(64), dim3(kClipBlock), 0, stream, g);
    }
  } else if (npairs > 0) {
    hipLaunchKernelGGL(k_clip_direct_tet, dim3(grid_for(npairs, kWalkBlock)), dim3(kWalkBlock), 0, stream, g);
  }
  hipLaunchKernelGGL(k_reduce_pairs, dim3(kClipGrid), dim3(kClipBlock), 0, stream, npairs, nullptr, counts, measure, parts);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kClipBlock), 0, stream, kClipGrid, parts);
  hipLaunchKernelGGL(k_noalloc_sum, dim3(1), dim3(64), 0, stream, parts, sum);
  LSM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Mixed mass assembly: (node_a, node_b, value) entries of non-empty pairs,
// sorted by key (stable, so equal keys keep pair order), then summed per key
// by a sequential walk -- deterministic.
// ---------------------------------------------------------------------------

template <int LOC>
__global__ void __launch_bounds__(256)
k_mass_entries(long npairs, const int* __restrict__ counts, const long* __restrict__ slot,
               const int* __restrict__ rows, const int* __restrict__ cols,
               const int* __restrict__ enla, const int* __restrict__ enlb,
               const double* __restrict__ blocks, long nb, uint64_t* __restrict__ keys,
               double* __restrict__ vals) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs || counts[p] == 0) return;
  long s = slot[p] - 1;  // inclusive scan of the non-empty flags
  long a = rows[p], b = cols[p];
#pragma unroll
  for (int i = 0; i < LOC; i++) {
    uint64_t ni = (uint64_t)enla[a * LOC + i];
#pragma unroll
    for (int j = 0; j < LOC; j++) {
      uint64_t nj = (uint64_t)enlb[b * LOC + j];
      long idx = (s * LOC + i) * LOC + j;
      keys[idx] = ni * (uint64_t)nb + nj;
      vals[idx] = blocks[p * LOC * LOC + i * LOC + j];
    }
  }
}

__global__ void __launch_bounds__(256)
k_nonempty_flag(long n, const int* __restrict__ counts, long* __restrict__ flag) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = counts[i] > 0;
}

__global__ void __launch_bounds__(256)
k_heads(long n, const uint64_t* __restrict__ keys, long* __restrict__ head) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void __launch_bounds__(256)
k_segment_sum(long n, const uint64_t* __restrict__ keys, const double* __restrict__ vals,
              const long* __restrict__ pos, long nb, int* __restrict__ mind,
              double* __restrict__ mval, unsigned long long* __restrict__ rowcnt) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = keys[i];
  if (i > 0 && keys[i - 1] == k) return;
  double s = 0.0;
  for (long j = i; j < n && keys[j] == k; j++) s = s + vals[j];
  long o = pos[i] - 1;
  mind[o] = (int)(k % (uint64_t)nb);
  mval[o] = s;
  atomicAdd(&rowcnt[k / (uint64_t)nb], 1ULL);
}

// ---------------------------------------------------------------------------
// Row-wise mixed mass assembly.  Row i of M (A node i) collects, from every A
// element e containing i (local vertex li) and every non-empty pair p of e,
// the entries (B node of local vertex j, blocks[p][li][j]).  One wave owns a
// row: it gathers the row's entries into LDS in a fixed order (elements of i
// in ascending (e, li), pairs of e in CSR order, j ascending), finds the
// distinct B nodes with an LDS hash table, sorts them, and each lane sums
// its columns' entries in gather order -- no global sort, and the same result
// on every run.  Rows too long for LDS take the global sort path.
// ---------------------------------------------------------------------------


__global__ void __launch_bounds__(256)
k_node_count(long ne, int loc, const int* __restrict__ enl, unsigned int* __restrict__ ncount) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * loc) return;
  atomicAdd(&ncount[enl[t]], 1u);
}

__global__ void __launch_bounds__(256)
k_node_fill(long ne, int loc, const int* __restrict__ enl, const long* __restrict__ nptr,
            unsigned int* __restrict__ cursor, int* __restrict__ nlist) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * loc) return;
  int v = enl[t];
  unsigned int k = atomicAdd(&cursor[v], 1u);
  nlist[nptr[v] + k] = (int)t;  // t = e*loc + li
}

// Sorts each node's entries (arrival order of the atomics is arbitrary) and
// counts, per entry (node, e, li), the row entries it contributes: loc per
// non-empty pair of e.
__global__ void __launch_bounds__(256)
k_node_rows(long nn, int loc, const long* __restrict__ nptr, int* __restrict__ nlist,
            const int* __restrict__ elem_nonempty, long* __restrict__ contrib) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  long b0 = nptr[i], b1 = nptr[i + 1];
  for (long x = b0; x < b1; x++) {
    int v = nlist[x];
    long y = x - 1;
    while (y >= b0 && nlist[y] > v) {
      nlist[y + 1] = nlist[y];
      y--;
    }
    nlist[y + 1] = v;
  }
  for (long x = b0; x < b1; x++) contrib[x] = (long)elem_nonempty[nlist[x] / loc] * loc;
}

__global__ void __launch_bounds__(256)
k_row_bounds(long nn, const long* __restrict__ nptr, const long* __restrict__ slot,
             long* __restrict__ raw_off, long* __restrict__ raw) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nn) return;
  raw_off[i] = slot[nptr[i]];
  if (i < nn) raw[i] = slot[nptr[i + 1]] - slot[nptr[i]];
}

__global__ void __launch_bounds__(256)
k_elem_nonempty(long ne, const long* __restrict__ indptr, const int* __restrict__ counts,
                int* __restrict__ out) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int c = 0;
  for (long p = indptr[e]; p < indptr[e + 1]; p++) c += counts[p] > 0;
  out[e] = c;
}

// Moves each row's reduced (column, value) list from its slot range to the
// CSR: 16 lanes per row (rows hold a few dozen columns).
constexpr int kCopyLanes = 16;
__global__ void __launch_bounds__(256)
k_mass_copy(long nn, const long* __restrict__ raw_off, const long* __restrict__ mptr,
            const int* __restrict__ tkey, const double* __restrict__ tval, int* __restrict__ mind,
            double* __restrict__ mval) {
  const long i = ((long)blockIdx.x * blockDim.x + threadIdx.x) / kCopyLanes;
  if (i >= nn) return;
  const long o = mptr[i], n = mptr[i + 1] - o, r = raw_off[i];
  for (long t = threadIdx.x % kCopyLanes; t < n; t += kCopyLanes) {
    mind[o + t] = tkey[r + t];
    mval[o + t] = tval[r + t];
  }
}


static void assemble_mass_sorted(Context& C, const DevMesh& A, const DevMesh& B, long npairs,
                   libsupermesh_gpu_result* out) {
  const int loc = A.dim + 1;
  const long na_nodes = A.nnodes, nb_nodes = B.nnodes;
  long* mindptr = C.mindptr.as<long>(na_nodes + 1);
  LSM_HIP(hipMemsetAsync(mindptr, 0, sizeof(long) * (na_nodes + 1), C.stream));
  out->mass_nnz = 0;
  out->mass_indptr = mindptr;
  out->mass_indices = static_cast<const int*>(C.mindices.get(4));
  out->mass_values = static_cast<const double*>(C.uvals.get(8));
  if (out->nonempty == 0) {
    LSM_HIP(hipStreamSynchronize(C.stream));
    return;
  }
  long* flag = C.mflag.as<long>(npairs);
  long* slot = C.row_count.as<long>(npairs);
  hipLaunchKernelGGL(k_nonempty_flag, dim3(grid_for(npairs, 256)), dim3(256), 0, C.stream, npairs,
                     out->counts, flag);
  inclusive_scan_long(C, flag, slot, npairs);
  const long nent = out->nonempty * loc * loc;
  uint64_t* keys = C.mkeys.as<uint64_t>(nent);
  double* vals = C.mvals.as<double>(nent);
  uint64_t* keys2 = C.mkeys_alt.as<uint64_t>(nent);
  double* vals2 = C.mvals_alt.as<double>(nent);
  if (loc == 3)
    hipLaunchKernelGGL(k_mass_entries<3>, dim3(grid_for(npairs, 256)), dim3(256), 0, C.stream, npairs,
                       out->counts, slot, out->rows, out->indices, A.enlist_dev, B.enlist_dev,
                       out->blocks, nb_nodes, keys, vals);
  else
    hipLaunchKernelGGL(k_mass_entries<4>, dim3(grid_for(npairs, 256)), dim3(256), 0, C.stream, npairs,
                       out->counts, slot, out->rows, out->indices, A.enlist_dev, B.enlist_dev,
                       out->blocks, nb_nodes, keys, vals);
  LSM_HIP(hipGetLastError());
  unsigned __int128 span = (unsigned __int128)na_nodes * (unsigned __int128)nb_nodes;
  unsigned end_bit = 1;
  while (end_bit < 64 && ((unsigned __int128)1 << end_bit) < span) end_bit++;
  size_t bytes = 0;
  LSM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys, keys2, vals, vals2, (size_t)nent, 0u, end_bit,
                                    C.stream));
  void* tmp = C.sort_tmp.get(bytes);
  LSM_HIP(rocprim::radix_sort_pairs(tmp, bytes, keys, keys2, vals, vals2, (size_t)nent, 0u, end_bit,
                                    C.stream));
  long* head = C.ukeys.as<long>(nent);
  long* pos = C.nuniq.as<long>(nent);
  hipLaunchKernelGGL(k_heads, dim3(grid_for(nent, 256)), dim3(256), 0, C.stream, nent, keys2, head);
  inclusive_scan_long(C, head, pos, nent);
  long nnz = 0;
  LSM_HIP(hipMemcpyAsync(&nnz, pos + nent - 1, sizeof(long), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  int* mind = C.mindices.as<int>(nnz);
  double* mval = C.uvals.as<double>(nnz);
  unsigned long long* rowcnt = reinterpret_cast<unsigned long long*>(C.mkeys.as<uint64_t>(na_nodes + 1));
  LSM_HIP(hipMemsetAsync(rowcnt, 0, sizeof(uint64_t) * (na_nodes + 1), C.stream));
  hipLaunchKernelGGL(k_segment_sum, dim3(grid_for(nent, 256)), dim3(256), 0, C.stream, nent, keys2, vals2,
                     pos, nb_nodes, mind, mval, rowcnt);
  LSM_HIP(hipGetLastError());
  inclusive_scan_long(C, reinterpret_cast<const long*>(rowcnt), mindptr + 1, na_nodes);
  LSM_HIP(hipStreamSynchronize(C.stream));
  out->mass_nnz = nnz;
  out->mass_indices = mind;
  out->mass_values = mval;
}

// Bitonic sort of up to 64*EPL int keys in registers (element t = lane + 64*r),
// keys only.
template <int EPL>
__device__ __forceinline__ void wave_bitonic_keys(int* key, int lane) {
  constexpr int N2 = 64 * EPL;
  for (int k = 2; k <= N2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j < 64) {
#pragma unroll
        for (int r = 0; r < EPL; r++) {
          const int t = lane + 64 * r;
          const int ok = __shfl_xor(key[r], j);
          const bool take_min = ((t & j) == 0) == ((t & k) == 0);
          key[r] = take_min ? (ok < key[r] ? ok : key[r]) : (ok > key[r] ? ok : key[r]);
        }
      } else {
#pragma unroll
        for (int r = 0; r < EPL; r++) {
#pragma unroll
          for (int r2 = 0; r2 < EPL; r2++) {
            if (r2 == (r ^ (j >> 6)) && r2 > r) {
              const int t = lane + 64 * r;
              const bool up = (t & k) == 0;
              if ((key[r] > key[r2]) == up) {
                const int kt = key[r];
                key[r] = key[r2];
                key[r2] = kt;
              }
            }
          }
        }
      }
    }
  }
}

constexpr int kHashSlots = 512;   // per-row table of distinct B nodes
constexpr int kMaxDistinct = 256;  // columns a row may hold on this path

// One wave per row of the mixed mass: gathers the row's contributions (its
// incident A elements in node-list order, their non-empty pairs in CSR order,
// B's local nodes in order) into LDS, finds the distinct columns with an LDS
// hash table, sorts them, and sums each column's contributions in gather
// order -- a fixed order, so the result is reproducible.  Writes the row's
// (column, value) list at raw_off[i] and its length to ucount[i]; rows whose
// distinct columns exceed kMaxDistinct raise *overflow (the caller then uses
// the global sort path).
// Deduplicates and sums one row held in LDS (rk/rv, n entries): distinct
// columns through an LDS hash table, sorted, each summed in entry order.
// Writes them at tkey/tval + o; returns the count, or -1 if the row has more
// than kMaxDistinct columns.  One wave; `ht` has 1 << HBITS entries (at
// least kMaxDistinct): the distinct columns are compacted into its front.
template <int HBITS = 9>
__device__ __forceinline__ int row_reduce(const int* rk, const double* rv, int n, int* ht, int lane,
                                          int* __restrict__ tkey, double* __restrict__ tval, long o) {
  // table of 128, 256 or 512 slots: at least 4/3 of the entries (n is the
  // same on every lane), so short rows clear and scan less; at most
  // 1 << HBITS (a table that fills sends the row to the sorted path)
  const int tb = n <= 96 ? 7 : (n <= 192 ? 8 : 9);
  const int tbits = tb < HBITS ? tb : HBITS;
  const int tsize = 1 << tbits;
  for (int t = lane; t < tsize; t += 64) ht[t] = -1;
  __syncthreads();
  bool full = false;
  for (int t = lane; t < n; t += 64) {
    const int key = rk[t];
    unsigned h = ((unsigned)key * 2654435761u) >> (32 - tbits);
    int tries = 0;
    while (true) {
      const int prev = atomicCAS(&ht[h], -1, key);
      if (prev == -1 || prev == key) break;
      h = (h + 1) & (tsize - 1);
      if (++tries >= tsize) {
        full = true;
        break;
      }
    }
  }
  __syncthreads();
  // compaction in place: slot pos <= s0 + lane, and every slot below s0 + 64
  // was read before the ballot that places the stores
  int nd = 0;
  for (int s0 = 0; s0 < tsize; s0 += 64) {
    const int key = ht[s0 + lane];
    const bool occ = key != -1;
    const unsigned long long m = __ballot(occ);
    const int pos = nd + __popcll(m & ((1ULL << lane) - 1));
    if (occ && pos < kMaxDistinct) ht[pos] = key;
    nd += __popcll(m);
  }
  const bool bad = __ballot(full) != 0ULL || nd > kMaxDistinct;
  __syncthreads();
  if (bad) return -1;
  int key[4];
#pragma unroll
  for (int r = 0; r < 4; r++) key[r] = lane + 64 * r < nd ? ht[lane + 64 * r] : 0x7fffffff;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // entries are read 4 at a time up to a multiple of 8 (the tail past n holds
  // key -2, which matches no column) so that the LDS loads overlap
  const int n8 = (n + 7) & ~7;
  if (nd <= 64) {
    wave_bitonic_keys<1>(key, lane);
    // few columns: the entries split into G consecutive parts; lane group g
    // sums part g for the column of rank lane % (64 / G), and the parts are
    // then added in part order -- a fixed order, so still reproducible
    const int G = nd <= 8 ? 8 : 64 / nd;
    const int W = 64 / G;
    const int r = lane % W;
    const int mykey = __shfl(key[0], r, 64);
    const int chunk = ((n8 / 8 + G - 1) / G) * 8;
    const int t0 = (lane / W) * chunk;
    const int t1 = t0 + chunk < n8 ? t0 + chunk : n8;
    double part = 0.0;
    for (int t = t0; t < t1; t += 4) {
      const int4 k0 = *reinterpret_cast<const int4*>(rk + t);
      const double2 v0 = *reinterpret_cast<const double2*>(rv + t);
      const double2 v1 = *reinterpret_cast<const double2*>(rv + t + 2);
      const int kk[4] = {k0.x, k0.y, k0.z, k0.w};
      const double vv[4] = {v0.x, v0.y, v1.x, v1.y};
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (kk[u] == mykey) part = part + vv[u];
    }
    double tot = part;
    for (int q = 1; q < G; q++) tot = tot + __shfl(part, r + q * W, 64);
    acc[0] = tot;  // lanes below W (all of rank < nd among them) hold their column's total
  } else {
    wave_bitonic_keys<4>(key, lane);
    for (int t = 0; t < n8; t += 4) {
      const int4 k0 = *reinterpret_cast<const int4*>(rk + t);
      const double2 v0 = *reinterpret_cast<const double2*>(rv + t);
      const double2 v1 = *reinterpret_cast<const double2*>(rv + t + 2);
      const int kk[4] = {k0.x, k0.y, k0.z, k0.w};
      const double vv[4] = {v0.x, v0.y, v1.x, v1.y};
#pragma unroll
      for (int u = 0; u < 4; u++) {
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (kk[u] == key[r]) acc[r] = acc[r] + vv[u];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int t = lane + 64 * r;
    if (t < nd) {
      tkey[o + t] = key[r];
      tval[o + t] = acc[r];
    }
  }
  return nd;
}

template <int LOC, int CAP, int WAVES, bool SLOTS = false>
__global__ void __launch_bounds__(64 * WAVES)
k_mass_rows_gather(long nn, const long* __restrict__ nptr, const int* __restrict__ nlist,
                   const long* __restrict__ indptr, const int* __restrict__ cols,
                   const int* __restrict__ counts, const double* __restrict__ blocks,
                   const int* __restrict__ enlb, const long* __restrict__ raw_off, int* __restrict__ tkey,
                   double* __restrict__ tval, long* __restrict__ ucount, int* __restrict__ overflow) {
  __shared__ __align__(16) int s_rk[WAVES][CAP + 8];
  __shared__ __align__(16) double s_rv[WAVES][CAP + 8];
  __shared__ int s_ht[WAVES][kHashSlots];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int* rk = s_rk[wv];
  double* rv = s_rv[wv];
  int* ht = s_ht[wv];
  const long nwaves = (long)gridDim.x * WAVES;
  // the same trip count and barrier sequence for every wave of the block
  for (long base = (long)blockIdx.x * WAVES; base < nn; base += nwaves) {
    const long i = base + wv;
    const bool valid = i < nn;
    // 1. gather: up to 64 incident elements at a time, one per lane; their
    // pair ranges are laid end to end and the lanes walk that flat list.  With
    // SLOTS the clip pass already wrote the row in that order (empty pairs as
    // node -1): it is read as is.
    int n = 0;
    if (SLOTS && valid) {
      const long r0 = raw_off[i], r1 = raw_off[i + 1];
      for (long rb = r0; rb < r1; rb += 64) {
        const long r = rb + lane;
        const int key = r < r1 ? tkey[r] : -1;
        const double v = key >= 0 ? tval[r] : 0.0;
        const unsigned long long m = __ballot(key >= 0);
        if (key >= 0) {
          const int pos = n + __popcll(m & ((1ULL << lane) - 1));
          rk[pos] = key;
          rv[pos] = v;
        }
        n += __popcll(m);
      }
    }
    if (!SLOTS && valid) {
      const long x0 = nptr[i], x1 = nptr[i + 1];
      for (long xb = x0; xb < x1; xb += 64) {
        const long x = xb + lane;
        long p0 = 0;
        int np = 0, li = 0;
        if (x < x1) {
          const int el = nlist[x];
          const int e = el / LOC;
          li = el - e * LOC;
          p0 = indptr[e];
          np = (int)(indptr[e + 1] - p0);
        }
        int incl = np;
        for (int o = 1; o < 64; o <<= 1) {
          const int v = __shfl_up(incl, o, 64);
          if (lane >= o) incl += v;
        }
        const int excl = incl - np;
        const int total = __shfl(incl, 63, 64);
        for (int f0 = 0; f0 < total; f0 += 64) {
          const int f = f0 + lane;
          // owner: the last lane whose range starts at or before f
          int own = 0;
          for (int step = 32; step > 0; step >>= 1) {
            const int cand = own + step;
            const int ev = __shfl(excl, cand < 64 ? cand : 63, 64);
            if (cand < 64 && ev <= f) own = cand;
          }
          const long p = __shfl(p0, own, 64) + (f - __shfl(excl, own, 64));
          const int lio = __shfl(li, own, 64);
          const bool ne = f < total && counts[p] > 0;
          const unsigned long long m = __ballot(ne);
          if (ne) {
            const int pos = n + LOC * __popcll(m & ((1ULL << lane) - 1));
            const long b = cols[p];
#pragma unroll
            for (int j = 0; j < LOC; j++) {
              rk[pos + j] = enlb[b * LOC + j];
              rv[pos + j] = blocks[p * LOC * LOC + lio * LOC + j];
            }
          }
          n += LOC * __popcll(m);
        }
      }
    }
    if (lane < 8) rk[n + lane] = -2;  // tail of the last 8-entry group
    __syncthreads();
    // 2-4. distinct columns, sorted, summed in gather order (every wave of the
    // block calls this, valid row or not, so that the barriers match)
    const int nd = row_reduce(rk, rv, valid ? n : 0, ht, lane, tkey, tval, valid ? raw_off[i] : 0);
    if (valid && lane == 0) {
      if (nd < 0) atomicOr(overflow, 1);
      else ucount[i] = nd;
    }
    __syncthreads();  // the LDS rows are refilled next
  }
}

// Rows laid out by the clip pass (2D): one wave per block, rows i, i + grid,
// ...; while a row is reduced, the next row's entries (up to 64 * PF) and the
// one after's bounds are already being loaded.  Rows hold at most CAP entries
// and the hash table 1 << HBITS slots: the smaller variant takes less LDS, so
// more waves share a CU.
template <int CAP, int PF, int HBITS>
__global__ void __launch_bounds__(64)
k_mass_rows_slots(long nn, const long* __restrict__ raw_off, int* __restrict__ tkey, double* __restrict__ tval,
                  long* __restrict__ ucount, int* __restrict__ overflow) {
  static_assert(HBITS >= 8 && HBITS <= 9, "hash table of 256 or 512 slots, room for kMaxDistinct columns");
  __shared__ __align__(16) int rk[CAP + 8];
  __shared__ __align__(16) double rv[CAP + 8];
  __shared__ int ht[1 << HBITS];
  const int lane = threadIdx.x;
  const long stride = gridDim.x;
  long i = blockIdx.x;
  if (i >= nn) return;
  long a0 = raw_off[i], a1 = raw_off[i + 1];
  int ck[PF];
  double cv[PF];
#pragma unroll
  for (int q = 0; q < PF; q++) {
    const long r = a0 + lane + 64 * q;
    ck[q] = r < a1 ? tkey[r] : -1;
    cv[q] = ck[q] >= 0 ? tval[r] : 0.0;
  }
  long j = i + stride;
  long b0 = 0, b1 = 0;
  if (j < nn) {
    b0 = raw_off[j];
    b1 = raw_off[j + 1];
  }
  for (; i < nn; i = j, j += stride) {
    // row i into LDS, skipping empty pairs (node -1), in entry order
    int n = 0;
#pragma unroll
    for (int q = 0; q < PF; q++) {
      const unsigned long long m = __ballot(ck[q] >= 0);
      if (ck[q] >= 0) {
        const int pos = n + __popcll(m & ((1ULL << lane) - 1));
        rk[pos] = ck[q];
        rv[pos] = cv[q];
      }
      n += __popcll(m);
    }
    for (long rb = a0 + 64 * PF; rb < a1; rb += 64) {  // the rest of a long row
      const long r = rb + lane;
      const int key = r < a1 ? tkey[r] : -1;
      const double v = key >= 0 ? tval[r] : 0.0;
      const unsigned long long m = __ballot(key >= 0);
      if (key >= 0) {
        const int pos = n + __popcll(m & ((1ULL << lane) - 1));
        rk[pos] = key;
        rv[pos] = v;
      }
      n += __popcll(m);
    }
    // next row's entries and the bounds of the one after
    const long jn = j + stride;
    long c0 = 0, c1 = 0;
    if (j < nn) {
#pragma unroll
      for (int q = 0; q < PF; q++) {
        const long r = b0 + lane + 64 * q;
        ck[q] = r < b1 ? tkey[r] : -1;
        cv[q] = ck[q] >= 0 ? tval[r] : 0.0;
      }
      if (jn < nn) {
        c0 = raw_off[jn];
        c1 = raw_off[jn + 1];
      }
    }
    if (lane < 8) rk[n + lane] = -2;  // tail of the last 8-entry group
    __syncthreads();
    const int nd = row_reduce<HBITS>(rk, rv, n, ht, lane, tkey, tval, a0);
    if (nd < 0) {
      if (lane == 0) atomicOr(overflow, 1);
    } else if (lane == 0) {
      ucount[i] = nd;
    }
    __syncthreads();
    a0 = b0;
    a1 = b1;
    b0 = c0;
    b1 = c1;
  }
}

// Per-pair blocks from the row slots (2D), for the sorted assembly when the
// count pass wrote rows only: block[p][li][j] is row (a, li)'s slot j of the
// pair; empty pairs get zeros.
__global__ void __launch_bounds__(256)
k_blocks_from_rows(long npairs, const int* __restrict__ rows, const long* __restrict__ indptr,
                   const int* __restrict__ counts, const long* __restrict__ rbase, const double* __restrict__ rval,
                   double* __restrict__ blocks) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const long a = rows[p];
  const bool ne = counts[p] > 0;
  const long r = (p - indptr[a]) * 3;
#pragma unroll
  for (int li = 0; li < 3; li++) {
    const long o = rbase[a * 3 + li] + r;
#pragma unroll
    for (int j = 0; j < 3; j++) blocks[p * 9 + li * 3 + j] = ne ? rval[o + j] : 0.0;
  }
}

static void blocks_from_rows(Context& C, const DevMesh& A, libsupermesh_gpu_result* out) {
  if (out->blocks) return;
  const long npairs = out->npairs;
  double* blocks = C.blocks.as<double>((size_t)(npairs > 0 ? npairs : 1) * 9);
  if (npairs > 0)
    hipLaunchKernelGGL(k_blocks_from_rows, dim3(grid_for(npairs, 256)), dim3(256), 0, C.stream, npairs, out->rows,
                       out->indptr, out->counts, static_cast<const long*>(C.mrbase.ptr),
                       static_cast<const double*>(C.mvals.ptr), blocks);
  LSM_HIP(hipGetLastError());
  out->blocks = blocks;
}

// After the clip pass filled the row slots (mass_prepare_rows).
static void assemble_mass_slots(Context& C, const DevMesh& A, const DevMesh& B, libsupermesh_gpu_result* out) {
  const long nn = A.nnodes;
  long* mindptr = C.mindptr.as<long>(nn + 1);
  LSM_HIP(hipMemsetAsync(mindptr, 0, sizeof(long) * (nn + 1), C.stream));
  out->mass_nnz = 0;
  out->mass_indptr = mindptr;
  out->mass_indices = static_cast<const int*>(C.mindices.get(4));
  out->mass_values = static_cast<const double*>(C.uvals.get(8));
  long* nptr = static_cast<long*>(C.nuniq.ptr);
  const long* raw_off = nptr + 2 * (nn + 1);
  int* tkey = static_cast<int*>(C.mkeys.ptr);
  double* tval = static_cast<double*>(C.mvals.ptr);
  long* ucount = C.mvals_alt.as<long>(nn + 1);
  int* ovf = C.scalars.as<int>(64) + 4;
  LSM_HIP(hipMemsetAsync(ovf, 0, sizeof(int), C.stream));
  // the longest row, copied back during the clip (the clip synchronised since)
  C.mass_maxraw = C.hstat[1];
  LSM_REQUIRE(C.mass_maxraw >= 0, LIBSUPERMESH_GPU_ERROR_STATE, "mass row bound not read back");
  if (C.mass_maxraw > 2048) {  // rows too long for the row kernels' LDS
    blocks_from_rows(C, A, out);
    assemble_mass_sorted(C, A, B, out->npairs, out);
    return;
  }
  int cus = 0;
  LSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, C.device));
  const long ncu = cus > 0 ? cus : 256;
  const dim3 grid((unsigned)std::max<long>(1, std::min<long>(ncu * 16, nn)));
  // one wave per block, a few rows each: many short-lived blocks hide the row
  // loads better than a resident grid that loops (1024^2 workload: 1.00 ms
  // with 512 blocks per CU, 1.09 ms with 64, 1.33 ms with 28)
  const dim3 short_grid((unsigned)std::max<long>(1, std::min<long>(ncu * 512, nn)));
  if (C.mass_maxraw <= 320)
    hipLaunchKernelGGL((k_mass_rows_slots<320, 4, 8>), short_grid, dim3(64), 0, C.stream, nn, raw_off, tkey, tval,
                       ucount, ovf);
  else if (C.mass_maxraw <= 512)
    hipLaunchKernelGGL((k_mass_rows_slots<512, 4, 9>), dim3(grid.x * 4), dim3(64), 0, C.stream, nn, raw_off, tkey,
                       tval, ucount, ovf);
  else
    hipLaunchKernelGGL((k_mass_rows_gather<3, 2048, 2, true>), grid, dim3(128), 0, C.stream, nn, nptr, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr, raw_off, tkey, tval, ucount, ovf);
  LSM_HIP(hipGetLastError());
  // row lengths scanned before the overflow flag is known: one read-back for
  // both (an overflowing row leaves the scan meaningless, and it is unused)
  inclusive_scan_long(C, ucount, mindptr + 1, nn);
  C.hstat[2] = 0;
  C.hstat[3] = 0;
  LSM_HIP(hipMemcpyAsync(C.hstat + 2, ovf, sizeof(int), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipMemcpyAsync(C.hstat + 3, mindptr + nn, sizeof(long), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  if (*reinterpret_cast<const int*>(C.hstat + 2)) {  // a row with more distinct columns than the kernel holds
    blocks_from_rows(C, A, out);
    assemble_mass_sorted(C, A, B, out->npairs, out);
    return;
  }
  const long nnz = C.hstat[3];
  int* mind = C.mindices.as<int>(nnz);
  double* mval = C.uvals.as<double>(nnz);
  hipLaunchKernelGGL(k_mass_copy, dim3(grid_for(nn, 256 / kCopyLanes)), dim3(256), 0, C.stream, nn, raw_off,
                     mindptr, tkey, tval, mind, mval);
  LSM_HIP(hipGetLastError());
  out->mass_nnz = nnz;
  out->mass_indices = mind;
  out->mass_values = mval;
}

__global__ void __launch_bounds__(256)
k_elem_pairs(long ne, const long* __restrict__ indptr, int* __restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) out[e] = (int)(indptr[e + 1] - indptr[e]);
}

__global__ void __launch_bounds__(256)
k_elem_base(long nent, const int* __restrict__ nlist, const long* __restrict__ slot, long* __restrict__ rbase) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < nent) rbase[nlist[x]] = slot[x];
}

// 2D mass rows laid out before the clip: row i of A node i gets, for each
// incident element (e, li) in ascending order, LOC slots per candidate pair of
// e; the clip pass fills them.  Returns false (and lays out nothing) when the
// longest row would not fit the row kernel.
bool mass_prepare_rows(Context& C, const DevMesh& A, const DevMesh& B, const long* indptr, ClipArgs& g) {
  const int loc = A.dim + 1;
  const long na = A.nelements, nn = A.nnodes;
  if (A.dim != 2 || na == 0 || nn == 0) return false;
  unsigned int* ncount = reinterpret_cast<unsigned int*>(C.mflag.as<long>(2 * nn + 2));
  unsigned int* cursor = ncount + nn + 1;
  LSM_HIP(hipMemsetAsync(ncount, 0, sizeof(unsigned int) * (2 * nn + 2), C.stream));
  const long nent = na * loc;
  hipLaunchKernelGGL(k_node_count, dim3(grid_for(nent, 256)), dim3(256), 0, C.stream, na, loc, A.enlist_dev,
                     ncount);
  long* nptr = C.nuniq.as<long>(3 * (nn + 1));
  long* raw = nptr + (nn + 1);
  long* raw_off = raw + (nn + 1);
  LSM_HIP(hipMemsetAsync(nptr, 0, sizeof(long), C.stream));
  LSM_HIP(hipMemsetAsync(raw_off, 0, sizeof(long), C.stream));
  inclusive_scan_long(C, ncount, nptr + 1, nn);
  int* nlist = C.ukeys.as<int>(nent + 1);
  hipLaunchKernelGGL(k_node_fill, dim3(grid_for(nent, 256)), dim3(256), 0, C.stream, na, loc, A.enlist_dev, nptr,
                     cursor, nlist);
  int* elem_np = C.mkeys_alt.as<int>(na + 1);
  hipLaunchKernelGGL(k_elem_pairs, dim3(grid_for(na, 256)), dim3(256), 0, C.stream, na, indptr, elem_np);
  long* contrib = C.mslot.as<long>(2 * nent + 2);
  long* slot = contrib + nent + 1;
  hipLaunchKernelGGL(k_node_rows, dim3(grid_for(nn, 256)), dim3(256), 0, C.stream, nn, loc, nptr, nlist, elem_np,
                     contrib);
  LSM_HIP(hipMemsetAsync(slot, 0, sizeof(long), C.stream));
  inclusive_scan_long(C, contrib, slot + 1, nent);
  hipLaunchKernelGGL(k_row_bounds, dim3(grid_for(nn + 1, 256)), dim3(256), 0, C.stream, nn, nptr, slot, raw_off,
                     raw);
  long* rbase = C.mrbase.as<long>(nent + 1);
  hipLaunchKernelGGL(k_elem_base, dim3(grid_for(nent, 256)), dim3(256), 0, C.stream, nent, nlist, slot, rbase);
  LSM_HIP(hipGetLastError());
  // every (element, local vertex) holds LOC slots per candidate pair of the
  // element, so the slots number LOC^2 per pair; the longest row decides the
  // row kernel and is copied back for the assembly (read after the clip's
  // synchronisation, no wait here)
  const long total_raw = (long)loc * loc * g.npairs;
  {
    size_t bytes = 0;
    long* dmax = reinterpret_cast<long*>(C.scalars.as<long>(64)) + 8;
    LSM_HIP(rocprim::reduce(nullptr, bytes, raw, dmax, 0L, (size_t)nn, rocprim::maximum<long>(), C.stream));
    void* tmp = C.sort_tmp.get(bytes);
    LSM_HIP(rocprim::reduce(tmp, bytes, raw, dmax, 0L, (size_t)nn, rocprim::maximum<long>(), C.stream));
    C.hstat[1] = -1;
    LSM_HIP(hipMemcpyAsync(C.hstat + 1, dmax, sizeof(long), hipMemcpyDeviceToHost, C.stream));
  }
  C.mass_rows_ready = true;
  C.mass_maxraw = -1;
  g.rbase = rbase;
  g.indptr = indptr;
  g.enlb = B.enlist_dev;
  g.rkey = C.mkeys.as<int>(total_raw + 1);
  g.rval = C.mvals.as<double>(total_raw + 1);
  return true;
}

void assemble_mass(Context& C, const DevMesh& A, const DevMesh& B, long npairs,
                   libsupermesh_gpu_result* out) {
  if (C.mass_rows_ready) {
    C.mass_rows_ready = false;
    assemble_mass_slots(C, A, B, out);
    return;
  }
  const int loc = A.dim + 1;
  const long na = A.nelements, nn = A.nnodes;
  long* mindptr = C.mindptr.as<long>(nn + 1);
  LSM_HIP(hipMemsetAsync(mindptr, 0, sizeof(long) * (nn + 1), C.stream));
  out->mass_nnz = 0;
  out->mass_indptr = mindptr;
  out->mass_indices = static_cast<const int*>(C.mindices.get(4));
  out->mass_values = static_cast<const double*>(C.uvals.get(8));
  if (out->nonempty == 0 || nn == 0) {
    LSM_HIP(hipStreamSynchronize(C.stream));
    return;
  }
  // A's node -> (element, local vertex) lists
  unsigned int* ncount = reinterpret_cast<unsigned int*>(C.mflag.as<long>(2 * nn + 2));
  unsigned int* cursor = ncount + nn + 1;
  LSM_HIP(hipMemsetAsync(ncount, 0, sizeof(unsigned int) * (2 * nn + 2), C.stream));
  const long nent = na * loc;
  hipLaunchKernelGGL(k_node_count, dim3(grid_for(nent, 256)), dim3(256), 0, C.stream, na, loc,
                     A.enlist_dev, ncount);
  long* nptr = C.nuniq.as<long>(3 * (nn + 1));
  long* raw = nptr + (nn + 1);
  long* raw_off = raw + (nn + 1);
  LSM_HIP(hipMemsetAsync(nptr, 0, sizeof(long), C.stream));
  LSM_HIP(hipMemsetAsync(raw_off, 0, sizeof(long), C.stream));
  inclusive_scan_long(C, ncount, nptr + 1, nn);
  int* nlist = C.ukeys.as<int>(nent + 1);
  hipLaunchKernelGGL(k_node_fill, dim3(grid_for(nent, 256)), dim3(256), 0, C.stream, na, loc,
                     A.enlist_dev, nptr, cursor, nlist);
  int* elem_ne = C.mkeys_alt.as<int>(na + 1);
  hipLaunchKernelGGL(k_elem_nonempty, dim3(grid_for(na, 256)), dim3(256), 0, C.stream, na, out->indptr,
                     out->counts, elem_ne);
  long* contrib = C.mslot.as<long>(2 * nent + 2);
  long* slot = contrib + nent + 1;
  hipLaunchKernelGGL(k_node_rows, dim3(grid_for(nn, 256)), dim3(256), 0, C.stream, nn, loc, nptr, nlist,
                     elem_ne, contrib);
  LSM_HIP(hipGetLastError());
  LSM_HIP(hipMemsetAsync(slot, 0, sizeof(long), C.stream));
  inclusive_scan_long(C, contrib, slot + 1, nent);
  hipLaunchKernelGGL(k_row_bounds, dim3(grid_for(nn + 1, 256)), dim3(256), 0, C.stream, nn, nptr, slot,
                     raw_off, raw);
  LSM_HIP(hipGetLastError());
  long maxraw = 0;
  {
    // the longest row decides the LDS capacity of the row kernel
    size_t bytes = 0;
    long* dmax = reinterpret_cast<long*>(C.scalars.as<long>(64)) + 8;
    LSM_HIP(rocprim::reduce(nullptr, bytes, raw, dmax, 0L, (size_t)nn, rocprim::maximum<long>(), C.stream));
    void* tmp = C.sort_tmp.get(bytes);
    LSM_HIP(rocprim::reduce(tmp, bytes, raw, dmax, 0L, (size_t)nn, rocprim::maximum<long>(), C.stream));
    LSM_HIP(hipMemcpyAsync(&maxraw, dmax, sizeof(long), hipMemcpyDeviceToHost, C.stream));
  }
  long total_raw = 0;
  LSM_HIP(hipMemcpyAsync(&total_raw, raw_off + nn, sizeof(long), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  if (maxraw > 2048) {  // rows too long for a register-resident sort: global sort path
    assemble_mass_sorted(C, A, B, npairs, out);
    return;
  }
  int* tkey = C.mkeys.as<int>(total_raw + 1);
  double* tval = C.mvals.as<double>(total_raw + 1);
  long* ucount = C.mvals_alt.as<long>(nn + 1);
  int* ovf = C.scalars.as<int>(64) + 4;
  LSM_HIP(hipMemsetAsync(ovf, 0, sizeof(int), C.stream));
  {
    // rows gathered, deduplicated and summed in one pass (LDS capacity by the
    // longest row)
    int cus = 0;
    LSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, C.device));
    const dim3 grid((unsigned)std::max<long>(1, std::min<long>((long)(cus > 0 ? cus : 256) * 16, nn)));
#define LSM_ROWS(L, CAPV, W)                                                                           \
  hipLaunchKernelGGL((k_mass_rows_gather<L, CAPV, W>), grid, dim3(64 * W), 0, C.stream, nn, nptr, nlist,  \
                     out->indptr, out->indices, out->counts, out->blocks, B.enlist_dev, raw_off, tkey,     \
                     tval, ucount, ovf)
    if (loc == 3) {
      if (maxraw <= 512) LSM_ROWS(3, 512, 4);
      else LSM_ROWS(3, 2048, 2);
    } else {
      if (maxraw <= 512) LSM_ROWS(4, 512, 4);
      else LSM_ROWS(4, 2048, 2);
    }
#undef LSM_ROWS
  }
  LSM_HIP(hipGetLastError());
  int hovf = 0;
  LSM_HIP(hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  if (hovf) {  // a row with more distinct columns than the row kernel holds
    assemble_mass_sorted(C, A, B, npairs, out);
    return;
  }
  inclusive_scan_long(C, ucount, mindptr + 1, nn);
  long nnz = 0;
  LSM_HIP(hipMemcpyAsync(&nnz, mindptr + nn, sizeof(long), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  int* mind = C.mindices.as<int>(nnz);
  double* mval = C.uvals.as<double>(nnz);
  hipLaunchKernelGGL(k_mass_copy, dim3(grid_for(nn, 256 / kCopyLanes)), dim3(256), 0, C.stream, nn, raw_off, mindptr, tkey,
                     tval, mind, mval);
  LSM_HIP(hipGetLastError());
  LSM_HIP(hipStreamSynchronize(C.stream));
  out->mass_nnz = nnz;
  out->mass_indices = mind;
  out->mass_values = mval;
}

}  // namespace lsm
