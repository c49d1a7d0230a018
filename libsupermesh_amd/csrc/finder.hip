// finder.hip -- candidate pair search on the device.
//
// Reference behaviour (what the result must equal): the quadtree / octree
// finder (src/Quadtree_Intersection_Finder.F90, src/Octree_...F90) returns,
// for every A element, the B elements whose bounding boxes overlap its own
// strictly in every dimension (bboxes_intersect, Quadtree_...F90:411-420) --
// the same set as brute force (src/Intersection_Finder.F90:1032-1050).  The
// R-tree finder uses closed intervals.  Both predicates run here in FP64 on
// exact min/max boxes.
//
// Design: B's element boxes are sorted by the leading 32 bits of the Morton code of their
// centres and grouped into an implicit F-ary hierarchy (node i of level l
// covers children [F*i, F*i+F) of level l-1; level 0 = elements).  Internal
// boxes are exact unions, so pruning with the same predicate never loses a
// pair.  A elements are visited in Morton order, one per lane, with a
// stackless depth-first walk (level, index) that needs no per-lane stack.
// A count pass sizes the CSR rows, a fill pass writes them, and each row is
// then sorted by B id so the output is canonical.
#include <cstring>
#include <utility>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "lsm_geometry.hpp"
#include "lsm_runtime.hpp"

namespace lsm {

constexpr int kFanout = 8;
constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Mesh staging
// ---------------------------------------------------------------------------

// Elements are simplices (LOC = DIM + 1) or, for the candidate finder only,
// quads / hexahedra (LOC = 2^DIM): those get coordinates, boxes and scale.
template <int DIM, int LOC>
__global__ void __launch_bounds__(kBlock)
k_stage(long ne, long nnodes, const double* __restrict__ pos, const int* __restrict__ enl,
        double* __restrict__ coords, double* __restrict__ bbox, double* __restrict__ sp,
        double* __restrict__ scale, double* __restrict__ planes, double* __restrict__ affine,
        int* __restrict__ bad) {
  constexpr bool SIMPLEX = LOC == DIM + 1;
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double c[LOC * DIM];
#pragma unroll
  for (int j = 0; j < LOC; j++) {
    long v = enl[e * LOC + j];
    if (v < 0 || v >= nnodes) {
      atomicOr(bad, 1);
      v = 0;
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) c[j * DIM + d] = pos[v * DIM + d];
  }
  double sc = 1.0;
#pragma unroll
  for (int k = 0; k < LOC * DIM; k++) {
    coords[e * LOC * DIM + k] = c[k];
    sc = fmax_f(sc, __builtin_fabs(c[k]));
  }
  scale[e] = sc;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    double lo = c[d], hi = c[d];
#pragma unroll
    for (int j = 1; j < LOC; j++) {
      lo = fmin_f(lo, c[j * DIM + d]);
      hi = fmax_f(hi, c[j * DIM + d]);
    }
    bbox[(long)d * ne + e] = lo;
    bbox[(long)(DIM + d) * ne + e] = hi;
  }
  if constexpr (!SIMPLEX) {
    sp[e] = 0.0;
    return;
  } else {
  double m;
  if constexpr (DIM == 2) m = triangle_area(c);
  else m = tet_volume_tet(c);
  sp[e] = fspacing(m);
  if constexpr (DIM == 3) {
   if (planes) {
    Planes P;
    tet_planes(c, P);
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
      for (int k = 0; k < 3; k++) planes[e * 16 + 4 * i + k] = P.n[i][k];
      planes[e * 16 + 4 * i + 3] = P.c[i];
    }
   }
  }
  if (affine) {
    // Barycentric coordinates as affine maps: lambda_{1..dim}(x) = Minv (x - v1),
    // lambda_0 = 1 - sum.  Stored as [v1 (dim), Minv (dim x dim, row-major)].
    double* out = affine + e * (DIM + DIM * DIM);
    if constexpr (DIM == 2) {
      double e1x = c[2] - c[0], e1y = c[3] - c[1];
      double e2x = c[4] - c[0], e2y = c[5] - c[1];
      double det = e2y * e1x - (-e2x) * (-e1y);
      out[0] = c[0];
      out[1] = c[1];
      out[2] = e2y / det;
      out[3] = -e2x / det;
      out[4] = -e1y / det;
      out[5] = e1x / det;
    } else {
      double m00 = c[3] - c[0], m01 = c[6] - c[0], m02 = c[9] - c[0];
      double m10 = c[4] - c[1], m11 = c[7] - c[1], m12 = c[10] - c[1];
      double m20 = c[5] - c[2], m21 = c[8] - c[2], m22 = c[11] - c[2];
      double c00 = m11 * m22 - m12 * m21, c01 = m02 * m21 - m01 * m22, c02 = m01 * m12 - m02 * m11;
      double c10 = m12 * m20 - m10 * m22, c11 = m00 * m22 - m02 * m20, c12 = m02 * m10 - m00 * m12;
      double c20 = m10 * m21 - m11 * m20, c21 = m01 * m20 - m00 * m21, c22 = m00 * m11 - m01 * m10;
      double det = m00 * c00 + m01 * c10 + m02 * c20;
      out[0] = c[0];
      out[1] = c[1];
      out[2] = c[2];
      out[3] = c00 / det;
      out[4] = c01 / det;
      out[5] = c02 / det;
      out[6] = c10 / det;
      out[7] = c11 / det;
      out[8] = c12 / det;
      out[9] = c20 / det;
      out[10] = c21 / det;
      out[11] = c22 / det;
    }
  }
  }
}

int* stage_flag(Context& C) { return C.scalars.as<int>(64); }

void stage_mesh(Context& C, DevMesh& M, int dim, int loc, long nnodes, long ne,
                const double* positions, const int* enlist, bool need_planes, bool need_affine,
                bool defer_check) {
  LSM_REQUIRE(dim == 2 || dim == 3, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension");
  const bool simplex = loc == dim + 1;
  LSM_REQUIRE(simplex || loc == (1 << dim), LIBSUPERMESH_GPU_ERROR_INVALID,
              "Unsupported element type: simplices, quads or hexahedra only");
  LSM_REQUIRE(simplex || (!need_planes && !need_affine), LIBSUPERMESH_GPU_ERROR_INVALID,
              "Unsupported element type: clipping and mass take simplices only");
  LSM_REQUIRE(nnodes >= 0 && ne >= 0, LIBSUPERMESH_GPU_ERROR_INVALID, "Negative mesh size");
  LSM_REQUIRE(ne < (1L << 31) && nnodes < (1L << 31), LIBSUPERMESH_GPU_ERROR_OVERFLOW,
              "Mesh larger than 2^31 elements or nodes");
  LSM_REQUIRE(ne == 0 || nnodes > 0, LIBSUPERMESH_GPU_ERROR_INVALID, "No nodes");
  M.dim = dim;
  M.loc = loc;
  M.nnodes = nnodes;
  M.nelements = ne;
  M.ngroups = 0;
  M.enlist_dev = enlist;
  double* coords = M.coords.as<double>((size_t)ne * loc * dim);
  double* bbox = M.bbox.as<double>((size_t)ne * 2 * dim);
  double* sp = M.meas_sp.as<double>((size_t)ne);
  double* scale = M.scale.as<double>((size_t)ne);
  double* planes = need_planes && dim == 3 ? M.planes.as<double>((size_t)ne * 16) : nullptr;
  double* affine = need_affine ? M.affine.as<double>((size_t)ne * (dim + dim * dim)) : nullptr;
  int* bad = stage_flag(C);
  if (!defer_check) LSM_HIP(hipMemsetAsync(bad, 0, sizeof(int), C.stream));
  if (ne > 0) {
#define LSM_STAGE(D, L)                                                                         \
  hipLaunchKernelGGL((k_stage<D, L>), dim3(grid_for(ne, kBlock)), dim3(kBlock), 0, C.stream, ne,  \
                     nnodes, positions, enlist, coords, bbox, sp, scale, planes, affine, bad)
    if (dim == 2) {
      if (simplex) LSM_STAGE(2, 3);
      else LSM_STAGE(2, 4);
    } else {
      if (simplex) LSM_STAGE(3, 4);
      else LSM_STAGE(3, 8);
    }
#undef LSM_STAGE
    LSM_HIP(hipGetLastError());
  }
  if (defer_check) return;
  int hbad = 0;
  LSM_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  LSM_REQUIRE(hbad == 0, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid node: element-node index out of range");
}

__global__ void __launch_bounds__(kBlock)
k_groups(const long* __restrict__ offsets, int* __restrict__ group) {
  const long g = blockIdx.x;
  for (long e = offsets[g] + threadIdx.x; e < offsets[g + 1]; e += kBlock) group[e] = (int)g;
}

void stage_groups(Context& C, DevMesh& M, long ngroups, const long* offsets) {
  M.ngroups = 0;
  if (ngroups <= 0) return;
  LSM_REQUIRE(ngroups < (1L << 24), LIBSUPERMESH_GPU_ERROR_INVALID, "too many meshes in a batch");
  std::vector<long> h((size_t)ngroups + 1);
  LSM_HIP(hipMemcpyAsync(h.data(), offsets, sizeof(long) * (ngroups + 1), hipMemcpyDeviceToHost,
                         C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  LSM_REQUIRE(h[0] == 0 && h[ngroups] == M.nelements, LIBSUPERMESH_GPU_ERROR_INVALID,
              "batch element offsets must run from 0 to the element count");
  for (long g = 0; g < ngroups; g++)
    LSM_REQUIRE(h[g + 1] >= h[g], LIBSUPERMESH_GPU_ERROR_INVALID, "batch element offsets decrease");
  int* group = M.group.as<int>((size_t)M.nelements);
  hipLaunchKernelGGL(k_groups, dim3((unsigned)ngroups), dim3(kBlock), 0, C.stream, offsets, group);
  LSM_HIP(hipGetLastError());
  M.ngroups = ngroups;
}

// ---------------------------------------------------------------------------
// Morton codes and bounds
// ---------------------------------------------------------------------------

template <int DIM>
__global__ void __launch_bounds__(kBlock)
k_center_bounds(long ne, const double* __restrict__ bbox, double* __restrict__ partial) {
  __shared__ double red[2 * DIM][kBlock];
  double lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    lo[d] = __builtin_inf();
    hi[d] = -__builtin_inf();
  }
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      double c = 0.5 * (bbox[(long)d * ne + e] + bbox[(long)(DIM + d) * ne + e]);
      lo[d] = fmin_f(lo[d], c);
      hi[d] = fmax_f(hi[d], c);
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    red[d][threadIdx.x] = lo[d];
    red[DIM + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        red[d][threadIdx.x] = fmin_f(red[d][threadIdx.x], red[d][threadIdx.x + s]);
        red[DIM + d][threadIdx.x] = fmax_f(red[DIM + d][threadIdx.x], red[DIM + d][threadIdx.x + s]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 2 * DIM; k++) partial[blockIdx.x * 2 * DIM + k] = red[k][0];
  }
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_bounds_final(int nparts, double* __restrict__ partial) {
  __shared__ double red[2 * DIM][kBlock];
  double b[2 * DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    b[d] = __builtin_inf();
    b[DIM + d] = -__builtin_inf();
  }
  for (int p = threadIdx.x; p < nparts; p += kBlock) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      b[d] = fmin_f(b[d], partial[p * 2 * DIM + d]);
      b[DIM + d] = fmax_f(b[DIM + d], partial[p * 2 * DIM + DIM + d]);
    }
  }
#pragma unroll
  for (int k = 0; k < 2 * DIM; k++) red[k][threadIdx.x] = b[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        red[d][threadIdx.x] = fmin_f(red[d][threadIdx.x], red[d][threadIdx.x + s]);
        red[DIM + d][threadIdx.x] = fmax_f(red[DIM + d][threadIdx.x], red[DIM + d][threadIdx.x + s]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 2 * DIM; k++) partial[k] = red[k][0];
  }
}

__device__ __forceinline__ uint64_t spread2(uint64_t x) {  // 32 bits -> even bits
  x &= 0xffffffffULL;
  x = (x | (x << 16)) & 0x0000ffff0000ffffULL;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffULL;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0fULL;
  x = (x | (x << 2)) & 0x3333333333333333ULL;
  x = (x | (x << 1)) & 0x5555555555555555ULL;
  return x;
}

__device__ __forceinline__ uint64_t spread3(uint64_t x) {  // 21 bits -> every third bit
  x &= 0x1fffffULL;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}

// Sort key of element e: its Morton code, or in a batch the mesh id in the
// top gbits bits above the code's leading bits, so that each mesh is one
// contiguous run of the order.
template <int DIM>
__global__ void __launch_bounds__(kBlock)
k_morton(long ne, const double* __restrict__ bbox, const double* __restrict__ bounds,
         const int* __restrict__ group, int gbits, uint32_t* __restrict__ keys,
         int* __restrict__ vals) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  constexpr int BITS = DIM == 2 ? 32 : 21;
  const double maxq = (double)((1ULL << BITS) - 1);
  uint64_t code = 0;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    double lo = bounds[d], hi = bounds[DIM + d];
    double c = 0.5 * (bbox[(long)d * ne + e] + bbox[(long)(DIM + d) * ne + e]);
    double t = hi > lo ? (c - lo) / (hi - lo) : 0.0;
    t = fmin_f(fmax_f(t, 0.0), 1.0);
    uint64_t q = (uint64_t)(t * maxq);
    code |= (DIM == 2 ? spread2(q) : spread3(q)) << d;
  }
  constexpr int KEYBITS = DIM == 2 ? 64 : 63;
  if (group && gbits > 0) code = ((uint64_t)group[e] << (KEYBITS - gbits)) | (code >> gbits);
  // the leading 32 bits (16 per axis in 2D, 10-11 in 3D, below a batch's
  // mesh id) order the elements finely enough for the hierarchy; ties keep
  // element order (the sort is stable), and the candidate sets do not depend
  // on the order at all
  keys[e] = (uint32_t)(code >> (KEYBITS - 32));
  vals[e] = (int)e;
}


// Sorts the elements of M by Morton code; returns the permutation in `perm`.
static void morton_order(Context& C, const DevMesh& M, DevBuf& perm) {
  long ne = M.nelements;
  int dim = M.dim;
  int nparts = (int)std::min<long>(1024, std::max<long>(1, (ne + kBlock - 1) / kBlock));
  double* part = C.partials.as<double>((size_t)nparts * 6 + 6);
  uint32_t* keys = C.keys.as<uint32_t>(ne);
  uint32_t* keys2 = C.keys_alt.as<uint32_t>(ne);
  int* vals = C.vals.as<int>(ne);
  int* out = perm.as<int>(ne);
  const double* bbox = static_cast<const double*>(M.bbox.ptr);
  const int* group = M.ngroups > 0 ? static_cast<const int*>(M.group.ptr) : nullptr;
  int gbits = 0;
  while (group && (1L << gbits) < M.ngroups) gbits++;
  if (dim == 2) {
    hipLaunchKernelGGL(k_center_bounds<2>, dim3(nparts), dim3(kBlock), 0, C.stream, ne, bbox, part);
    hipLaunchKernelGGL(k_bounds_final<2>, dim3(1), dim3(kBlock), 0, C.stream, nparts, part);
    hipLaunchKernelGGL(k_morton<2>, dim3(grid_for(ne, kBlock)), dim3(kBlock), 0, C.stream, ne, bbox,
                       part, group, gbits, keys, vals);
  } else {
    hipLaunchKernelGGL(k_center_bounds<3>, dim3(nparts), dim3(kBlock), 0, C.stream, ne, bbox, part);
    hipLaunchKernelGGL(k_bounds_final<3>, dim3(1), dim3(kBlock), 0, C.stream, nparts, part);
    hipLaunchKernelGGL(k_morton<3>, dim3(grid_for(ne, kBlock)), dim3(kBlock), 0, C.stream, ne, bbox,
                       part, group, gbits, keys, vals);
  }
  LSM_HIP(hipGetLastError());
  const unsigned begin_bit = 0, end_bit = 32;
  size_t tmp_bytes = 0;
  LSM_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, vals, out, (size_t)ne, begin_bit,
                                    end_bit, C.stream));
  void* tmp = C.sort_tmp.get(tmp_bytes);
  LSM_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys2, vals, out, (size_t)ne, begin_bit, end_bit,
                                    C.stream));
}

// ---------------------------------------------------------------------------
// Implicit hierarchy
// ---------------------------------------------------------------------------

struct LevelInfo {
  int levels;
  long count[32];
  long offset[32];
};

// Level-0 boxes (node-major: [lo_0..lo_{dim-1}, hi_0..hi_{dim-1}]) in sorted order.
template <int DIM>
__global__ void __launch_bounds__(kBlock)
k_leaf_boxes(long ne, const double* __restrict__ bbox, const int* __restrict__ perm,
             double* __restrict__ boxes, const int* __restrict__ group, int* __restrict__ granges) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  long e = perm[i];
#pragma unroll
  for (int k = 0; k < 2 * DIM; k++) boxes[i * 2 * DIM + k] = bbox[(long)k * ne + e];
  if (group) {
    granges[2 * i] = group[e];
    granges[2 * i + 1] = group[e];
  }
}

template <int DIM>
__global__ void __launch_bounds__(kBlock)
k_parent_boxes(long nchild, long nparent, const double* __restrict__ child, double* __restrict__ parent,
               const int* __restrict__ cg, int* __restrict__ pg) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nparent) return;
  double b[2 * DIM];
  long c0 = i * kFanout;
  long c1 = c0 + kFanout < nchild ? c0 + kFanout : nchild;
#pragma unroll
  for (int k = 0; k < 2 * DIM; k++) b[k] = child[c0 * 2 * DIM + k];
  for (long c = c0 + 1; c < c1; c++) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      b[d] = fmin_f(b[d], child[c * 2 * DIM + d]);
      b[DIM + d] = fmax_f(b[DIM + d], child[c * 2 * DIM + DIM + d]);
    }
  }
#pragma unroll
  for (int k = 0; k < 2 * DIM; k++) parent[i * 2 * DIM + k] = b[k];
  if (cg) {
    // children are in mesh order: the first one holds the lowest id
    pg[2 * i] = cg[2 * c0];
    pg[2 * i + 1] = cg[2 * (c1 - 1) + 1];
  }
}

void build_bvh(Context& C, const DevMesh& B) {
  DevBVH& H = C.bvh;
  H.dim = B.dim;
  H.n = B.nelements;
  if (H.n == 0) {
    H.levels = 0;
    return;
  }
  morton_order(C, B, H.perm);
  long n = H.n;
  int L = 0;
  long total = 0;
  while (true) {
    H.count[L] = n;
    H.offset[L] = total;
    total += n;
    L++;
    if (n == 1) break;
    n = (n + kFanout - 1) / kFanout;
    LSM_REQUIRE(L < 31, LIBSUPERMESH_GPU_ERROR_OVERFLOW, "hierarchy too deep");
  }
  H.levels = L;
  H.total = total;
  int w = 2 * B.dim;
  double* boxes = H.boxes.as<double>((size_t)total * w);
  const double* bbox = static_cast<const double*>(B.bbox.ptr);
  const int* perm = static_cast<const int*>(H.perm.ptr);
  H.grouped = B.ngroups > 0;
  const int* group = H.grouped ? static_cast<const int*>(B.group.ptr) : nullptr;
  int* gr = H.grouped ? H.granges.as<int>((size_t)total * 2) : nullptr;
  if (B.dim == 2)
    hipLaunchKernelGGL(k_leaf_boxes<2>, dim3(grid_for(H.n, kBlock)), dim3(kBlock), 0, C.stream, H.n,
                       bbox, perm, boxes, group, gr);
  else
    hipLaunchKernelGGL(k_leaf_boxes<3>, dim3(grid_for(H.n, kBlock)), dim3(kBlock), 0, C.stream, H.n,
                       bbox, perm, boxes, group, gr);
  for (int l = 1; l < L; l++) {
    const double* ch = boxes + H.offset[l - 1] * w;
    double* pa = boxes + H.offset[l] * w;
    const int* cg = gr ? gr + 2 * H.offset[l - 1] : nullptr;
    int* pg = gr ? gr + 2 * H.offset[l] : nullptr;
    if (B.dim == 2)
      hipLaunchKernelGGL(k_parent_boxes<2>, dim3(grid_for(H.count[l], kBlock)), dim3(kBlock), 0,
                         C.stream, H.count[l - 1], H.count[l], ch, pa, cg, pg);
    else
      hipLaunchKernelGGL(k_parent_boxes<3>, dim3(grid_for(H.count[l], kBlock)), dim3(kBlock), 0,
                         C.stream, H.count[l - 1], H.count[l], ch, pa, cg, pg);
  }
  LSM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Traversal
// ---------------------------------------------------------------------------

template <int DIM, bool CLOSED>
__device__ __forceinline__ bool overlaps(const double* __restrict__ box, const double* qlo,
                                         const double* qhi) {
  bool hit = true;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    double lo = box[d], hi = box[DIM + d];
    // strict: b.max > a.min and b.min < a.max (Quadtree_...F90:411-420)
    hit = hit && (CLOSED ? (hi >= qlo[d] && lo <= qhi[d]) : (hi > qlo[d] && lo < qhi[d]));
  }
  return hit;
}

template <int DIM, bool CLOSED, bool FILL, bool GROUPED>
__global__ void __launch_bounds__(kBlock)
k_query(long na, const int* __restrict__ qperm, const double* __restrict__ abox, LevelInfo info,
        const double* __restrict__ boxes, const int* __restrict__ perm, long* __restrict__ row_count,
        const long* __restrict__ indptr, int* __restrict__ indices, int* __restrict__ rows,
        const int* __restrict__ agroup, const int* __restrict__ granges, long skip_upto) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= na) return;
  const int a = qperm[t];
  // fill pass after the single pass: only the rows that outgrew its scratch
  if (FILL && row_count[a] <= skip_upto) return;
  // in a batch, only nodes whose mesh range holds A's mesh are visited
  const int g = GROUPED ? agroup[a] : 0;
  double qlo[DIM], qhi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    qlo[d] = abox[(long)d * na + a];
    qhi[d] = abox[(long)(DIM + d) * na + a];
  }
  const int top = info.levels - 1;
  long base = FILL ? indptr[a] : 0;
  long cnt = 0;
  int l = top;
  long i = 0;
  while (true) {
    bool hit = false;
    if (i < info.count[l]) {
      hit = overlaps<DIM, CLOSED>(boxes + (info.offset[l] + i) * 2 * DIM, qlo, qhi);
      if (GROUPED) {
        const int* r = granges + 2 * (info.offset[l] + i);
        hit = hit && r[0] <= g && g <= r[1];
      }
    }
    if (hit) {
      if (l == 0) {
        if (FILL) {
          indices[base + cnt] = perm[i];
          rows[base + cnt] = a;
        }
        cnt++;
      } else {
        l--;
        i *= kFanout;
        continue;
      }
    }
    // advance to the next sibling, climbing while at the end of a group
    bool done = false;
    while (true) {
      if (l == top) {
        done = true;
        break;
      }
      if ((i % kFanout) != kFanout - 1 && i + 1 < info.count[l]) {
        i++;
        break;
      }
      i /= kFanout;
      l++;
    }
    if (done) break;
  }
  if (!FILL) row_count[a] = cnt;
}

// Boxes of the up to F children first .. first+F-1 of level K that overlap
// the query, as a bit mask.
template <int DIM, bool CLOSED, bool GROUPED>
__device__ __forceinline__ unsigned child_mask(const LevelInfo& info, const double* __restrict__ boxes,
                                               const int* __restrict__ granges, int K, long first,
                                               const double* qlo, const double* qhi, int g) {
  const long n = info.count[K];
  const long base = info.offset[K] + first;
  unsigned m = 0;
#pragma unroll
  for (int c = 0; c < kFanout; c++) {
    if (first + c < n) {
      bool hit = overlaps<DIM, CLOSED>(boxes + (base + c) * 2 * DIM, qlo, qhi);
      if (GROUPED) {
        const int* r = granges + 2 * (base + c);
        hit = hit && r[0] <= g && g <= r[1];
      }
      m |= (unsigned)hit << c;
    }
  }
  return m;
}

// OR of an 8-bit mask over the wave's active lanes (uniform result).
__device__ __forceinline__ unsigned wave_or8(unsigned x) {
  unsigned r = 0;
#pragma unroll
  for (int b = 0; b < 8; b++) r |= (__ballot((x >> b) & 1u) != 0ULL ? 1u : 0u) << b;
  return r;
}

// Single-pass search: the same depth-first order as k_query, but a node's
// children are tested together (their boxes are contiguous) and the walk
// keeps, per level, a byte of the hit children not yet visited; the node
// index is implicit (parent = child / F).  The first `cap` results of a query
// go to scratch (slot-major: scratch[k * na + t]); all are counted.  The wave
// walks one path: the union of its 64 queries' paths (consecutive queries are
// close in Morton order, so their paths mostly coincide).  Every lane tests the children of the wave's current node
// (the same boxes: one broadcast load per box, no divergent gathers) and
// keeps its own per-level hit masks, so its hits, and their order, are those
// of its own depth-first walk alone: each level's union mask is
// consumed lowest child first, which is each lane's own order.
template <int DIM, bool CLOSED, bool GROUPED>
__global__ void __launch_bounds__(kBlock)
k_query_wave(long na, const int* __restrict__ qperm, const double* __restrict__ abox, LevelInfo info,
             const double* __restrict__ boxes, const int* __restrict__ perm, int cap,
             int* __restrict__ scratch, long* __restrict__ row_count, const int* __restrict__ agroup,
             const int* __restrict__ granges) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= na) return;  // the rest of the wave walks on without these lanes
  const int a = qperm[t];
  const int g = GROUPED ? agroup[a] : 0;
  double qlo[DIM], qhi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    qlo[d] = abox[(long)d * na + a];
    qhi[d] = abox[(long)(DIM + d) * na + a];
  }
  long cnt = 0;
  auto emit = [&](long leaf) {
    if (cnt < cap) {
      // 2D: streaming stores keep the hierarchy's boxes in L2 (0.96 -> 0.93 ms
      // on tri 1024^2 vs 1000^2); in 3D they made the search slower
      if constexpr (DIM == 2) __builtin_nontemporal_store(perm[leaf], scratch + cnt * na + t);
      else scratch[cnt * na + t] = perm[leaf];
    }
    cnt++;
  };
  const int top = info.levels - 1;
  if (top == 0) {
    if (child_mask<DIM, CLOSED, GROUPED>(info, boxes, granges, 0, 0, qlo, qhi, g) & 1u) emit(0);
  } else {
    const bool root = child_mask<DIM, CLOSED, GROUPED>(info, boxes, granges, top, 0, qlo, qhi, g) & 1u;
    // per lane: masks of levels 1..16, a byte each; per wave: their union
    unsigned long long mlo = 0, mhi = 0, ulo = 0, uhi = 0;
    auto get = [](unsigned long long lo, unsigned long long hi, int L) -> unsigned {
      return L <= 8 ? (unsigned)(lo >> (8 * (L - 1))) & 0xffu : (unsigned)(hi >> (8 * (L - 9))) & 0xffu;
    };
    auto put = [](unsigned long long& lo, unsigned long long& hi, int L, unsigned m) {
      if (L <= 8) {
        const int sh = 8 * (L - 1);
        lo = (lo & ~(0xffULL << sh)) | ((unsigned long long)m << sh);
      } else {
        const int sh = 8 * (L - 9);
        hi = (hi & ~(0xffULL << sh)) | ((unsigned long long)m << sh);
      }
    };
    int L = top;
    long node = 0;
    {
      unsigned m = child_mask<DIM, CLOSED, GROUPED>(info, boxes, granges, top - 1, 0, qlo, qhi, g);
      if (!root) m = 0;
      put(mlo, mhi, L, m);
      put(ulo, uhi, L, wave_or8(m));
    }
    while (true) {
      const unsigned um = get(ulo, uhi, L);
      if (um == 0) {
        if (L == top) break;
        node /= kFanout;
        L++;
        continue;
      }
      if (L == 1) {  // children are elements: each lane its own hits
        unsigned m = get(mlo, mhi, 1);
        while (m) {
          const int c = __builtin_ctz(m);
          m &= m - 1;
          emit(node * kFanout + c);
        }
        put(mlo, mhi, 1, 0);
        put(ulo, uhi, 1, 0);
        continue;
      }
      const int c = __builtin_ctz(um);
      put(ulo, uhi, L, um & (um - 1));
      const unsigned mine = get(mlo, mhi, L);
      put(mlo, mhi, L, mine & ~(1u << c));
      const long child = node * kFanout + c;
      unsigned mm = child_mask<DIM, CLOSED, GROUPED>(info, boxes, granges, L - 2, child * kFanout, qlo, qhi, g);
      if (!((mine >> c) & 1u)) mm = 0;
      const unsigned umm = wave_or8(mm);
      if (umm) {
        node = child;
        L--;
        put(mlo, mhi, L, mm);
        put(ulo, uhi, L, umm);
      }
    }
  }
  row_count[a] = cnt;
}

// Rows that fit the scratch: copy them into the CSR arrays.  A wave takes 64
// consecutive queries and transposes their slot-major scratch through LDS in
// J-slot tiles, so that both the scratch reads and the row writes are
// contiguous across the wave.  The tile is 16 slots tall when that holds the
// longest row, else 32 (longer rows take several tiles): a small tile lets
// more waves share a CU (8 and 4 per SIMD; 2 with 64 slots, which measured
// slower on 3D rows).
template <int J>
__global__ void __launch_bounds__(kBlock)
k_rows_from_scratch(long na, const int* __restrict__ qperm, int cap, const int* __restrict__ scratch,
                    const long* __restrict__ row_count, const long* __restrict__ indptr,
                    int* __restrict__ indices, int* __restrict__ rows) {
  static_assert(J == 16 || J == 32, "tile height");
  __shared__ int tile[kBlock / 64][J][65];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long base = ((long)blockIdx.x * (kBlock / 64) + wv) * 64;
  if (base >= na) return;  // whole waves only; no block barriers below
  const long t = base + lane;
  int a = 0;
  long n = 0, o = 0;
  if (t < na) {
    a = qperm[t];
    n = row_count[a];
    o = indptr[a];
    if (n > cap) n = 0;  // left to the fill pass
  }
  long nmax = n;
  for (int off = 32; off > 0; off >>= 1) {
    const long v = __shfl_xor(nmax, off, 64);
    nmax = v > nmax ? v : nmax;
  }
  int (*T)[65] = tile[wv];
  for (long s0 = 0; s0 < nmax; s0 += J) {
    // tile[j][q] = slot s0 + j of query base + q
    const int nj = nmax - s0 < J ? (int)(nmax - s0) : J;
    for (int j = 0; j < nj; j++) {
      const long s = s0 + j;
      T[j][lane] = (s < cap && t < na) ? scratch[s * na + t] : 0;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // 64 / J queries at a time: lane l writes slot l % J of query q + l / J
    for (int q0 = 0; q0 < 64; q0 += 64 / J) {
      const int q = q0 + lane / J, sl = lane % J;
      const long nq = __shfl(n, q, 64), oq = __shfl(o, q, 64);
      const int aq = __shfl(a, q, 64);
      const long s = s0 + sl;
      if (s < nq) {
        indices[oq + s] = T[sl][q];
        rows[oq + s] = aq;
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  }
}

// Insertion sort of each row by B id (rows are short: tens of entries).
__global__ void __launch_bounds__(kBlock)
k_sort_rows(long na, const int* __restrict__ qperm, const long* __restrict__ indptr,
            int* __restrict__ indices) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= na) return;
  int a = qperm[t];
  long b0 = indptr[a], b1 = indptr[a + 1];
  for (long i = b0 + 1; i < b1; i++) {
    int v = indices[i];
    long j = i - 1;
    while (j >= b0 && indices[j] > v) {
      indices[j + 1] = indices[j];
      j--;
    }
    indices[j + 1] = v;
  }
}

long find_pairs(Context& C, const DevMesh& A, const DevMesh& B, int finder, int sort_rows) {
  long na = A.nelements;
  long* row_count = C.row_count.as<long>(na + 1);
  long* indptr = C.indptr.as<long>(na + 1);
  LSM_HIP(hipMemsetAsync(indptr, 0, sizeof(long) * (na + 1), C.stream));
  if (na == 0 || B.nelements == 0) {
    LSM_HIP(hipStreamSynchronize(C.stream));
    return 0;
  }
  LSM_REQUIRE(A.dim == B.dim, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension: meshes differ");
  const DevBVH& H = C.bvh;
  LevelInfo info;
  info.levels = H.levels;
  for (int l = 0; l < 32; l++) {
    info.count[l] = l < H.levels ? H.count[l] : 0;
    info.offset[l] = l < H.levels ? H.offset[l] : 0;
  }
  morton_order(C, A, C.qperm);
  const int* qperm = static_cast<const int*>(C.qperm.ptr);
  const double* abox = static_cast<const double*>(A.bbox.ptr);
  const double* boxes = static_cast<const double*>(H.boxes.ptr);
  const int* perm = static_cast<const int*>(H.perm.ptr);
  dim3 grid(grid_for(na, kBlock)), block(kBlock);
  LSM_REQUIRE((A.ngroups > 0) == H.grouped && (!H.grouped || A.ngroups == B.ngroups),
              LIBSUPERMESH_GPU_ERROR_INVALID, "batches of A and B hold different mesh counts");
  const int* agroup = H.grouped ? static_cast<const int*>(A.group.ptr) : nullptr;
  const int* granges = H.grouped ? static_cast<const int*>(H.granges.ptr) : nullptr;
#define LSM_QUERY(DIM, CL, FILL, SKIP)                                                              \
  do {                                                                                            \
    if (H.grouped)                                                                                \
      hipLaunchKernelGGL((k_query<DIM, CL, FILL, true>), grid, block, 0, C.stream, na, qperm, abox, \
                         info, boxes, perm, row_count, indptr, indices, rows, agroup, granges,    \
                         (long)(SKIP));                                                           \
    else                                                                                          \
      hipLaunchKernelGGL((k_query<DIM, CL, FILL, false>), grid, block, 0, C.stream, na, qperm,      \
                         abox, info, boxes, perm, row_count, indptr, indices, rows, agroup,        \
                         granges, (long)(SKIP));                                                  \
  } while (0)
#define LSM_MASK(DIM, CL)                                                                           \
  do {                                                                                            \
    if (H.grouped)                                                                                \
      hipLaunchKernelGGL((k_query_wave<DIM, CL, true>), grid, block, 0, C.stream, na, qperm, abox,  \
                         info, boxes, perm, cap, scratch, row_count, agroup, granges);            \
    else                                                                                          \
      hipLaunchKernelGGL((k_query_wave<DIM, CL, false>), grid, block, 0, C.stream, na, qperm, abox, \
                         info, boxes, perm, cap, scratch, row_count, agroup, granges);            \
  } while (0)
  int* indices = nullptr;
  int* rows = nullptr;
  bool closed = finder == LIBSUPERMESH_GPU_FINDER_RTREE;
  // single pass into a scratch of `cap` results per query; longer rows are
  // searched again by the fill pass
  const int cap = A.dim == 2 ? 32 : 160;
  LSM_REQUIRE(H.levels <= 17, LIBSUPERMESH_GPU_ERROR_OVERFLOW, "hierarchy too deep");
  int* scratch = C.fscratch.as<int>((size_t)cap * na);
  if (A.dim == 2) {
    if (closed) LSM_MASK(2, true);
    else LSM_MASK(2, false);
  } else {
    if (closed) LSM_MASK(3, true);
    else LSM_MASK(3, false);
  }
  LSM_HIP(hipGetLastError());
  size_t tmp_bytes = 0;
  LSM_HIP(rocprim::inclusive_scan(nullptr, tmp_bytes, row_count, indptr + 1, (size_t)na,
                                  rocprim::plus<long>(), C.stream));
  void* tmp = C.scan_tmp.get(tmp_bytes);
  LSM_HIP(rocprim::inclusive_scan(tmp, tmp_bytes, row_count, indptr + 1, (size_t)na,
                                  rocprim::plus<long>(), C.stream));
  // longest row: decides whether the fill pass is needed
  long* dmax = C.scalars.as<long>(64) + 16;
  {
    size_t bytes = 0;
    LSM_HIP(rocprim::reduce(nullptr, bytes, row_count, dmax, 0L, (size_t)na, rocprim::maximum<long>(), C.stream));
    void* t2 = C.sort_tmp.get(bytes);
    LSM_HIP(rocprim::reduce(t2, bytes, row_count, dmax, 0L, (size_t)na, rocprim::maximum<long>(), C.stream));
  }
  long hdr[2] = {0, 0};
  LSM_HIP(hipMemcpyAsync(&hdr[0], indptr + na, sizeof(long), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipMemcpyAsync(&hdr[1], dmax, sizeof(long), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  const long npairs = hdr[0], maxrow = hdr[1];
  LSM_REQUIRE(npairs < (1L << 40), LIBSUPERMESH_GPU_ERROR_OVERFLOW, "too many candidate pairs");
  indices = C.indices.as<int>(npairs);
  rows = C.rows.as<int>(npairs);
  if (npairs > 0) {
    const long tall = maxrow < cap ? maxrow : cap;  // the most slots a copied row holds
    if (tall <= 16)
      hipLaunchKernelGGL(k_rows_from_scratch<16>, grid, block, 0, C.stream, na, qperm, cap, scratch, row_count,
                         indptr, indices, rows);
    else
      hipLaunchKernelGGL(k_rows_from_scratch<32>, grid, block, 0, C.stream, na, qperm, cap, scratch, row_count,
                         indptr, indices, rows);
    if (maxrow > cap) {
      if (A.dim == 2) {
        if (closed) LSM_QUERY(2, true, true, cap);
        else LSM_QUERY(2, false, true, cap);
      } else {
        if (closed) LSM_QUERY(3, true, true, cap);
        else LSM_QUERY(3, false, true, cap);
      }
    }
    LSM_HIP(hipGetLastError());
    if (sort_rows) {
      if (maxrow <= 64) {
        hipLaunchKernelGGL(k_sort_rows, grid, block, 0, C.stream, na, qperm, indptr, indices);
      } else {
        // long rows (a coarse A element over a fine B mesh): segmented radix
        // sort of the rows, then the sorted copy becomes the indices buffer
        LSM_REQUIRE(npairs < (1L << 32) && na < (1L << 32), LIBSUPERMESH_GPU_ERROR_OVERFLOW,
                    "too many pairs to sort rows");
        int* sorted = C.vals_alt.as<int>(npairs);
        unsigned end_bit = 1;
        while (end_bit < 31 && (1L << end_bit) < B.nelements) end_bit++;
        size_t bytes = 0;
        LSM_HIP(rocprim::segmented_radix_sort_keys(nullptr, bytes, indices, sorted, (unsigned)npairs, (unsigned)na,
                                                   indptr, indptr + 1, 0u, end_bit, C.stream));
        void* t3 = C.sort_tmp.get(bytes);
        LSM_HIP(rocprim::segmented_radix_sort_keys(t3, bytes, indices, sorted, (unsigned)npairs, (unsigned)na,
                                                   indptr, indptr + 1, 0u, end_bit, C.stream));
        std::swap(C.indices, C.vals_alt);
      }
      LSM_HIP(hipGetLastError());
    }
  }
#undef LSM_MASK
#undef LSM_QUERY
  return npairs;
}

// indptr over A from the pairs sorted by A id (arow): indptr[r] = the first
// position of row r, indptr[na] = n
__global__ void __launch_bounds__(kBlock)
k_rows_to_indptr(long n, const unsigned* __restrict__ arow, long na, long* __restrict__ indptr) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const long lo = i == 0 ? 0 : (long)arow[i - 1] + 1;
  const long hi = i == n ? na : (long)arow[i];
  for (long r = lo; r <= hi; r++) indptr[r] = i;
}

long find_pairs_from_b(Context& C, const DevMesh& A, const DevMesh& B, int finder) {
  // B's elements query the hierarchy over A (built by the caller): a CSR over
  // B, in B id order, whose entries are A ids.  The overlap tests are
  // symmetric in the two boxes, so the pair set is the one A's queries find.
  const long npairs = find_pairs(C, B, A, finder, 0);
  const long na = A.nelements;
  long* indptr = C.indptr.as<long>(na + 1);
  if (npairs == 0) {
    LSM_HIP(hipMemsetAsync(indptr, 0, sizeof(long) * (na + 1), C.stream));
    return 0;
  }
  LSM_REQUIRE(npairs < (1L << 32), LIBSUPERMESH_GPU_ERROR_OVERFLOW, "too many candidate pairs to regroup");
  unsigned* a_sorted = C.keys_alt.as<unsigned>(npairs);
  int* b_sorted = C.vals_alt.as<int>(npairs);
  const unsigned* a_ids = static_cast<const unsigned*>(C.indices.ptr);
  const int* b_ids = static_cast<const int*>(C.rows.ptr);
  // regrouped by A id with a stable sort: each A row keeps the B id order of
  // the input, ascending
  unsigned end_bit = 1;
  while (end_bit < 31 && (1L << end_bit) < na) end_bit++;
  size_t bytes = 0;
  LSM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, a_ids, a_sorted, b_ids, b_sorted, (size_t)npairs, 0u, end_bit,
                                    C.stream));
  void* tmp = C.sort_tmp.get(bytes);
  LSM_HIP(rocprim::radix_sort_pairs(tmp, bytes, a_ids, a_sorted, b_ids, b_sorted, (size_t)npairs, 0u, end_bit,
                                    C.stream));
  hipLaunchKernelGGL(k_rows_to_indptr, dim3(grid_for(npairs + 1, kBlock)), dim3(kBlock), 0, C.stream, npairs,
                     a_sorted, na, indptr);
  LSM_HIP(hipGetLastError());
  std::swap(C.rows, C.keys_alt);     // the A id of each pair
  std::swap(C.indices, C.vals_alt);  // its B id, ascending within each row
  return npairs;
}

}  // namespace lsm
