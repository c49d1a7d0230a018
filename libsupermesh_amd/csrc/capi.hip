// capi.hip -- the C ABI: libsupermesh-c.h drop-in symbols and the bulk GPU
// entry points of libsupermesh_gpu.h.
//
// The reference exports void functions and aborts on bad input
// (src/Debug.F90:78-94).  Here the drop-in symbols keep their void
// signatures and never abort: failures are recorded and readable through
// libsupermesh_gpu_last_error() / _code().  The reference keeps the finder's
// CSR in module globals (src/C_bindings.F90:23-32) and frees it in
// get_output; the same sequence is kept, guarded by a mutex.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/libsupermesh-c.h"
#include "../../include/libsupermesh_gpu.h"
#include "lsm_geometry.hpp"
#include "lsm_runtime.hpp"
#include "predictor.hpp"

struct libsupermesh_gpu_context {
  lsm::Context ctx;
};

// A search hierarchy kept on the device for one-element queries: B staged
// and its hierarchy built once, in a context of its own.
struct libsupermesh_gpu_tree {
  lsm::Context ctx;
  int finder = 0;
  int dim = 0;
  std::vector<int> result;
};

namespace lsm {

static thread_local int tl_code = 0;
static thread_local std::string tl_msg;

void set_last_error(int code, const char* msg) {
  tl_code = code;
  tl_msg = msg ? msg : "";
}

static void clear_error() {
  if (tl_code != 0 || !tl_msg.empty()) {
    tl_code = 0;
    tl_msg.clear();
  }
}

static void init_context(Context& C, int device) {
  int n = 0;
  LSM_HIP(hipGetDeviceCount(&n));
  LSM_REQUIRE(n > 0, LIBSUPERMESH_GPU_ERROR_HIP, "no HIP device");
  LSM_REQUIRE(device >= 0 && device < n, LIBSUPERMESH_GPU_ERROR_INVALID, "invalid device index");
  C.device = device;
  LSM_HIP(hipSetDevice(device));
  LSM_HIP(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
  for (int i = 0; i < 16; i++) LSM_HIP(hipEventCreate(&C.ev[i]));
  C.events = true;
  LSM_HIP(hipHostMalloc(reinterpret_cast<void**>(&C.hstat), 64 * sizeof(long), hipHostMallocDefault));
}

static void free_context(Context& C) {
  hipSetDevice(C.device);
  if (C.stream) hipStreamSynchronize(C.stream);
  DevBuf* bufs[] = {&C.mesh_a.enl_copy, &C.mesh_a.pos_copy, &C.mesh_a.coords, &C.mesh_a.bbox,
                    &C.mesh_a.meas_sp, &C.mesh_a.planes, &C.mesh_a.affine, &C.mesh_b.enl_copy,
                    &C.mesh_b.pos_copy, &C.mesh_b.coords, &C.mesh_b.bbox, &C.mesh_b.meas_sp,
                    &C.mesh_b.planes, &C.mesh_b.affine, &C.bvh.boxes, &C.bvh.perm, &C.keys,
                    &C.keys_alt, &C.vals, &C.vals_alt, &C.sort_tmp, &C.scan_tmp, &C.qperm,
                    &C.row_count, &C.indptr, &C.indices, &C.rows, &C.counts, &C.measure,
                    &C.partials, &C.scalars, &C.offsets, &C.elements, &C.blocks, &C.mflag, &C.mslot, &C.work, &C.nwork,
                    &C.mkeys, &C.mvals, &C.mkeys_alt, &C.mvals_alt, &C.ukeys, &C.uvals, &C.nuniq,
                    &C.mindptr, &C.mindices, &C.mrbase, &C.mesh_a.scale,
                    &C.mesh_b.scale, &C.mesh_a.group, &C.mesh_b.group, &C.bvh.granges, &C.gpairs,
                    &C.gsum, &C.predictor.weights, &C.predictor.bias, &C.predictor.master,
                    &C.predictor.adam_m, &C.predictor.adam_v, &C.predictor.grad,
                    &C.predictor.features, &C.predictor.labels, &C.predictor.act,
                    &C.predictor.loss, &C.predictor.index, &C.predictor.budget,
                    &C.predictor.overflow, &C.predictor.spill, &C.predictor.aoff, &C.offsets_dense, &C.queue,
                    &C.elements_dense, &C.dense_ctl, &C.fscratch};
  release_graph(C);
  for (DevBuf* b : bufs) b->release();
  if (C.hstat) hipHostFree(C.hstat);
  C.hstat = nullptr;
  if (C.events)
    for (int i = 0; i < 16; i++) hipEventDestroy(C.ev[i]);
  C.events = false;
  if (C.stream) hipStreamDestroy(C.stream);
  C.stream = nullptr;
}

Context* default_context() {
  static std::unique_ptr<libsupermesh_gpu_context> g;
  static std::mutex m;
  std::lock_guard<std::mutex> lk(m);
  if (!g) {
    auto c = std::make_unique<libsupermesh_gpu_context>();
    init_context(c->ctx, 0);
    g = std::move(c);
  }
  return &g->ctx;
}

template <class F>
static int guarded(F&& f) {
  clear_error();
  try {
    f();
    return LIBSUPERMESH_GPU_OK;
  } catch (const Error& e) {
    set_last_error(e.code, e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error(LIBSUPERMESH_GPU_ERROR_HIP, "host allocation failed");
    return LIBSUPERMESH_GPU_ERROR_HIP;
  } catch (const std::exception& e) {
    set_last_error(LIBSUPERMESH_GPU_ERROR_HIP, e.what());
    return LIBSUPERMESH_GPU_ERROR_HIP;
  }
}

static void check_mesh(const libsupermesh_gpu_mesh* m, const char* which) {
  LSM_REQUIRE(m != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, std::string("null mesh ") + which);
  LSM_REQUIRE(m->dim == 2 || m->dim == 3, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension");
  LSM_REQUIRE(m->nelements == 0 || (m->positions && m->enlist), LIBSUPERMESH_GPU_ERROR_INVALID,
              std::string("null arrays in mesh ") + which);
}

struct Phase {
  Context& C;
  bool on;
  int n = 0;
  explicit Phase(Context& c, bool enabled) : C(c), on(enabled) {}
  void mark() {
    if (on && n < 16) LSM_HIP(hipEventRecord(C.ev[n++], C.stream));
  }
  double between(int i, int j) {
    float ms = 0.f;
    if (!on || j >= n) return 0.0;
    LSM_HIP(hipEventElapsedTime(&ms, C.ev[i], C.ev[j]));
    return ms;
  }
};

// Whole pipeline on device meshes.  In a batch, a_off / b_off (device,
// [ngroups + 1]) split A and B into ngroups meshes, paired by index.
static void run_supermesh(Context& C, const libsupermesh_gpu_mesh* a, const libsupermesh_gpu_mesh* b,
                          int finder, int flags, bool clip, libsupermesh_gpu_result* out,
                          long ngroups = 0, const long* a_off = nullptr, const long* b_off = nullptr,
                          int loc_a = 0, int loc_b = 0) {
  check_mesh(a, "a");
  check_mesh(b, "b");
  LSM_REQUIRE(a->dim == b->dim, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension: meshes differ");
  LSM_REQUIRE(finder == LIBSUPERMESH_GPU_FINDER_TREE || finder == LIBSUPERMESH_GPU_FINDER_RTREE,
              LIBSUPERMESH_GPU_ERROR_INVALID, "unknown finder");
  LSM_REQUIRE(out != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null result");
  LSM_HIP(hipSetDevice(C.device));
  std::memset(out, 0, sizeof(*out));
  const bool mass = clip && (flags & LIBSUPERMESH_GPU_MASS);
  if (!loc_a) loc_a = a->dim + 1;
  if (!loc_b) loc_b = b->dim + 1;
  // quads / hexahedra on either side: the mixed clip kernels, which need no
  // staged planes
  const bool mixed = loc_a != a->dim + 1 || loc_b != b->dim + 1;
  if (mixed && clip)
    LSM_REQUIRE(!(flags & (LIBSUPERMESH_GPU_MASS | LIBSUPERMESH_GPU_LEARNED)) && ngroups == 0,
                LIBSUPERMESH_GPU_ERROR_INVALID, "MASS, LEARNED and batches need simplex meshes");
  Phase ph(C, flags & LIBSUPERMESH_GPU_TIMING);
  C.timing = ph.on;
  ph.mark();
  // the node-index check of both meshes is read back with the candidate count
  // (a bad index is clamped on the device, so the search in between is safe)
  int* bad = stage_flag(C);
  LSM_HIP(hipMemsetAsync(bad, 0, sizeof(int), C.stream));
  stage_mesh(C, C.mesh_a, a->dim, loc_a, a->nnodes, a->nelements, a->positions, a->enlist, false, mass, true);
  stage_mesh(C, C.mesh_b, b->dim, loc_b, b->nnodes, b->nelements, b->positions, b->enlist,
             clip && !mixed, mass, true);
  C.hstat[0] = 0;
  LSM_HIP(hipMemcpyAsync(C.hstat, bad, sizeof(int), hipMemcpyDeviceToHost, C.stream));
  stage_groups(C, C.mesh_a, ngroups, a_off);
  stage_groups(C, C.mesh_b, ngroups, b_off);
  ph.mark();
  // a coarse A over a fine B: B's elements query a hierarchy over A, many
  // short walks instead of a few long ones (each of A's few queries would
  // walk thousands of leaves on one wave); the rows come out sorted
  const bool from_b = ngroups == 0 && C.mesh_a.nelements > 0 && C.mesh_a.nelements * 8 < C.mesh_b.nelements;
  build_bvh(C, from_b ? C.mesh_a : C.mesh_b);
  ph.mark();
  // find_pairs synchronises after reading the candidate count, which also
  // completes the copy of the flag queued before it
  long npairs = from_b ? find_pairs_from_b(C, C.mesh_a, C.mesh_b, finder)
                       : find_pairs(C, C.mesh_a, C.mesh_b, finder, flags & LIBSUPERMESH_GPU_SORTED);
  LSM_REQUIRE(*reinterpret_cast<const int*>(C.hstat) == 0, LIBSUPERMESH_GPU_ERROR_INVALID,
              "Invalid node: element-node index out of range");
  ph.mark();
  out->nelements_a = a->nelements;
  out->npairs = npairs;
  out->indptr = static_cast<const long*>(C.indptr.ptr);
  out->indices = static_cast<const int*>(C.indices.get(4));
  out->rows = static_cast<const int*>(C.rows.get(4));
  LSM_REQUIRE(!(clip && (flags & LIBSUPERMESH_GPU_GRAPH) &&
                (mixed || (flags & (LIBSUPERMESH_GPU_ELEMENTS | LIBSUPERMESH_GPU_MASS | LIBSUPERMESH_GPU_LEARNED)))),
              LIBSUPERMESH_GPU_ERROR_INVALID,
              "LIBSUPERMESH_GPU_GRAPH records the count pass of simplex meshes: not with ELEMENTS, MASS or LEARNED");
  if (clip && mixed) {
    clip_pairs_mixed(C, C.mesh_a, C.mesh_b, npairs, flags, out);
    ph.mark();
    ph.mark();
  } else if (clip && (flags & LIBSUPERMESH_GPU_GRAPH) && npairs > 0) {
    count_pass_graph(C, C.mesh_a, C.mesh_b, npairs, out, ngroups, a_off);
    ph.mark();
    ph.mark();
  } else if (clip) {
    clip_pairs(C, C.mesh_a, C.mesh_b, npairs, flags, out);
    if (ngroups > 0) batch_sums(C, ngroups, a_off, out->indptr, out->measure);
    ph.mark();
    if (mass) assemble_mass(C, C.mesh_a, C.mesh_b, npairs, out);
    ph.mark();
  }
  LSM_HIP(hipStreamSynchronize(C.stream));
  if (ph.on) {
    out->ms[0] = ph.between(0, 1);
    out->ms[1] = ph.between(1, 2);
    out->ms[2] = ph.between(2, 3);
    out->ms[3] = ph.between(3, 4);
    out->ms[4] = ph.between(4, 5);
    out->ms[6] = ph.between(0, ph.n - 1);
  }
  C.timing = false;
}

// Single-GPU form of the self-self loop of parallel_supermesh
// (Parallel_Supermesh.F90:436-453): closed-overlap candidates with A as the
// searched mesh, B elements in order, the A candidates of each in ascending
// id; every candidate pair, empty or not, is handed to the callback with
// the clipped elements of intersect_elements(A, B) (Supermesh.F90:162-209:
// simplices, quadrilaterals or hexahedra on either side, loc_a / loc_b
// nodes per element).  Pairs are clipped on the device in chunks and the
// callbacks run on the calling thread.
static void run_callback_driver(Context& C, const libsupermesh_gpu_mesh* a, int loc_a,
                                const libsupermesh_gpu_mesh* b, int loc_b, libsupermesh_gpu_pair_callback cb,
                                void* user) {
  check_mesh(a, "a");
  check_mesh(b, "b");
  const int dim = a->dim;
  LSM_REQUIRE((loc_a == dim + 1 || loc_a == 1 << dim) && (loc_b == dim + 1 || loc_b == 1 << dim),
              LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid element type: loc must be dim+1 or 2^dim");
  libsupermesh_gpu_result r;
  run_supermesh(C, b, a, LIBSUPERMESH_GPU_FINDER_RTREE, LIBSUPERMESH_GPU_SORTED, false, &r, 0, nullptr, nullptr,
                loc_b, loc_a);
  const long npairs = r.npairs;
  const int per = (dim + 1) * dim;
  const bool mixed = loc_a != dim + 1 || loc_b != dim + 1;
  // host copies of both meshes, for the element coordinates and nodes_b
  std::vector<double> pos_a((size_t)a->nnodes * dim), pos_b((size_t)b->nnodes * dim);
  std::vector<int> enl_a((size_t)a->nelements * loc_a), enl_b((size_t)b->nelements * loc_b);
  auto d2h = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) LSM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, C.stream));
  };
  d2h(pos_a.data(), a->positions, sizeof(double) * pos_a.size());
  d2h(pos_b.data(), b->positions, sizeof(double) * pos_b.size());
  d2h(enl_a.data(), a->enlist, sizeof(int) * enl_a.size());
  d2h(enl_b.data(), b->enlist, sizeof(int) * enl_b.size());
  LSM_HIP(hipStreamSynchronize(C.stream));
  // pair lists: B element (CSR row) and A element (CSR entry) of each pair
  const int* d_pair_b = r.rows;
  const int* d_pair_a = r.indices;
  stage_mesh(C, C.mesh_a, dim, loc_a, a->nnodes, a->nelements, a->positions, a->enlist, false, false);
  stage_mesh(C, C.mesh_b, dim, loc_b, b->nnodes, b->nelements, b->positions, b->enlist, !mixed, false);
  constexpr long kChunk = 1L << 21;
  std::vector<int> pa, pb, counts;
  std::vector<long> offs;
  std::vector<double> elems;
  double ea[8 * 3], eb[8 * 3];
  long nodes_b[8];
  for (long p0 = 0; p0 < npairs; p0 += kChunk) {
    const long n = std::min(kChunk, npairs - p0);
    libsupermesh_gpu_result cr;
    std::memset(&cr, 0, sizeof(cr));
    cr.npairs = n;
    cr.rows = d_pair_a + p0;
    cr.indices = d_pair_b + p0;
    if (mixed) clip_pairs_mixed(C, C.mesh_a, C.mesh_b, n, LIBSUPERMESH_GPU_ELEMENTS, &cr);
    else clip_pairs(C, C.mesh_a, C.mesh_b, n, LIBSUPERMESH_GPU_ELEMENTS, &cr);
    pa.resize(n);
    pb.resize(n);
    counts.resize(n);
    offs.resize(n + 1);
    elems.resize((size_t)std::max(cr.nout, 1L) * per);
    d2h(pa.data(), cr.rows, sizeof(int) * n);
    d2h(pb.data(), cr.indices, sizeof(int) * n);
    d2h(counts.data(), cr.counts, sizeof(int) * n);
    d2h(offs.data(), cr.offsets, sizeof(long) * (n + 1));
    d2h(elems.data(), cr.elements, sizeof(double) * (size_t)cr.nout * per);
    LSM_HIP(hipStreamSynchronize(C.stream));
    for (long i = 0; i < n; i++) {
      const long ia = pa[i], ib = pb[i];
      for (int v = 0; v < loc_a; v++) {
        const long na = enl_a[(size_t)ia * loc_a + v];
        for (int d = 0; d < dim; d++) ea[v * dim + d] = pos_a[(size_t)na * dim + d];
      }
      for (int v = 0; v < loc_b; v++) {
        const long nb = enl_b[(size_t)ib * loc_b + v];
        nodes_b[v] = nb;
        for (int d = 0; d < dim; d++) eb[v * dim + d] = pos_b[(size_t)nb * dim + d];
      }
      cb(ea, eb, elems.data() + (size_t)offs[i] * per, counts[i], nodes_b, ia, ib, user);
    }
  }
}

// Flags (bit 0) any pair whose A id is outside [0, na) or B id outside
// [0, nb).
__global__ void __launch_bounds__(256) k_check_pair_ids(long n, const int* __restrict__ pa,
                                                        const int* __restrict__ pb, long na, long nb,
                                                        int* __restrict__ flag) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (pa[i] < 0 || pa[i] >= na || pb[i] < 0 || pb[i] >= nb)) *flag = 1;
}

// Checks caller-supplied pair lists against the mesh sizes before any kernel
// reads coordinates through them.
static void check_pair_ids(Context& C, long n, const int* pa, const int* pb, long na, long nb) {
  if (n <= 0) return;
  int* flag = stage_flag(C);
  LSM_HIP(hipMemsetAsync(flag, 0, sizeof(int), C.stream));
  hipLaunchKernelGGL(k_check_pair_ids, dim3(grid_for(n, 256)), dim3(256), 0, C.stream, n, pa, pb, na, nb, flag);
  LSM_HIP(hipGetLastError());
  int h = 0;
  LSM_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  LSM_REQUIRE(h == 0, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid element: pair id out of range");
}

__global__ void __launch_bounds__(256) k_low_byte(long n, int* __restrict__ v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] &= 0xff;
}

// ---------------------------------------------------------------------------
// The reference's one-pair C functions (C_bindings.F90:42-64,134-174) run on
// the calling thread: the same __host__ __device__ clippers as the kernels,
// compiled for the host with the same flags (IEEE double, no contraction), so
// the results are bitwise those of the bulk path.  No device, no locks: the
// calls are re-entrant.  Each writes at most `cap` elements of `stride`
// doubles to out and returns the number the reference reports.
// ---------------------------------------------------------------------------

static int one_tris(const double* tri_a, const double* tri_b, double* out, int cap) {
  Line L[3];
  tri_lines(tri_b, L);
  const double tol = tri_tol(triangle_area(tri_a), triangle_area(tri_b));
  int k = 0;
  return clip_tri(tri_a, L, tol, [&](double x0, double y0, double x1, double y1, double x2, double y2, double) {
    if (k < cap) {
      double* o = out + 6 * k;
      o[0] = x0;
      o[1] = y0;
      o[2] = x1;
      o[3] = y1;
      o[4] = x2;
      o[5] = y2;
    }
    k++;
  });
}

static int one_tets(const double* tet_a, const double* tet_b, double* out, int cap) {
  Planes P;
  tet_planes(tet_b, P);
  Tet A;
  for (int i = 0; i < 12; i++) A.v[i] = tet_a[i];
  const double tol = tet_tol(tet_volume_tet(tet_a), tet_volume_tet(tet_b));
  int k = 0;
  return clip_tet(A, P, tol, [&](const Tet& T, double) {
    if (k < cap)
      for (int i = 0; i < 12; i++) out[12 * k + i] = T.v[i];
    k++;
  });
}

// intersect_tri_quad (Supermesh.F90:211-230): the triangle clipped by the
// quad's lines, tolerance from the triangle's area only.
static int one_tri_quad(const double* tri, const double* quad, double* out, int cap) {
  Line L[4];
  quad_lines(quad, L);
  const double tol = 10.0 * fspacing(triangle_area(tri));
  int k = 0;
  return clip_poly<4>(tri, L, tol, [&](double x0, double y0, double x1, double y1, double x2, double y2, double) {
    if (k < cap) {
      double* o = out + 6 * k;
      o[0] = x0;
      o[1] = y0;
      o[2] = x1;
      o[3] = y1;
      o[4] = x2;
      o[5] = y2;
    }
    k++;
  });
}

// intersect_quads (Supermesh.F90:231-248): quad A by quad B's lines, tolerance
// 10 * tiny.
static int one_quads(const double* qa, const double* qb, double* out, int cap) {
  Line L[4];
  quad_lines(qb, L);
  int k = 0;
  return clip_poly<4, 4>(qa, L, 10.0 * 2.2250738585072014e-308,
                         [&](double x0, double y0, double x1, double y1, double x2, double y2, double) {
                           if (k < cap) {
                             double* o = out + 6 * k;
                             o[0] = x0;
                             o[1] = y0;
                             o[2] = x1;
                             o[3] = y1;
                             o[4] = x2;
                             o[5] = y2;
                           }
                           k++;
                         });
}

static int one_tet_hex(const double* tet, const double* hex, double* out, int cap) {
  int k = 0;
  return clip_tet_hex(tet, hex, [&](const double* t, double) {
    if (k < cap)
      for (int i = 0; i < 12; i++) out[12 * k + i] = t[i];
    k++;
  });
}

// intersect_hexes (Supermesh.F90:303-334): A's five tets in turn.
static int one_hexes(const double* ha, const double* hb, double* out, int cap) {
  int k = 0;
  for (int s = 0; s < 5; s++) {
    double T[12];
    hex_split_tet(ha, s, T);
    clip_tet_hex(T, hb, [&](const double* t, double) {
      if (k < cap)
        for (int i = 0; i < 12; i++) out[12 * k + i] = t[i];
      k++;
    });
  }
  return k;
}

// ---------------------------------------------------------------------------
// Finder state of the reference's set_input / query_output / get_output
// ---------------------------------------------------------------------------

struct FinderState {
  std::mutex mu;
  bool ready = false;
  long nelements_a = 0;
  std::vector<long> indptr;
  std::vector<int> indices;
};

static FinderState g_tree, g_rtree, g_sort, g_front;

// State of the mesh_*list set_input / query_output / get_output calls
// (C_bindings.F90:176-344): the CSR computed by set_input, freed by
// get_output.
struct GraphState {
  std::mutex mu;
  bool ready = false;
  std::vector<long> indptr, indices;
};

static GraphState g_nelist, g_nnlist, g_eelist;

template <class F>
static void graph_set_input(GraphState& S, F&& compute) {
  std::lock_guard<std::mutex> lk(S.mu);
  S.ready = false;
  S.indptr.clear();
  S.indices.clear();
  Context& C = *default_context();
  std::lock_guard<std::mutex> lk2(C.mu);
  compute(C, S.indptr, S.indices);
  S.ready = true;
}

static void graph_query_output(GraphState& S, long* nindices) {
  std::lock_guard<std::mutex> lk(S.mu);
  LSM_REQUIRE(S.ready, LIBSUPERMESH_GPU_ERROR_STATE, "query_output called before set_input");
  *nindices = (long)S.indices.size();
}

static void graph_get_output(GraphState& S, long nrows, long nindices, long* indices, long* ind_ptr) {
  std::lock_guard<std::mutex> lk(S.mu);
  LSM_REQUIRE(S.ready, LIBSUPERMESH_GPU_ERROR_STATE, "get_output called before set_input");
  LSM_REQUIRE(nrows + 1 == (long)S.indptr.size() && nindices == (long)S.indices.size(),
              LIBSUPERMESH_GPU_ERROR_INVALID, "get_output sizes differ from set_input's result");
  std::memcpy(ind_ptr, S.indptr.data(), sizeof(long) * S.indptr.size());
  if (nindices > 0) std::memcpy(indices, S.indices.data(), sizeof(long) * nindices);
  S.ready = false;
  std::vector<long>().swap(S.indptr);
  std::vector<long>().swap(S.indices);
}

static void host_mesh_to_device(Context& C, DevMesh& M, long nnodes, int dim, long ne, int loc,
                                const double* pos, const long* enl, libsupermesh_gpu_mesh& view) {
  LSM_REQUIRE(dim == 2 || dim == 3, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension");
  LSM_REQUIRE(loc == dim + 1 || loc == (1 << dim), LIBSUPERMESH_GPU_ERROR_INVALID,
              "Unsupported element type: simplices, quads or hexahedra only");
  LSM_REQUIRE(nnodes >= 0 && ne >= 0 && nnodes < (1L << 31) && ne < (1L << 31),
              LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid mesh size");
  std::vector<int> e32((size_t)ne * loc);
  for (size_t i = 0; i < e32.size(); i++) {
    long v = enl[i];
    LSM_REQUIRE(v >= 0 && v < nnodes, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid node");
    e32[i] = (int)v;
  }
  double* dpos = M.pos_copy.as<double>((size_t)nnodes * dim);
  int* denl = M.enl_copy.as<int>((size_t)ne * loc);
  if (nnodes > 0)
    LSM_HIP(hipMemcpyAsync(dpos, pos, sizeof(double) * nnodes * dim, hipMemcpyHostToDevice, C.stream));
  if (ne > 0)
    LSM_HIP(hipMemcpyAsync(denl, e32.data(), sizeof(int) * ne * loc, hipMemcpyHostToDevice, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  view.dim = dim;
  view.nnodes = nnodes;
  view.nelements = ne;
  view.positions = dpos;
  view.enlist = denl;
}

static void finder_set_input(FinderState& S, int finder, long nnodes_a, int dim_a, long ne_a,
                             int loc_a, long nnodes_b, int dim_b, long ne_b, int loc_b,
                             const double* pos_a, const long* enl_a, const double* pos_b,
                             const long* enl_b) {
  std::lock_guard<std::mutex> lk(S.mu);
  S.ready = false;
  S.indptr.clear();
  S.indices.clear();
  LSM_REQUIRE(dim_a == dim_b, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension");
  Context& C = *default_context();
  std::lock_guard<std::mutex> lk2(C.mu);
  LSM_HIP(hipSetDevice(C.device));
  libsupermesh_gpu_mesh va, vb;
  host_mesh_to_device(C, C.mesh_a, nnodes_a, dim_a, ne_a, loc_a, pos_a, enl_a, va);
  host_mesh_to_device(C, C.mesh_b, nnodes_b, dim_b, ne_b, loc_b, pos_b, enl_b, vb);
  libsupermesh_gpu_result r;
  run_supermesh(C, &va, &vb, finder, LIBSUPERMESH_GPU_SORTED, false, &r, 0, nullptr, nullptr, loc_a,
                loc_b);
  S.nelements_a = ne_a;
  S.indptr.resize((size_t)ne_a + 1);
  S.indices.resize((size_t)r.npairs);
  LSM_HIP(hipMemcpy(S.indptr.data(), r.indptr, sizeof(long) * (ne_a + 1), hipMemcpyDeviceToHost));
  if (r.npairs > 0)
    LSM_HIP(hipMemcpy(S.indices.data(), r.indices, sizeof(int) * r.npairs, hipMemcpyDeviceToHost));
  S.ready = true;
}

static void finder_query_output(FinderState& S, long* nindices) {
  std::lock_guard<std::mutex> lk(S.mu);
  LSM_REQUIRE(S.ready, LIBSUPERMESH_GPU_ERROR_STATE, "query_output called before set_input");
  *nindices = (long)S.indices.size();
}

static void finder_get_output(FinderState& S, long nelements, long nindices, long* indices,
                              long* ind_ptr) {
  std::lock_guard<std::mutex> lk(S.mu);
  LSM_REQUIRE(S.ready, LIBSUPERMESH_GPU_ERROR_STATE, "get_output called before set_input");
  LSM_REQUIRE(nelements == S.nelements_a && nindices == (long)S.indices.size(),
              LIBSUPERMESH_GPU_ERROR_INVALID, "get_output sizes differ from query_output");
  for (long i = 0; i <= nelements; i++) ind_ptr[i] = S.indptr[i];
  for (long i = 0; i < nindices; i++) indices[i] = S.indices[i];
  S.ready = false;
  std::vector<long>().swap(S.indptr);
  std::vector<int>().swap(S.indices);
}

}  // namespace lsm

using namespace lsm;

extern "C" {

// ---- bulk API (include/libsupermesh_gpu.h) --------------------------------

int libsupermesh_gpu_context_create(int device, libsupermesh_gpu_context** ctx) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null context pointer");
    auto c = std::make_unique<libsupermesh_gpu_context>();
    init_context(c->ctx, device);
    *ctx = c.release();
  });
}

void libsupermesh_gpu_context_destroy(libsupermesh_gpu_context* ctx) {
  if (!ctx) return;
  free_context(ctx->ctx);
  delete ctx;
}

void* libsupermesh_gpu_context_stream(libsupermesh_gpu_context* ctx) {
  return ctx ? (void*)ctx->ctx.stream : nullptr;
}

int libsupermesh_gpu_malloc(int device, size_t bytes, void** ptr) {
  return guarded([&] {
    LSM_REQUIRE(ptr != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null pointer");
    *ptr = nullptr;
    LSM_HIP(hipSetDevice(device));
    LSM_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  });
}

int libsupermesh_gpu_free(void* ptr) {
  // a successful free leaves the last error as it was: callers free their
  // buffers after a failed call and still read its message
  if (!ptr) return LIBSUPERMESH_GPU_OK;
  const hipError_t e = hipFree(ptr);
  if (e == hipSuccess) return LIBSUPERMESH_GPU_OK;
  set_last_error(LIBSUPERMESH_GPU_ERROR_HIP, (std::string("hipFree failed: ") + hipGetErrorString(e)).c_str());
  return LIBSUPERMESH_GPU_ERROR_HIP;
}

int libsupermesh_gpu_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  return guarded([&] {
    LSM_REQUIRE(kind >= 1 && kind <= 3, LIBSUPERMESH_GPU_ERROR_INVALID, "invalid copy kind");
    if (bytes == 0) return;
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                                : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    LSM_HIP(hipMemcpy(dst, src, bytes, k));
  });
}

int libsupermesh_gpu_synchronize(int device) {
  return guarded([&] {
    LSM_HIP(hipSetDevice(device));
    LSM_HIP(hipDeviceSynchronize());
  });
}

const char* libsupermesh_gpu_last_error(void) { return tl_msg.c_str(); }
int libsupermesh_gpu_last_error_code(void) { return tl_code; }

int libsupermesh_gpu_intersection_finder(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                                         const libsupermesh_gpu_mesh* b, int finder, int flags,
                                         libsupermesh_gpu_result* out) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null context");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    run_supermesh(ctx->ctx, a, b, finder, flags, false, out);
  });
}

int libsupermesh_gpu_supermesh(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                               const libsupermesh_gpu_mesh* b, int finder, int flags,
                               libsupermesh_gpu_result* out) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null context");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    run_supermesh(ctx->ctx, a, b, finder, flags, true, out);
  });
}

int libsupermesh_gpu_supermesh_mixed(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a, int loc_a,
                                     const libsupermesh_gpu_mesh* b, int loc_b, int finder, int flags,
                                     libsupermesh_gpu_result* out) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && a != nullptr && b != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID,
                "null argument");
    const int d = a->dim;
    LSM_REQUIRE((loc_a == d + 1 || loc_a == 1 << d) && (loc_b == d + 1 || loc_b == 1 << d),
                LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid element type: loc must be dim+1 or 2^dim");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    run_supermesh(ctx->ctx, a, b, finder, flags, true, out, 0, nullptr, nullptr, loc_a, loc_b);
  });
}

int libsupermesh_gpu_supermesh_batched(libsupermesh_gpu_context* ctx,
                                       const libsupermesh_gpu_mesh_batch* a,
                                       const libsupermesh_gpu_mesh_batch* b, int finder, int flags,
                                       libsupermesh_gpu_result* out,
                                       libsupermesh_gpu_batch_result* batch) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && a != nullptr && b != nullptr && batch != nullptr,
                LIBSUPERMESH_GPU_ERROR_INVALID, "null argument");
    LSM_REQUIRE(a->nmeshes == b->nmeshes && a->nmeshes >= 1, LIBSUPERMESH_GPU_ERROR_INVALID,
                "batches of A and B hold different mesh counts");
    LSM_REQUIRE(a->element_offsets && b->element_offsets, LIBSUPERMESH_GPU_ERROR_INVALID,
                "null batch element offsets");
    Context& C = ctx->ctx;
    std::lock_guard<std::mutex> lk(C.mu);
    libsupermesh_gpu_mesh ma = {a->dim, a->nnodes, a->nelements, a->positions, a->enlist};
    libsupermesh_gpu_mesh mb = {b->dim, b->nnodes, b->nelements, b->positions, b->enlist};
    run_supermesh(C, &ma, &mb, finder, flags, true, out, a->nmeshes, a->element_offsets,
                  b->element_offsets);
    batch->nmeshes = a->nmeshes;
    batch->pair_offsets = static_cast<const long*>(C.gpairs.ptr);
    batch->sums = static_cast<const double*>(C.gsum.ptr);
  });
}

int libsupermesh_gpu_count_shape(libsupermesh_gpu_context* ctx, int dim, int* block, int* grid, int* lds_depths,
                                 float* ms) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && (dim == 2 || dim == 3), LIBSUPERMESH_GPU_ERROR_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    const CountShape& s = ctx->ctx.shape[dim - 2];
    if (block) *block = s.block;
    if (grid) *grid = s.grid;
    if (lds_depths) *lds_depths = s.block == 0 ? 0 : (dim == 3 ? 3 - s.reg : 0);
    if (ms) *ms = s.ms;
  });
}

int libsupermesh_gpu_compact(libsupermesh_gpu_context* ctx, libsupermesh_gpu_result* result) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && result != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null argument");
    Context& C = ctx->ctx;
    std::lock_guard<std::mutex> lk(C.mu);
    LSM_HIP(hipSetDevice(C.device));
    LSM_REQUIRE(result->elements != nullptr, LIBSUPERMESH_GPU_ERROR_STATE,
                "libsupermesh_gpu_compact needs a result with ELEMENTS");
    if (!C.runs_layout || result->elements == C.elements_dense.ptr || result->offsets == C.offsets_dense.ptr)
      return;  // the plain layout already
    LSM_REQUIRE(result->offsets != nullptr && result->offsets == C.offsets.ptr &&
                    result->elements == C.elements.ptr && result->counts == C.counts.ptr,
                LIBSUPERMESH_GPU_ERROR_STATE,
                "libsupermesh_gpu_compact needs the context's latest result with ELEMENTS");
    compact_result(C, C.mesh_a.dim, result);
  });
}

int libsupermesh_gpu_checksum(libsupermesh_gpu_context* ctx, const void* ptr, size_t bytes,
                              unsigned long long* out) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && out != nullptr && (bytes == 0 || ptr != nullptr),
                LIBSUPERMESH_GPU_ERROR_INVALID, "invalid argument");
    LSM_REQUIRE(bytes % 4 == 0, LIBSUPERMESH_GPU_ERROR_INVALID, "checksum size must be whole 32-bit words");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    LSM_HIP(hipSetDevice(ctx->ctx.device));
    *out = checksum(ctx->ctx, ptr, bytes);
  });
}

int libsupermesh_gpu_predictor_set(libsupermesh_gpu_context* ctx, const float* params) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && params != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    LSM_HIP(hipSetDevice(ctx->ctx.device));
    predictor_set(ctx->ctx, params);
  });
}

int libsupermesh_gpu_predictor_get(libsupermesh_gpu_context* ctx, float* params) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && params != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    LSM_HIP(hipSetDevice(ctx->ctx.device));
    predictor_get(ctx->ctx, params);
  });
}

int libsupermesh_gpu_predictor_quantile(libsupermesh_gpu_context* ctx, float q) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null context");
    LSM_REQUIRE(q > 0.f && q <= 1.f, LIBSUPERMESH_GPU_ERROR_INVALID, "quantile must be in (0, 1]");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    ctx->ctx.predictor.quantile = q;
  });
}

int libsupermesh_gpu_predictor_arena(libsupermesh_gpu_context* ctx, float fraction, long min_tets) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && fraction >= 0.f && min_tets >= 0, LIBSUPERMESH_GPU_ERROR_INVALID,
                "invalid arena size");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    ctx->ctx.predictor.arena_fraction = fraction;
    ctx->ctx.predictor.arena_min = min_tets;
  });
}

int libsupermesh_gpu_predictor_spill(libsupermesh_gpu_context* ctx, float fraction, long min_tets) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && fraction >= 0.f && min_tets >= 1, LIBSUPERMESH_GPU_ERROR_INVALID,
                "invalid spill reserve size");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    ctx->ctx.predictor.spill_fraction = fraction;
    ctx->ctx.predictor.spill_min = min_tets;
  });
}

int libsupermesh_gpu_predictor_collect(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                                       const libsupermesh_gpu_mesh* b, long max_samples,
                                       long* nsamples) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null context");
    LSM_REQUIRE(a != nullptr && b != nullptr && a->dim == 3 && b->dim == 3, LIBSUPERMESH_GPU_ERROR_INVALID,
                "the predictor learns 3D (tet) pairs");
    Context& C = ctx->ctx;
    std::lock_guard<std::mutex> lk(C.mu);
    libsupermesh_gpu_result r;
    run_supermesh(C, a, b, LIBSUPERMESH_GPU_FINDER_TREE, 0, true, &r);
    const int* work = nullptr;
    const long nwork = r.npairs > 0 ? last_work_list(C, &work) : 0;
    predictor_collect(C, nwork, work, r.rows, r.indices, r.counts,
                      static_cast<const double*>(C.mesh_a.coords.ptr),
                      static_cast<const double*>(C.mesh_b.coords.ptr), max_samples);
    if (nsamples) *nsamples = C.predictor.nsamples;
  });
}

int libsupermesh_gpu_predictor_train(libsupermesh_gpu_context* ctx, int epochs, int batch, float lr,
                                     unsigned seed, float* loss_per_epoch) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && epochs >= 0, LIBSUPERMESH_GPU_ERROR_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    LSM_HIP(hipSetDevice(ctx->ctx.device));
    predictor_train(ctx->ctx, epochs, batch, lr, seed, loss_per_epoch);
  });
}

int libsupermesh_gpu_predictor_infer(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                                     const libsupermesh_gpu_mesh* b, long npairs, const int* pair_a,
                                     const int* pair_b, int* budgets) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null context");
    check_mesh(a, "a");
    check_mesh(b, "b");
    LSM_REQUIRE(a->dim == 3 && b->dim == 3, LIBSUPERMESH_GPU_ERROR_INVALID, "the predictor takes tet pairs");
    LSM_REQUIRE(npairs >= 0 && (npairs == 0 || (pair_a && pair_b && budgets)),
                LIBSUPERMESH_GPU_ERROR_INVALID, "invalid pair list");
    Context& C = ctx->ctx;
    std::lock_guard<std::mutex> lk(C.mu);
    LSM_HIP(hipSetDevice(C.device));
    LSM_REQUIRE(C.predictor.ready, LIBSUPERMESH_GPU_ERROR_STATE, "no predictor weights loaded");
    check_pair_ids(C, npairs, pair_a, pair_b, a->nelements, b->nelements);
    stage_mesh(C, C.mesh_a, 3, 4, a->nnodes, a->nelements, a->positions, a->enlist, false, false);
    stage_mesh(C, C.mesh_b, 3, 4, b->nnodes, b->nelements, b->positions, b->enlist, false, false);
    PredictArgs pa = {};
    pa.nwork = npairs;
    pa.rows = pair_a;
    pa.cols = pair_b;
    pa.acoords = static_cast<const double*>(C.mesh_a.coords.ptr);
    pa.bcoords = static_cast<const double*>(C.mesh_b.coords.ptr);
    pa.weights = static_cast<const __bf16*>(C.predictor.weights.ptr);
    pa.bias = static_cast<const float*>(C.predictor.bias.ptr);
    budget_levels(C.predictor.quantile, pa.levels);
    // packed levels land in the caller's array; level 0 is the low byte
    pa.budget = reinterpret_cast<unsigned*>(budgets);
    pa.sums = C.scalars.as<unsigned long long>(16) + 8;
    pa.spare = reinterpret_cast<unsigned*>(pa.sums + 5);
    LSM_HIP(hipMemsetAsync(pa.sums, 0, 5 * sizeof(unsigned long long), C.stream));
    predict_budgets(C, pa);
    if (npairs > 0)
      hipLaunchKernelGGL(k_low_byte, dim3(grid_for(npairs, 256)), dim3(256), 0, C.stream, npairs, budgets);
    LSM_HIP(hipGetLastError());
    LSM_HIP(hipStreamSynchronize(C.stream));
  });
}

int libsupermesh_gpu_predictor_add_samples(libsupermesh_gpu_context* ctx, long n, const float* features,
                                           const int* counts) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr && n >= 0 && (n == 0 || (features && counts)), LIBSUPERMESH_GPU_ERROR_INVALID,
                "invalid samples");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    LSM_HIP(hipSetDevice(ctx->ctx.device));
    predictor_add_samples(ctx->ctx, n, features, counts);
  });
}

int libsupermesh_gpu_predictor_clear(libsupermesh_gpu_context* ctx) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null context");
    std::lock_guard<std::mutex> lk(ctx->ctx.mu);
    ctx->ctx.predictor.nsamples = 0;
  });
}

int libsupermesh_gpu_supermesh_callback(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                                        const libsupermesh_gpu_mesh* b,
                                        libsupermesh_gpu_pair_callback callback, void* user) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null context");
    LSM_REQUIRE(callback != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null callback");
    check_mesh(a, "a");
    check_mesh(b, "b");
    Context& C = ctx->ctx;
    std::lock_guard<std::mutex> lk(C.mu);
    run_callback_driver(C, a, a->dim + 1, b, b->dim + 1, callback, user);
  });
}

int libsupermesh_gpu_supermesh_callback_mixed(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                                              int loc_a, const libsupermesh_gpu_mesh* b, int loc_b,
                                              libsupermesh_gpu_pair_callback callback, void* user) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null context");
    LSM_REQUIRE(callback != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null callback");
    Context& C = ctx->ctx;
    std::lock_guard<std::mutex> lk(C.mu);
    run_callback_driver(C, a, loc_a, b, loc_b, callback, user);
  });
}

int libsupermesh_gpu_clip_pairs(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                                const libsupermesh_gpu_mesh* b, long npairs, const int* pair_a,
                                const int* pair_b, int flags, libsupermesh_gpu_result* out) {
  return guarded([&] {
    LSM_REQUIRE(ctx != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null context");
    check_mesh(a, "a");
    check_mesh(b, "b");
    LSM_REQUIRE(a->dim == b->dim, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension: meshes differ");
    LSM_REQUIRE(npairs >= 0 && (npairs == 0 || (pair_a && pair_b)), LIBSUPERMESH_GPU_ERROR_INVALID,
                "invalid pair list");
    LSM_REQUIRE(out != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null result");
    // the clip kernels index their work lists with 32-bit pair numbers
    LSM_REQUIRE(npairs < (1L << 31), LIBSUPERMESH_GPU_ERROR_OVERFLOW, "more than 2^31 - 1 pairs in one call");
    Context& C = ctx->ctx;
    std::lock_guard<std::mutex> lk(C.mu);
    LSM_HIP(hipSetDevice(C.device));
    std::memset(out, 0, sizeof(*out));
    check_pair_ids(C, npairs, pair_a, pair_b, a->nelements, b->nelements);
    const bool mass = flags & LIBSUPERMESH_GPU_MASS;
    stage_mesh(C, C.mesh_a, a->dim, a->dim + 1, a->nnodes, a->nelements, a->positions, a->enlist,
               false, mass);
    stage_mesh(C, C.mesh_b, b->dim, b->dim + 1, b->nnodes, b->nelements, b->positions, b->enlist,
               true, mass);
    out->npairs = npairs;
    out->rows = pair_a;
    out->indices = pair_b;
    clip_pairs(C, C.mesh_a, C.mesh_b, npairs, flags, out);
    LSM_HIP(hipStreamSynchronize(C.stream));
  });
}

int libsupermesh_gpu_clip_pairs_noalloc(int device, void* stream, const libsupermesh_gpu_mesh* a,
                                        const libsupermesh_gpu_mesh* b, long npairs, const int* pair_a,
                                        const int* pair_b, int* counts, double* measure, double* sum) {
  return guarded([&] {
    LSM_REQUIRE(a != nullptr && b != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null mesh");
    clip_pairs_noalloc(device, static_cast<hipStream_t>(stream), *a, *b, npairs, pair_a, pair_b, counts, measure,
                       sum);
  });
}

// ---- drop-in symbols (include/libsupermesh-c.h) ---------------------------

void libsupermesh_tree_intersection_finder_set_input(long* nnodes_a, int* dim_a, long* nelements_a,
                                                     int* loc_a, long* nnodes_b, int* dim_b,
                                                     long* nelements_b, int* loc_b,
                                                     double* positions_a, long* enlist_a,
                                                     double* positions_b, long* enlist_b) {
  guarded([&] {
    finder_set_input(g_tree, LIBSUPERMESH_GPU_FINDER_TREE, *nnodes_a, *dim_a, *nelements_a, *loc_a,
                     *nnodes_b, *dim_b, *nelements_b, *loc_b, positions_a, enlist_a, positions_b,
                     enlist_b);
  });
}

void libsupermesh_tree_intersection_finder_query_output(long* nindices) {
  guarded([&] { finder_query_output(g_tree, nindices); });
}

void libsupermesh_tree_intersection_finder_get_output(long* nelements, long* nindices, long* indices,
                                                      long* ind_ptr) {
  guarded([&] { finder_get_output(g_tree, *nelements, *nindices, indices, ind_ptr); });
}

void libsupermesh_rtree_intersection_finder_set_input(long* nnodes_a, int* dim_a, long* nelements_a,
                                                      int* loc_a, long* nnodes_b, int* dim_b,
                                                      long* nelements_b, int* loc_b,
                                                      double* positions_a, long* enlist_a,
                                                      double* positions_b, long* enlist_b) {
  guarded([&] {
    finder_set_input(g_rtree, LIBSUPERMESH_GPU_FINDER_RTREE, *nnodes_a, *dim_a, *nelements_a, *loc_a,
                     *nnodes_b, *dim_b, *nelements_b, *loc_b, positions_a, enlist_a, positions_b,
                     enlist_b);
  });
}

void libsupermesh_rtree_intersection_finder_query_output(long* nindices) {
  guarded([&] { finder_query_output(g_rtree, nindices); });
}

void libsupermesh_rtree_intersection_finder_get_output(long* nelements, long* nindices, long* indices,
                                                       long* ind_ptr) {
  guarded([&] { finder_get_output(g_rtree, *nelements, *nindices, indices, ind_ptr); });
}

void libsupermesh_intersect_tris_real(double* tri_a, double* tri_b, double* tris_c, int* n_tris_c) {
  clear_error();
  *n_tris_c = one_tris(tri_a, tri_b, tris_c, kTriBufSize);
}

void libsupermesh_intersect_tets_real(double* tet_a, double* tet_b, double* tets_c, int* n_tets_c) {
  clear_error();
  *n_tets_c = one_tets(tet_a, tet_b, tets_c, kTetBufSize);
}

void libsupermesh_triangle_area(double* tri, double* area) {
  clear_error();
  *area = triangle_area(tri);
}

void libsupermesh_tetrahedron_volume(double* tet, double* volume) {
  clear_error();
  *volume = tet_volume_real(tet);
}

void libsupermesh_intersect_tri_quad(double* tri_a, double* quad_b, double* tris_c, int* n_tris_c) {
  clear_error();
  *n_tris_c = one_tri_quad(tri_a, quad_b, tris_c, kTriQuadBufSize);
}

void libsupermesh_intersect_tet_hex(double* tet_a, double* hex_b, double* tets_c, int* n_tets_c) {
  clear_error();
  *n_tets_c = one_tet_hex(tet_a, hex_b, tets_c, kTetHexBufSize);
}

void libsupermesh_intersect_quads(double* quad_a, double* quad_b, double* tris_c, int* n_tris_c) {
  int rc = guarded([&] {
    const int n = one_quads(quad_a, quad_b, tris_c, kQuadsBufSize);
    LSM_REQUIRE(n <= kQuadsBufSize, LIBSUPERMESH_GPU_ERROR_OVERFLOW, "intersect_quads: more than 62 triangles");
    *n_tris_c = n;
  });
  if (rc != LIBSUPERMESH_GPU_OK) *n_tris_c = 0;
}

void libsupermesh_intersect_hexes(double* hex_a, double* hex_b, double* tets_c, int* n_tets_c) {
  int rc = guarded([&] {
    const int n = one_hexes(hex_a, hex_b, tets_c, kHexesBufSize);
    LSM_REQUIRE(n <= kHexesBufSize, LIBSUPERMESH_GPU_ERROR_OVERFLOW, "intersect_hexes: more than 3645 tets");
    *n_tets_c = n;
  });
  if (rc != LIBSUPERMESH_GPU_OK) *n_tets_c = 0;
}

void libsupermesh_interval_size(double* interval, double* size) {
  guarded([&] { *size = lsm::interval_size(interval); });
}

void libsupermesh_intersect_intervals(double* interval_a, double* interval_b, double* intervals_c,
                                      int* n_intervals_c) {
  int rc = guarded([&] { *n_intervals_c = lsm::intersect_intervals(interval_a, interval_b, intervals_c); });
  if (rc != LIBSUPERMESH_GPU_OK) *n_intervals_c = 0;
}

void libsupermesh_sort_intersection_finder_set_input(long* nnodes_a, int* dim_a, long* nelements_a,
                                                     int* loc_a, long* nnodes_b, int* dim_b,
                                                     long* nelements_b, int* loc_b, double* positions_a,
                                                     long* enlist_a, double* positions_b, long* enlist_b) {
  guarded([&] {
    std::lock_guard<std::mutex> lk(g_sort.mu);
    g_sort.ready = false;
    g_sort.indptr.clear();
    g_sort.indices.clear();
    LSM_REQUIRE(*dim_a == 1 && *dim_b == 1, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension");
    LSM_REQUIRE(*loc_a == 2 && *loc_b == 2, LIBSUPERMESH_GPU_ERROR_INVALID,
                "Unsupported element type: intervals have 2 vertices");
    lsm::sort_finder(*nnodes_a, *nelements_a, positions_a, enlist_a, *nnodes_b, *nelements_b,
                     positions_b, enlist_b, g_sort.indptr, g_sort.indices);
    g_sort.nelements_a = *nelements_a;
    g_sort.ready = true;
  });
}

void libsupermesh_sort_intersection_finder_query_output(long* nindices) {
  guarded([&] { finder_query_output(g_sort, nindices); });
}

void libsupermesh_sort_intersection_finder_get_output(long* nelements, long* nindices, long* indices,
                                                      long* ind_ptr) {
  guarded([&] { finder_get_output(g_sort, *nelements, *nindices, indices, ind_ptr); });
}

int libsupermesh_gpu_tree_create(int finder, int dim, long nnodes, const double* positions, long nelements,
                                 int loc, const long* enlist, libsupermesh_gpu_tree** tree) {
  return guarded([&] {
    LSM_REQUIRE(tree != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null tree");
    *tree = nullptr;
    LSM_REQUIRE(finder == LIBSUPERMESH_GPU_FINDER_TREE || finder == LIBSUPERMESH_GPU_FINDER_RTREE,
                LIBSUPERMESH_GPU_ERROR_INVALID, "unknown finder");
    LSM_REQUIRE(positions != nullptr && (nelements == 0 || enlist != nullptr), LIBSUPERMESH_GPU_ERROR_INVALID,
                "null mesh arrays");
    auto t = std::make_unique<libsupermesh_gpu_tree>();
    t->finder = finder;
    t->dim = dim;
    Context& C = t->ctx;
    init_context(C, 0);
    try {
      libsupermesh_gpu_mesh vb;
      host_mesh_to_device(C, C.mesh_b, nnodes, dim, nelements, loc, positions, enlist, vb);
      stage_mesh(C, C.mesh_b, dim, loc, nnodes, nelements, vb.positions, vb.enlist, false, false);
      build_bvh(C, C.mesh_b);
      LSM_HIP(hipStreamSynchronize(C.stream));
    } catch (...) {
      free_context(C);
      throw;
    }
    *tree = t.release();
  });
}

int libsupermesh_gpu_tree_query(libsupermesh_gpu_tree* tree, int loc, const double* element, long* neles) {
  return guarded([&] {
    LSM_REQUIRE(tree != nullptr && neles != nullptr && element != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID,
                "null argument");
    Context& C = tree->ctx;
    std::lock_guard<std::mutex> lk(C.mu);
    *neles = 0;
    tree->result.clear();
    LSM_HIP(hipSetDevice(C.device));
    const int dim = tree->dim;
    std::vector<long> enl((size_t)loc);
    for (int v = 0; v < loc; v++) enl[v] = v;
    libsupermesh_gpu_mesh va;
    host_mesh_to_device(C, C.mesh_a, loc, dim, 1, loc, element, enl.data(), va);
    stage_mesh(C, C.mesh_a, dim, loc, loc, 1, va.positions, va.enlist, false, false);
    const long n = find_pairs(C, C.mesh_a, C.mesh_b, tree->finder, LIBSUPERMESH_GPU_SORTED);
    tree->result.resize((size_t)n);
    if (n > 0)
      LSM_HIP(hipMemcpyAsync(tree->result.data(), C.indices.ptr, sizeof(int) * n, hipMemcpyDeviceToHost,
                             C.stream));
    LSM_HIP(hipStreamSynchronize(C.stream));
    *neles = n;
  });
}

int libsupermesh_gpu_tree_result(libsupermesh_gpu_tree* tree, long neles, long* eles) {
  return guarded([&] {
    LSM_REQUIRE(tree != nullptr, LIBSUPERMESH_GPU_ERROR_INVALID, "null tree");
    std::lock_guard<std::mutex> lk(tree->ctx.mu);
    LSM_REQUIRE(neles == (long)tree->result.size() && (neles == 0 || eles != nullptr),
                LIBSUPERMESH_GPU_ERROR_INVALID, "result size differs from the last query's");
    for (long i = 0; i < neles; i++) eles[i] = tree->result[i];
  });
}

void libsupermesh_gpu_tree_destroy(libsupermesh_gpu_tree* tree) {
  if (!tree) return;
  free_context(tree->ctx);
  delete tree;
}

void libsupermesh_advancing_front_intersection_finder_set_input(
    long* nnodes_a, int* dim_a, long* nelements_a, int* loc_a, long* nnodes_b, int* dim_b, long* nelements_b,
    int* loc_b, double* positions_a, long* enlist_a, double* positions_b, long* enlist_b) {
  guarded([&] {
    std::lock_guard<std::mutex> lk(g_front.mu);
    g_front.ready = false;
    g_front.indptr.clear();
    g_front.indices.clear();
    LSM_REQUIRE(*dim_a == *dim_b, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension");
    Context& C = *default_context();
    std::lock_guard<std::mutex> lk2(C.mu);
    LSM_HIP(hipSetDevice(C.device));
    lsm::advancing_front_finder(C, *dim_a, *nnodes_a, *nelements_a, *loc_a, positions_a, enlist_a, *nnodes_b,
                                *nelements_b, *loc_b, positions_b, enlist_b, g_front.indptr, g_front.indices);
    g_front.nelements_a = *nelements_a;
    g_front.ready = true;
  });
}

void libsupermesh_advancing_front_intersection_finder_query_output(long* nindices) {
  guarded([&] { finder_query_output(g_front, nindices); });
}

void libsupermesh_advancing_front_intersection_finder_get_output(long* nelements, long* nindices, long* indices,
                                                                 long* ind_ptr) {
  guarded([&] { finder_get_output(g_front, *nelements, *nindices, indices, ind_ptr); });
}

void libsupermesh_mesh_nelist_set_input(long* nnodes, long* nelements, long* loc, long* enlist) {
  guarded([&] {
    graph_set_input(g_nelist, [&](Context& C, std::vector<long>& ptr, std::vector<long>& idx) {
      lsm::mesh_nelist(C, *nnodes, *nelements, *loc, enlist, ptr, idx);
    });
  });
}

void libsupermesh_mesh_nelist_query_output(long* nindices) {
  guarded([&] { graph_query_output(g_nelist, nindices); });
}

void libsupermesh_mesh_nelist_get_output(long* nnodes, long* nindices, long* indices, long* ind_ptr) {
  guarded([&] { graph_get_output(g_nelist, *nnodes, *nindices, indices, ind_ptr); });
}

void libsupermesh_mesh_nnlist_set_input(long* nnodes, long* nelements, long* loc, long* enlist) {
  guarded([&] {
    graph_set_input(g_nnlist, [&](Context& C, std::vector<long>& ptr, std::vector<long>& idx) {
      lsm::mesh_nnlist(C, *nnodes, *nelements, *loc, enlist, ptr, idx);
    });
  });
}

void libsupermesh_mesh_nnlist_query_output(long* nindices) {
  guarded([&] { graph_query_output(g_nnlist, nindices); });
}

void libsupermesh_mesh_nnlist_get_output(long* nnodes, long* nindices, long* indices, long* ind_ptr) {
  guarded([&] { graph_get_output(g_nnlist, *nnodes, *nindices, indices, ind_ptr); });
}

void libsupermesh_mesh_eelist_set_input(long* nnodes, long* nelements, long* loc, long* sloc, long* enlist) {
  guarded([&] {
    // mesh_eelist without nneigh_ele: room for loc neighbours per element
    graph_set_input(g_eelist, [&](Context& C, std::vector<long>& ptr, std::vector<long>& idx) {
      lsm::mesh_eelist(C, *nnodes, *nelements, *loc, *sloc, *loc, enlist, ptr, idx);
    });
  });
}

void libsupermesh_mesh_eelist_query_output(long* nindices) {
  guarded([&] { graph_query_output(g_eelist, nindices); });
}

void libsupermesh_mesh_eelist_get_output(long* nelements, long* nindices, long* indices, long* ind_ptr) {
  guarded([&] { graph_get_output(g_eelist, *nelements, *nindices, indices, ind_ptr); });
}

void libsupermesh_transpose_sparsity(long* nnodes, long* nnodes_t, long* nindices, long* indices, long* ind_ptr,
                                     long* indices_t, long* ind_ptr_t) {
  guarded([&] {
    Context& C = *default_context();
    std::lock_guard<std::mutex> lk(C.mu);
    lsm::transpose_sparsity(C, *nnodes, *nnodes_t, *nindices, indices, ind_ptr, indices_t, ind_ptr_t);
  });
}

}  // extern "C"
