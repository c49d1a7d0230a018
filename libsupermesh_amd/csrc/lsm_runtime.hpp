// lsm_runtime.hpp -- context, device buffers and error reporting shared by
// the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <vector>
#include <stdexcept>
#include <string>

#include "../../include/libsupermesh_gpu.h"

namespace lsm {

// Raised inside the library, turned into a status code + last-error text at
// the C boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(int code, const char* msg);

#define LSM_HIP(call)                                                               \
  do {                                                                              \
    hipError_t _e = (call);                                                         \
    if (_e != hipSuccess)                                                           \
      throw ::lsm::Error(LIBSUPERMESH_GPU_ERROR_HIP,                                \
                         std::string(#call " failed: ") + hipGetErrorString(_e));   \
  } while (0)

#define LSM_REQUIRE(cond, code, msg)                \
  do {                                              \
    if (!(cond)) throw ::lsm::Error((code), (msg)); \
  } while (0)

// A device buffer that only grows; reused across calls on one context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void* get(size_t need) {
    if (need == 0) need = 1;
    if (need > bytes) {
      // queued kernels may still use the old allocation
      if (ptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(ptr);
      }
      ptr = nullptr;
      bytes = 0;
      size_t cap = need + std::min(need / 4, (size_t)256 << 20);
      LSM_HIP(hipMalloc(&ptr, cap));
      bytes = cap;
    }
    return ptr;
  }
  template <class T>
  T* as(size_t n) {
    return static_cast<T*>(get(n * sizeof(T)));
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

// Mesh data staged on the device in the layouts the kernels read.
struct DevMesh {
  int dim = 0, loc = 0;
  long nnodes = 0, nelements = 0;
  const int* enlist_dev = nullptr;  // element-node list on the device (borrowed)
  DevBuf enl_copy, pos_copy;        // owned copies when the caller passed host arrays
  DevBuf coords;   // double [ne][loc][dim]: element vertex coordinates
  DevBuf bbox;     // double [2*dim][ne]: lo_x.. lo_z, hi_x.. hi_z (SoA)
  DevBuf meas_sp;  // double [ne]: SPACING of the clipper's element measure
  DevBuf scale;    // double [ne]: max(1, max |vertex coordinate|)
  DevBuf planes;   // double [ne][16]: tet planes (normals, c) of B
  DevBuf affine;   // double [ne][loc][dim+1]: barycentric maps (mass)
  // batches: element e belongs to mesh group[e]; ngroups == 0 outside batches
  long ngroups = 0;
  DevBuf group;    // int [ne]
};

// Implicit F-ary bounding volume hierarchy over Morton-sorted elements.
struct DevBVH {
  int dim = 0;
  long n = 0;
  int levels = 0;           // level 0 = elements, levels-1 = root
  long count[32];           // nodes per level
  long offset[32];          // node offset of each level in the box arrays
  long total = 0;
  DevBuf boxes;             // double [total][2*dim]: lo_0.., hi_0.. per node
  DevBuf perm;              // int [n]: sorted position -> element id
  bool grouped = false;     // batches: granges holds each node's mesh range
  DevBuf granges;           // int [total][2]: lowest and highest mesh id below the node
};

// Learned output-size predictor state (predictor.hip).
struct Predictor {
  bool ready = false;
  long steps = 0;           // Adam steps taken
  float quantile = 0.9f;    // budget = this quantile of the predicted count
  int level = 0;            // quantile level the last single pass used (0 = quantile)
  float arena_fraction = 0.25f;  // arena size: this fraction of the expected output ...
  long arena_min = 1L << 20;     // ... plus this many tets (both 0: no arena)
  float spill_fraction = 0.25f;  // spill reserve: this fraction of the expected output ...
  long spill_min = 1L << 20;     // ... plus this many tets
  DevBuf weights, bias;     // BF16 layouts, FP32 biases [4][64]
  DevBuf master, adam_m, adam_v, grad;
  DevBuf features, labels;  // collected training samples
  long nsamples = 0;
  DevBuf act, loss, index;  // training scratch
  DevBuf budget, overflow, spill, aoff;  // single-pass clip: budgets, flags, spill offsets, arena offsets
};

// Launch shape of the count-pass kernel: block size, grid (blocks), and in
// 3D the walk's tet depths held in registers (the rest of 1..3 in LDS).
struct CountShape {
  int block = 0, grid = 0, reg = 0;
  float ms = 0.f;  // its time on the sample it was measured on
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  DevMesh mesh_a, mesh_b;
  DevBVH bvh;
  // finder
  DevBuf keys, keys_alt, vals, vals_alt, sort_tmp, scan_tmp;
  DevBuf qperm;                  // A elements in Morton order
  DevBuf row_count, indptr, indices, rows;
  DevBuf fscratch;               // single-pass search results (slot-major)
  // clipping
  DevBuf counts, measure, partials, scalars, offsets, elements, blocks;
  DevBuf offsets_dense, elements_dense;  // LIBSUPERMESH_GPU_DENSE: the single pass's runs compacted
  DevBuf dense_ctl;              // the dense move's fit flag and step cursor
  bool runs_layout = false;      // the latest result holds the single pass's runs apart
  DevBuf work, nwork;            // compacted list of pairs that need the full clip
  DevBuf queue;                  // the tet walk's work-item counters, one per XCD share
  // mass assembly
  DevBuf mflag, mslot, mkeys, mvals, mkeys_alt, mvals_alt, ukeys, uvals, nuniq, mindptr, mindices;
  // 2D row slots: each (element, vertex) row's first slot
  DevBuf mrbase;
  // batches: per-mesh pair offsets and measure sums
  DevBuf gpairs, gsum;
  // pinned status words the pipeline copies back asynchronously and reads
  // after its next synchronisation: [0] staging flag, [1] longest mass row
  long* hstat = nullptr;
  Predictor predictor;
  // 2D mixed mass: rows laid out before the clip and filled by it
  bool mass_rows_ready = false;
  long mass_maxraw = 0;
  // timing
  hipEvent_t ev[16];
  bool events = false;
  bool timing = false;  // record per-phase / clip-kernel events this call
  // count pass, per dimension ([0] 2D, [1] 3D): the launch shape measured
  // fastest, and the pass recorded as a graph (LIBSUPERMESH_GPU_GRAPH) with
  // the arguments it was recorded for
  CountShape shape[2];
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  std::vector<const void*> graph_key[2];
  long graph_cap = 0;          // pairs the recorded passes' buffers hold
  bool graph_refused = false;  // a recording held other than kernel nodes: not used again
};

Context* default_context();

// Mesh staging + BVH (finder.hip).
// With defer_check the node-index check is left to the caller: it clears the
// flag (stage_flag) before staging and reads it back later.
void stage_mesh(Context& C, DevMesh& M, int dim, int loc, long nnodes, long nelements,
                const double* positions, const int* enlist, bool need_planes, bool need_affine,
                bool defer_check = false);
int* stage_flag(Context& C);
// Mesh id of every element of a batch from its element offsets (device,
// [ngroups + 1]); validates the offsets.
void stage_groups(Context& C, DevMesh& M, long ngroups, const long* offsets);
void build_bvh(Context& C, const DevMesh& B);
// Count / fill candidate pairs of A against the BVH of B.  Returns npairs.
long find_pairs(Context& C, const DevMesh& A, const DevMesh& B, int finder, int sort_rows);
// The same pairs found the other way round (B's elements query a hierarchy
// over A), regrouped as a CSR over A with every row in ascending B id order.
long find_pairs_from_b(Context& C, const DevMesh& A, const DevMesh& B, int finder);

// Clipping (supermesh.hip).
void clip_pairs(Context& C, const DevMesh& A, const DevMesh& B, long npairs, int flags,
                libsupermesh_gpu_result* out);
void assemble_mass(Context& C, const DevMesh& A, const DevMesh& B, long npairs,
                   libsupermesh_gpu_result* out);
// Mesh adjacency (graphs.hip): host element-node lists in, host CSR out.
void mesh_nelist(Context& C, long nnodes, long ne, long loc, const long* enlist, std::vector<long>& indptr,
                 std::vector<long>& indices);
void mesh_nnlist(Context& C, long nnodes, long ne, long loc, const long* enlist, std::vector<long>& indptr,
                 std::vector<long>& indices);
void mesh_eelist(Context& C, long nnodes, long ne, long loc, long sloc, long nneigh, const long* enlist,
                 std::vector<long>& indptr, std::vector<long>& indices);
void transpose_sparsity(Context& C, long nrows, long ncols, long nindices, const long* indices, const long* ind_ptr,
                        long* indices_t, long* ind_ptr_t);

// The advancing front finder (advancing_front.cpp): host arrays, 0-based.
void advancing_front_finder(Context& C, int dim, long nn_a, long ne_a, int loc_a, const double* pos_a,
                            const long* enl_a, long nn_b, long ne_b, int loc_b, const double* pos_b,
                            const long* enl_b, std::vector<long>& indptr, std::vector<int>& indices);

// Clipping of candidate pairs where A or B is not a simplex mesh (quads in
// 2D, hexahedra in 3D).
// Explicit pair list, no allocation: asynchronous on `stream`, results and
// the measure sum written to the caller's device arrays (supermesh.hip).
void clip_pairs_noalloc(int device, hipStream_t stream, const libsupermesh_gpu_mesh& a,
                        const libsupermesh_gpu_mesh& b, long npairs, const int* pair_a, const int* pair_b,
                        int* counts, double* measure, double* sum);
void clip_pairs_mixed(Context& C, const DevMesh& A, const DevMesh& B, long npairs, int flags,
                      libsupermesh_gpu_result* out);
// Order-sensitive checksum of a device buffer (whole 32-bit words).
unsigned long long checksum(Context& C, const void* ptr, size_t bytes);
void compact_result(Context& C, int dim, libsupermesh_gpu_result* out);
// The count pass of a simplex supermesh (counts, measures, sums, per-mesh
// sums of a batch) in the launch shape measured fastest, as one recorded
// graph (LIBSUPERMESH_GPU_GRAPH).
void count_pass_graph(Context& C, const DevMesh& A, const DevMesh& B, long npairs, libsupermesh_gpu_result* out,
                      long ngroups, const long* a_off);
void release_graph(Context& C);
// Work list (pairs that needed the full clip) of the last 3D clip_pairs call.
long last_work_list(Context& C, const int** work);
// Batches: per-mesh candidate pair offsets (gpairs, from the CSR row pointers
// at the A mesh offsets) and compensated measure sums per mesh (gsum).
void batch_sums(Context& C, long ngroups, const long* a_offsets, const long* indptr,
                const double* measure);

inline unsigned grid_for(long n, int block) {
  long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 0x7fffffffL) g = 0x7fffffffL;
  return (unsigned)g;
}

// 1D entry points (interval.cpp), host code.
double interval_size(const double* v);
int intersect_intervals(const double* a, const double* b, double* c);
void sort_finder(long nnodes_a, long ne_a, const double* pos_a, const long* enl_a, long nnodes_b,
                 long ne_b, const double* pos_b, const long* enl_b, std::vector<long>& indptr,
                 std::vector<int>& indices);

}  // namespace lsm
