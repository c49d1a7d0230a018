// Mesh adjacency lists and CSR transposes on the device: node -> element
// (mesh_nelist, Graphs.F90:76-116), node -> node (mesh_nnlist,
// C_bindings.F90:215-255), element -> element (mesh_eelist, Graphs.F90:118-187)
// and transpose_sparsity (C_bindings.F90:397-430).  Each result equals the
// reference's, order included:
//   * nelist: per node, the elements holding it in ascending element order
//     (one entry per occurrence) -- a stable radix sort of the flattened
//     element-node list by node id;
//   * nnlist: per node, the ascending set of nodes sharing an element with it
//     (std::set in Data_Structures_C++.cpp) -- sort + unique of (node, node)
//     keys;
//   * eelist: per element, its neighbours in the order the reference finds
//     them -- in local-node order j, then ascending element id, a neighbour
//     is listed when its sloc-th shared node is met, at most nneigh per
//     element (the release build's early exit, Graphs.F90:170-175);
//   * transpose: per column, the rows holding it in ascending row order
//     (stable sort of the entries by column).
// These are integer gather / sort kernels bound by memory traffic.

#include <algorithm>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include "lsm_runtime.hpp"

namespace lsm {

namespace {

struct ScopedBuf : DevBuf {
  ScopedBuf() = default;
  ScopedBuf(const ScopedBuf&) = delete;
  ~ScopedBuf() { release(); }
};

int bits_for(long n) {
  int b = 1;
  while (b < 63 && (1L << b) < n) b++;
  return b;
}

__global__ void __launch_bounds__(256)
k_range_check(long n, const long* __restrict__ v, long hi, int* __restrict__ bad) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (v[i] < 0 || v[i] >= hi)) atomicOr(bad, 1);
}

__global__ void __launch_bounds__(256)
k_node_keys(long n, int loc, const long* __restrict__ enl, unsigned* __restrict__ key, int* __restrict__ val) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = (unsigned)enl[i];
  val[i] = (int)(i / loc);
}

// indptr[r] = first position of row r in the sorted row ids (rows in
// [0, nrows)), indptr[nrows] = n.
template <class Row>
__global__ void __launch_bounds__(256)
k_row_bounds(long n, Row row, long nrows, long* __restrict__ indptr) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const long lo = i == 0 ? 0 : (long)row(i - 1) + 1;
  const long hi = i == n ? nrows : (long)row(i);
  for (long r = lo; r <= hi && r <= nrows; r++) indptr[r] = i;
}

struct RowOfU32 {
  const unsigned* k;
  __device__ unsigned operator()(long i) const { return k[i]; }
};
struct RowOfU64 {
  const unsigned long long* k;
  __device__ unsigned operator()(long i) const { return (unsigned)(k[i] >> 32); }
};

__global__ void __launch_bounds__(256)
k_int_to_long(long n, const int* __restrict__ in, long* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

__global__ void __launch_bounds__(256)
k_low_to_long(long n, const unsigned long long* __restrict__ in, long* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (long)(in[i] & 0xffffffffull);
}

__global__ void __launch_bounds__(256)
k_fill_long(long n, long v, long* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}

// Both directions of every local node pair (i < j) of every element:
// key = (node_i << 32) | node_j.
__global__ void __launch_bounds__(256)
k_pair_keys(long ne, int loc, const long* __restrict__ enl, unsigned long long* __restrict__ key) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  unsigned long long* o = key + e * (long)loc * (loc - 1);
  int k = 0;
  for (int i = 0; i < loc; i++) {
    const unsigned long long a = (unsigned long long)enl[e * loc + i];
    for (int j = i + 1; j < loc; j++) {
      const unsigned long long b = (unsigned long long)enl[e * loc + j];
      o[k++] = (a << 32) | b;
      o[k++] = (b << 32) | a;
    }
  }
}

// Neighbours of element i in the reference's discovery order, up to nneigh.
// Element e's entries in a node's list are consecutive (the list is in
// element order), so e is handled once per node, at its first entry: with S
// shared nodes among local nodes 0..j-1 and m occurrences of node j in e, the
// reference's counter passes sloc at node j iff S < sloc <= S + m.
__global__ void __launch_bounds__(256)
k_eelist(long ne, int loc, int sloc, int nneigh, const long* __restrict__ enl,
         const long* __restrict__ nel_ptr, const int* __restrict__ nel, int* __restrict__ out,
         int* __restrict__ count) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  int n = 0;
  for (int j = 0; j < loc && n < nneigh; j++) {
    const long x = enl[i * loc + j];
    const long t0 = nel_ptr[x], t1 = nel_ptr[x + 1];
    for (long t = t0; t < t1 && n < nneigh; t++) {
      const int e = nel[t];
      if (e == i || (t > t0 && nel[t - 1] == e)) continue;
      int m = 0, s = 0;
      for (int l = 0; l < loc; l++) {
        const long y = enl[(long)e * loc + l];
        m += y == x;
        for (int k = 0; k < j; k++) s += y == enl[i * loc + k];
      }
      if (s < sloc && sloc <= s + m) out[i * nneigh + n++] = e;
    }
  }
  count[i] = n;
}

__global__ void __launch_bounds__(256)
k_eelist_compact(long ne, int nneigh, const int* __restrict__ tmp, const int* __restrict__ count,
                 const long* __restrict__ ptr, long* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  for (int k = 0; k < count[i]; k++) out[ptr[i] + k] = tmp[i * nneigh + k];
}

__global__ void __launch_bounds__(256)
k_count_long(long n, const int* __restrict__ c, long* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = c[i];
}

// Row id of every entry of a CSR whose row r spans [ptr[r], ptr[r+1]),
// positions taken relative to ptr[0]; columns as 32-bit keys.
__global__ void __launch_bounds__(256)
k_entry_rows(long nrows, const long* __restrict__ ptr, const long* __restrict__ cols, int* __restrict__ rowid,
             unsigned* __restrict__ key) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const long base = ptr[0];
  for (long t = ptr[r]; t < ptr[r + 1]; t++) {
    rowid[t - base] = (int)r;
    key[t - base] = (unsigned)cols[t];
  }
}

template <class T>
T* upload(ScopedBuf& b, const T* h, size_t n, hipStream_t s) {
  T* d = b.as<T>(n);
  if (n) LSM_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}

template <class T>
void download(std::vector<T>& h, const T* d, size_t n, hipStream_t s) {
  h.resize(n);
  if (n) LSM_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

void check_nodes(Context& C, long n, const long* d, long hi, const char* what) {
  if (n == 0) return;
  int* bad = C.scalars.as<int>(64);
  LSM_HIP(hipMemsetAsync(bad, 0, sizeof(int), C.stream));
  hipLaunchKernelGGL(k_range_check, dim3(grid_for(n, 256)), dim3(256), 0, C.stream, n, d, hi, bad);
  LSM_HIP(hipGetLastError());
  int h = 0;
  LSM_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
  LSM_REQUIRE(h == 0, LIBSUPERMESH_GPU_ERROR_INVALID, what);
}

void check_sizes(long nnodes, long ne, long loc) {
  LSM_REQUIRE(nnodes >= 0 && ne >= 0 && loc >= 1, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid mesh size");
  LSM_REQUIRE(nnodes < (1L << 31) && ne * loc < (1L << 31), LIBSUPERMESH_GPU_ERROR_OVERFLOW,
              "Mesh larger than 2^31 nodes or element-node entries");
}

// Node -> element CSR of a device element-node list (stream-ordered; the
// results stay in ptr / idx).
void nelist_device(Context& C, long nnodes, long ne, int loc, const long* denl, ScopedBuf& ptr, ScopedBuf& idx) {
  const long n = ne * loc;
  ScopedBuf keys, keys2, vals, tmp;
  unsigned* k1 = keys.as<unsigned>(n);
  unsigned* k2 = keys2.as<unsigned>(n);
  int* v1 = vals.as<int>(2 * n);
  int* v2 = v1 + n;
  long* p = ptr.as<long>(nnodes + 1);
  int* out = idx.as<int>(n);
  if (n == 0) {
    hipLaunchKernelGGL(k_fill_long, dim3(grid_for(nnodes + 1, 256)), dim3(256), 0, C.stream, nnodes + 1, 0L, p);
    LSM_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_node_keys, dim3(grid_for(n, 256)), dim3(256), 0, C.stream, n, loc, denl, k1, v1);
  LSM_HIP(hipGetLastError());
  size_t bytes = 0;
  const unsigned end_bit = (unsigned)bits_for(nnodes);
  LSM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k1, k2, v1, v2, (size_t)n, 0u, end_bit, C.stream));
  void* t = tmp.get(bytes);
  LSM_HIP(rocprim::radix_sort_pairs(t, bytes, k1, k2, v1, v2, (size_t)n, 0u, end_bit, C.stream));
  hipLaunchKernelGGL(k_row_bounds<RowOfU32>, dim3(grid_for(n + 1, 256)), dim3(256), 0, C.stream, n,
                     RowOfU32{k2}, nnodes, p);
  LSM_HIP(hipMemcpyAsync(out, v2, sizeof(int) * n, hipMemcpyDeviceToDevice, C.stream));
  LSM_HIP(hipGetLastError());
}

}  // namespace

void mesh_nelist(Context& C, long nnodes, long ne, long loc, const long* enlist, std::vector<long>& indptr,
                 std::vector<long>& indices) {
  check_sizes(nnodes, ne, loc);
  LSM_HIP(hipSetDevice(C.device));
  const long n = ne * loc;
  ScopedBuf denl, ptr, idx, wide;
  const long* d = upload(denl, enlist, (size_t)n, C.stream);
  check_nodes(C, n, d, nnodes, "Invalid node: element-node index out of range");
  nelist_device(C, nnodes, ne, (int)loc, d, ptr, idx);
  long* w = wide.as<long>(n);
  if (n) hipLaunchKernelGGL(k_int_to_long, dim3(grid_for(n, 256)), dim3(256), 0, C.stream, n,
                            static_cast<const int*>(idx.ptr), w);
  LSM_HIP(hipGetLastError());
  download(indptr, static_cast<const long*>(ptr.ptr), (size_t)nnodes + 1, C.stream);
  download(indices, static_cast<const long*>(w), (size_t)n, C.stream);
  LSM_HIP(hipStreamSynchronize(C.stream));
}

void mesh_nnlist(Context& C, long nnodes, long ne, long loc, const long* enlist, std::vector<long>& indptr,
                 std::vector<long>& indices) {
  check_sizes(nnodes, ne, loc);
  LSM_HIP(hipSetDevice(C.device));
  const long n = ne * loc, np = ne * loc * (loc - 1);
  LSM_REQUIRE(np < (1L << 31), LIBSUPERMESH_GPU_ERROR_OVERFLOW, "too many node pairs");
  ScopedBuf denl, keys, keys2, tmp, uniq, nuniq, ptr, out;
  const long* d = upload(denl, enlist, (size_t)n, C.stream);
  check_nodes(C, n, d, nnodes, "Invalid node: element-node index out of range");
  long* p = ptr.as<long>(nnodes + 1);
  long m = 0;
  if (np > 0) {
    unsigned long long* k1 = keys.as<unsigned long long>(np);
    unsigned long long* k2 = keys2.as<unsigned long long>(np);
    hipLaunchKernelGGL(k_pair_keys, dim3(grid_for(ne, 256)), dim3(256), 0, C.stream, ne, (int)loc, d, k1);
    LSM_HIP(hipGetLastError());
    size_t bytes = 0;
    const unsigned end_bit = 32u + (unsigned)bits_for(nnodes);
    LSM_HIP(rocprim::radix_sort_keys(nullptr, bytes, k1, k2, (size_t)np, 0u, end_bit, C.stream));
    void* t = tmp.get(bytes);
    LSM_HIP(rocprim::radix_sort_keys(t, bytes, k1, k2, (size_t)np, 0u, end_bit, C.stream));
    unsigned long long* u = uniq.as<unsigned long long>(np);
    size_t* cnt = nuniq.as<size_t>(1);
    bytes = 0;
    LSM_HIP(rocprim::unique(nullptr, bytes, k2, u, cnt, (size_t)np, rocprim::equal_to<unsigned long long>(),
                            C.stream));
    t = tmp.get(bytes);
    LSM_HIP(rocprim::unique(t, bytes, k2, u, cnt, (size_t)np, rocprim::equal_to<unsigned long long>(),
                            C.stream));
    size_t hm = 0;
    LSM_HIP(hipMemcpyAsync(&hm, cnt, sizeof(size_t), hipMemcpyDeviceToHost, C.stream));
    LSM_HIP(hipStreamSynchronize(C.stream));
    m = (long)hm;
    hipLaunchKernelGGL(k_row_bounds<RowOfU64>, dim3(grid_for(m + 1, 256)), dim3(256), 0, C.stream, m,
                       RowOfU64{u}, nnodes, p);
    long* o = out.as<long>(m);
    hipLaunchKernelGGL(k_low_to_long, dim3(grid_for(m, 256)), dim3(256), 0, C.stream, m, u, o);
  } else {
    hipLaunchKernelGGL(k_fill_long, dim3(grid_for(nnodes + 1, 256)), dim3(256), 0, C.stream, nnodes + 1, 0L, p);
  }
  LSM_HIP(hipGetLastError());
  download(indptr, p, (size_t)nnodes + 1, C.stream);
  download(indices, static_cast<const long*>(out.get(8)), (size_t)m, C.stream);
  LSM_HIP(hipStreamSynchronize(C.stream));
}

void mesh_eelist(Context& C, long nnodes, long ne, long loc, long sloc, long nneigh, const long* enlist,
                 std::vector<long>& indptr, std::vector<long>& indices) {
  check_sizes(nnodes, ne, loc);
  LSM_REQUIRE(nneigh >= 0 && nneigh <= 64, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid neighbour count");
  LSM_HIP(hipSetDevice(C.device));
  const long n = ne * loc;
  ScopedBuf denl, nptr, nidx, tmp, cnt, wide, ptr, out, scan_tmp;
  const long* d = upload(denl, enlist, (size_t)n, C.stream);
  check_nodes(C, n, d, nnodes, "Invalid node: element-node index out of range");
  nelist_device(C, nnodes, ne, (int)loc, d, nptr, nidx);
  long* p = ptr.as<long>(ne + 1);
  LSM_HIP(hipMemsetAsync(p, 0, sizeof(long), C.stream));
  long total = 0;
  if (ne > 0) {
    const int s = (int)std::max(-1L, std::min(sloc, loc * loc + 1));  // counts never exceed loc^2
    int* t = tmp.as<int>((size_t)ne * std::max(1L, nneigh));
    int* c = cnt.as<int>(ne);
    hipLaunchKernelGGL(k_eelist, dim3(grid_for(ne, 256)), dim3(256), 0, C.stream, ne, (int)loc, s, (int)nneigh, d,
                       static_cast<const long*>(nptr.ptr), static_cast<const int*>(nidx.ptr), t, c);
    LSM_HIP(hipGetLastError());
    long* w = wide.as<long>(ne);
    hipLaunchKernelGGL(k_count_long, dim3(grid_for(ne, 256)), dim3(256), 0, C.stream, ne, c, w);
    size_t bytes = 0;
    LSM_HIP(rocprim::inclusive_scan(nullptr, bytes, w, p + 1, (size_t)ne, rocprim::plus<long>(), C.stream));
    void* st = scan_tmp.get(bytes);
    LSM_HIP(rocprim::inclusive_scan(st, bytes, w, p + 1, (size_t)ne, rocprim::plus<long>(), C.stream));
    LSM_HIP(hipMemcpyAsync(&total, p + ne, sizeof(long), hipMemcpyDeviceToHost, C.stream));
    LSM_HIP(hipStreamSynchronize(C.stream));
    long* o = out.as<long>(total);
    hipLaunchKernelGGL(k_eelist_compact, dim3(grid_for(ne, 256)), dim3(256), 0, C.stream, ne, (int)nneigh, t, c,
                       p, o);
    LSM_HIP(hipGetLastError());
  }
  download(indptr, p, (size_t)ne + 1, C.stream);
  download(indices, static_cast<const long*>(out.get(8)), (size_t)total, C.stream);
  LSM_HIP(hipStreamSynchronize(C.stream));
}

void transpose_sparsity(Context& C, long nrows, long ncols, long nindices, const long* indices, const long* ind_ptr,
                        long* indices_t, long* ind_ptr_t) {
  LSM_REQUIRE(nrows >= 0 && ncols >= 0 && nindices >= 0, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid sizes");
  LSM_REQUIRE(nrows < (1L << 31) && ncols < (1L << 31) && nindices < (1L << 31), LIBSUPERMESH_GPU_ERROR_OVERFLOW,
              "Sparsity larger than 2^31 rows, columns or entries");
  LSM_REQUIRE(ind_ptr[0] >= 0 && ind_ptr[nrows] <= nindices, LIBSUPERMESH_GPU_ERROR_INVALID,
              "Invalid sparsity: ind_ptr outside the indices");
  for (long r = 0; r < nrows; r++)
    LSM_REQUIRE(ind_ptr[r] <= ind_ptr[r + 1], LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid sparsity: ind_ptr decreases");
  LSM_HIP(hipSetDevice(C.device));
  const long base = ind_ptr[0], n = ind_ptr[nrows] - ind_ptr[0];
  ScopedBuf dcol, dptr, rows, rows2, keys, keys2, tmp, pt, wide;
  const long* c = upload(dcol, indices, (size_t)nindices, C.stream);
  const long* p = upload(dptr, ind_ptr, (size_t)nrows + 1, C.stream);
  check_nodes(C, n, c + base, ncols, "Invalid sparsity: column index out of range");
  long* q = pt.as<long>(ncols + 1);
  long* w = wide.as<long>(n);
  if (n > 0) {
    int* r1 = rows.as<int>(n);
    int* r2 = rows2.as<int>(n);
    unsigned* k1 = keys.as<unsigned>(n);
    unsigned* k2 = keys2.as<unsigned>(n);
    hipLaunchKernelGGL(k_entry_rows, dim3(grid_for(nrows, 256)), dim3(256), 0, C.stream, nrows, p, c, r1, k1);
    LSM_HIP(hipGetLastError());
    size_t bytes = 0;
    const unsigned end_bit = (unsigned)bits_for(ncols);
    LSM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k1, k2, r1, r2, (size_t)n, 0u, end_bit, C.stream));
    void* t = tmp.get(bytes);
    LSM_HIP(rocprim::radix_sort_pairs(t, bytes, k1, k2, r1, r2, (size_t)n, 0u, end_bit, C.stream));
    hipLaunchKernelGGL(k_row_bounds<RowOfU32>, dim3(grid_for(n + 1, 256)), dim3(256), 0, C.stream, n,
                       RowOfU32{k2}, ncols, q);
    hipLaunchKernelGGL(k_int_to_long, dim3(grid_for(n, 256)), dim3(256), 0, C.stream, n, r2, w);
  } else {
    hipLaunchKernelGGL(k_fill_long, dim3(grid_for(ncols + 1, 256)), dim3(256), 0, C.stream, ncols + 1, 0L, q);
  }
  LSM_HIP(hipGetLastError());
  LSM_HIP(hipMemcpyAsync(ind_ptr_t, q, sizeof(long) * (ncols + 1), hipMemcpyDeviceToHost, C.stream));
  if (n > 0) LSM_HIP(hipMemcpyAsync(indices_t, w, sizeof(long) * n, hipMemcpyDeviceToHost, C.stream));
  LSM_HIP(hipStreamSynchronize(C.stream));
}

}  // namespace lsm
