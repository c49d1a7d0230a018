// advancing_front.cpp -- the advancing front intersection finder
// (advancing_front_intersection_finder, Intersection_Finder.F90:270-454).
//
// Each A element's candidate list starts from the list of the A neighbour
// that reached it first and grows through B's face neighbours while their
// boxes meet A's (closed test, Intersection_Finder.F90:429-450).  Its rows,
// order included, depend on the order the front visits A, a depth-first
// chain as long as A itself, so the front runs here on the host; the face
// neighbour lists of both meshes (mesh_eelist) come from the device.
#include <algorithm>
#include <vector>

#include "../../include/libsupermesh_gpu.h"
#include "lsm_runtime.hpp"

namespace lsm {

namespace {

// Face-sharing node count and face-neighbour capacity of an element type
// (sloc / nneigh_ele, Graphs.F90:189-223).
int shared_nodes(int dim, int loc) { return loc == dim + 1 ? loc - 1 : loc / 2; }
int face_neighbours(int dim, int loc) { return loc == dim + 1 ? loc : 2 * dim; }

struct Boxes {
  int dim = 0;
  std::vector<double> lo, hi;  // [n][dim]
  void build(int d, long n, int loc, const double* pos, const long* enl) {
    dim = d;
    lo.resize((size_t)n * d);
    hi.resize((size_t)n * d);
    for (long e = 0; e < n; e++)
      for (int k = 0; k < d; k++) {
        double l = pos[enl[e * loc] * d + k], h = l;
        for (int v = 1; v < loc; v++) {
          const double x = pos[enl[e * loc + v] * d + k];
          l = std::min(l, x);
          h = std::max(h, x);
        }
        lo[e * d + k] = l;
        hi[e * d + k] = h;
      }
  }
  // B's box b meets the box (alo, ahi) unless it lies wholly below or above
  // it on some axis.
  bool meets(const double* alo, const double* ahi, long b) const {
    for (int k = 0; k < dim; k++)
      if (hi[b * dim + k] < alo[k] || lo[b * dim + k] > ahi[k]) return false;
    return true;
  }
};

}  // namespace

void advancing_front_finder(Context& C, int dim, long nn_a, long ne_a, int loc_a, const double* pos_a,
                            const long* enl_a, long nn_b, long ne_b, int loc_b, const double* pos_b,
                            const long* enl_b, std::vector<long>& indptr, std::vector<int>& indices) {
  LSM_REQUIRE(dim >= 1 && dim <= 3, LIBSUPERMESH_GPU_ERROR_INVALID, "Invalid dimension");
  LSM_REQUIRE((loc_a == dim + 1 || loc_a == (1 << dim)) && (loc_b == dim + 1 || loc_b == (1 << dim)),
              LIBSUPERMESH_GPU_ERROR_INVALID, "Unsupported element type");
  LSM_REQUIRE(ne_b < (1L << 31), LIBSUPERMESH_GPU_ERROR_OVERFLOW, "Mesh larger than 2^31 elements");
  indptr.assign((size_t)ne_a + 1, 0);
  indices.clear();
  if (ne_a == 0 || ne_b == 0) return;
  // neighbour lists on the device (these also reject out-of-range node ids)
  std::vector<long> nptr_a, nidx_a, nptr_b, nidx_b;
  mesh_eelist(C, nn_b, ne_b, loc_b, shared_nodes(dim, loc_b), face_neighbours(dim, loc_b), enl_b, nptr_b, nidx_b);
  mesh_eelist(C, nn_a, ne_a, loc_a, shared_nodes(dim, loc_a), face_neighbours(dim, loc_a), enl_a, nptr_a, nidx_a);
  Boxes bb, ba;
  bb.build(dim, ne_b, loc_b, pos_b, enl_b);
  ba.build(dim, ne_a, loc_a, pos_a, enl_a);

  // rows are appended to `pool` in the order the front visits A
  std::vector<int> pool;
  std::vector<long> start((size_t)ne_a, -1), len((size_t)ne_a, 0);
  std::vector<char> in_front_b((size_t)ne_b, 0), visited_a((size_t)ne_a, 0);
  std::vector<int> front_b;
  std::vector<long> stack_a;  // (element, clue) pairs
  front_b.reserve(1024);

  auto push_neighbours = [&](long a) {
    for (long k = nptr_a[a]; k < nptr_a[a + 1]; k++) {
      const long n = nidx_a[k];
      if (!visited_a[n]) {
        visited_a[n] = 1;
        stack_a.push_back(n);
        stack_a.push_back(a);
      }
    }
  };

  long seed = 0;
  while (true) {
    // a new connected part of A: its first element against all of B
    visited_a[seed] = 1;
    start[seed] = (long)pool.size();
    const double* slo = &ba.lo[seed * dim];
    const double* shi = &ba.hi[seed * dim];
    for (long b = 0; b < ne_b; b++)
      if (bb.meets(slo, shi, b)) pool.push_back((int)b);
    len[seed] = (long)pool.size() - start[seed];
    push_neighbours(seed);
    while (!stack_a.empty()) {
      const long clue = stack_a.back();
      stack_a.pop_back();
      const long a = stack_a.back();
      stack_a.pop_back();
      const double* lo = &ba.lo[a * dim];
      const double* hi = &ba.hi[a * dim];
      front_b.assign(pool.begin() + start[clue], pool.begin() + start[clue] + len[clue]);
      for (int b : front_b) in_front_b[b] = 1;
      start[a] = (long)pool.size();
      for (size_t i = 0; i < front_b.size(); i++) {
        const int b = front_b[i];
        if (!bb.meets(lo, hi, b)) continue;
        pool.push_back(b);
        for (long k = nptr_b[b]; k < nptr_b[b + 1]; k++) {
          const int n = (int)nidx_b[k];
          if (!in_front_b[n]) {
            in_front_b[n] = 1;
            front_b.push_back(n);
          }
        }
      }
      for (int b : front_b) in_front_b[b] = 0;
      len[a] = (long)pool.size() - start[a];
      push_neighbours(a);
    }
    while (seed < ne_a && visited_a[seed]) seed++;
    if (seed >= ne_a) break;
  }
  for (long a = 0; a < ne_a; a++) indptr[a + 1] = indptr[a] + len[a];
  indices.resize((size_t)indptr[ne_a]);
  for (long a = 0; a < ne_a; a++)
    std::copy(pool.begin() + start[a], pool.begin() + start[a] + len[a], indices.begin() + indptr[a]);
}

}  // namespace lsm
