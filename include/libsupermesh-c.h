/*
 * libsupermesh-c.h -- drop-in C ABI of the MI355X supermeshing engine.
 *
 * Same names, argument meaning and layouts as the reference's
 * include/libsupermesh-c.h (double precision build): positions are
 * double[nnodes][dim], element-node lists long[nelements][loc] with 0-based
 * node ids, triangles double[3][2], tetrahedra double[4][3], quads double[4][2],
 * hexahedra double[8][3] (Gmsh vertex order), intervals double[2].  The 2D and
 * 3D functions run on the GPU (device 0); the 1D ones (intervals and the sort
 * finder) are host code.  See libsupermesh_gpu.h for the bulk entry points
 * that a GPU caller should prefer over per-pair calls.
 *
 * Errors: the reference aborts the process (src/Debug.F90:78-94); these
 * functions instead return with the outputs zeroed and record the failure,
 * readable through libsupermesh_gpu_last_error() / _code().
 */
#ifndef LIBSUPERMESH_C_H
#define LIBSUPERMESH_C_H

#define LIBSUPERMESH_DOUBLE_PRECISION
#define TRI_BUF_SIZE 22
#define TET_BUF_SIZE 81

typedef double c_kind;

#ifdef __cplusplus
#include <cmath>
extern "C" {
#else
#include <math.h>
#endif

/* Tree (quadtree / octree) candidate finder, strict AABB overlap.
 * set_input runs the whole search; query_output returns the number of pairs;
 * get_output fills indices[nindices] (B ids, ascending within each row) and
 * ind_ptr[nelements + 1] and releases the result.
 * Replaces src/C_bindings.F90:346-391 (declared at include/libsupermesh-c.h:21-23). */
void libsupermesh_tree_intersection_finder_set_input(long* nnodes_a, int* dim_a, long* nelements_a, int* loc_a, long* nnodes_b, int* dim_b, long* nelements_b, int* loc_b, c_kind* positions_a, long* enlist_a, c_kind* positions_b, long* enlist_b);
void libsupermesh_tree_intersection_finder_query_output(long* nindices);
void libsupermesh_tree_intersection_finder_get_output(long* nelements, long* nindices, long* indices, long* ind_ptr);

/* 1D sort finder (intervals, loc 2): sorts both meshes by left end point and
 * sweeps A with one cursor into B, as the reference does; the rows hold B ids
 * in sweep order.  Host code.
 * Replaces src/C_bindings.F90:66-111 (include/libsupermesh-c.h:18-20). */
void libsupermesh_sort_intersection_finder_set_input(long* nnodes_a, int* dim_a, long* nelements_a, int* loc_a, long* nnodes_b, int* dim_b, long* nelements_b, int* loc_b, c_kind* positions_a, long* enlist_a, c_kind* positions_b, long* enlist_b);
void libsupermesh_sort_intersection_finder_query_output(long* nindices);
void libsupermesh_sort_intersection_finder_get_output(long* nelements, long* nindices, long* indices, long* ind_ptr);

/* R-tree candidate finder semantics (closed AABB overlap), same calling
 * sequence.  The reference reaches it only from Fortran
 * (rtree_intersection_finder, src/Intersection_Finder.F90:577-643); these
 * C symbols expose it to the Fortran module of this library. */
void libsupermesh_rtree_intersection_finder_set_input(long* nnodes_a, int* dim_a, long* nelements_a, int* loc_a, long* nnodes_b, int* dim_b, long* nelements_b, int* loc_b, c_kind* positions_a, long* enlist_a, c_kind* positions_b, long* enlist_b);
void libsupermesh_rtree_intersection_finder_query_output(long* nindices);
void libsupermesh_rtree_intersection_finder_get_output(long* nelements, long* nindices, long* indices, long* ind_ptr);

/* Triangle-triangle intersection: tris_c[TRI_BUF_SIZE][3][2], *n_tris_c outputs.
 * Replaces src/C_bindings.F90:164-174 (include/libsupermesh-c.h:24). */
void libsupermesh_intersect_tris_real(c_kind* tri_a, c_kind* tri_b, c_kind* tris_c, int* n_tris_c);
/* Tet-tet intersection: tets_c[TET_BUF_SIZE][4][3].
 * Replaces src/C_bindings.F90:152-162 (include/libsupermesh-c.h:25). */
void libsupermesh_intersect_tets_real(c_kind* tet_a, c_kind* tet_b, c_kind* tets_c, int* n_tets_c);
/* Replaces src/C_bindings.F90:143-150 (include/libsupermesh-c.h:26). */
void libsupermesh_triangle_area(c_kind* tri, c_kind* area);
/* Edges from the first vertex, times 1/6.
 * Replaces src/C_bindings.F90:134-141 (include/libsupermesh-c.h:27). */
void libsupermesh_tetrahedron_volume(c_kind* tet, c_kind* volume);

/* |interval[1] - interval[0]|.  Host code.
 * Replaces src/C_bindings.F90:125-132 (include/libsupermesh-c.h:28). */
void libsupermesh_interval_size(c_kind* interval, c_kind* size);
/* Interval-interval intersection: *n_intervals_c is 0 or 1.  Host code.
 * Replaces src/C_bindings.F90:113-123 (include/libsupermesh-c.h:29). */
void libsupermesh_intersect_intervals(c_kind* interval_a, c_kind* interval_b, c_kind* intervals_c, int* n_intervals_c);
/* Tet clipped by the 6 face planes of a hexahedron: tets_c[729][4][3].
 * Replaces src/C_bindings.F90:42-52 (include/libsupermesh-c.h:30). */
void libsupermesh_intersect_tet_hex(c_kind* tet_a, c_kind* hex_b, c_kind* tets_c, int* n_tets_c);
/* Triangle clipped by the 4 edge lines of a quad: tris_c[46][3][2].
 * Replaces src/C_bindings.F90:54-64 (include/libsupermesh-c.h:31). */
void libsupermesh_intersect_tri_quad(c_kind* tri_a, c_kind* quad_b, c_kind* tris_c, int* n_tris_c);

/* Quad-quad and hex-hex intersection (intersect_quads / intersect_hexes,
 * src/Supermesh.F90:231-248, 303-334; the reference exports them from Fortran
 * only): tris_c[62][3][2], tets_c[5 * 729][4][3]. */
void libsupermesh_intersect_quads(c_kind* quad_a, c_kind* quad_b, c_kind* tris_c, int* n_tris_c);
void libsupermesh_intersect_hexes(c_kind* hex_a, c_kind* hex_b, c_kind* tets_c, int* n_tets_c);

/* The advancing front finder, in the tree finder's set_input / query_output /
 * get_output form.  The reference has it in Fortran only
 * (advancing_front_intersection_finder, src/Intersection_Finder.F90:138-141,
 * 270-454, 456-476); these are its C form, and the Fortran module's
 * advancing_front_intersection_finder calls them. */
void libsupermesh_advancing_front_intersection_finder_set_input(long* nnodes_a, int* dim_a, long* nelements_a, int* loc_a, long* nnodes_b, int* dim_b, long* nelements_b, int* loc_b, c_kind* positions_a, long* enlist_a, c_kind* positions_b, long* enlist_b);
void libsupermesh_advancing_front_intersection_finder_query_output(long* nindices);
void libsupermesh_advancing_front_intersection_finder_get_output(long* nelements, long* nindices, long* indices, long* ind_ptr);

/* Mesh adjacency, exported by the reference's C_bindings.F90 but not declared
 * in its header.  All ids 0-based, CSR index pointers starting at 0.
 * node -> elements, ascending element order, one entry per occurrence:
 * replaces src/C_bindings.F90:176-213 (mesh_nelist, src/Graphs.F90:76-116). */
void libsupermesh_mesh_nelist_set_input(long* nnodes, long* nelements, long* loc, long* enlist);
void libsupermesh_mesh_nelist_query_output(long* nindices);
void libsupermesh_mesh_nelist_get_output(long* nnodes, long* nindices, long* indices, long* ind_ptr);
/* node -> nodes sharing an element, ascending: replaces src/C_bindings.F90:215-295. */
void libsupermesh_mesh_nnlist_set_input(long* nnodes, long* nelements, long* loc, long* enlist);
void libsupermesh_mesh_nnlist_query_output(long* nindices);
void libsupermesh_mesh_nnlist_get_output(long* nnodes, long* nindices, long* indices, long* ind_ptr);
/* element -> elements sharing sloc nodes, in the reference's discovery order,
 * at most loc per element: replaces src/C_bindings.F90:297-344
 * (mesh_eelist, src/Graphs.F90:118-187). */
void libsupermesh_mesh_eelist_set_input(long* nnodes, long* nelements, long* loc, long* sloc, long* enlist);
void libsupermesh_mesh_eelist_query_output(long* nindices);
void libsupermesh_mesh_eelist_get_output(long* nelements, long* nindices, long* indices, long* ind_ptr);
/* CSR transpose (rows of each column in ascending order): replaces
 * src/C_bindings.F90:397-430.  indices_t gets ind_ptr[nnodes] - ind_ptr[0]
 * entries. */
void libsupermesh_transpose_sparsity(long* nnodes, long* nnodes_t, long* nindices, long* indices, long* ind_ptr,
                                     long* indices_t, long* ind_ptr_t);

/* Error of the last call on this thread: text ("" if none) and code
 * (0 = ok; see libsupermesh_gpu.h).  No counterpart in the reference, which
 * aborts instead. */
const char* libsupermesh_gpu_last_error(void);
int libsupermesh_gpu_last_error_code(void);

#ifdef __cplusplus
}
#endif

static inline int libsupermesh_max_n_tris_c(int n_lines_b, int n_lines_a) {
  return n_lines_a * (1 << n_lines_b) - 2;
}

static inline int libsupermesh_max_n_tets_c(int n_planes_b) {
  int out = 1;
  for (int i = 0; i < n_planes_b; i++) out *= 3;
  return out;
}

#endif
