/*
 * libsupermesh_gpu.h -- bulk entry points of the MI355X supermeshing engine.
 *
 * These have no one-call counterpart in the reference: there a caller loops
 * over candidate pairs and calls intersect_tris / intersect_tets once per pair
 * (doc/manual.tex:2064-2082, src/tests/test_tri_intersector.F90:89-104).  A
 * GPU cannot amortise a launch per pair, so the whole loop -- candidate search,
 * clipping, measures and the P1 x P1 mixed mass -- is one call here, on device
 * arrays.  Results are bit-identical to the per-pair loop for candidate sets,
 * per-pair output counts, measures and output elements.
 *
 * Memory: results live in buffers owned by the context and stay valid until
 * the next call on the same context.  All device pointers are HIP device
 * memory on the context's device.  Calls on one context are serialised.
 */
#ifndef LIBSUPERMESH_GPU_H
#define LIBSUPERMESH_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LIBSUPERMESH_GPU_OK 0
#define LIBSUPERMESH_GPU_ERROR_INVALID 1   /* bad argument (dimension, loc, sizes) */
#define LIBSUPERMESH_GPU_ERROR_HIP 2       /* a HIP call failed (no device, OOM) */
#define LIBSUPERMESH_GPU_ERROR_STATE 3     /* call sequence (get before set_input) */
#define LIBSUPERMESH_GPU_ERROR_OVERFLOW 4  /* result larger than a 32-bit index */

/* Candidate finder semantics. */
#define LIBSUPERMESH_GPU_FINDER_TREE 0   /* strict AABB overlap: the quadtree /
                                            octree finder and brute force
                                            (Quadtree_Intersection_Finder.F90:411-420) */
#define LIBSUPERMESH_GPU_FINDER_RTREE 1  /* closed AABB overlap: the R-tree finder
                                            (R-Tree_Intersection_Finder_C++.h:233-251) */

/* Flags for libsupermesh_gpu_supermesh. */
#define LIBSUPERMESH_GPU_ELEMENTS 1   /* materialise the supermesh elements */
#define LIBSUPERMESH_GPU_MASS 2       /* P1 x P1 mixed mass blocks + CSR */
#define LIBSUPERMESH_GPU_TIMING 4     /* fill ms[] with per-phase times */
#define LIBSUPERMESH_GPU_SORTED 8     /* sort each candidate row by B id (otherwise
                                         rows keep the hierarchy's traversal
                                         order, which is deterministic; when
                                         A has under 1/8 of B's elements and
                                         the call is not batched, B's elements
                                         query a hierarchy over A and the rows
                                         come out sorted either way) */
#define LIBSUPERMESH_GPU_LEARNED 16   /* 3D with ELEMENTS: place each pair's run
                                         of output elements from the learned
                                         predictor's budget and clip once
                                         (outputs beyond a budget go to an
                                         arena and the run is then joined in a
                                         spill reserve; pairs that outgrow the
                                         arena are clipped again).  Every
                                         pair's elements are those of the
                                         plain path, in the same order, but
                                         the runs lie apart: pair p's are
                                         elements[offsets[p] .. offsets[p] +
                                         counts[p]) */
#define LIBSUPERMESH_GPU_BLOCKS 32    /* with MASS: also keep the per-pair
                                         blocks (result.blocks); 3D, and the
                                         explicit pair lists, always keep them */
#define LIBSUPERMESH_GPU_DENSE 64     /* with LEARNED: compact the runs into the
                                         plain path's layout (offsets = prefix
                                         sums of the counts), byte for byte the
                                         plain path's arrays; moved within the
                                         runs' own buffer (no second copy) */
#define LIBSUPERMESH_GPU_GRAPH 128    /* the count pass (counts, measures, sums;
                                         not with ELEMENTS, MASS or LEARNED)
                                         of a simplex supermesh or batch in
                                         the launch shape measured fastest on
                                         the context's first such call, run as
                                         a recorded hipGraph that later calls
                                         with the same buffers replay; results
                                         are those of the plain count pass */

typedef struct libsupermesh_gpu_context libsupermesh_gpu_context;

/* A simplex mesh in device memory: positions double[nnodes][dim],
 * enlist int[nelements][dim+1], 0-based. */
typedef struct {
  int dim;
  long nnodes;
  long nelements;
  const double* positions;
  const int* enlist;
} libsupermesh_gpu_mesh;

typedef struct {
  /* candidate pairs, CSR over A elements */
  long nelements_a;
  long npairs;
  const long* indptr;        /* [nelements_a + 1] */
  const int* indices;        /* [npairs] B element ids */
  const int* rows;           /* [npairs] A element id of each pair */
  /* per-pair clip results */
  const int* counts;         /* [npairs] output elements per pair */
  const double* measure;     /* [npairs] sum of output areas / volumes */
  /* materialised supermesh (LIBSUPERMESH_GPU_ELEMENTS): pair p's output
   * elements are elements[offsets[p] .. offsets[p] + counts[p]).  Plain
   * layout: offsets are the prefix sums of the counts (offsets[npairs] =
   * nout) and elements_size = nout.  LIBSUPERMESH_GPU_LEARNED without
   * LIBSUPERMESH_GPU_DENSE: runs lie apart (budgets, arena, spill) in an
   * array of elements_size elements; offsets[npairs] is then unspecified. */
  const long* offsets;       /* [npairs + 1] first output element of each pair */
  const double* elements;    /* [elements_size][dim+1][dim] */
  /* P1 x P1 mixed mass (LIBSUPERMESH_GPU_MASS):
   * blocks[p][i][j] = int over the pair's supermesh of phi^A_i phi^B_j
   * (null for a 2D supermesh unless LIBSUPERMESH_GPU_BLOCKS),
   * and the assembled matrix M[nnodes_a][nnodes_b] in CSR. */
  const double* blocks;      /* [npairs][dim+1][dim+1] */
  long mass_nnz;
  const long* mass_indptr;   /* [nnodes_a + 1] */
  const int* mass_indices;   /* [mass_nnz] */
  const double* mass_values; /* [mass_nnz] */
  /* reductions */
  double sum;                /* compensated sum of all output measures */
  long nonempty;             /* pairs with at least one output element */
  long nout;                 /* output elements */
  int max_count;             /* largest per-pair output count */
  /* phase times in ms (LIBSUPERMESH_GPU_TIMING): [0] staging, [1] hierarchy
   * build, [2] candidate search, [3] clipping (count pass + reductions +
   * element fill), [4] mass assembly, [5] the predictor kernel (LEARNED),
   * [6] total, [7] the clip count kernel alone (the single pass with LEARNED) */
  double ms[8];
  /* LIBSUPERMESH_GPU_LEARNED: output elements placed from the predicted
   * budgets (their sum), pairs whose output outgrew the budget (their runs
   * joined from the arena, or clipped again when it ran out), and the budget
   * rule used: the quantile whose runs fit in device memory (the configured
   * one or a lower level), or 0 when none fit and the plain count-then-fill
   * path ran */
  long learned_slots;
  long learned_overflow;
  double learned_quantile;
  long elements_size;        /* elements in the `elements` array */
  int count_graph;           /* LIBSUPERMESH_GPU_GRAPH: 1 when the count pass ran
                                as the recorded graph (0: eagerly, in the same
                                launch shape) */
  int dense_route;           /* LIBSUPERMESH_GPU_LEARNED | LIBSUPERMESH_GPU_DENSE:
                                how the plain layout was made: 1 the runs moved
                                within their buffer, 2 copied to a second one
                                (some pair's plain start lay beyond its run's),
                                3 every pair filled again (0 otherwise) */
} libsupermesh_gpu_result;

/* Context on HIP device `device` with its own stream (default priority). */
int libsupermesh_gpu_context_create(int device, libsupermesh_gpu_context** ctx);
void libsupermesh_gpu_context_destroy(libsupermesh_gpu_context* ctx);
/* The context's hipStream_t, as void*. */
void* libsupermesh_gpu_context_stream(libsupermesh_gpu_context* ctx);

/* Text of the last error on this thread and its code: every call but
 * libsupermesh_gpu_free clears them when it succeeds ("" and 0), so after a
 * failed call its message survives the clean-up of the caller's buffers. */
const char* libsupermesh_gpu_last_error(void);
int libsupermesh_gpu_last_error_code(void);

/* Device memory helpers on the library's own HIP runtime, for callers that
 * hold no other GPU runtime (Fortran, ctypes).  copy kinds: 1 host->device,
 * 2 device->host, 3 device->device. */
int libsupermesh_gpu_malloc(int device, size_t bytes, void** ptr);
int libsupermesh_gpu_free(void* ptr);
int libsupermesh_gpu_memcpy(void* dst, const void* src, size_t bytes, int kind);
int libsupermesh_gpu_synchronize(int device);

/* The count pass's launch shape the context measured fastest for dimension
 * dim (LIBSUPERMESH_GPU_GRAPH; zeros before the first such call): threads per
 * block, blocks in the grid (3D: the walk's persistent lanes / block; 2D:
 * blocks of the grid-stride loop, i.e. the pairs each wave takes, or 0: one
 * block per `block` pairs, each lane one pair), the tet
 * depths whose stack lives in LDS (3D: 2 or 3; the rest in registers) and the
 * time of that shape on the sample it was measured on. */
int libsupermesh_gpu_count_shape(libsupermesh_gpu_context* ctx, int dim, int* block, int* grid, int* lds_depths,
                                 float* ms);

/* The plain layout of a LIBSUPERMESH_GPU_LEARNED result (what
 * LIBSUPERMESH_GPU_DENSE gives in the same call): rewrites result->offsets,
 * ->elements and ->elements_size to arrays of the context holding the runs
 * end to end.  The result must be the context's latest; the buffers it
 * pointed to are left as they were. */
int libsupermesh_gpu_compact(libsupermesh_gpu_context* ctx, libsupermesh_gpu_result* result);

/* Order-sensitive checksum of `bytes` (a multiple of 4) of device memory:
 * the sum of its 32-bit words w_i times (2 i + 1), modulo 2^64.  For comparing
 * large results without copying them to the host. */
int libsupermesh_gpu_checksum(libsupermesh_gpu_context* ctx, const void* ptr, size_t bytes,
                              unsigned long long* out);

/* Candidate pairs only (finder = LIBSUPERMESH_GPU_FINDER_*). */
int libsupermesh_gpu_intersection_finder(libsupermesh_gpu_context* ctx,
                                         const libsupermesh_gpu_mesh* a,
                                         const libsupermesh_gpu_mesh* b, int finder, int flags,
                                         libsupermesh_gpu_result* out);

/* Candidate pairs + clipping + reductions (+ elements / mass per flags). */
int libsupermesh_gpu_supermesh(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                               const libsupermesh_gpu_mesh* b, int finder, int flags,
                               libsupermesh_gpu_result* out);

/* The same on meshes of any element type: loc_a / loc_b are dim+1 (triangles,
 * tetrahedra) or 2^dim (quadrilaterals, hexahedra; nodes in the
 * reference's order), and each pair is clipped as intersect_elements does
 * (Supermesh.F90:162-209): tri x quad as intersect_tri_quad, hex x hex as
 * intersect_hexes (Supermesh.F90:303-334), and so on.  The output elements are
 * always simplices.  Flags: ELEMENTS, SORTED, TIMING (MASS and LEARNED need
 * simplex meshes: use libsupermesh_gpu_supermesh).  enlist holds loc ids per
 * element. */
int libsupermesh_gpu_supermesh_mixed(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                                     int loc_a, const libsupermesh_gpu_mesh* b, int loc_b, int finder,
                                     int flags, libsupermesh_gpu_result* out);

/* A search hierarchy over mesh B kept on the device for queries one element
 * at a time: the reference's tree_type / rtree_type allocate, query and
 * deallocate (src/Intersection_Finder.F90:481-541, 645-712), which the
 * Fortran module's stateful *_intersection_finder_set_input / _find /
 * _query_output / _get_output / _reset are built on.  finder =
 * LIBSUPERMESH_GPU_FINDER_TREE (strict overlap) or _RTREE (closed); host
 * arrays, enlist long[nelements][loc] 0-based.  A query of one element
 * (element double[loc][dim]) returns the count; _result copies the B ids,
 * ascending, 0-based.  Each query is a few device launches: for many
 * elements use libsupermesh_gpu_intersection_finder. */
typedef struct libsupermesh_gpu_tree libsupermesh_gpu_tree;
int libsupermesh_gpu_tree_create(int finder, int dim, long nnodes, const double* positions, long nelements,
                                 int loc, const long* enlist, libsupermesh_gpu_tree** tree);
int libsupermesh_gpu_tree_query(libsupermesh_gpu_tree* tree, int loc, const double* element, long* neles);
int libsupermesh_gpu_tree_result(libsupermesh_gpu_tree* tree, long neles, long* eles);
void libsupermesh_gpu_tree_destroy(libsupermesh_gpu_tree* tree);

/* A batch of independent meshes of one dimension, concatenated: mesh m owns
 * elements element_offsets[m] .. element_offsets[m+1]-1 (device array,
 * [nmeshes + 1], from 0 to nelements).  enlist holds node ids into the whole
 * positions array (0-based).  Mesh m of batch A is paired with mesh m of
 * batch B only. */
typedef struct {
  int dim;
  long nmeshes;
  long nnodes;
  long nelements;
  const double* positions;     /* [nnodes][dim] */
  const int* enlist;           /* [nelements][dim+1] */
  const long* element_offsets; /* [nmeshes + 1] */
} libsupermesh_gpu_mesh_batch;

typedef struct {
  long nmeshes;
  const long* pair_offsets;    /* [nmeshes + 1]: candidate pairs of mesh m are
                                  pair_offsets[m] .. pair_offsets[m+1]-1 of the
                                  result's pair arrays */
  const double* sums;          /* [nmeshes]: compensated measure sum per mesh */
} libsupermesh_gpu_batch_result;

/* Batched supermesh: every mesh pair of the batch in one pass (one candidate
 * search over all meshes, one clip over all candidate pairs, per-mesh sums).
 * `out` holds the pairs of all meshes in batch order, with the same per-pair
 * content as separate libsupermesh_gpu_supermesh calls (element ids are
 * batch-global). */
int libsupermesh_gpu_supermesh_batched(libsupermesh_gpu_context* ctx,
                                       const libsupermesh_gpu_mesh_batch* a,
                                       const libsupermesh_gpu_mesh_batch* b, int finder, int flags,
                                       libsupermesh_gpu_result* out,
                                       libsupermesh_gpu_batch_result* batch);

/* Learned output-size predictor (a BF16 MLP on the MFMA units; see
 * LIBSUPERMESH_GPU_LEARNED).  Parameters are LIBSUPERMESH_GPU_PREDICTOR_PARAMS
 * floats: W1[64][24] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[33][64]
 * b4[33] (W[out][in]); inputs are the pair's 8 vertices normalised to the
 * pair's bounding box, outputs the logits of counts 0..32 (32 = "32 or more").
 * set/get copy host arrays; quantile sets the budget rule (the smallest count
 * whose predicted cumulative probability reaches q, default 0.9); collect
 * clips A x B with the plain path and keeps up to max_samples (features,
 * count) of its fully clipped pairs as training data; train runs `epochs`
 * passes of Adam over the collected samples (batch a multiple of 64),
 * writing each epoch's mean cross-entropy to loss_per_epoch; clear drops the
 * samples.  No counterpart in the reference. */
#define LIBSUPERMESH_GPU_PREDICTOR_PARAMS 12065
int libsupermesh_gpu_predictor_set(libsupermesh_gpu_context* ctx, const float* params);
int libsupermesh_gpu_predictor_get(libsupermesh_gpu_context* ctx, float* params);
int libsupermesh_gpu_predictor_quantile(libsupermesh_gpu_context* ctx, float q);
/* Outputs beyond a pair's budget go to an arena of fraction x the predicted
 * output plus min_tets tets (defaults 0.25 and 2^20; 0 and 0: no arena)
 * split among the clip's lanes; such runs are then joined in the spill
 * reserve, and pairs that find their lane's share full (every pair beyond
 * its budget, without an arena) are clipped again into it. */
int libsupermesh_gpu_predictor_arena(libsupermesh_gpu_context* ctx, float fraction, long min_tets);
/* The spill reserve: fraction x the predicted output plus min_tets tets
 * (defaults 0.25 and 2^20).  When the runs to place there outgrow it every
 * pair is filled again in the plain layout (elements_size = nout). */
int libsupermesh_gpu_predictor_spill(libsupermesh_gpu_context* ctx, float fraction, long min_tets);
int libsupermesh_gpu_predictor_collect(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                                       const libsupermesh_gpu_mesh* b, long max_samples,
                                       long* nsamples);
int libsupermesh_gpu_predictor_train(libsupermesh_gpu_context* ctx, int epochs, int batch, float lr,
                                     unsigned seed, float* loss_per_epoch);
int libsupermesh_gpu_predictor_clear(libsupermesh_gpu_context* ctx);
/* Adds n training samples from host arrays: features[n][24] (as the
 * predictor's inputs) and their output counts[n] (clamped to 0..32). */
int libsupermesh_gpu_predictor_add_samples(libsupermesh_gpu_context* ctx, long n, const float* features,
                                           const int* counts);
/* Predicted output budget of each listed tet pair (device arrays pair_a[n],
 * pair_b[n] -> budgets[n]), by the budget rule above. */
int libsupermesh_gpu_predictor_infer(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                                     const libsupermesh_gpu_mesh* b, long npairs, const int* pair_a,
                                     const int* pair_b, int* budgets);

/* Callback driver: the single-GPU form of the self-self loop of the
 * reference's parallel_supermesh (src/Parallel_Supermesh.F90:436-453, its
 * intersection_calculation interface :97-113; the MPI exchange is left out).
 * Candidates use closed overlap (R-tree semantics).  For each B element in
 * order, and each of its A candidates in ascending id, the callback receives
 * host pointers: element_a[dim+1][dim], element_b[dim+1][dim], the clipped
 * elements_c[n_elements_c][dim+1][dim] (n_elements_c may be 0), nodes_b[dim+1]
 * (B's 0-based node ids), and the 0-based element ids.  The pointers are valid
 * only during the call.  Clipping runs on the device in bulk; the callbacks run
 * on the calling thread. */
typedef void (*libsupermesh_gpu_pair_callback)(const double* element_a, const double* element_b,
                                               const double* elements_c, int n_elements_c,
                                               const long* nodes_b, long ele_a, long ele_b,
                                               void* user);
int libsupermesh_gpu_supermesh_callback(libsupermesh_gpu_context* ctx,
                                        const libsupermesh_gpu_mesh* a,
                                        const libsupermesh_gpu_mesh* b,
                                        libsupermesh_gpu_pair_callback callback, void* user);
/* The same on meshes of any element type, as parallel_supermesh takes them
 * (its intersection_calculation gets element_a(dim, loc_a) and
 * element_b(dim, loc_b), Parallel_Supermesh.F90:105-120): loc_a / loc_b are
 * dim+1 or 2^dim nodes per element (enlist holds loc ids per element), each
 * pair is clipped as intersect_elements does (Supermesh.F90:162-209), and the
 * callback receives element_a[loc_a][dim], element_b[loc_b][dim] and
 * nodes_b[loc_b]; the clipped elements are simplices. */
int libsupermesh_gpu_supermesh_callback_mixed(libsupermesh_gpu_context* ctx,
                                              const libsupermesh_gpu_mesh* a, int loc_a,
                                              const libsupermesh_gpu_mesh* b, int loc_b,
                                              libsupermesh_gpu_pair_callback callback, void* user);

/* Clip an explicit list of pairs (device arrays pair_a[n], pair_b[n], n <
 * 2^31; ids are checked against the meshes' element counts);
 * fills counts / measure (and elements, blocks per flags; no mass CSR). */
int libsupermesh_gpu_clip_pairs(libsupermesh_gpu_context* ctx, const libsupermesh_gpu_mesh* a,
                                const libsupermesh_gpu_mesh* b, long npairs, const int* pair_a,
                                const int* pair_b, int flags, libsupermesh_gpu_result* out);

/* libsupermesh_gpu_clip_pairs with no device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every argument is the caller's: device arrays pair_a[n],
 * pair_b[n] in, counts[n], measure[n] and sum[1] (the compensated sum of the
 * measures) out, all written asynchronously on `stream` (a hipStream_t; null
 * for the null stream) of GPU `device`, which becomes the calling thread's
 * current device.  Simplex meshes (dim 2 or 3) as libsupermesh_gpu_mesh,
 * read directly (nothing is staged).  Counts and measures equal
 * libsupermesh_gpu_clip_pairs's bit for bit; a pair whose element or node ids
 * are out of range gets count -1 and measure 0.  n is at most 4194304
 * (2^22): longer lists are refused (split them).  The pass's working memory
 * is one fixed 16.1 MiB array per GPU, kept until the process ends and
 * shared by every call on that GPU: calls that overlap on one GPU race. */
int libsupermesh_gpu_clip_pairs_noalloc(int device, void* stream, const libsupermesh_gpu_mesh* a,
                                        const libsupermesh_gpu_mesh* b, long npairs, const int* pair_a,
                                        const int* pair_b, int* counts, double* measure, double* sum);

#ifdef __cplusplus
}
#endif
#endif
