/* freud_amd.h -- C-ABI of the MI355X implementation of freud's pair-analysis path.
 *
 * Plain pointers and sizes only.  Arrays named d_* live in device memory (HBM);
 * everything else is host memory.  Every function returns FA_OK (0) or a
 * negative FA_ERR_* status; fa_last_error() returns the message of the last
 * failure on the calling thread.  The status codes map onto the exception types
 * the reference raises through nanobind: FA_ERR_INVALID -> ValueError
 * (std::invalid_argument / std::domain_error), FA_ERR_RUNTIME -> RuntimeError
 * (std::runtime_error), FA_ERR_RANGE -> IndexError (std::out_of_range).
 *
 * Citations name files of glotzerlab/freud 3.5.0: each entry point replaces the
 * C++ seam cited beside it (which the reference binds with nanobind in the
 * export-*.cc file named there).
 */
#ifndef FREUD_AMD_H
#define FREUD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FA_OK 0
#define FA_ERR_INVALID (-1) /* ValueError   */
#define FA_ERR_RUNTIME (-2) /* RuntimeError */
#define FA_ERR_RANGE (-3)   /* IndexError   */
#define FA_ERR_DEVICE (-4)  /* RuntimeError: a HIP call failed */

#define FA_NQ_AABB 0     /* freud.locality.AABBQuery: r = p_j - (q_i + image) */
#define FA_NQ_LINKCELL 1 /* freud.locality.LinkCell:  r = box.wrap(p_j - q_i)  */
#define FA_NQ_CELLQUERY 2 /* freud.locality.CellQuery: r = (p_j + ghost shift) - q_i, ball queries only */

#define FA_MODE_NONE 0
#define FA_MODE_BALL 1
#define FA_MODE_NEAREST 2

/* Steinhardt flags (freud/order/Steinhardt.h:59-90 constructor arguments) */
#define FA_ST_AVERAGE 1
#define FA_ST_WL 2
#define FA_ST_WEIGHTED 4
#define FA_ST_WL_NORMALIZE 8

/* freud::box::Box (freud/box/Box.h:44-606): lengths, tilt factors, 2D flag.
 * Every dimension is periodic (the reference's pair queries reject other boxes,
 * freud/locality/NeighborQuery.h:134-138). */
typedef struct fa_box
{
    float Lx, Ly, Lz, xy, xz, yz;
    int32_t is2D;
} fa_box;

/* freud::locality::QueryArgs (freud/locality/NeighborQuery.h:37-73) with the
 * reference's sentinels: num_neighbors 0xffffffff, r_max -1, r_guess -1,
 * scale -1 mean "unset". */
typedef struct fa_query_args
{
    int32_t mode;
    uint32_t num_neighbors;
    float r_max;
    float r_min;
    float r_guess;
    float scale;
    int32_t exclude_ii;
} fa_query_args;

/* freud::locality::NeighborList storage (freud/locality/NeighborList.h:29-158),
 * device resident, in the reference's layout:
 *   neighbors  uint32 (num_bonds, 2) row-major  [i0, j0, i1, j1, ...]
 *   distances  float  (num_bonds)
 *   weights    float  (num_bonds)
 *   vectors    float  (num_bonds, 3)
 *   counts     uint32 (num_query_points)
 *   segments   uint32 (num_query_points) -- first bond of each query point, 0
 *              for query points without bonds (NeighborList.cc:199-232)
 *   offsets    uint32 (num_query_points + 1) -- exclusive scan of counts
 * Buffers are owned by the library and released with fa_nlist_free. */
typedef struct fa_nlist
{
    uint32_t* neighbors;
    float* distances;
    float* weights;
    float* vectors;
    uint32_t* counts;
    uint32_t* segments;
    uint32_t* offsets;
    uint64_t num_bonds;
    uint32_t num_query_points;
    uint32_t num_points;
    uint64_t capacity;
} fa_nlist;

typedef struct fa_nq fa_nq;

/* ---------------------------------------------------------------- runtime */
const char* fa_last_error(void);
int fa_init(void);
int fa_device_name(char* buf, int len);
/* Device memory from the library's caching allocator (no counterpart in the
 * reference).  fa_free returns the block to the cache without synchronising;
 * work that uses it must be ordered on the library stream (all entry points
 * are) or have completed. */
int fa_malloc(void** d_ptr, size_t bytes);
int fa_free(void* d_ptr);
int fa_memcpy_h2d(void* d_dst, const void* src, size_t bytes);
int fa_memcpy_d2h(void* dst, const void* d_src, size_t bytes);
int fa_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes);
int fa_memset(void* d_dst, int value, size_t bytes);
int fa_synchronize(void);
/* Event timing on the library's stream (milliseconds between two records). */
int fa_timer_start(void);
int fa_timer_stop(float* ms);
/* Per-kernel timing: when enabled, every kernel launch of the library is
 * bracketed by HIP events on the library's stream; fa_profile_get returns the
 * launch count and summed milliseconds per kernel name.  (No counterpart in
 * the reference; used by the benchmark to time the dominant kernel.) */
int fa_profile_enable(int on);
int fa_profile_reset(void);
int fa_profile_count(void);
int fa_profile_get(int idx, char* name, int len, uint64_t* launches, double* total_ms);

/* Autotuner (no counterpart in the reference, which has one CPU code path).
 * Analyses with several bit-identical launch variants (cell size relative to
 * r_max) time each variant on their first call for an unseen problem class
 * and record the fastest.  fa_tune_get/fa_tune_set read and preset the table
 * (analysis, key, variant) so a caller can persist it; variant < 0 in
 * fa_tune_set removes an entry.  Tuning is on by default. */
int fa_tune_enable(int on);
int fa_tune_count(void);
int fa_tune_get(int idx, int* analysis, int* key, int* variant);
int fa_tune_set(int analysis, int key, int variant);
int fa_tune_reset(void);
/* Number of launch variants the tuner chooses among for (analysis, key):
 * variants 0 .. count-1 are valid in fa_tune_set. */
int fa_tune_variants(int analysis, int key);

/* ---------------------------------------------------------------- batched trajectories
 * No counterpart in the reference, which analyses one frame per call (a
 * Python loop of compute(..., reset=False) over a trajectory; per frame
 * RDF.cc:101-110, NeighborQuery.h:434-481, Steinhardt.cc:85-222).  A batch
 * holds many frames of any box shape (cubic, triclinic, 2D) and size; their
 * points are concatenated, frame f at rows point_offsets[f] ..
 * point_offsets[f+1]-1 of the batch's point array (fa_batch_points).  One run
 * computes, for every frame, the planned analyses with the single-frame
 * semantics: an AABBQuery ball query (query points = points, exclude_ii) ->
 * the frame's NeighborList, the RDF bin counts (0 .. rdf_r_max), and
 * Steinhardt q_l over the ball bonds with the frame's system order.  Runs
 * after the first replay a captured HIP graph when use_graph is set. */
typedef struct fa_batch fa_batch;
typedef struct fa_batch_plan
{
    float ball_r_max;  /* > 0: ball query radius; 0: no ball query (and no q_l) */
    uint32_t rdf_bins; /* 0: no RDF; at most 2048 */
    float rdf_r_max;
    uint32_t l; /* Steinhardt l over the ball bonds; 0xffffffff: none */
} fa_batch_plan;
int fa_batch_create(uint32_t n_frames, const fa_box* boxes, const uint64_t* point_offsets, const fa_batch_plan* plan,
                    fa_batch** out);
int fa_batch_destroy(fa_batch* b);
/* The batch's device point array (total x 3 float32) for the caller to fill. */
int fa_batch_points(fa_batch* b, float** d_points);
/* Runs every planned analysis on the current points and synchronises. */
int fa_batch_run(fa_batch* b, int use_graph);
/* Bonds per frame (host array of n_frames). */
int fa_batch_num_bonds(fa_batch* b, uint64_t* per_frame);
/* Copies frame f's NeighborList (frame-local indices) into *out. */
int fa_batch_nlist(fa_batch* b, uint32_t frame, fa_nlist* out);
/* Device pointer to the n_frames x rdf_bins counts of the last run. */
int fa_batch_rdf_counts(fa_batch* b, uint32_t** d_counts);
/* Device pointers to q_l (total) and q_lm (total x (2l+1) complex, m order
 * 0..l, -1..-l) of the last run; order (host, n_frames) = each frame's
 * system q_l.  Any output pointer may be NULL. */
int fa_batch_ql(fa_batch* b, float** d_ql, float** d_qlm, float* order);
/* The launch shape the autotuner chose (cells r/R wide; RDF cells rdf_F times
 * narrower along x), the bond capacity and whether a graph is captured. */
int fa_batch_config(fa_batch* b, int* ball_R, int* rdf_R, int* rdf_F, uint64_t* capacity, int* graph_ready);

/* ---------------------------------------------------------------- box
 * freud/box/Box.h:212-329: op 0 makeFractional, 1 makeAbsolute, 2 wrap over n
 * device vec3s (in place allowed). */
int fa_box_transform(const fa_box* box, int op, const float* d_in, float* d_out, uint64_t n);

/* ---------------------------------------------------------------- locality
 * NeighborQuery constructors: AABBQuery (freud/locality/AABBQuery.cc:18-26),
 * LinkCell (freud/locality/LinkCell.cc:222-260; cell_width 0 = the reference's
 * default width) and CellQuery (freud/locality/CellQuery.h:29-40; its ghost
 * grid is a function of each query's r_max, CellQuery.cc:24-34).  d_points (n x 3) is caller-owned and must outlive the handle
 * (the reference also keeps a raw pointer, NeighborQuery.h:93-95). */
int fa_nq_create(const fa_box* box, const float* d_points, uint32_t n, int kind, float cell_width, fa_nq** out);
int fa_nq_destroy(fa_nq* nq);
/* LinkCell::getCellWidth (LinkCell.h) */
int fa_nq_cell_width(const fa_nq* nq, float* width);
/* The cell grid a CellQuery ball query of radius r_cut uses
 * (freud/locality/CellQuery.h:136-160 setupGrid: getNx/getNy/getNz,
 * getMinPos, getCellInverseWidth).  Host arithmetic only; no device work. */
int fa_cellquery_grid(const fa_box* box, float r_cut, int32_t* n_xyz, float* min_pos, float* inverse_width);

/* NeighborQuery::query(...)->toNeighborList(sort_by_distance)
 * (freud/locality/NeighborQuery.h:131-142, 434-481; binding
 * export-NeighborQuery.cc:26-41).  On entry *out may hold buffers of an
 * earlier list (capacity reused) or be zeroed. */
int fa_query(fa_nq* nq, const float* d_query_points, uint32_t n_query, const fa_query_args* args,
             int sort_by_distance, fa_nlist* out);

int fa_nlist_free(fa_nlist* nl);
/* NeighborList(num_bonds, query_point_index, ...) from device arrays
 * (freud/locality/NeighborList.cc:51-82): validates sorting and index ranges,
 * distances = sqrt(dot(v, v)). */
int fa_nlist_from_arrays(uint32_t num_query_points, uint32_t num_points, const uint32_t* d_query_point_indices,
                         const uint32_t* d_point_indices, const float* d_vectors, const float* d_weights,
                         uint64_t num_bonds, fa_nlist* out);
/* NeighborList::NeighborList(points, query_points, box, exclude_ii, ...)
 * all_pairs (freud/locality/NeighborList.cc:84-128). */
int fa_nlist_all_pairs(const fa_box* box, const float* d_points, uint32_t n_points, const float* d_query_points,
                       uint32_t n_query, int exclude_ii, fa_nlist* out);
/* NeighborList::filter (NeighborList.cc:239-277): keep bonds whose d_keep byte is nonzero. */
int fa_nlist_filter(fa_nlist* nl, const uint8_t* d_keep);
/* NeighborList::filter_r (NeighborList.cc:286-307): keep r_min <= d < r_max. */
int fa_nlist_filter_r(fa_nlist* nl, float r_max, float r_min);
/* NeighborList::sort (NeighborList.cc:371-399). */
int fa_nlist_sort(fa_nlist* nl, int by_distance);
/* NeighborList::copy (NeighborList.cc:360-370) */
int fa_nlist_copy(const fa_nlist* src, fa_nlist* dst);

/* NeighborList filters over an existing list nl (left unchanged: the
 * reference's unfiltered_nlist).  *out (a different list; zeroed or holding
 * buffers of an earlier list) receives the kept bonds in (query point,
 * distance, point, weight) order, with num_query_points / num_points one past
 * the largest kept indices (NeighborList.cc:130-173; an empty result keeps
 * nl's).  d_unfilled (nl->num_query_points device bytes) gets 1 for each query
 * point whose shell the list could not fill and *n_unfilled their number: the
 * reference raises or warns over these indices (Filter.h:80-102), which the
 * caller does from the mask.
 * FilterSANN::compute (freud/locality/FilterSANN.cc:22-104): kept bonds keep
 * their vectors and weights. */
int fa_filter_sann(const fa_nlist* nl, fa_nlist* out, uint8_t* d_unfilled, uint32_t* n_unfilled);
/* FilterRAD::compute (freud/locality/FilterRAD.cc:22-131): d_points (n_points
 * x 3) and d_query_points (nl->num_query_points x 3) are the positions nl
 * indexes; d_points_radii (n_points, NULL = all 1) the per-point radii.  Kept
 * bonds carry box.wrap(q_i - p_j) and weight 1.  terminate_after_blocked
 * nonzero is RAD-closed, zero RAD-open. */
int fa_filter_rad(const fa_box* box, const float* d_points, uint32_t n_points, const float* d_query_points,
                  const float* d_points_radii, const fa_nlist* nl, int terminate_after_blocked, fa_nlist* out,
                  uint8_t* d_unfilled, uint32_t* n_unfilled);

/* ---------------------------------------------------------------- density
 * RDF::accumulate without a NeighborList (freud/density/RDF.cc:101-110 via
 * BondHistogramCompute::accumulateGeneral, BondHistogramCompute.h:115-126): a
 * fused ball query (query radius q_r_max/q_r_min) + RegularAxis binning
 * (freud/util/Histogram.h:152-174) into d_counts (bins uint32, accumulated). */
/* The host-built bin table the RDF kernels use (no counterpart in the
 * reference; exposed so that its exactness can be checked on the host):
 * bins + 1 r_sq thresholds T[k] = the smallest float r_sq whose
 * RegularAxis::bin (Histogram.h:152-174) of sqrtf(r_sq) is >= k, then
 * inv_delta, then x_fast, then a 4096-entry uint16 table (cell c =
 * uint32(r_sq * inv_delta), below 4096 for every r_sq < T[bins]; entry k: the
 * cell's bins are k or k + 1; 0xffff: more, only in cells below x_fast).
 * Returns the number of floats (out may be NULL to query it). */
int fa_rdf_bin_table(uint32_t bins, float r_min, float r_max, float* out, uint64_t cap);
int fa_rdf_accumulate_query(fa_nq* nq, const float* d_query_points, uint32_t n_query, float q_r_max, float q_r_min,
                            int exclude_ii, uint32_t bins, float r_min, float r_max, uint32_t* d_counts);
/* No counterpart in the reference: a variant of fa_rdf_accumulate_query for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It computes what
 * fa_rdf_accumulate_query computes for an AABBQuery over d_points with the
 * query points equal to the points and exclude_ii set (the default of
 * RDF.compute((box, points))), adding into d_counts (bins uint32), and calls
 * no hipMalloc, hipFree or allocating library routine: the cell grid, bin
 * table, cluster table and scan storage live in a 64 MiB array that each GPU
 * holds from module load to process exit.  Limits: n <= 1,048,576 points,
 * 1 <= bins <= 4096, and a cell grid that fits the array (an error says when
 * it does not).  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (from several host threads) race. */
int fa_rdf_accumulate_self_noalloc(const fa_box* box, const float* d_points, uint32_t n, float q_r_max,
                                   float q_r_min, uint32_t bins, float r_min, float r_max, uint32_t* d_counts);
/* Same, binning the distances of an existing NeighborList. */
int fa_rdf_accumulate_nlist(const fa_nlist* nl, uint32_t bins, float r_min, float r_max, uint32_t* d_counts);
/* RDF::reduce (freud/density/RDF.cc:73-99) and the shell volumes of the RDF
 * constructor (RDF.cc:44-63), host arrays.  Arithmetic in the reference's float
 * order. */
int fa_rdf_reduce(const fa_box* box, const uint32_t* counts, uint32_t bins, float r_min, float r_max,
                  uint32_t n_points, uint32_t n_query_points, uint32_t n_frames, int finite_size, float* rdf,
                  float* n_r);

/* LocalDensity::compute over a NeighborList (freud/density/LocalDensity.cc:39-84;
 * binding export-LocalDensity.cc): per query point of nl, num_neighbors sums
 * 1 for bonds with d < r_max - diameter/2 and 1 + (r_max - (d + diameter/2)) /
 * diameter otherwise, in bond order; density = num_neighbors / (pi r_max^2) in
 * 2D, / (4/3 pi r_max^3) in 3D.  d_density and d_num_neighbors hold
 * num_query_points floats. */
int fa_local_density(const fa_nlist* nl, float r_max, float diameter, int is2D, float* d_density,
                     float* d_num_neighbors);
/* PMFTXY::accumulate over a NeighborList (freud/pmft/PMFTXY.cc:69-88 via
 * BondHistogramCompute::accumulateGeneral): every bond vector (x, y) rotated by
 * minus its query point's orientation angle (query_orientations: host,
 * radians, num_query_points floats, as the reference takes them) is binned
 * into the n_x x n_y RegularAxis histogram over [-x_max, x_max) x
 * [-y_max, y_max), row-major, added to d_counts.  The rotation's cos and sin
 * come from the host C library, as in the reference. */
int fa_pmftxy_accumulate(const fa_nlist* nl, const float* query_orientations, float x_max, float y_max,
                         uint32_t n_x, uint32_t n_y, uint32_t* d_counts);
/* PMFTXY::reduce (PMFTXY.cc:62-66, PMFT.h:76-87): pcf (host, n_x * n_y) from
 * host counts, in the reference's float order. */
int fa_pmftxy_reduce(const fa_box* box, const uint32_t* counts, uint32_t n_x, uint32_t n_y, float x_max, float y_max,
                     uint32_t n_points, uint32_t n_query_points, uint32_t n_frames, float* pcf);

/* BondOrder::accumulate (freud/environment/BondOrder.cc:87-136; binding
 * export-BondOrder.cc): per bond the direction (mode 0 bod: the bond vector;
 * 1 lbod: rotated by conj(q_j); 2 obcd: by conj(q_j) then q_i; 3 oocd: the
 * query point's director q_i z rotated by conj(q_j)) binned by azimuth
 * theta in [0, 2 pi) and polar angle phi in [0, pi) into n_theta x n_phi
 * (row-major).  orientations: host (num_points, 4) quaternions (s, x, y, z);
 * query_orientations: host (num_query_points, 4).  The angles' atan2 and acos
 * run in double rounded to float (the reference's come from the host C
 * library). */
int fa_bond_order_accumulate(const fa_nlist* nl, const float* orientations, const float* query_orientations, int mode,
                             uint32_t n_theta, uint32_t n_phi, uint32_t* d_counts);
/* BondOrder::reduce (BondOrder.cc:26-53, 68-74): bond_order (host) =
 * count / surface area / n_frames, surface areas dt (cos phi - cos(phi + dp))
 * with the host C library's cos, in the reference's float order. */
int fa_bond_order_reduce(const uint32_t* counts, uint32_t n_theta, uint32_t n_phi, uint32_t n_frames, float* bond_order);
/* CorrelationFunction::accumulate (freud/density/CorrelationFunction.cc:84-98;
 * binding export-CorrelationFunction.cc): per bond of nl with distance in
 * [0, r_max) (RegularAxis(bins, 0, r_max)), d_counts[bin] += 1 and
 * d_sums[bin] += conj(values[j]) * query_values[i].  values / query_values:
 * host complex128 (re, im interleaved), num_points / num_query_points of them;
 * d_sums: 2 * bins doubles (device).  The order of the sums is unspecified,
 * as in the reference's thread-local histograms. */
int fa_correlation_accumulate(const fa_nlist* nl, const double* values, const double* query_values, uint32_t bins,
                              float r_max, uint32_t* d_counts, double* d_sums);
/* CorrelationFunction::reduce (CorrelationFunction.cc:50-60): host arrays;
 * correlation (2 * bins doubles) = sums / counts where counts != 0. */
int fa_correlation_reduce(const uint32_t* counts, const double* sums, uint32_t bins, double* correlation);
/* PMFTXYT::accumulate (freud/pmft/PMFTXYT.cc:76-100): per bond, (x, y)
 * rotated by minus the query point's angle (host cos / sin, as PMFTXY) and
 * t = modulusPositive(theta_j - atan2(-y, -x), 2 pi), binned into
 * n_x x n_y x n_t over [-x_max, x_max) x [-y_max, y_max) x [0, 2 pi).
 * orientations: host, num_points angles; query_orientations: host,
 * num_query_points angles.  The reference's angle comes from the host C
 * library's float atan2; here it is atan2 in double rounded to float. */
int fa_pmftxyt_accumulate(const fa_nlist* nl, const float* orientations, const float* query_orientations, float x_max,
                          float y_max, uint32_t n_x, uint32_t n_y, uint32_t n_t, uint32_t* d_counts);
/* PMFTR12::accumulate (freud/pmft/PMFTR12.cc:90-111): per bond (d, t1, t2)
 * with t1 = modulusPositive(theta_j - atan2(y, x), 2 pi) and
 * t2 = modulusPositive(theta_i - atan2(-y, -x), 2 pi), binned into
 * n_r x n_t1 x n_t2 over [0, r_max) x [0, 2 pi)^2.  Host angle arrays as
 * fa_pmftxyt_accumulate. */
int fa_pmftr12_accumulate(const fa_nlist* nl, const float* orientations, const float* query_orientations, float r_max,
                          uint32_t n_r, uint32_t n_t1, uint32_t n_t2, uint32_t* d_counts);
/* PMFTXYZ::accumulate (freud/pmft/PMFTXYZ.cc:106-145): per bond and per
 * equivalent orientation k, v = rotate(equiv_k, rotate(conj(q_i), r)) binned
 * into n_x x n_y x n_z over [-x_max, x_max) x [-y_max, y_max) x
 * [-z_max, z_max).  query_orientations: host quaternions (s, x, y, z), one
 * per query point; equiv_orientations: host, n_equiv quaternions. */
int fa_pmftxyz_accumulate(const fa_nlist* nl, const float* query_orientations, const float* equiv_orientations,
                          uint32_t n_equiv, float x_max, float y_max, float z_max, uint32_t n_x, uint32_t n_y,
                          uint32_t n_z, uint32_t* d_counts);
#define FA_PMFT_XYT 0
#define FA_PMFT_R12 1
#define FA_PMFT_XYZ 2
/* PMFT::reduce (freud/pmft/PMFT.h:76-87 with the Jacobians of PMFTXYT.cc:52-57,
 * PMFTR12.cc:58-77) and PMFTXYZ::reduce (PMFTXYZ.cc:92-103, normalised also by
 * n_equiv): pcf (host, nbins[0] * nbins[1] * nbins[2]) from host counts.
 * maxes: (x_max, y_max, -) for XYT, (r_max, -, -) for R12, (x_max, y_max,
 * z_max) for XYZ. */
int fa_pmft_reduce(int kind, const fa_box* box, const uint32_t* counts, const uint32_t* nbins, const float* maxes,
                   uint32_t n_points, uint32_t n_query_points, uint32_t n_frames, uint32_t n_equiv, float* pcf);

/* ---------------------------------------------------------------- order
 * Steinhardt::compute over a NeighborList (freud/order/Steinhardt.cc:85-359;
 * binding export-Steinhardt.cc:19-34).  ls[n_l] are the l values.  Outputs
 * (device, caller-allocated):
 *   d_qlm     complex64 interleaved, l-major blocks: block k is (n, 2l_k+1) with
 *             m = 0..l, -1..-l
 *   d_ql      float (n, n_l)
 *   d_qlave   float (n, n_l)            (FA_ST_AVERAGE; may be NULL otherwise)
 *   d_qlmave  like d_qlm                (FA_ST_AVERAGE; may be NULL)
 *   d_wl      float (n, n_l)            (FA_ST_WL; may be NULL otherwise)
 *   order     host float (n_l)          system-wide order (normalizeSystem) */
int fa_steinhardt(const fa_box* box, const float* d_points, uint32_t n, const fa_nlist* nl, const uint32_t* ls,
                  int n_l, int flags, float* d_qlm, float* d_ql, float* d_qlave, float* d_qlmave, float* d_wl,
                  float* order);

/* SolidLiquid::compute (freud/order/SolidLiquid.cc:31-114; binding
 * export-SolidLiquid.cc:11-28) over a NeighborList: q_lm of Steinhardt(l),
 * per-bond ql_ij, num_connections and solid clusters.  d_qlm (n, 2l+1)
 * complex64 and d_ql (n) are written too. */
int fa_solidliquid(const fa_box* box, const float* d_points, uint32_t n, const fa_nlist* nl, uint32_t l,
                   float q_threshold, uint32_t solid_threshold, int normalize_q, float* d_qlm, float* d_ql,
                   float* d_ql_ij, uint32_t* d_num_connections, uint32_t* d_cluster_idx, uint32_t* d_cluster_sizes,
                   uint32_t* num_clusters);

/* Cluster::compute over the bonds of a NeighborList (freud/cluster/Cluster.cc:21-128):
 * labels sorted by cluster size (descending) then smallest member index.
 * d_keep (optional, per bond) selects the bonds used.  d_cluster_sizes
 * (optional, n entries) receives the sizes in label order
 * (SolidLiquid.h:95-109 getClusterSizes). */
int fa_cluster(uint32_t n, const fa_nlist* nl, const uint8_t* d_keep, uint32_t* d_cluster_idx,
               uint32_t* d_cluster_sizes, uint32_t* num_clusters);

/* Hexatic::compute (freud/order/HexaticTranslational.cc:22-74): psi_k per
 * particle, complex64 interleaved. */
/* Learned SolidLiquid features (no counterpart in the reference; they feed the
 * classifier below).  Per particle, over the bonds of nl (unweighted):
 * d_features (n x 12 float32) = q_l for l = 4, 6, 8, 10, 12, normalised w_l
 * for the same l (Steinhardt.cc:120-222, 291-359 definitions), the bond count,
 * and a zero pad; with lmax = 8 the columns of l = 10 and 12 are 0 (lmax is 8
 * or 12).  d_qlm6 (n x 13 complex64, m order 0..6, -1..-6) and d_ql6 (n)
 * equal fa_steinhardt's q_6m and q_6 bit for bit. */
int fa_sl_features(const fa_box* box, const float* d_points, uint32_t n, const fa_nlist* nl, int lmax,
                   float* d_features, float* d_qlm6, float* d_ql6);
/* Learned SolidLiquid labels (no counterpart in the reference): particles
 * whose classifier logit lies at least `margin` from 0 take the classifier's
 * label (logit > 0: solid); the others, and NaN logits, count their solid-like
 * bonds exactly as SolidLiquid.cc:51-96 does (q_ij from d_qlm6 / d_ql6 of
 * fa_sl_features, strict q_threshold, at least solid_threshold bonds).
 * d_solid and d_fallback are n bytes; *n_fallback (host, may be NULL) is the
 * number of exact-path particles. */
int fa_sl_decide(uint32_t n, const fa_nlist* nl, const float* d_logits, float margin, const float* d_qlm6,
                 const float* d_ql6, int normalize_q, float q_threshold, uint32_t solid_threshold, uint8_t* d_solid,
                 uint8_t* d_fallback, uint32_t* n_fallback);

/* The classifier network (no counterpart in the reference): 11 -> 64 -> 64 ->
 * 1, ReLU, BF16 MFMA with FP32 accumulation.  Parameters are one float32
 * vector of fa_mlp_num_params() values: W1 (64 x 16, columns 11..15 unused),
 * b1 (64), W2 (64 x 64), b2 (64), w3 (64), b3 (1).  Inputs are the first 11
 * columns of fa_sl_features rows, standardised with (x - mu) / sigma. */
typedef struct fa_mlp fa_mlp;
int fa_mlp_create(uint64_t seed, fa_mlp** out);
int fa_mlp_destroy(fa_mlp* m);
int fa_mlp_num_params(void);
int fa_mlp_set_params(fa_mlp* m, const float* params);
int fa_mlp_get_params(fa_mlp* m, float* params);
int fa_mlp_set_norm(fa_mlp* m, const float* mu, const float* sigma);
/* logits (n) of feature rows d_features (n x 12) */
int fa_mlp_forward(fa_mlp* m, const float* d_features, uint32_t n, float* d_logits);
/* Gradient (host, fa_mlp_num_params() floats) and loss of the mean
 * BCE-with-logits loss over the batch rows d_idx[0..batch) with 0/1 labels. */
int fa_mlp_gradient(fa_mlp* m, const float* d_features, const uint8_t* d_labels, const uint32_t* d_idx,
                    uint32_t batch, float* grads, float* loss);
/* Adam training over shuffled mini-batches; epoch_loss (host, epochs) may be NULL. */
int fa_mlp_train(fa_mlp* m, const float* d_features, const uint8_t* d_labels, uint32_t n, uint32_t epochs,
                 uint32_t batch, float lr, uint64_t seed, float* epoch_loss);
int fa_hexatic(const fa_box* box, const float* d_points, uint32_t n, const fa_nlist* nl, uint32_t k, int weighted,
               float* d_psi);

#ifdef __cplusplus
}
#endif

#endif /* FREUD_AMD_H */
