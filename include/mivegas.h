/*
 * mivegas.h - C ABI of the MI355X (gfx950) VEGAS engine, libmivegas.so.
 *
 * The library replaces the TensorFlow graph bodies of vegasflow's hot path
 * (vegasflow 1.4.0, src/vegasflow/...). Each entry point names the reference
 * function it stands in for. All pointers are device pointers owned by the
 * caller (the PyTorch caching allocator in the Python host code); the library
 * never allocates or frees device memory and never synchronises: every call
 * enqueues work on the given HIP stream and returns.
 *
 * Return value: >= 0 on success (some calls return a count), or a negative
 * mv_status code; mv_last_error() then holds a message (thread-local).
 *
 * Conventions
 *   - grids are (n_dim, MV_BINS + 1) FP64 row-major, edges in [0, 1];
 *   - bin indices are uint8 (n_events, n_dim) row-major;
 *   - partial sums travel in rows of mv_partial_stride(...) doubles
 *     [sum_k, sum_k^2, histogram]. A producer (fused / digest call) has
 *     every workgroup reduce its events to one row and add it, with FP64
 *     atomics, into group row (workgroup mod MV_ROWS) of the caller's
 *     mv_partials. Producers return the number of group rows they touched;
 *     successive producers of one iteration (steps) add into the same rows.
 *     The finalize calls reduce the rows in a fixed order and zero them
 *     again, so a workspace is all zero between iterations. Sums are
 *     reproducible to rounding (the atomic order varies), not bitwise.
 */
#ifndef MIVEGAS_H
#define MIVEGAS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MV_ABI_VERSION 3
#define MV_BINS 50             /* BINS_MAX, configflow.py:13 */
#define MV_FUSED_MAX_DIM 16    /* largest n_dim the fused kernels take */
#define MV_MAX_COMP 4          /* components of the built-in mixture family */
#define MV_MAX_OUT 16          /* outputs of a vector integrand (digest) */
#define MV_MAX_BLOCKS 2048     /* workgroups of one producer launch */
#define MV_ROWS 32             /* group rows of a partials workspace */
#define MV_DIGEST_DIMS 16      /* dimensions per grid.y slice of the digest */

/* Bits of the status word of the finalize calls (mv_finalize.status,
 * mv_vplus_finalize, mv_vegas_iteration_static). A non-finite w*f at any
 * event (an integrand returning NaN or +-Inf) makes the iteration's sum of
 * w*f or of (w*f)^2 non-finite, since NaN and Inf survive additions and
 * squares, so the finalize kernel tests those sums once per iteration
 * rather than every event. It then ORs MV_STATUS_NONFINITE into the word
 * (the caller zeroes it) and leaves the grid unrefined, so the grid stays
 * finite for the next iteration; res and sigma are reported as computed
 * (NaN or Inf, as the reference's arithmetic gives them). */
#define MV_STATUS_NONFINITE 1u

typedef enum {
  MV_OK = 0,
  MV_ERR_ARG = -1,
  MV_ERR_UNSUPPORTED = -2,
  MV_ERR_LAUNCH = -3,
  MV_ERR_WORKSPACE = -4,
  MV_ERR_DEVICE = -5
} mv_status;

typedef enum { MV_F64 = 0, MV_MIXED = 1 } mv_precision;

typedef enum {
  MV_FN_CONSTANT = 0, /* f = c                                               */
  MV_FN_GAUSSIAN = 1, /* (1/(a sqrt(pi)))^d exp(-sum((x_j - c_j)/a)^2)       */
  MV_FN_PRODUCT = 2,  /* prod_j x_j                                          */
  MV_FN_GAUSSMIX = 3, /* sum_k w_k gaussian(x; a_k, c_k)                     */
  MV_FN_BWPOWER = 4   /* (G/2pi)/((x_0-M)^2 + G^2/4) * x_1^p, rest flat      */
} mv_family;

/* Built-in integrand: evaluated in registers inside the fused kernels.
 * comp[k] = {weight, 1/a, prefactor, unused} for the Gaussian families,
 * comp[0] = {mass, width/2, (width/2)/pi, power} for MV_FN_BWPOWER,
 * comp[0][0] = c for MV_FN_CONSTANT. centre[k][j] is component k's centre
 * in dimension j. */
typedef struct {
  int32_t family;
  int32_t n_comp;
  int32_t reserved[2];
  double comp[MV_MAX_COMP][4];
  double centre[MV_MAX_COMP][MV_FUSED_MAX_DIM];
} mv_integrand;

/* One launch worth of events of an iteration.
 * Uniform draws: Philox4x32-10, key = seed, counter = (event_lo, event_hi,
 * draw_id, block) with event = event_offset + local index. In MV_F64 a draw
 * is r = ((hi:lo) >> 12) * 2^-52 mapped to [1e-8, 1 - 1e-8]
 * (monte_carlo.py:264-266); in MV_MIXED it is the 32-bit word hi alone.
 * ``inject`` (nullable) replaces the generator: FP64 uniforms (n_events,
 * n_dim) in MV_F64, uint32 hi words (n_events, n_dim) in MV_MIXED.
 * ``draw_dev`` (nullable) points to a device uint32 that holds the draw id
 * instead of ``draw_id``: with the counter of mv_finalize / mv_vplus_finalize
 * a captured iteration can be replayed with a new draw each time. */
typedef struct {
  int32_t n_dim;
  int32_t precision;     /* mv_precision */
  int32_t train;         /* accumulate the bin histogram (vflow.py:376) */
  int32_t reserved;
  int64_t n_events;      /* events in this launch */
  int64_t event_offset;  /* global index of its first event */
  uint64_t seed;
  uint32_t draw_id;
  uint32_t reserved2;
  double xjac;           /* weight factor: 1/N (vegas, plain), 1/n_cubes (plus) */
  double xdeltajac;      /* prod(xmax - xmin), 1 without limits */
  const double* xmin;    /* nullable: integration limits (n_dim) */
  const double* xdelta;  /* nullable: xmax - xmin (n_dim) */
  const void* inject;    /* nullable, see above */
  const uint32_t* draw_dev;  /* nullable, see above */
} mv_step;

/* Caller-owned partial-sum workspace: MV_ROWS rows of ``stride`` doubles,
 * zero when first used (e.g. allocated zeroed). Calls that share one
 * workspace must be stream-ordered. */
typedef struct {
  double* rows;
  int64_t stride;       /* doubles per row, >= mv_partial_stride(...) */
} mv_partials;

/* Stratification of VegasFlowPlus (vflowplus.py:113-139). Hypercube c has
 * coordinates given by the base-n_strat digits of c, last dimension fastest
 * (itertools.product order); offsets is the exclusive prefix sum of n_ev
 * (n_cubes + 1 entries, offsets[n_cubes] = events of the iteration). */
typedef struct {
  int32_t n_strat;
  int32_t n_cubes;
  const int32_t* n_ev;
  const int32_t* offsets;
} mv_strata;

/* Reduction of partial rows into iteration results (vflow.py:432-442,
 * plain.py:37-43): res_k = sum tmp_k, sigma_k from N and sum tmp_k^2. */
typedef struct {
  int32_t n_dim;
  int32_t n_out;         /* outputs per event (1 for a scalar integrand) */
  int32_t train;         /* rows carry a histogram; refine the grid */
  int32_t refine;        /* apply the refinement to ``grid`` */
  int32_t raw;           /* write sum tmp^2 instead of sigma */
  int32_t reserved;
  double n_events;       /* N of the sigma formula */
  /* Iteration counter (nullable): draw_dev[0] is the draw id of the
   * iteration being finalized, draw_dev[1] a base. After the results,
   * {res, sigma, events} of output 0 go to history[3 k .. 3 k + 2] with
   * k = draw_dev[0] - draw_dev[1] when 0 <= k < history_len (history
   * nullable; events = n_events), and draw_dev[0] is incremented. */
  uint32_t* draw_dev;
  double* history;
  int32_t history_len;
  int32_t reserved2;
  uint32_t* status;      /* nullable: MV_STATUS_* bits are ORed in */
} mv_finalize;

/* ---- library ---------------------------------------------------------- */
int mv_abi_version(void);
const char* mv_last_error(void);
/* sizeof of the ABI structs, for binding-layout checks:
 * 0 mv_integrand, 1 mv_step, 2 mv_partials, 3 mv_strata, 4 mv_finalize,
 * 5 mv_batch_item, 6 mv_batch_tile, 7 mv_batch_vplus_item. */
int64_t mv_struct_size(int32_t which);
/* Doubles per partial row: 2 * n_out + (train ? n_dim * MV_BINS : 0). */
int64_t mv_partial_stride(int32_t n_dim, int32_t n_out, int32_t train);
/* Upper bound of the workgroups a producer launches for ``n_events``
 * events (the launch also caps them at what is resident on the device at
 * once), and the group rows it then touches. */
int32_t mv_producer_blocks(int64_t n_events);
int32_t mv_producer_rows(int64_t n_events);

/* ---- VegasFlow ---------------------------------------------------------- */
/* Fused iteration body for built-in integrands: draws, grid map and
 * Jacobian (vflow.py:87-126 + 33-83, monte_carlo.py:249-275), integrand,
 * w*f sums (vflow.py:416-421) and the per-dimension (w*f)^2 bin histogram
 * (vflow.py:370-387, utils.py:10-44), without any per-event HBM traffic.
 * Returns the number of group rows. Requires n_dim <= MV_FUSED_MAX_DIM. */
int mv_vegas_fused(const mv_step* step, const mv_integrand* fn, const double* grid,
                   const mv_partials* parts, void* stream);

/* Sampling half of the unfused path (vflow.py:87-126, monte_carlo.py:249-275):
 * x (n, d) and w (n,) in the step precision (double or float), bin indices
 * idx (n, d) uint8 (nullable when not training). Any n_dim. */
int mv_vegas_sample(const mv_step* step, const double* grid, void* x, void* w,
                    uint8_t* idx, void* stream);

/* Digest half of the unfused path (vflow.py:414-430): tmp = w * f for a
 * (n, n_out) integrand f with row stride f_stride (elements), histogram of
 * tmp[:, main_dim]^2 by idx. f and w are double (MV_F64) or float
 * (MV_MIXED); w_stride 0 broadcasts w[0] (plain.py:18-35). Any n_dim.
 * Returns the number of group rows. */
int mv_vegas_digest(const mv_step* step, const void* f, int64_t f_stride, int32_t n_out,
                    int32_t main_dim, const void* w, int64_t w_stride, const uint8_t* idx,
                    const mv_partials* parts, void* stream);

/* Reduce ``n_rows`` (1..MV_ROWS, as returned by the producer) group rows
 * (fixed order, then zero them) into
 * out[0:n_out] = res and
 * out[n_out:2 n_out] = sigma (vflow.py:437-438, plain.py:37-43),
 * hist_out (nullable, (n_dim, MV_BINS)) = arr_res2, and when fz->refine
 * refine ``grid`` in place (vflow.py:349-362 + 129-211). */
int mv_vegas_finalize(const mv_finalize* fz, double* rows, int64_t row_stride,
                      int32_t n_rows, double* grid, double* out, double* hist_out,
                      void* stream);

/* refine_grid_per_dimension (vflow.py:129-211) for every dimension:
 * hist (n_dim, MV_BINS) with row stride hist_stride, grid (n_dim,
 * MV_BINS + 1) refined in place. */
int mv_refine_grid(int32_t n_dim, const double* hist, int64_t hist_stride, double* grid,
                   void* stream);

/* consume_array_into_indices (utils.py:10-44): out[b] = sum of input[i]
 * with indices[i] == b, b < result_size; out is overwritten. */
int mv_consume_array_into_indices(const double* input, const int32_t* indices, int64_t n,
                                  int32_t result_size, double* out, void* stream);

/* ---- PlainFlow ---------------------------------------------------------- */
/* Fused plain MC (plain.py:18-35) for built-in integrands; rows without a
 * histogram. */
int mv_plain_fused(const mv_step* step, const mv_integrand* fn, const mv_partials* parts,
                   void* stream);
/* Uniform points (monte_carlo.py:264-274) for the unfused plain path:
 * x (n, d) in the step precision. */
int mv_plain_sample(const mv_step* step, void* x, void* stream);

/* ---- VegasFlowPlus ------------------------------------------------------ */
/* Fused VEGAS+ body (vflowplus.py:37-80 + 187-220) for built-in integrands.
 * Adds per-cube sums into ress / ress2 (n_cubes; zero before the first
 * iteration, zeroed again by mv_vplus_finalize) and writes histogram rows
 * (n_out = 1). step->n_events is an upper bound of the events; the kernel
 * takes the actual count from strata->offsets[n_cubes]. */
int mv_vplus_fused(const mv_step* step, const mv_strata* strata, const mv_integrand* fn,
                   const double* grid, double* ress, double* ress2, const mv_partials* parts,
                   void* stream);
/* Sampling half (vflowplus.py:37-80): x, w, idx as mv_vegas_sample plus the
 * cube of each event seg (n,) int32. */
int mv_vplus_sample(const mv_step* step, const mv_strata* strata, const double* grid,
                    void* x, void* w, uint8_t* idx, int32_t* seg, void* stream);
/* Digest half (vflowplus.py:209-218): per-cube sums of w*f and (w*f)^2 into
 * ress / ress2, histogram rows. */
int mv_vplus_digest(const mv_step* step, const mv_strata* strata, const void* f,
                    const void* w, const uint8_t* idx, const int32_t* seg, double* ress,
                    double* ress2, const mv_partials* parts, void* stream);
/* Doubles of ``out`` for mv_vplus_finalize: 3 results, then 4 doubles per
 * 256 cubes that the call uses as scratch. */
int64_t mv_vplus_out_len(int32_t n_cubes);
/* Iteration end (vflowplus.py:222-242, 153-163): out (mv_vplus_out_len
 * doubles) = {res, sigma, events of the next iteration, scratch}; when
 * adaptive, n_ev and offsets are
 * rewritten for the next iteration (damped = max(arr_var, 0)^(beta/2),
 * n_ev = int(max(min_neval, damped * init_calls / 2 / sum damped))); ress
 * and ress2 are zeroed; with train the histogram rows are reduced into
 * hist_out (nullable) and, with refine, the grid is refined. arr_var_out
 * (nullable, n_cubes) receives ress2 * n_ev - ress^2. draw_dev, history,
 * history_len: the iteration counter of mv_finalize (nullable); the events
 * recorded are those of the iteration just finalized, offsets[n_cubes].
 * status (nullable): as mv_finalize.status (a non-finite res also leaves
 * n_ev unchanged: the reallocation's denominator is then not positive). */
int mv_vplus_finalize(int32_t n_dim, const mv_strata* strata, int32_t* n_ev, int32_t* offsets,
                      double* ress, double* ress2, int32_t adaptive, int32_t min_neval,
                      double init_calls, double beta, double* rows, int64_t row_stride,
                      int32_t n_rows, int32_t train, int32_t refine, double* grid, double* out,
                      double* arr_var_out, double* hist_out, uint32_t* draw_dev,
                      double* history, int32_t history_len, uint32_t* status, void* stream);

/* One VegasFlow iteration (mv_vegas_fused then mv_vegas_finalize with
 * n_out = 1, N = step->n_events) that allocates nothing: no caller
 * workspace, and no hipMalloc / hipFree or allocating library call. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid it). The partial rows live in a static device array
 * of MV_ROWS x (2 + MV_FUSED_MAX_DIM * MV_BINS) doubles (205 KB) that each
 * GPU the library runs on holds until the process ends, so n_dim is at most
 * MV_FUSED_MAX_DIM (larger is rejected with MV_ERR_UNSUPPORTED). out (2
 * doubles) = {res, sigma}; hist_out (nullable, (n_dim, MV_BINS)) and
 * refine (grid refined in place) need step->train; status (nullable) as
 * mv_finalize.status. Every call on a GPU
 * shares that GPU's array and leaves it zeroed: calls that overlap on one
 * GPU (different streams, threads) race and must not be issued. */
int mv_vegas_iteration_static(const mv_step* step, const mv_integrand* fn, double* grid,
                              int32_t refine, double* out, double* hist_out, uint32_t* status,
                              void* stream);

/* ---- batched independent integrals ------------------------------------- */
/* No counterpart in the reference (each VegasFlow instance there runs its
 * own graph): many VegasFlow integrals of different n_dim (<=
 * MV_FUSED_MAX_DIM), event count and built-in integrand advance one
 * iteration each in one persistent launch, with their grids, accumulators
 * and iteration counters resident on the device. Item i draws exactly what
 * VegasFlow(n_dim_i, n_events_i, seed = seed_i) would at iteration
 * draws[i], so its results equal a lone run up to summation order. */
typedef struct {
  int32_t n_dim;
  int32_t train;         /* refine this item's grid in mv_batch_finalize */
  int64_t n_events;      /* events per iteration (xjac = 1 / n_events) */
  uint64_t seed;
  int64_t grid_offset;   /* doubles from ``grids`` to its (n_dim, 51) grid */
  int64_t row_offset;    /* doubles from ``rows`` to its 2 + 50 n_dim row */
  mv_integrand fn;
} mv_batch_item;

/* A tile: ``count`` consecutive events of one item, from event ``first``.
 * A tile must not cross a multiple of 2^32 events (first mod 2^32 + count
 * <= 2^32): the kernel takes the high Philox counter word of its first event
 * for the whole tile. Tiles cut at multiples of a power-of-two size never
 * cross one. */
typedef struct {
  int32_t item;
  int32_t count;
  int64_t first;
} mv_batch_tile;

/* Per-item device state: draws[i] = iteration (Philox draw id) of item i;
 * stats[4 i ..] = {last integral, last sigma, sum 1/sigma^2,
 * sum integral/sigma^2} (the weighted average of monte_carlo.py:713-732 is
 * stats[3]/stats[2], its error sqrt(1/stats[2])). */

/* One iteration of the events listed in ``tiles`` (device array) for the
 * items (device array): fused draws/map/integrand/sums/histograms as
 * mv_vegas_fused, added with FP64 atomics into each item's row. ``blocks``
 * persistent workgroups stride over the tiles. precision: mv_precision;
 * family: the mv_family shared by every tiled item, or -1 for a mix;
 * max_dim: the largest n_dim among the tiled items. */
int mv_batch_fused(const mv_batch_item* items, int32_t n_items, const mv_batch_tile* tiles,
                   int32_t n_tiles, int32_t precision, int32_t family, int32_t max_dim,
                   const int32_t* draws, const double* grids, double* rows, int32_t blocks,
                   void* stream);

/* As mv_batch_fused in MV_MIXED, with the histograms of training items kept
 * as pair tables (dimensions 2p and 2p + 1 share a 50 x 50 table: one LDS
 * atomic per pair of dimensions) in 1024-thread workgroups; frozen items
 * run as in mv_batch_fused. Requires max_dim <= 14. ``blocks`` persistent
 * workgroups of 1024 threads; mv_batch_pairs_blocks(family, max_dim) is the
 * number resident on the current device at once (0: unsupported). No
 * counterpart in the reference (batching is an addition). */
int mv_batch_fused_pairs(const mv_batch_item* items, int32_t n_items, const mv_batch_tile* tiles,
                         int32_t n_tiles, int32_t family, int32_t max_dim, const int32_t* draws,
                         const double* grids, double* rows, int32_t blocks, void* stream);
int32_t mv_batch_pairs_blocks(int32_t family, int32_t max_dim);
/* The same in MV_F64 (the FP64 event body in the reference's operation
 * order, pair tables of FP64 cells); mv_batch_pairs_blocks_f64 gives its
 * resident workgroups. No counterpart in the reference (batching is an
 * addition). */
int mv_batch_fused_pairs_f64(const mv_batch_item* items, int32_t n_items, const mv_batch_tile* tiles,
                             int32_t n_tiles, int32_t family, int32_t max_dim, const int32_t* draws,
                             const double* grids, double* rows, int32_t blocks, void* stream);
int32_t mv_batch_pairs_blocks_f64(int32_t family, int32_t max_dim);

/* Iteration end for the items named in ``jobs`` (device int32 pairs
 * {item, dim}: dim = -1 reduces the item's sums into stats (vflow.py:437-438,
 * monte_carlo.py:718-720), stores {integral, sigma} at history[2 (item *
 * max_iter + draws[item])] while draws[item] < max_iter (history nullable)
 * and advances draws[item]; dim = j refines grid row j of the item from its
 * histogram (vflow.py:129-211) when the item trains). Consumed row columns
 * are zeroed. */
int mv_batch_finalize(const mv_batch_item* items, const int32_t* jobs, int32_t n_jobs,
                      double* grids, double* rows, int32_t* draws, double* stats,
                      double* history, int32_t max_iter, void* stream);

/* ---- batched VegasFlowPlus integrals ----------------------------------- */
/* No counterpart in the reference (batching is an addition): many
 * VegasFlowPlus integrals (vflowplus.py:88-247) of n_dim <= MV_FUSED_MAX_DIM
 * advance one iteration each in one persistent producer launch and three
 * iteration-end launches, with their grids, event allocations, per-cube sums
 * and iteration counters resident on the device. Item i draws exactly what
 * VegasFlowPlus(n_dim_i, init_calls_i, adaptive = adaptive_i, seed = seed_i)
 * would at iteration draws[i] (event index within the iteration as the
 * Philox counter), so its results equal a lone run up to summation order. */
typedef struct {
  int32_t n_dim;
  int32_t train;         /* refine this item's grid */
  int32_t n_strat;       /* stratification of vflowplus.py:113-137 */
  int32_t n_cubes;       /* n_strat^n_dim */
  int32_t adaptive;      /* reallocate events (vflowplus.py:153-163) */
  int32_t min_neval;
  double init_calls;     /* events requested: N0 of the reallocation */
  uint64_t seed;
  int64_t grid_offset;   /* doubles from ``grids`` to its (n_dim, 51) grid */
  int64_t row_offset;    /* doubles from ``rows`` to its 2 + 50 n_dim row */
  int64_t cube_offset;   /* entries from n_ev, ress, ress2 (n_cubes each) */
  int64_t off_offset;    /* entries from offsets (n_cubes + 1, exclusive
                          * prefix sum of n_ev; offsets[n_cubes] = events
                          * of the iteration) */
  int64_t out_offset;    /* doubles from ``outs``: mv_vplus_out_len(n_cubes),
                          * {res, sigma, events of the next iteration,
                          * scratch} after each iteration end */
  mv_integrand fn;
} mv_batch_vplus_item;

/* One iteration of the items' events listed in ``tiles`` (mv_batch_tile over
 * event indices of an item's iteration; events at or beyond the item's
 * offsets[n_cubes] are skipped, so tiles may cover an upper bound of the
 * events, e.g. n_cubes * min_neval + init_calls / 2 + 1 when adaptive):
 * draws in the events' hypercubes, grid map, integrand, per-cube sums into
 * ress / ress2 and the (w f)^2 bin histogram into each item's row (columns
 * 2 ..). Every tiled item has integrand family ``family``; min_dim and
 * max_dim bound their n_dim. When 2 <= min_dim and max_dim <= 14 the
 * histograms are pair tables (dimensions 2p and 2p + 1 share a 50 x 50 table)
 * in 1024-thread workgroups, at most ``blocks`` and the number resident at
 * once; otherwise per-wave histograms in ``blocks`` 256-thread workgroups. */
int mv_batch_vplus_fused(const mv_batch_vplus_item* items, const mv_batch_tile* tiles, int32_t n_tiles,
                         int32_t precision, int32_t family, int32_t min_dim, int32_t max_dim,
                         const int32_t* draws, const double* grids, const int32_t* n_ev, const int32_t* offsets,
                         double* ress, double* ress2, double* rows, int32_t blocks, void* stream);

/* Iteration end of the items (vflowplus.py:222-242, 153-163 and vflow.py:
 * 349-362 per item, as mv_vplus_finalize): ``chunks`` lists device int32
 * pairs {item, chunk} with chunk = 0 .. ceil(n_cubes / 256) - 1 for every
 * item, ``dims`` pairs {item, dim} for the histogram columns of every item
 * (refined when the item trains). Per item: outs = {res, sigma, events of
 * the next iteration}, n_ev and offsets reallocated when adaptive, ress and
 * ress2 and the row zeroed, stats[4 i ..] = {integral, sigma, sum 1/sigma^2,
 * sum integral/sigma^2} (the running weighted average of monte_carlo.py:
 * 713-732), draws[i] advanced. */
int mv_batch_vplus_finalize(const mv_batch_vplus_item* items, const int32_t* chunks, int32_t n_chunks,
                            const int32_t* dims, int32_t n_dims, double beta, double* grids, int32_t* n_ev,
                            int32_t* offsets, double* ress, double* ress2, double* rows, double* outs,
                            int32_t* draws, double* stats, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MIVEGAS_H */
