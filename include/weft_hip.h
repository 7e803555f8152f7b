/*
 * weft_hip.h -- C ABI of libweft_hip.so, the MI355X (gfx950) implementation of the
 * Weft encoder tile-fit scoring path and the primitive-walk renderer it shares.
 *
 * Every entry point returns 0 on success and a negative WEFT_E* code (or a negated
 * hipError_t) on failure; weft_last_error() describes the last failure.  No C++
 * exception crosses this boundary.  Pointers named d_* are DEVICE pointers (for
 * example torch tensors' data_ptr()); `stream` is a hipStream_t (NULL = default
 * stream).  All launches are asynchronous on `stream`.
 *
 * Reference interfaces replaced (jarmak-personal/weft, paths relative to its root):
 *   weft_init / weft_shutdown     gpu_context.get_cuda_state / reset_cuda_state
 *                                 (src/weft/gpu_context.py:16-62) and the NVRTC build
 *                                 + cuModuleGetFunction of gpu_render._get_render_state
 *                                 (src/weft/gpu_render.py:696-734)
 *   weft_score_evals              kernel batch_tile_objective
 *                                 (src/weft/gpu_render.py:113-152) as launched by
 *                                 gpu_batch_score_prepacked (gpu_render.py:455-566) and
 *                                 gpu_batch_objectives (gpu_render.py:737-845); MSE per
 *                                 eval, but with the CPU scorer's semantics
 *                                 (render.py:127-234 + encoder.py:271-285)
 *   weft_fit_*                    the greedy round of encoder._fit_adaptive_state
 *                                 (src/weft/encoder.py:2745-2811) / _greedy_add
 *                                 (encoder.py:752-817) with the current stack kept as a
 *                                 cached canvas on the device
 *   weft_render_tiles             kernel decode_tile_pixels (gpu_render.py:194-237) as
 *                                 launched by gpu_render_tiles_to_image
 *                                 (gpu_render.py:594-690), with render.py's alpha clamp
 *   weft_render_scene             render.render_scene_adaptive (src/weft/render.py:270-358)
 *                                 as decoder._render_primitive_stack_gpu
 *                                 (src/weft/decoder.py:231-390) places its tiles, RES1 included
 *   weft_template_masks /         kernel weft_tile_fit_candidates and its launcher
 *   weft_template_fit             fit_tiles_gpu_constant (src/weft/gpu_encoder.py:42-251,
 *                                 411-562), the stage-3 baseline fit
 *   weft_gen_candidates           encoder._generate_candidates (src/weft/encoder.py:535-745)
 *                                 for every tile of one size
 *   weft_gen_residual_candidates  encoder._generate_residual_candidates
 *                                 (src/weft/encoder.py:1255-1366) for every tile of one size
 *   weft_gen_edge_candidates      edge_analysis.generate_edge_driven_candidates
 *                                 (src/weft/edge_analysis.py:139-315) for every tile of one size
 *   weft_merge_pools              _dedup_candidates(edge + template), as encoder._fit_tile
 *                                 (encoder.py:1383-1385) and _worker_gen_cands (2513-2515) form it
 *   weft_argmin_f64, weft_exclusive_scan_i64, weft_compact_flags, weft_segmented_topk
 *                                 cccl_argmin, cccl_exclusive_scan, cccl_compact,
 *                                 cccl_segmented_topk (src/weft/cuda_backend.py:80-181)
 */
#ifndef WEFT_HIP_H
#define WEFT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Packed primitive, 60 bytes: the reference's PrimData / _PRIM_DTYPE
 * (src/weft/gpu_render.py:26-32, 241-247).  Kinds 0-4 as in src/weft/constants.py:38-42;
 * kind 6 = rotated ellipse, 7 = rotated rectangle (geom = cx, cy, rx, ry, cos t, sin t);
 * kind 8 = triangle drawn from (cx, cy, rx, ry, t), geom = its three vertices. */
typedef struct WeftPrim {
    int32_t kind;
    float geom[7];
    float color0[3];
    float color1[3];
    float alpha;
} WeftPrim;

/* One evaluation: the reference's EvalDesc (src/weft/gpu_render.py:106-111). */
typedef struct WeftEval {
    int32_t prim_offset;
    int32_t prim_count;
    int32_t tile_offset; /* in floats into the packed tile buffer */
    int32_t tile_size;
} WeftEval;

/* Decode descriptor: the reference's TileDecodeDesc (src/weft/gpu_render.py:179-192). */
typedef struct WeftTileDecodeDesc {
    int32_t prim_offset;
    int32_t prim_count;
    int32_t target_x0;
    int32_t target_y0;
    int32_t target_w;
    int32_t target_h;
    int32_t source_size;
    int32_t _pad;
    float res0_r;
    float res0_g;
    float res0_b;
    float _pad2;
} WeftTileDecodeDesc;

/* Quantized candidate code (no reference counterpart; the configurations' random
 * candidates).  Field meanings are documented at weft_draw_codes. */
typedef struct WeftCode {
    uint16_t family, cx, cy, rx, ry, theta, r, g, b, alpha;
} WeftCode;

/* Device-resident state of a plain tile fit: one entry per tile, SoA. */
typedef struct WeftFitState {
    const float* d_target;   /* packed tiles, tile k at float offset 3*d_tile_pix[k] */
    float* d_canvas;         /* unclipped composite of the tile's accepted stack */
    const int64_t* d_tile_pix;
    const int32_t* d_tile_ts;
    int32_t* d_tile_nsel;    /* primitives accepted so far */
    int32_t* d_tile_bits;    /* _estimate_bits of the accepted stack */
    double* d_tile_obj;      /* current objective (MSE + lambda * bits) */
    int32_t n_tiles;
    int32_t depth;           /* maximum stack depth */
    double lam;
    int32_t obj_mode;        /* 0: float64 objective (_greedy_add), 1: float32 (_fit_adaptive_state) */
    int32_t _pad;
} WeftFitState;

/* Work list for one tile-size class: work item i scores candidates
 * [d_work_cand0[i], d_work_cand0[i] + block) of tile d_work_tile[i].
 *
 * Split items (64-px tiles only; n_split > 0): a 64-px tile's candidates may be scored by two
 * work items, rows 0-31 and rows 32-63.  Their 128-element pairwise leaves split exactly there
 * (48 + 48 leaves), so each half's partial subtree sums (two floats per candidate, into
 * d_split_sse) combine into the single-pass SSE bit for bit; a combine launch after scoring
 * forms the SSE, the tables and the work item's first minimum.  d_work_split[i] is -1 for an
 * unsplit item, else 2 * pair + half; d_split_pairs[2 * pair + half] is the pair's work items;
 * d_split_sse holds n_split * 2 * 2 * block floats.  Few candidates per 64-px tile leave few
 * wavefronts for them, each walking 4096 pixels alone: the split halves that walk. */
typedef struct WeftWork {
    const int32_t* d_work_tile;
    const int32_t* d_work_cand0;
    int32_t n_work;
    int32_t tile_size;       /* 8, 16, 32 or 64: every tile in the list has this size;
                                0: tiles of any supported size (one launch for all) */
    int32_t block;           /* threads per workgroup (= candidates per work item), 64..1024 */
    int32_t n_split;         /* split pairs (0: none; the fields below may then be NULL) */
    const int32_t* d_work_split;
    const int32_t* d_split_pairs;
    float* d_split_sse;
} WeftWork;

/* Candidate source: explicit records or drawn on the device from (seed, step).
 * Explicit pools may be ragged and shrink: with d_tile_ncand, tile t has d_tile_ncand[t]
 * live candidates (at most n_cand, which is then each pool's capacity and the row stride of
 * the class order and the score tables); with d_tile_active, only tiles whose entry is
 * nonzero take part in the step (the others are handled like tiles whose stack is full);
 * with WEFT_CANDS_REMOVE_WINNER in flags, weft_fit_select removes an accepted winner from its
 * pool as the reference does (src/weft/encoder.py:2805-2808: the records behind it move
 * down one place in order and d_tile_ncand[t] decreases) and stores it at slot
 * n_cand - 1 - nsel of the tile's region (the accepted stack grows down from the top), so
 * d_prims is written; with
 * WEFT_CANDS_FORCE_ACCEPT it accepts the winner whatever its objective (the reference's
 * initial pick, encoder.py:2700-2715).  The extra fields are for explicit candidates. */
#define WEFT_CANDS_REMOVE_WINNER 1
#define WEFT_CANDS_FORCE_ACCEPT 2
typedef struct WeftCands {
    const WeftPrim* d_prims;     /* NULL: draw codes on the device */
    const int64_t* d_tile_cand;  /* per tile, offset of its first candidate in d_prims */
    uint64_t seed;
    uint32_t step;
    int32_t n_cand;              /* candidates per tile (with d_tile_ncand: pool capacity) */
    int32_t has_general;         /* nonzero if any explicit candidate has kind 1, 2 or 3 */
    int32_t flags;               /* WEFT_CANDS_* */
    const int32_t* d_tile_active;  /* NULL: every tile takes part */
    int32_t* d_tile_ncand;       /* NULL: n_cand candidates in every pool */
} WeftCands;

enum {
    WEFT_OK = 0,
    WEFT_E_NOT_INITIALISED = -1000,
    WEFT_E_INVALID = -1001,
    WEFT_E_UNSUPPORTED_TILE = -1002,
    WEFT_E_MODULE = -1003,
    WEFT_E_TOO_LARGE = -1004,
};

/* Load the gfx950 code object embedded in the library on `device` and build the
 * candidate tables.  Idempotent per device. */
int weft_init(int device);
/* Unload the module and drop every cached handle (reset_cuda_state). */
int weft_shutdown(void);
const char* weft_last_error(void);
/* Library and kernel ABI version (for the Python loader's sanity check). */
int weft_abi_version(void);
/* 0 if one workgroup may hold `bytes` of LDS (static + dynamic): at most
 * hipDeviceAttributeMaxSharedMemoryPerBlock of the initialised device, or gfx950's 160 KiB
 * before weft_init; otherwise WEFT_E_INVALID.  Every launch of the library is checked this
 * way (its kernel's static LDS plus the dynamic request) before it is issued.  No
 * reference counterpart: the reference launched through the CUDA driver unchecked
 * (src/weft/gpu_render.py:549-550, 670-675). */
int weft_check_lds(int64_t bytes);

/* Reference-semantics scoring of explicit primitive stacks (one lane per eval):
 * out_mse[i] = float32(SSE_i / (3 * ts^2)), SSE in NumPy's pairwise float32 order
 * over clip(render_tile(stack_i)) - target.  d_out_sse may be NULL.
 * Tile sizes 8, 16, 32, 64. */
int weft_score_evals(const WeftEval* d_evals, int n_evals, const WeftPrim* d_prims,
                     const float* d_tiles, float* d_out_mse, float* d_out_sse, void* stream);

/* Replaces the same batch_tile_objective launch (src/weft/gpu_render.py:113-152, launched
 * at 455-566) for the evals d_order[0 .. n) that all have tile size `tile_size` (8, 16,
 * 32, 64): each eval's pixels split over 2, 8, 32 or 32 lanes along NumPy's pairwise
 * tree, the same values as weft_score_evals bit for bit, for large tiles many times
 * faster.  Results go to the evals' own indices; an eval of another size gets NaN. */
int weft_score_evals_split(const WeftEval* d_evals, const int32_t* d_order, int n, int tile_size,
                           const WeftPrim* d_prims, const float* d_tiles, float* d_out_mse,
                           float* d_out_sse, void* stream);

/* Plain fit, part 1: objective of every tile's current canvas (call once after the
 * canvas is set).  d_tile_obj[t] = MSE(clip(canvas_t), target_t) + lam * bits_t. */
int weft_fit_init(const WeftFitState* st, void* stream);

/* Plain fit, part 1 from scratch: every tile back to an empty canvas and stack (canvas 0,
 * nsel 0, bits = bits0) and d_tile_obj[t] = MSE(0, target_t) + lam * bits0, in one launch.
 * The state an encode starts from in `_fit_adaptive_state` (`encoder.py:2745-2811`): the
 * same result as zeroing the canvas and nsel, setting bits and calling weft_fit_init. */
int weft_fit_reset(const WeftFitState* st, int32_t bits0, void* stream);

/* Plain fit, part 2a: per tile, the candidate indices sorted (stably) by coverage class
 * (ellipse, rectangle, triangle, constant, generic), so that each wavefront of the
 * scoring kernel runs one class.  d_order: n_tiles * n_cand int32.  Needed for explicit
 * candidates only: drawn candidates come in class order by construction (families in
 * aligned groups of 64), and weft_fit_score forms their order itself. */
int weft_fit_order(const WeftFitState* st, const WeftCands* cands, int32_t* d_order, void* stream);

/* Plain fit, part 2b: score candidates of the tiles in `work` against their cached
 * canvases, taking slots [cand0, cand0 + block) of each tile's class order (d_order for
 * explicit candidates; for drawn ones d_order is not read).  Per candidate k
 * of tile t (index t * n_cand + k) writes SSE / MSE when the tables are given, and per
 * work item the first minimum of the objective into d_part_obj / d_part_idx.  Tiles
 * whose stack is full are skipped. */
int weft_fit_score(const WeftFitState* st, const WeftWork* work, const WeftCands* cands,
                   const int32_t* d_order, float* d_out_sse, float* d_out_mse, double* d_part_obj,
                   int32_t* d_part_idx, void* stream);

/* Plain fit, part 3: per tile reduce its work items' partial minima (work items
 * d_tile_work0[t] + j * d_tile_wstride[t], j < d_tile_nwork[t], in the concatenated
 * partial arrays; d_tile_wstride may be NULL for stride 1), accept the
 * winner when its objective beats the tile's current one, composite it onto the canvas
 * and update the state.  Records per tile: winner index, objective, accepted flag and,
 * when candidates are drawn on the device, the winner's code (d_win_code may be NULL). */
int weft_fit_select(const WeftFitState* st, const WeftCands* cands, const int32_t* d_tile_work0,
                    const int32_t* d_tile_nwork, const int32_t* d_tile_wstride, const double* d_part_obj,
                    const int32_t* d_part_idx, int32_t* d_win_idx, double* d_win_obj,
                    int32_t* d_win_acc, WeftCode* d_win_code, void* stream);

/* ---- The reference's batched greedy rounds (src/weft/encoder.py:2682-2831) ----------
 * Round control of encoder._fit_adaptive_state on the device, for explicit ragged pools
 * (WeftCands with d_tile_ncand / d_tile_active pointing at this struct's arrays): a round is
 * weft_fit_order + weft_fit_score + weft_fit_select (float32 objective, obj_mode 1,
 * WEFT_CANDS_REMOVE_WINNER), then weft_adaptive_gate decides what the reference's loop does
 * next.  One workgroup walks all tiles, so every mode is one small launch:
 *   WEFT_ADAPT_START  complex[t] = obj[t] > threshold && ncand[t] > 0 (encoder.py:2737-2739)
 *   WEFT_ADAPT_COUNT  counts[1] = tiles whose winner was accepted (d_win_acc)
 *   WEFT_ADAPT_KEEP   complex[t] = win_acc[t]: only improved tiles stay (encoder.py:2831)
 *   WEFT_ADAPT_RESEED active[t] = complex[t] && nsel[t] < max_prims[t]: the tiles whose pools
 *                     a stall re-seeds (encoder.py:2813-2828)
 *   WEFT_ADAPT_NEXT   (also after START and KEEP) active[t] = complex[t] && ncand[t] > 0 &&
 *                     nsel[t] < max_prims[t] (encoder.py:2753-2757)
 * counts[0] receives the number of active tiles. */
#define WEFT_ADAPT_START 0
#define WEFT_ADAPT_COUNT 1
#define WEFT_ADAPT_KEEP 2
#define WEFT_ADAPT_RESEED 3
#define WEFT_ADAPT_NEXT 4
typedef struct WeftAdaptive {
    int32_t* d_complex;          /* per tile: still searched (complex_tiles) */
    int32_t* d_active;           /* per tile: takes part in the next round / re-seed */
    int32_t* d_tile_ncand;       /* per tile: live pool size */
    const int32_t* d_max_prims;  /* per tile: QuadTile.max_primitives */
    const int32_t* d_win_acc;    /* per tile: the last select's accepted flags */
    int32_t* d_counts;           /* [2]: active tiles, improved tiles (or records dropped) */
    double threshold;            /* complexity threshold of the initial objective */
    int32_t n_tiles;
    int32_t _pad;
} WeftAdaptive;

int weft_adaptive_gate(const WeftFitState* st, const WeftAdaptive* a, int mode, void* stream);

/* Re-seed: append the residual candidates of list entry i (d_new[i * max_new ..], d_new_count[i]
 * records, from weft_gen_residual_candidates over the tiles d_tiles[0 .. n)) to the pool of
 * tile d_tiles[i] when a->d_active marks it (after WEFT_ADAPT_RESEED), in order behind its live
 * candidates.  Pools hold `pool_cap` records each (the WeftCands n_cand of the rounds), tile t's
 * at d_pools + t * pool_cap, with the tile's accepted stack at the top of the region (see
 * WEFT_CANDS_REMOVE_WINNER), so live + accepted records never exceed pool_cap: records that
 * would not fit are dropped and counted in a->d_counts[1] (which WEFT_ADAPT_RESEED sets to 0). */
int weft_adaptive_append(const WeftFitState* st, const WeftAdaptive* a, WeftPrim* d_pools, int pool_cap,
                         const WeftPrim* d_new, const int32_t* d_new_count, int max_new, const int32_t* d_tiles,
                         int n, void* stream);

/* Allocation-free plain fit step.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  One call = weft_fit_order (explicit candidates only)
 * + weft_fit_score + weft_fit_select with the same arguments, the same results bit for
 * bit, but the class order and the partial minima live in a static device array of
 * WEFT_STATIC_SCRATCH_BYTES that the library holds on the device from weft_init until
 * the module is unloaded, instead of caller buffers.  `work` is one work list covering
 * every tile (as in one merged weft_fit_score launch), with d_tile_work0 / d_tile_nwork /
 * d_tile_wstride indexing into it from 0.  The scratch needs 4 * n_tiles * n_cand bytes
 * for explicit candidates (rounded up to 256) plus 12 * n_work bytes; a larger request
 * returns WEFT_E_TOO_LARGE and launches nothing.  The call allocates and frees nothing.
 * Every call on a device shares that device's one array: calls whose launches overlap
 * on one device (other streams, other threads) race and must be serialised by the
 * caller.  Nothing in the array has to start at zero (every entry a call reads, it wrote
 * first in the same call).  Work lists with split 64-px items (n_split > 0) are refused with
 * WEFT_E_INVALID: their partial sums would be scratch too. */
#define WEFT_STATIC_SCRATCH_BYTES (16u << 20)
int weft_fit_step_static(const WeftFitState* st, const WeftWork* work, const WeftCands* cands,
                         const int32_t* d_tile_work0, const int32_t* d_tile_nwork,
                         const int32_t* d_tile_wstride, int32_t* d_win_idx, double* d_win_obj,
                         int32_t* d_win_acc, WeftCode* d_win_code, void* stream);

/* Draw candidate codes / records for the given tiles exactly as weft_fit_score does on
 * the device (for checking and serialization).  d_codes / d_prims: n_tiles * n_cand.
 *   family : 0 ellipse (kind 6), 1 rectangle (kind 7), 2 triangle (kind 8)
 *   cx, cy : code / 1024 px, tile-local, in [0, ts)
 *   rx, ry : code / 1024 px, log-uniform in [4, 64)
 *   theta  : angle = code * pi / 4096, code < 4096
 *   r,g,b  : code / 255 ; alpha : code / 65535 in [0.3, 1] */
int weft_draw_codes(uint64_t seed, uint32_t step, const int32_t* d_tile_ids,
                    const int32_t* d_tile_ts, int n_tiles, int n_cand, WeftCode* d_codes,
                    WeftPrim* d_prims, void* stream);

/* Persistent batched fit (tiles of many images in one launch).  Workgroups take tiles
 * from a device queue (d_tile_order[0 .. n_queue)) and run each tile's whole greedy fit:
 * n_steps rounds of drawing n_cand candidates from (seed, step0 + s, tile, index),
 * scoring them against the tile's canvas, keeping the first minimum of the objective and
 * accepting it when it improves the tile.  Per round and tile the winner index,
 * objective, accepted flag and code go to d_win_*[s * n_tiles + t].  When the last
 * tile of an image finishes, d_image_done[image] receives the device real-time counter
 * (weft_wallclock_rate_khz), and d_t_start the earliest workgroup start.  Drawn
 * candidates only; kinds are ellipse / rectangle / triangle.  The launch is preceded
 * by a small reset kernel (queue, counters), so the pair can be captured in a graph. */
typedef struct WeftBatch {
    const int32_t* d_tile_order;
    const int32_t* d_tile_image;
    const int32_t* d_image_tiles;
    int32_t* d_image_left;
    uint64_t* d_image_done;
    uint64_t* d_t_start;
    int32_t* d_queue;
    int32_t* d_win_idx;
    double* d_win_obj;
    int32_t* d_win_acc;
    WeftCode* d_win_code;
    int32_t n_queue;
    int32_t n_images;
    uint32_t step0;
    int32_t n_steps;
    int32_t grid;   /* persistent workgroups */
    int32_t block;  /* threads per workgroup: 64, 128, 192 or 256 */
    /* weft_batch_refine_fit only (weft_batch_fit ignores them): the queue hands out groups
     * of tiles, group g = d_tile_order[d_group_first[g] .. d_group_first[g + 1]) (n_groups + 1
     * entries), whose runs of at most group_max tiles of one size are refined and re-scored
     * together; NULL: one tile per queue entry. */
    const int32_t* d_group_first;
    int32_t n_groups;
    int32_t group_max;  /* 1 .. 64 / k */
} WeftBatch;

int weft_batch_fit(const WeftFitState* st, const WeftCands* cands, const WeftBatch* b, void* stream);

/* Frequency of the counter behind d_image_done / d_t_start, in kHz. */
int weft_wallclock_rate_khz(int* khz);

/* ---- Differentiable refinement (no reference counterpart) -----------------------
 * Candidates of the three drawn families as continuous parameters, 9 floats each:
 * (cx, cy, rx, ry, theta, r, g, b, a), with the hard coverage replaced by a soft one
 * of edge width w pixels: ellipse sigma((1 - rho) * sqrt(rx*ry) / w) with rho the
 * elliptic radius; rectangle and triangle the product of sigma(d_i / w) over their
 * edges' signed distances d_i.  The candidate is composited over the tile's canvas
 * with alpha a * coverage, clipped to [0, 1] and compared with the target.  As w -> 0
 * the soft SSE tends to the hard-edge SSE of the scoring kernels. */
typedef struct WeftSoftBatch {
    const int32_t* d_group_tile;  /* tile of each group of k candidates */
    const int32_t* d_kinds;       /* [n_groups * k]: 6 ellipse, 7 rectangle, 8 triangle */
    const float* d_params;        /* [n_groups * k][9] */
    float* d_out_sse;             /* [n_groups * k] soft SSE */
    float* d_out_grad;            /* [n_groups * k][9] d SSE / d params, or NULL */
    int32_t n_groups;
    int32_t k;                    /* candidates per group: 1, 2, 4, ..., 64 */
    float w;                      /* edge width in pixels (> 0) */
    int32_t waves;                /* 0 or 1: 64 / k lanes per candidate; 2, 4, 8 (k = 1): the
                                     candidate's pixels over the 64 * waves lanes of that many
                                     waves, as weft_refine splits them for cpw = -waves */
} WeftSoftBatch;

int weft_soft_eval(const WeftFitState* st, const WeftSoftBatch* b, void* stream);

/* Per tile, the k smallest values of d_table[t * n .. t * n + n) as indices, ascending
 * by (value, index), i.e. the first k of a stable sort (NaN last).  Tiles whose stack
 * is full (d_tile_nsel >= depth) are skipped.  n <= 8192, k <= min(n, 256). */
int weft_topk(const WeftFitState* st, const float* d_table, int n, int k, int32_t* d_out_idx,
              void* stream);

/* Adam refinement of each tile's top-k drawn candidates (d_topk[t * k + j], drawn
 * from cands' seed and step), n_iter steps on the soft SSE, candidates spread over the
 * work items of WeftRefineOpts.  Step i uses edge width
 * d_sched[4 i], reciprocal bias corrections d_sched[4 i + 1] = 1 / (1 - beta1^(i+1)) and
 * d_sched[4 i + 2] = 1 / (1 - beta2^(i+1)) (m_hat = m1 * d_sched[4 i + 1]).  After each step the parameters are projected
 * onto the code ranges.  Outputs per tile k + 1 slots: slot 0 the best drawn
 * candidate unchanged, slot j + 1 candidate j refined and quantized to a code, as
 * codes and as records (exactly as the scoring kernels decode codes). */
#define WEFT_WORK_MAX_TILES 32  /* tile slots per work item (weft_refine, weft_fit_score_split) */
typedef struct WeftRefineOpts {
    float lr[9];
    float beta1, one_minus_beta1, beta2, one_minus_beta2, eps;
    int32_t k;       /* 1, 2, 4, ..., 64 */
    int32_t n_iter;  /* >= 0 */
    const float* d_sched;  /* [n_iter][4] */
    /* Work items, one workgroup each, holding cpw = cpw_by_size[log2(ts) - 3] candidates
     * (a power of two that divides k or is a multiple of it); the 64 / cpw lanes of a
     * candidate split its pixels, which fixes the order of the soft sums, so cpw is part
     * of the refinement's arithmetic.  cpw <= k: candidates d_work_j0[i] .. + cpw - 1 of
     * tile d_work_tile[i * WEFT_WORK_MAX_TILES]; cpw > k: all k candidates of each of the
     * cpw / k (<= WEFT_WORK_MAX_TILES) tiles d_work_tile[i * WEFT_WORK_MAX_TILES + g] (-1
     * for an unused slot; slot 0 is always a tile).  cpw = -W (W = 2, 4 or 8; one W for
     * every size that uses it): one candidate per item, its pixels over the 64 W lanes of
     * W waves; every item's workgroup then has 64 W lanes (an item of another size uses
     * its first wave).  It shortens the longest queue entries of weft_batch_refine_fit.
     * Every (tile, candidate) must be covered once; tiles whose stack is full are skipped.
     * weft_batch_refine_fit reads only cpw_by_size (one tile at a time, 64 / cpw lanes per
     * candidate, idle lanes when cpw > k; cpw = -W takes W of its workgroup's waves per
     * candidate, 64 W <= its block). */
    const int32_t* d_work_tile;
    const int32_t* d_work_j0;
    int32_t n_work;
    int32_t cpw_by_size[4];  /* tiles of 8, 16, 32, 64 px */
    int32_t _pad;
} WeftRefineOpts;

int weft_refine(const WeftFitState* st, const WeftCands* cands, const int32_t* d_topk,
                const WeftRefineOpts* o, WeftCode* d_out_codes, WeftPrim* d_out_prims,
                float* d_out_params, void* stream);

/* Persistent refined fit: weft_batch_fit whose every round also takes the tile's o->k
 * best drawn candidates (by MSE, then index), refines them by o->n_iter Adam steps on the
 * soft SSE (cpw_by_size candidates per wave, as weft_refine), scores the refined codes and
 * the unrefined best exactly, and accepts the first minimum if it beats the tile's
 * objective.  Records: win_idx = slot (0 the unrefined best, j + 1 refined candidate j),
 * win_code = the accepted code.  Equals the stepwise refined round (weft_fit_score,
 * weft_topk, weft_refine, weft_fit_score_split, weft_fit_select) with the same cpw.
 * The work-list fields of o are not used.  No reference counterpart. */
int weft_batch_refine_fit(const WeftFitState* st, const WeftCands* cands, const WeftBatch* b,
                          const WeftRefineOpts* o, void* stream);

/* Exact scoring (the same values as weft_fit_score) of a few candidates per tile, for
 * refined candidates (no reference counterpart; the candidates' SSE follows NumPy's
 * pairwise float32 order exactly, see weft_score_evals).  Work item i (one 64-lane
 * workgroup) holds tpw tiles, d_work_tile[i * WEFT_WORK_MAX_TILES + g] for g < tpw (-1
 * for an unused slot; slot 0 is always a tile), and scores candidates d_work_c0[i] ..
 * + cpw - 1 of each (those below cands->n_cand), with cpw = cpw_by_size[log2(ts) - 3] and
 * tpw = tpw_by_size[log2(ts) - 3]: each candidate's pixels are split over
 * 64 / (cpw * tpw) lanes along the pairwise tree, which must be a power of two at most
 * 2, 8, 32, 32 for tiles of 8, 16, 32, 64 px (several small tiles share a workgroup).
 * Outputs as weft_fit_score: the first minimum of the objective of tile slot g of item i
 * at d_part_obj / d_part_idx[i * WEFT_WORK_MAX_TILES + g] (for weft_fit_select), and
 * optional SSE / MSE tables. */
typedef struct WeftSplitWork {
    const int32_t* d_work_tile;  /* [n_work * WEFT_WORK_MAX_TILES] */
    const int32_t* d_work_c0;    /* [n_work] */
    int32_t n_work;
    int32_t cpw_by_size[4];
    int32_t tpw_by_size[4];
    int32_t _pad;
} WeftSplitWork;

int weft_fit_score_split(const WeftFitState* st, const WeftCands* cands, const WeftSplitWork* w,
                         float* d_out_sse, float* d_out_mse, double* d_part_obj, int32_t* d_part_idx,
                         void* stream);

/* ---- Chunked PRIM entropy coding (static order-0 rANS) ----------------------------
 * One chunk = the PRIM bytes of consecutive tiles, coded independently in the format of
 * src/weft/entropy.py: raw [0][u32 n][bytes] or rANS [1][u32 n][u16 n_pairs]
 * [(u8 sym, u16 freq)...][u32 state][u32 n_emit][emitted], 12-bit frequencies. */
typedef struct WeftChunkDesc {       /* ChunkDesc of src/weft/gpu_entropy.py:72-79 */
    int32_t start_tile;
    int32_t tile_count;
    int32_t raw_offset;
    int32_t raw_length;
    int32_t coded_offset;
    int32_t coded_length;
} WeftChunkDesc;

/* Decode every chunk (replaces weft_decode_chunks, src/weft/gpu_entropy.py:90-211):
 * d_status[i] = 0, or the reference decoder's error code for chunk i. */
int weft_rans_decode_chunks(const uint8_t* d_coded, const WeftChunkDesc* d_descs, int n_chunks,
                            uint8_t* d_out_raw, int32_t* d_status, void* stream);

/* Encode every chunk as encode_bytes would (src/weft/entropy.py:190-198, called per chunk
 * by build_prim_chunks, src/weft/prim_chunks.py:23-78) and pack the payloads back to
 * back.  In d_descs the caller fills start_tile, tile_count, raw_offset and raw_length;
 * the call fills coded_offset and coded_length.  Chunk i is staged at
 * d_stage + d_stage_off[i] (at least weft_rans_stage_bytes(raw_length) bytes); d_coded
 * receives the packed stream (at most the sum of the staging sizes) and *d_coded_total
 * its length.  d_status[i] = 0, or 12 if a symbol's frequency came out zero. */
int64_t weft_rans_stage_bytes(int64_t raw_length);
int weft_rans_encode_chunks(const uint8_t* d_raw, WeftChunkDesc* d_descs, int n_chunks,
                            uint8_t* d_stage, const int64_t* d_stage_off, uint8_t* d_coded,
                            int64_t* d_coded_total, int32_t* d_status, void* stream);

/* Host-side copies of the candidate tables (log-radius codes, cos, sin: 4096 each). */
int weft_get_tables(uint16_t* logr, float* cosv, float* sinv);

/* Decode: render every tile's stack into an H x W x 3 float32 image (RES0 bias, clip
 * to [0, 1]); pixels no tile covers are left untouched. */
int weft_render_tiles(const WeftTileDecodeDesc* d_descs, int n_tiles, const WeftPrim* d_prims,
                      float* d_out, int width, int height, void* stream);

/* Scale-independent scene render: render.render_scene_adaptive (src/weft/render.py:270-358)
 * for the descriptors its placement gives, as decoder._render_primitive_stack_gpu
 * (src/weft/decoder.py:231-390) does with the reference decode kernel, plus its RES1 step.
 * Tile i is rendered as a square of d_descs[i]._pad pixels sampling [0, source_size) with
 * step source_size / _pad on both axes (the pixel-centre grid when they are equal), cut to
 * its top-left target_w x target_h, given the RES0 bias and clipped (render_tile); tiles
 * i < n_res1 then add the bilinear upsample of their res1_grid x res1_grid x 3 float32 grid
 * d_res1[i] (upsample_residual_map, render.py:474-515, over target_h pixels: the caller
 * passes RES1 only for square rects) and are clipped again.  Each tile's rect
 * (target_x0, target_y0, target_w, target_h) is written where it falls inside the
 * width x height image; rects must not overlap (their writes are unordered). */
int weft_render_scene(const WeftTileDecodeDesc* d_descs, int n_tiles, const WeftPrim* d_prims,
                      const float* d_res1, int n_res1, int res1_grid, float* d_out, int width, int height,
                      void* stream);

/* Stage-3 template fit (src/weft/gpu_encoder.py:42-251, weft_model_bucket and
 * weft_tile_fit_candidates; launched by fit_tiles_gpu_constant, gpu_encoder.py:411-562).
 * 18 templates: 0 constant, 1-8 half-planes through the tile centre at angles (m-1)*pi/8,
 * 9-11 bands of half-width 1.5 (vertical, horizontal, diagonal), 12-14 corner triangles,
 * 15-17 bands of half-width 1.25 around quadratic curves; positions sampled at integer
 * coordinates.
 *
 * weft_template_masks: d_mask holds ts^2 + 18 uint32: d_mask[y * ts + x] gets bit m set
 * when position (x, y) falls in bucket 1 of template m, and d_mask[ts^2 + m] is the
 * number of positions in bucket 0 of template m.  Depends only on tile_size (1..64).
 *
 * weft_template_fit: the image is H x W x 3 float32 (row-major, RGB interleaved), cut into
 * ceil(W / ts) x ceil(H / ts) tiles in row-major order; pixels past the right / bottom
 * edge repeat the last column / row.  Per tile t and template m: bucket means
 * (sum * (1 / max(count, 1)), float32; template 0's colour1 = colour0) in
 * d_c0_all / d_c1_all[(t * 18 + m) * 3 + c], MSE of the two-colour prediction in
 * d_mse[t * 18 + m]; the first template of least MSE in d_model[t] and its colours in
 * d_c0 / d_c1[t * 3 + c].  Colours are not clipped.  Sums are float64 in a fixed order,
 * so the result is deterministic. */
#define WEFT_TEMPLATE_MODELS 18
int weft_template_masks(int tile_size, uint32_t* d_mask, void* stream);
int weft_template_fit(const float* d_image, int width, int height, int tile_size, const uint32_t* d_mask,
                      int32_t* d_model, float* d_c0, float* d_c1, float* d_mse, float* d_c0_all,
                      float* d_c1_all, void* stream);

/* Selection primitives (src/weft/cuda_backend.py:80-181).  Each takes a caller-owned
 * device workspace of at least weft_select_work_bytes(op, n) bytes; calls sharing one
 * workspace must not overlap.  Orders are NumPy's: NaN after +inf when sorting and first
 * in argmin, -0.0 equal to +0.0, equal keys in index order. */
#define WEFT_SELECT_ARGMIN 0
#define WEFT_SELECT_SCAN 1
#define WEFT_SELECT_COMPACT 2
#define WEFT_SELECT_TOPK 3
int64_t weft_select_work_bytes(int op, int64_t n);

/* cccl_argmin (cuda_backend.py:80-98): *d_out_index = np.argmin(values) -- the first
 * index of the least value, or of the first NaN.  n > 0. */
int weft_argmin_f64(const double* d_values, int64_t n, int64_t* d_out_index, void* d_work, int64_t work_bytes,
                    void* stream);

/* cccl_exclusive_scan (cuda_backend.py:101-118): d_out[i] = sum of d_in[0 .. i) (int64,
 * wrapping on overflow); the grand total goes to *d_total when it is not NULL. */
int weft_exclusive_scan_i64(const int64_t* d_in, int64_t n, int64_t* d_out, int64_t* d_total, void* d_work,
                            int64_t work_bytes, void* stream);

/* cccl_compact (cuda_backend.py:121-140): the indices of the nonzero flags in increasing
 * order into d_out_indices (room for n), their number into *d_out_count. */
int weft_compact_flags(const uint8_t* d_flags, int64_t n, int64_t* d_out_indices, int64_t* d_out_count,
                       void* d_work, int64_t work_bytes, void* stream);

/* cccl_segmented_topk (cuda_backend.py:143-181): the indices of the k smallest scores of
 * each segment (largest != 0: the k largest), listed in the order of a stable ascending
 * sort of all scores (largest: that order reversed), into d_out_indices (room for n); the
 * number written into *d_out_count.  n <= 2^30. */
int weft_segmented_topk(const double* d_scores, const int64_t* d_segments, int64_t n, int64_t k, int largest,
                        int64_t* d_out_indices, int64_t* d_out_count, void* d_work, int64_t work_bytes,
                        void* stream);

/* Template candidate pools (src/weft/encoder.py:535-745, _generate_candidates, with
 * _kmeans_colors 414-441, _split_by_lines_batched 343-404 and _dedup_candidates 447-530)
 * for the tiles d_tiles[0 .. n_tiles) of one size (8, 16, 32, 64 px) of a packed pixel
 * buffer (tile t's RGB pixels start at float 3 * d_tile_pix[t], row-major).  The pool of
 * list entry i -- the reference's candidates in its order, de-duplicated by its keys --
 * goes to d_out[i * max_cands ...] as packed records (geometry zero-padded, colour1 zero
 * except for linear patches) and its length to d_out_count[i]; when d_out_f64 is not NULL
 * it receives, per record, the same values before rounding to float32 (7 geometry, 3 + 3
 * colour, 1 alpha doubles).  max_cands >= WEFT_GEN_MAX_CANDS.  Where the reference depends
 * on unspecified NumPy behaviour the pool uses: a stable luminance sort for the palette,
 * row-major side sums for the linear patches, and the six brightest / darkest pixels in
 * ascending (luminance, index) order. */
#define WEFT_GEN_MAX_CANDS 352
int weft_gen_candidates(const float* d_pixels, const int64_t* d_tile_pix, const int32_t* d_tiles, int n_tiles,
                        int tile_size, int quality, WeftPrim* d_out, double* d_out_f64, int32_t* d_out_count,
                        int max_cands, void* stream);

/* Residual candidate pools (src/weft/encoder.py:1255-1366, _generate_residual_candidates,
 * with _split_by_line 325-340): candidates aimed at the error between each tile and its
 * prediction d_pred (same layout as d_pixels; clip_pred != 0 clips it to [0, 1] first, so
 * an unclipped canvas can be passed).  Output as for weft_gen_candidates, max_cands >=
 * WEFT_GEN_RES_MAX_CANDS.  The error pixels are ranked by x with a stable order (the
 * reference's np.argsort does not promise one for equal x), which picks the curve points. */
#define WEFT_GEN_RES_MAX_CANDS 128
int weft_gen_residual_candidates(const float* d_pixels, const float* d_pred, int clip_pred, const int64_t* d_tile_pix,
                                 const int32_t* d_tiles, int n_tiles, int tile_size, WeftPrim* d_out,
                                 double* d_out_f64, int32_t* d_out_count, int max_cands, void* stream);

/* Edge-driven candidate pools (src/weft/edge_analysis.py:139-315,
 * generate_edge_driven_candidates, with _sobel_edges 17-41, _dominant_orientations 44-68,
 * _extract_edge_positions 71-80 and _ct 133-136; encoder._dedup_candidates 447-530 over the
 * result).  Layout and output as for weft_gen_candidates, max_cands >= WEFT_GEN_EDGE_MAX_CANDS.
 * The Sobel gradients are formed as the reference's slices form them (every tap reads the
 * luminance one row up and one column left, so interior taps cancel); the float32 angle is
 * the arctangent of min / max rounded to float32 with float32 quadrant steps and
 * atan(1) = 0.7853981.  The line and curve sections, reached only when the 60th percentile of
 * the magnitudes is at least 1e-6, are not built: a tile with fewer than
 * floor(0.6 (n - 1)) + 2 magnitudes below 9.99e-7 (values far outside [0, 1]) gets the count
 * -1 and no pool.  Pixels must be finite: the reference's merged pool raises on NaN, and NaN
 * distances are not ranked as np.argmin ranks them. */
#define WEFT_GEN_EDGE_MAX_CANDS 114
int weft_gen_edge_candidates(const float* d_pixels, const int64_t* d_tile_pix, const int32_t* d_tiles, int n_tiles,
                             int tile_size, int quality, WeftPrim* d_out, double* d_out_f64, int32_t* d_out_count,
                             int max_cands, void* stream);

/* Merged pools: pool i of d_out is _dedup_candidates(a_i + b_i) (src/weft/encoder.py:447-530
 * over the concatenation that _fit_tile 1383-1385 and _worker_gen_cands 2513-2515 form from the
 * edge-driven and the template pool).  Pools as weft_gen_*_candidates write them: a_cap / b_cap
 * records and 14 doubles each per pool, counts in d_a_count / d_b_count; the keys are formed
 * from the doubles.  a_cap + b_cap <= WEFT_MERGE_MAX_CANDS <= out_cap is required; a count
 * outside [0, cap] (an edge pool's -1 among them) gives the merged count -1.  d_out_f64 may be
 * NULL. */
#define WEFT_MERGE_MAX_CANDS 512
int weft_merge_pools(const WeftPrim* d_a, const double* d_a_f64, const int32_t* d_a_count, int a_cap,
                     const WeftPrim* d_b, const double* d_b_f64, const int32_t* d_b_count, int b_cap, int n_pools,
                     WeftPrim* d_out, double* d_out_f64, int32_t* d_out_count, int out_cap, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WEFT_HIP_H */
