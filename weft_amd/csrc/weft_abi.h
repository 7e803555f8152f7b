// Kernel-argument blocks shared by the gfx950 code object (weft_kernels.hip) and the
// host launcher (weft_host.cpp).  Each kernel takes one of these by value.
#pragma once

#include "weft_hip.h"

#define WEFT_ABI_VERSION 22
#define WEFT_MAX_BLOCK 1024
#define WEFT_N_CLASSES 6
#define WEFT_TABLE_N 4096

struct WeftScoreEvalsArgs {
    const WeftEval* evals;
    const WeftPrim* prims;
    const float* tiles;
    float* out_mse;
    float* out_sse;
    int32_t n_evals;
    int32_t _pad;
};

struct WeftFitInitArgs {
    WeftFitState st;
    int32_t reset;  // go back to an empty canvas and stack first
    int32_t bits0;  // the bits of an empty stack (with reset)
};

struct WeftFitOrderArgs {
    WeftFitState st;
    WeftCands cd;
    int32_t* order;
};

struct WeftFitScoreArgs {
    WeftFitState st;
    WeftWork wk;
    WeftCands cd;
    const int32_t* order;
    float* out_sse;
    float* out_mse;
    double* part_obj;
    int32_t* part_idx;
};

struct WeftFitSelectArgs {
    WeftFitState st;
    WeftCands cd;
    const int32_t* tile_work0;
    const int32_t* tile_nwork;
    const int32_t* tile_wstride;
    const double* part_obj;
    const int32_t* part_idx;
    int32_t* win_idx;
    double* win_obj;
    int32_t* win_acc;
    WeftCode* win_code;
};

struct WeftDrawArgs {
    const int32_t* tile_ids;
    const int32_t* tile_ts;
    WeftCode* codes;
    WeftPrim* prims;
    uint64_t seed;
    uint32_t step;
    int32_t n_tiles;
    int32_t n_cand;
    int32_t _pad;
};

struct WeftRenderArgs {
    const WeftTileDecodeDesc* descs;
    const WeftPrim* prims;
    float* out;
    const float* res1;   // RES1 grids of tiles 0 .. n_res1-1 (g x g x 3 float32 each), or NULL
    int32_t n_tiles;
    int32_t width;
    int32_t height;
    int32_t scene;       // 1: desc._pad is the tile's square render size (render_scene_adaptive)
    int32_t n_res1;
    int32_t res1_grid;
};

struct WeftBatchArgs {
    WeftFitState st;
    WeftCands cd;
    const int32_t* tile_order;   // queue: tile indices in processing order
    const int32_t* tile_image;   // image of each tile
    const int32_t* image_tiles;  // tiles per image (reset value of image_left)
    int32_t* image_left;         // tiles of the image not yet finished
    unsigned long long* image_done;  // device real-time stamp when the image's last tile finished
    unsigned long long* t_start;     // earliest workgroup start stamp
    int32_t* queue;              // next queue position
    int32_t* win_idx;            // [n_steps][n_tiles]
    double* win_obj;
    int32_t* win_acc;
    WeftCode* win_code;
    int32_t n_queue;
    int32_t n_images;
    uint32_t step0;
    int32_t n_steps;
};

struct WeftSoftArgs {
    WeftFitState st;
    WeftSoftBatch b;
};

struct WeftTopkArgs {
    WeftFitState st;
    const float* table;
    int32_t* out_idx;
    int32_t n;
    int32_t k;
    int32_t n_pow2;
    int32_t _pad;
};

struct WeftRefineArgs {
    WeftFitState st;
    WeftCands cd;
    WeftRefineOpts o;
    const int32_t* topk;
    WeftCode* out_codes;
    WeftPrim* out_prims;
    float* out_params;
};

struct WeftSplitScoreArgs {
    WeftFitState st;
    WeftCands cd;
    const int32_t* work_tile;
    const int32_t* work_c0;
    int32_t cpw[4];
    int32_t tpw[4];
    float* out_sse;
    float* out_mse;
    double* part_obj;
    int32_t* part_idx;
};

struct WeftRansEncodeArgs {
    const uint8_t* raw;
    WeftChunkDesc* descs;
    uint8_t* stage;
    const int64_t* stage_off;
    uint8_t* coded;
    int64_t* coded_total;
    int32_t* status;
    int32_t n_chunks;
    int32_t _pad;
};

struct WeftRansDecodeArgs {
    const uint8_t* coded;
    const WeftChunkDesc* descs;
    uint8_t* out_raw;
    int32_t* status;
    int32_t n_chunks;
    int32_t _pad;
};

struct WeftBatchRefineArgs {
    WeftBatchArgs b;
    WeftRefineOpts o;
    const int32_t* group_first;  // queue entries: groups of tile_order (nullptr: one tile each)
    int32_t n_groups;
    int32_t group_max;           // tiles refined together, at most 64 / k
    int32_t split_lanes[4];
    int32_t n_pow2;
    int32_t _pad;
};

struct WeftScoreEvalsSplitArgs {
    const WeftEval* evals;
    const int32_t* order;
    const WeftPrim* prims;
    const float* tiles;
    float* out_mse;
    float* out_sse;
    int32_t n;
    int32_t tile_size;
};

struct WeftTemplateMaskArgs {
    uint32_t* mask;
    int32_t tile_size;
    int32_t _pad;
};

struct WeftTemplateFitArgs {
    const float* image;
    const uint32_t* mask;
    int32_t* out_model;
    float* out_c0;
    float* out_c1;
    float* out_mse;
    float* out_c0_all;
    float* out_c1_all;
    int32_t width;
    int32_t height;
    int32_t tile_size;
    int32_t tile_cols;
    int32_t tile_rows;
    int32_t _pad;
};

struct WeftSelectArgs {
    const double* f64_in;
    const int64_t* i64_in;
    const uint8_t* u8_in;
    int64_t* out_i64;
    int64_t* out_total;
    double* work_f64;
    int64_t* work_i64;
    uint8_t* flags;
    const int64_t* positions;
    const int64_t* count_in;
    const uint64_t* rs_keys_in;   // radix pass input / output
    const uint32_t* rs_vals_in;
    uint64_t* rs_keys_out;
    uint32_t* rs_vals_out;
    int64_t* rs_table;            // digit-major per-workgroup digit counts
    const int64_t* rs_prefix;     // their exclusive scan
    const uint32_t* rs_order;     // element index at each position of the score order
    int64_t n;
    int64_t n_blocks;
    int64_t k;
    int32_t largest;
    int32_t rs_shift;
};

struct WeftCandGenArgs {
    const float* pixels;
    const int64_t* tile_pix;
    const int32_t* tiles;
    WeftPrim* out;
    double* out_f64;
    int32_t* out_count;
    int32_t n_tiles;
    int32_t tile_size;
    int32_t quality;
    int32_t max_cands;
    double cos_ang[16];
    double sin_ang[16];
    double cos6[6];
    double sin6[6];
};

struct WeftEdgeGenArgs {
    const float* pixels;
    const int64_t* tile_pix;
    const int32_t* tiles;
    WeftPrim* out;
    double* out_f64;
    int32_t* out_count;
    int32_t n_tiles;
    int32_t tile_size;
    int32_t quality;
    int32_t max_cands;
    int32_t min_small;  // floor(0.6 (n - 1)) + 2 magnitudes below 9.99e-7: no edge points
    int32_t _pad;
    double edges[13];   // np.linspace(0, pi, 13)
    double cos_o[28];   // 12 bin midpoints, 0, pi/4, pi/2, 3pi/4, 12 bin starts + pi/24
    double sin_o[28];
};

struct WeftMergePoolsArgs {
    const WeftPrim* a;
    const double* a_f64;
    const int32_t* a_count;
    const WeftPrim* b;
    const double* b_f64;
    const int32_t* b_count;
    WeftPrim* out;
    double* out_f64;
    int32_t* out_count;
    int32_t a_cap;
    int32_t b_cap;
    int32_t out_cap;
    int32_t n_pools;
};

struct WeftResGenArgs {
    const float* pixels;
    const float* pred;
    const int64_t* tile_pix;
    const int32_t* tiles;
    WeftPrim* out;
    double* out_f64;
    int32_t* out_count;
    int32_t n_tiles;
    int32_t tile_size;
    int32_t clip_pred;
    int32_t max_cands;
};

struct WeftAdaptGateArgs {
    WeftFitState st;
    WeftAdaptive ad;
    int32_t mode;
    int32_t _pad;
};

struct WeftAdaptAppendArgs {
    WeftFitState st;
    WeftAdaptive ad;
    WeftPrim* pools;
    const WeftPrim* fresh;
    const int32_t* fresh_count;
    const int32_t* tiles;
    int32_t pool_cap;
    int32_t max_new;
    int32_t n;
    int32_t _pad;
};
