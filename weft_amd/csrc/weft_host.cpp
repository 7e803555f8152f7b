// weft_host.cpp -- libweft_hip.so: the C ABI declared in include/weft_hip.h.
//
// Loads the gfx950 code object that the build embeds in this library
// (weft_kernels.hsaco, linked in by weft_embed.S) with hipModuleLoadData and launches
// its kernels with hipModuleLaunchKernel.  The reference compiled CUDA source with
// NVRTC at run time and kept module handles in a Python dict
// (src/weft/gpu_render.py:696-734, src/weft/gpu_context.py:16-62); here the code object
// is built offline and the handles live in this file.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

#include "weft_abi.h"
#include "weft_hip.h"

extern "C" const unsigned char weft_hsaco_start[];
extern "C" const unsigned char weft_hsaco_end[];

namespace {

struct Module {
    int device = -1;
    hipModule_t mod = nullptr;
    hipFunction_t f_evals = nullptr;
    hipFunction_t f_init = nullptr;
    hipFunction_t f_score[4] = {nullptr, nullptr, nullptr, nullptr};
    hipFunction_t f_score_gen[4] = {nullptr, nullptr, nullptr, nullptr};
    hipFunction_t f_score_mixed = nullptr;
    hipFunction_t f_score_gen_mixed = nullptr;
    hipFunction_t f_score_combine = nullptr;
    hipFunction_t f_order = nullptr;
    hipFunction_t f_select = nullptr;
    hipFunction_t f_draw = nullptr;
    hipFunction_t f_batch = nullptr;
    hipFunction_t f_soft = nullptr;
    hipFunction_t f_topk = nullptr;
    hipFunction_t f_refine = nullptr;
    hipFunction_t f_refine_mw = nullptr;
    hipFunction_t f_split = nullptr;
    hipFunction_t f_batch_refine = nullptr;
    hipFunction_t f_evals_split = nullptr;
    hipFunction_t f_tmpl_masks = nullptr;
    hipFunction_t f_tmpl_fit = nullptr;
    hipFunction_t f_sel_argmin_partial = nullptr;
    hipFunction_t f_sel_argmin_final = nullptr;
    hipFunction_t f_sel_tile_sums = nullptr;
    hipFunction_t f_sel_scan_tiles = nullptr;
    hipFunction_t f_sel_scan_apply = nullptr;
    hipFunction_t f_sel_rs_score_keys = nullptr;
    hipFunction_t f_sel_rs_hist = nullptr;
    hipFunction_t f_sel_rs_scatter = nullptr;
    hipFunction_t f_sel_rs_segment_keys = nullptr;
    hipFunction_t f_sel_rs_keep = nullptr;
    hipFunction_t f_sel_rs_emit = nullptr;
    hipFunction_t f_gen_cands = nullptr;
    hipFunction_t f_gen_res = nullptr;
    hipFunction_t f_gen_edge = nullptr;
    hipFunction_t f_merge_pools = nullptr;
    hipFunction_t f_rans_enc = nullptr;
    hipFunction_t f_rans_scan = nullptr;
    hipFunction_t f_rans_pack = nullptr;
    hipFunction_t f_rans_dec = nullptr;
    hipFunction_t f_batch_reset = nullptr;
    hipFunction_t f_render = nullptr;
    hipFunction_t f_adapt_gate = nullptr;
    hipFunction_t f_adapt_append = nullptr;
    int64_t max_lds = 0;                            // hipDeviceAttributeMaxSharedMemoryPerBlock
    unsigned char* scratch = nullptr;               // device address of weft_scratch on `device`
    std::unordered_map<hipFunction_t, int> static_lds;  // __shared__ bytes of each kernel
};

// LDS per workgroup on gfx950 (the limit checked before a device is initialised).
constexpr int64_t kGfx950LdsBytes = 160 * 1024;

Module g_mod;
std::mutex g_mu;
uint16_t g_logr[WEFT_TABLE_N];
float g_cos[WEFT_TABLE_N];
float g_sin[WEFT_TABLE_N];
float g_u8[256];
bool g_tables = false;
thread_local std::string g_err;

int set_err(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return -(int)e;
}

#define WEFT_HIP(call, what)                          \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

void build_tables() {
    if (g_tables) return;
    for (int k = 0; k < WEFT_TABLE_N; ++k) {
        // log-uniform radius over [4, 64) px, stored as code = round(r * 1024)
        double r = 4.0 * std::pow(16.0, ((double)k + 0.5) / (double)WEFT_TABLE_N);
        g_logr[k] = (uint16_t)std::llround(r * 1024.0);
        double th = (double)k * 3.14159265358979323846 / (double)WEFT_TABLE_N;
        g_cos[k] = (float)std::cos(th);
        if (k < 256) g_u8[k] = (float)((double)k / 255.0);
        g_sin[k] = (float)std::sin(th);
    }
    g_tables = true;
}

int upload_global(const char* name, const void* src, size_t bytes) {
    hipDeviceptr_t dptr = nullptr;
    size_t sz = 0;
    WEFT_HIP(hipModuleGetGlobal(&dptr, &sz, g_mod.mod, name), name);
    if (sz != bytes) return set_err(WEFT_E_MODULE, std::string("size mismatch for ") + name);
    WEFT_HIP(hipMemcpyHtoD(dptr, const_cast<void*>(src), bytes), name);
    return WEFT_OK;
}

int require_init() {
    if (!g_mod.mod) return set_err(WEFT_E_NOT_INITIALISED, "weft_init has not been called");
    return WEFT_OK;
}

// A workgroup's LDS (static + dynamic) against the device limit.
int lds_check(int64_t bytes, const char* what) {
    const int64_t limit = g_mod.mod ? g_mod.max_lds : kGfx950LdsBytes;
    if (bytes < 0 || bytes > limit)
        return set_err(WEFT_E_INVALID, std::string(what) + ": " + std::to_string(bytes) +
                                           " bytes of LDS per workgroup exceed the device limit of " +
                                           std::to_string(limit));
    return WEFT_OK;
}

int launch(hipFunction_t f, unsigned grid, unsigned block, unsigned shmem, void* stream, void* args) {
    if (grid == 0) return WEFT_OK;
    const auto it = g_mod.static_lds.find(f);
    const int64_t static_bytes = it == g_mod.static_lds.end() ? 0 : it->second;
    if (int rc = lds_check(static_bytes + (int64_t)shmem, "kernel launch")) return rc;
    void* params[] = {args};
    WEFT_HIP(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, shmem, (hipStream_t)stream, params,
                                   nullptr),
             "hipModuleLaunchKernel");
    return WEFT_OK;
}

int ts_slot(int ts) {
    switch (ts) {
    case 8: return 0;
    case 16: return 1;
    case 32: return 2;
    case 64: return 3;
    default: return -1;
    }
}

unsigned score_lds_bytes(int ts) {
    // one band of at most 512 pixels (36 bytes each: two float4 + one float) + the
    // per-wave argmin reduction; a mixed-size launch (ts == 0) sizes for the largest band
    if (ts == 0) ts = 64;
    const int rows = ts * ts <= 512 ? ts : 512 / ts;
    return (unsigned)(36 * rows * ts + 8 * 16 + 4 * 16);
}

}  // namespace

extern "C" {

int weft_abi_version(void) { return WEFT_ABI_VERSION; }

const char* weft_last_error(void) { return g_err.c_str(); }

int weft_check_lds(int64_t bytes) { return lds_check(bytes, "weft_check_lds"); }

// The Python binding (weft_amd/_lib.py) mirrors these layouts.
static_assert(sizeof(WeftFitState) == 80, "WeftFitState layout");
static_assert(sizeof(WeftWork) == 56, "WeftWork layout");
static_assert(sizeof(WeftCands) == 56, "WeftCands layout");
static_assert(sizeof(WeftAdaptive) == 64, "WeftAdaptive layout");
static_assert(sizeof(WeftBatch) == 128, "WeftBatch layout");
static_assert(sizeof(WeftSoftBatch) == 56, "WeftSoftBatch layout");
static_assert(sizeof(WeftRefineOpts) == 112, "WeftRefineOpts layout");
static_assert(sizeof(WeftCode) == 20, "WeftCode layout");
static_assert(sizeof(WeftSplitWork) == 56, "WeftSplitWork layout");
static_assert(sizeof(WeftChunkDesc) == 24, "WeftChunkDesc layout");
static_assert(sizeof(WeftPrim) == 60, "WeftPrim layout");

int weft_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_mod.mod && g_mod.device == device) return WEFT_OK;
    if (g_mod.mod) return set_err(WEFT_E_INVALID, "weft_init: already initialised on another device");
    int n = 0;
    WEFT_HIP(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) return set_err(WEFT_E_INVALID, "weft_init: no such device");
    WEFT_HIP(hipSetDevice(device), "hipSetDevice");
    Module m;
    m.device = device;
    WEFT_HIP(hipModuleLoadData(&m.mod, weft_hsaco_start), "hipModuleLoadData");
    g_mod = m;
    struct {
        hipFunction_t* f;
        const char* name;
    } fns[] = {
        {&g_mod.f_evals, "weft_score_evals_kernel"},
        {&g_mod.f_init, "weft_fit_init_kernel"},
        {&g_mod.f_score[0], "weft_fit_score_ts8"},
        {&g_mod.f_score[1], "weft_fit_score_ts16"},
        {&g_mod.f_score[2], "weft_fit_score_ts32"},
        {&g_mod.f_score[3], "weft_fit_score_ts64"},
        {&g_mod.f_score_gen[0], "weft_fit_score_gen_ts8"},
        {&g_mod.f_score_gen[1], "weft_fit_score_gen_ts16"},
        {&g_mod.f_score_gen[2], "weft_fit_score_gen_ts32"},
        {&g_mod.f_score_gen[3], "weft_fit_score_gen_ts64"},
        {&g_mod.f_score_mixed, "weft_fit_score_mixed"},
        {&g_mod.f_score_gen_mixed, "weft_fit_score_gen_mixed"},
        {&g_mod.f_score_combine, "weft_fit_score_combine_kernel"},
        {&g_mod.f_order, "weft_fit_order_kernel"},
        {&g_mod.f_select, "weft_fit_select_kernel"},
        {&g_mod.f_draw, "weft_draw_kernel"},
        {&g_mod.f_batch, "weft_batch_fit_kernel"},
        {&g_mod.f_soft, "weft_soft_eval_kernel"},
        {&g_mod.f_topk, "weft_topk_kernel"},
        {&g_mod.f_refine, "weft_refine_kernel"},
        {&g_mod.f_refine_mw, "weft_refine_mw_kernel"},
        {&g_mod.f_split, "weft_fit_score_split_kernel"},
        {&g_mod.f_batch_refine, "weft_batch_refine_kernel"},
        {&g_mod.f_evals_split, "weft_score_evals_split_kernel"},
        {&g_mod.f_tmpl_masks, "weft_template_masks_kernel"},
        {&g_mod.f_tmpl_fit, "weft_template_fit_kernel"},
        {&g_mod.f_sel_argmin_partial, "weft_sel_argmin_partial_kernel"},
        {&g_mod.f_sel_argmin_final, "weft_sel_argmin_final_kernel"},
        {&g_mod.f_sel_tile_sums, "weft_sel_tile_sums_kernel"},
        {&g_mod.f_sel_scan_tiles, "weft_sel_scan_tiles_kernel"},
        {&g_mod.f_sel_scan_apply, "weft_sel_scan_apply_kernel"},
        {&g_mod.f_sel_rs_score_keys, "weft_sel_rs_score_keys_kernel"},
        {&g_mod.f_sel_rs_hist, "weft_sel_rs_hist_kernel"},
        {&g_mod.f_sel_rs_scatter, "weft_sel_rs_scatter_kernel"},
        {&g_mod.f_sel_rs_segment_keys, "weft_sel_rs_segment_keys_kernel"},
        {&g_mod.f_sel_rs_keep, "weft_sel_rs_keep_kernel"},
        {&g_mod.f_sel_rs_emit, "weft_sel_rs_emit_kernel"},
        {&g_mod.f_gen_cands, "weft_gen_candidates_kernel"},
        {&g_mod.f_gen_res, "weft_gen_residual_candidates_kernel"},
        {&g_mod.f_gen_edge, "weft_gen_edge_candidates_kernel"},
        {&g_mod.f_merge_pools, "weft_merge_pools_kernel"},
        {&g_mod.f_rans_enc, "weft_rans_encode_kernel"},
        {&g_mod.f_rans_scan, "weft_rans_scan_kernel"},
        {&g_mod.f_rans_pack, "weft_rans_pack_kernel"},
        {&g_mod.f_rans_dec, "weft_rans_decode_kernel"},
        {&g_mod.f_batch_reset, "weft_batch_reset_kernel"},
        {&g_mod.f_render, "weft_render_tiles_kernel"},
        {&g_mod.f_adapt_gate, "weft_adaptive_gate_kernel"},
        {&g_mod.f_adapt_append, "weft_adaptive_append_kernel"},
    };
    for (auto& fn : fns) {
        hipError_t e = hipModuleGetFunction(fn.f, g_mod.mod, fn.name);
        int lds = 0;
        if (e == hipSuccess) e = hipFuncGetAttribute(&lds, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, *fn.f);
        if (e != hipSuccess) {
            (void)hipModuleUnload(g_mod.mod);
            g_mod = Module();
            return hip_fail(e, fn.name);
        }
        g_mod.static_lds[*fn.f] = lds;
    }
    int max_lds = 0;
    hipError_t e_lds = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    if (e_lds != hipSuccess) {
        (void)hipModuleUnload(g_mod.mod);
        g_mod = Module();
        return hip_fail(e_lds, "hipDeviceGetAttribute(MaxSharedMemoryPerBlock)");
    }
    g_mod.max_lds = max_lds;
    // The allocation-free entry point's scratch: the code object's static array on this
    // device (the device is current), looked up once and kept with the module.
    hipDeviceptr_t scratch = nullptr;
    size_t scratch_bytes = 0;
    hipError_t e_scr = hipModuleGetGlobal(&scratch, &scratch_bytes, g_mod.mod, "weft_scratch");
    if (e_scr != hipSuccess || scratch_bytes != WEFT_STATIC_SCRATCH_BYTES) {
        (void)hipModuleUnload(g_mod.mod);
        g_mod = Module();
        if (e_scr != hipSuccess) return hip_fail(e_scr, "hipModuleGetGlobal(weft_scratch)");
        return set_err(WEFT_E_MODULE, "weft_scratch has an unexpected size");
    }
    g_mod.scratch = static_cast<unsigned char*>(scratch);
    build_tables();
    int rc = upload_global("weft_tab_logr", g_logr, sizeof(g_logr));
    if (rc == WEFT_OK) rc = upload_global("weft_tab_cos", g_cos, sizeof(g_cos));
    if (rc == WEFT_OK) rc = upload_global("weft_tab_sin", g_sin, sizeof(g_sin));
    if (rc == WEFT_OK) rc = upload_global("weft_tab_u8", g_u8, sizeof(g_u8));
    if (rc != WEFT_OK) {
        (void)hipModuleUnload(g_mod.mod);
        g_mod = Module();
    }
    return rc;
}

int weft_shutdown(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_mod.mod) {
        hipError_t e = hipModuleUnload(g_mod.mod);
        g_mod = Module();
        if (e != hipSuccess) return hip_fail(e, "hipModuleUnload");
    }
    return WEFT_OK;
}

int weft_get_tables(uint16_t* logr, float* cosv, float* sinv) {
    build_tables();
    if (logr) std::memcpy(logr, g_logr, sizeof(g_logr));
    if (cosv) std::memcpy(cosv, g_cos, sizeof(g_cos));
    if (sinv) std::memcpy(sinv, g_sin, sizeof(g_sin));
    return WEFT_OK;
}

int weft_score_evals(const WeftEval* d_evals, int n_evals, const WeftPrim* d_prims,
                     const float* d_tiles, float* d_out_mse, float* d_out_sse, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n_evals < 0) return set_err(WEFT_E_INVALID, "weft_score_evals: n_evals < 0");
    if (n_evals == 0) return WEFT_OK;
    if (!d_evals || !d_prims || !d_tiles || !d_out_mse)
        return set_err(WEFT_E_INVALID, "weft_score_evals: null pointer");
    WeftScoreEvalsArgs a{d_evals, d_prims, d_tiles, d_out_mse, d_out_sse, n_evals, 0};
    return launch(g_mod.f_evals, (unsigned)((n_evals + 255) / 256), 256, 0, stream, &a);
}

int weft_score_evals_split(const WeftEval* d_evals, const int32_t* d_order, int n, int tile_size,
                           const WeftPrim* d_prims, const float* d_tiles, float* d_out_mse, float* d_out_sse,
                           void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n < 0) return set_err(WEFT_E_INVALID, "weft_score_evals_split: n < 0");
    if (tile_size != 8 && tile_size != 16 && tile_size != 32 && tile_size != 64)
        return set_err(WEFT_E_UNSUPPORTED_TILE, "weft_score_evals_split: tile size must be 8, 16, 32 or 64");
    if (n == 0) return WEFT_OK;
    if (!d_evals || !d_order || !d_prims || !d_tiles || !d_out_mse)
        return set_err(WEFT_E_INVALID, "weft_score_evals_split: null pointer");
    const int lanes = tile_size == 8 ? 2 : (tile_size == 16 ? 8 : 32);
    const int per = 64 / lanes;
    WeftScoreEvalsSplitArgs a{d_evals, d_order, d_prims, d_tiles, d_out_mse, d_out_sse, n, tile_size};
    return launch(g_mod.f_evals_split, (unsigned)((n + per - 1) / per), 64, 0, stream, &a);
}

int weft_fit_init(const WeftFitState* st, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || st->n_tiles < 0) return set_err(WEFT_E_INVALID, "weft_fit_init: bad state");
    if (st->n_tiles == 0) return WEFT_OK;
    WeftFitInitArgs a{*st, 0, 0};
    return launch(g_mod.f_init, (unsigned)st->n_tiles, 64, 0, stream, &a);
}

int weft_fit_reset(const WeftFitState* st, int32_t bits0, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || st->n_tiles < 0) return set_err(WEFT_E_INVALID, "weft_fit_reset: bad state");
    if (st->n_tiles == 0) return WEFT_OK;
    WeftFitInitArgs a{*st, 1, bits0};
    return launch(g_mod.f_init, (unsigned)st->n_tiles, 64, 0, stream, &a);
}

int weft_fit_order(const WeftFitState* st, const WeftCands* cands, int32_t* d_order, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !cands || !d_order) return set_err(WEFT_E_INVALID, "weft_fit_order: null argument");
    if (cands->n_cand < 1) return set_err(WEFT_E_INVALID, "weft_fit_order: n_cand < 1");
    WeftFitOrderArgs a{*st, *cands, d_order};
    return launch(g_mod.f_order, (unsigned)st->n_tiles, 256, 0, stream, &a);
}

int weft_fit_score(const WeftFitState* st, const WeftWork* work, const WeftCands* cands,
                   const int32_t* d_order, float* d_out_sse, float* d_out_mse, double* d_part_obj,
                   int32_t* d_part_idx, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !work || !cands || !d_order || !d_part_obj || !d_part_idx)
        return set_err(WEFT_E_INVALID, "weft_fit_score: null argument");
    const int slot = work->tile_size == 0 ? 4 : ts_slot(work->tile_size);
    if (slot < 0)
        return set_err(WEFT_E_UNSUPPORTED_TILE, "weft_fit_score: tile size must be 0 (mixed), 8, 16, 32 or 64");
    if (work->block < 64 || work->block > WEFT_MAX_BLOCK || (work->block % 64) != 0)
        return set_err(WEFT_E_INVALID, "weft_fit_score: block must be a multiple of 64 in [64, 1024]");
    if (cands->n_cand < 1) return set_err(WEFT_E_INVALID, "weft_fit_score: n_cand < 1");
    if (work->n_split < 0 || work->n_split > work->n_work)
        return set_err(WEFT_E_INVALID, "weft_fit_score: n_split must be in [0, n_work]");
    if (work->n_split > 0) {
        if (!work->d_work_split || !work->d_split_pairs || !work->d_split_sse)
            return set_err(WEFT_E_INVALID, "weft_fit_score: split items need d_work_split, d_split_pairs and d_split_sse");
        if (work->tile_size != 0 && work->tile_size != 64)
            return set_err(WEFT_E_INVALID, "weft_fit_score: only 64-px tiles are split");
    }
    if (work->n_work <= 0) return WEFT_OK;
    WeftFitScoreArgs a{*st, *work, *cands, d_order, d_out_sse, d_out_mse, d_part_obj, d_part_idx};
    if (work->n_split == 0) a.wk.d_work_split = nullptr;
    hipFunction_t f = slot == 4 ? (cands->has_general ? g_mod.f_score_gen_mixed : g_mod.f_score_mixed)
                                : (cands->has_general ? g_mod.f_score_gen[slot] : g_mod.f_score[slot]);
    if ((rc = launch(f, (unsigned)work->n_work, (unsigned)work->block, score_lds_bytes(work->tile_size), stream,
                     &a)))
        return rc;
    if (work->n_split == 0) return WEFT_OK;
    // the split pairs' partial sums -> SSE, tables and each pair's first minimum
    return launch(g_mod.f_score_combine, (unsigned)work->n_split, (unsigned)work->block, 0, stream, &a);
}

int weft_fit_select(const WeftFitState* st, const WeftCands* cands, const int32_t* d_tile_work0,
                    const int32_t* d_tile_nwork, const int32_t* d_tile_wstride, const double* d_part_obj,
                    const int32_t* d_part_idx, int32_t* d_win_idx, double* d_win_obj,
                    int32_t* d_win_acc, WeftCode* d_win_code, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !cands || !d_tile_work0 || !d_tile_nwork || !d_part_obj || !d_part_idx ||
        !d_win_idx || !d_win_obj || !d_win_acc)
        return set_err(WEFT_E_INVALID, "weft_fit_select: null argument");
    if ((cands->flags & WEFT_CANDS_REMOVE_WINNER) && (!cands->d_prims || !cands->d_tile_ncand))
        return set_err(WEFT_E_INVALID, "weft_fit_select: removing winners needs explicit pools with live counts");
    WeftFitSelectArgs a{*st, *cands, d_tile_work0, d_tile_nwork, d_tile_wstride, d_part_obj, d_part_idx,
                        d_win_idx, d_win_obj, d_win_acc, d_win_code};
    return launch(g_mod.f_select, (unsigned)st->n_tiles, 256, 0, stream, &a);
}

int weft_adaptive_gate(const WeftFitState* st, const WeftAdaptive* a, int mode, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !a || !a->d_complex || !a->d_active || !a->d_tile_ncand || !a->d_max_prims || !a->d_counts)
        return set_err(WEFT_E_INVALID, "weft_adaptive_gate: null argument");
    if (mode < WEFT_ADAPT_START || mode > WEFT_ADAPT_NEXT) return set_err(WEFT_E_INVALID, "weft_adaptive_gate: bad mode");
    if ((mode == WEFT_ADAPT_COUNT || mode == WEFT_ADAPT_KEEP) && !a->d_win_acc)
        return set_err(WEFT_E_INVALID, "weft_adaptive_gate: accepted flags needed");
    if (a->n_tiles != st->n_tiles || a->n_tiles < 0) return set_err(WEFT_E_INVALID, "weft_adaptive_gate: tile counts differ");
    WeftAdaptGateArgs g{*st, *a, mode, 0};
    return launch(g_mod.f_adapt_gate, 1, 1024, 0, stream, &g);
}

int weft_adaptive_append(const WeftFitState* st, const WeftAdaptive* a, WeftPrim* d_pools, int pool_cap,
                         const WeftPrim* d_new, const int32_t* d_new_count, int max_new, const int32_t* d_tiles,
                         int n, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !a || !a->d_active || !a->d_tile_ncand || !a->d_counts || n < 0 || pool_cap < 1 || max_new < 1)
        return set_err(WEFT_E_INVALID, "weft_adaptive_append: bad arguments");
    if (n == 0) return WEFT_OK;
    if (!d_pools || !d_new || !d_new_count || !d_tiles) return set_err(WEFT_E_INVALID, "weft_adaptive_append: null pointer");
    WeftAdaptAppendArgs g{*st, *a, d_pools, d_new, d_new_count, d_tiles, pool_cap, max_new, n, 0};
    return launch(g_mod.f_adapt_append, (unsigned)n, 256, 0, stream, &g);
}

int weft_fit_step_static(const WeftFitState* st, const WeftWork* work, const WeftCands* cands,
                         const int32_t* d_tile_work0, const int32_t* d_tile_nwork,
                         const int32_t* d_tile_wstride, int32_t* d_win_idx, double* d_win_obj,
                         int32_t* d_win_acc, WeftCode* d_win_code, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !work || !cands) return set_err(WEFT_E_INVALID, "weft_fit_step_static: null argument");
    if (st->n_tiles < 0 || work->n_work < 0 || cands->n_cand < 1)
        return set_err(WEFT_E_INVALID, "weft_fit_step_static: bad sizes");
    if (work->n_split != 0)
        return set_err(WEFT_E_INVALID, "weft_fit_step_static: work lists with split items are not supported");
    // Scratch layout: [class order, explicit candidates only][partial objectives][partial indices].
    const bool explicit_cands = cands->d_prims != nullptr;
    const uint64_t order_bytes =
        explicit_cands ? ((4ull * (uint64_t)st->n_tiles * (uint64_t)cands->n_cand + 255ull) & ~255ull) : 0ull;
    const uint64_t obj_bytes = (8ull * (uint64_t)work->n_work + 255ull) & ~255ull;
    const uint64_t need = order_bytes + obj_bytes + 4ull * (uint64_t)work->n_work;
    if (need > WEFT_STATIC_SCRATCH_BYTES)
        return set_err(WEFT_E_TOO_LARGE, "weft_fit_step_static: " + std::to_string(need) +
                                             " bytes of scratch needed, the static array holds " +
                                             std::to_string(WEFT_STATIC_SCRATCH_BYTES));
    unsigned char* base = g_mod.scratch;
    int32_t* order = reinterpret_cast<int32_t*>(base);
    double* part_obj = reinterpret_cast<double*>(base + order_bytes);
    int32_t* part_idx = reinterpret_cast<int32_t*>(base + order_bytes + obj_bytes);
    if (explicit_cands && (rc = weft_fit_order(st, cands, order, stream))) return rc;
    if ((rc = weft_fit_score(st, work, cands, order, nullptr, nullptr, part_obj, part_idx, stream))) return rc;
    return weft_fit_select(st, cands, d_tile_work0, d_tile_nwork, d_tile_wstride, part_obj, part_idx, d_win_idx,
                           d_win_obj, d_win_acc, d_win_code, stream);
}

int weft_draw_codes(uint64_t seed, uint32_t step, const int32_t* d_tile_ids,
                    const int32_t* d_tile_ts, int n_tiles, int n_cand, WeftCode* d_codes,
                    WeftPrim* d_prims, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n_tiles < 0 || n_cand < 0 || !d_tile_ids || !d_tile_ts)
        return set_err(WEFT_E_INVALID, "weft_draw_codes: bad arguments");
    const long long total = (long long)n_tiles * n_cand;
    if (total == 0) return WEFT_OK;
    if (total > (1ll << 31) * 255ll) return set_err(WEFT_E_TOO_LARGE, "weft_draw_codes: too many");
    WeftDrawArgs a{d_tile_ids, d_tile_ts, d_codes, d_prims, seed, step, n_tiles, n_cand, 0};
    return launch(g_mod.f_draw, (unsigned)((total + 255) / 256), 256, 0, stream, &a);
}

int weft_batch_fit(const WeftFitState* st, const WeftCands* cands, const WeftBatch* b, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !cands || !b || !b->d_tile_order || !b->d_queue || !b->d_win_idx || !b->d_win_obj ||
        !b->d_win_acc)
        return set_err(WEFT_E_INVALID, "weft_batch_fit: null argument");
    if (cands->d_prims) return set_err(WEFT_E_INVALID, "weft_batch_fit: drawn candidates only");
    if (cands->n_cand < 1 || cands->n_cand > 8192)
        return set_err(WEFT_E_INVALID, "weft_batch_fit: n_cand must be in [1, 8192]");
    if (b->block < 64 || b->block > 256 || b->block % 64)
        return set_err(WEFT_E_INVALID, "weft_batch_fit: block must be 64, 128, 192 or 256");
    if (b->grid < 1 || b->n_queue < 0 || b->n_steps < 0 || b->n_images < 0)
        return set_err(WEFT_E_INVALID, "weft_batch_fit: bad sizes");
    if (b->n_images > 0 && (!b->d_tile_image || !b->d_image_tiles || !b->d_image_left || !b->d_image_done))
        return set_err(WEFT_E_INVALID, "weft_batch_fit: image counters missing");
    WeftBatchArgs a{*st, *cands, b->d_tile_order, b->d_tile_image, b->d_image_tiles, b->d_image_left,
                    reinterpret_cast<unsigned long long*>(b->d_image_done),
                    reinterpret_cast<unsigned long long*>(b->d_t_start), b->d_queue, b->d_win_idx,
                    b->d_win_obj, b->d_win_acc, b->d_win_code, b->n_queue,
                    b->n_images > 0 ? b->n_images : 0, b->step0, b->n_steps};
    if (b->n_images == 0) {
        a.image_left = nullptr;
    }
    const int reset_n = b->n_images > 1 ? b->n_images : 1;
    rc = launch(g_mod.f_batch_reset, (unsigned)((reset_n + 255) / 256), 256, 0, stream, &a);
    if (rc) return rc;
    const unsigned lds = (unsigned)(36 * 512);
    return launch(g_mod.f_batch, (unsigned)b->grid, (unsigned)b->block, lds, stream, &a);
}

static bool pow2_le64(int k) { return k >= 1 && k <= 64 && (k & (k - 1)) == 0; }

// Candidates per refinement workgroup: a power of two <= 64 dividing k (one tile), or a
// multiple of k by at most WEFT_WORK_MAX_TILES (that many tiles).
static bool refine_cpw_ok(int cpw, int k) {
    return pow2_le64(cpw) && k >= 1 && (k % cpw == 0 || (cpw % k == 0 && cpw / k <= WEFT_WORK_MAX_TILES));
}

int weft_soft_eval(const WeftFitState* st, const WeftSoftBatch* b, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !b || !b->d_group_tile || !b->d_kinds || !b->d_params || !b->d_out_sse)
        return set_err(WEFT_E_INVALID, "weft_soft_eval: null argument");
    if (!pow2_le64(b->k)) return set_err(WEFT_E_INVALID, "weft_soft_eval: k must be a power of two <= 64");
    if (b->n_groups < 0) return set_err(WEFT_E_INVALID, "weft_soft_eval: n_groups < 0");
    if (!(b->w > 0.0f)) return set_err(WEFT_E_INVALID, "weft_soft_eval: edge width must be > 0");
    const int waves = b->waves > 1 ? b->waves : 1;
    if (b->waves < 0 || (waves > 1 && (!pow2_le64(waves) || waves > 8 || b->k != 1)))
        return set_err(WEFT_E_INVALID, "weft_soft_eval: waves must be 0, 1, 2, 4 or 8, and k 1 when above 1");
    if (b->n_groups == 0) return WEFT_OK;
    WeftSoftArgs a{*st, *b};
    return launch(g_mod.f_soft, (unsigned)b->n_groups, (unsigned)(64 * waves), 0, stream, &a);
}

int weft_topk(const WeftFitState* st, const float* d_table, int n, int k, int32_t* d_out_idx, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !d_table || !d_out_idx) return set_err(WEFT_E_INVALID, "weft_topk: null argument");
    if (n < 1 || n > 8192) return set_err(WEFT_E_INVALID, "weft_topk: n must be in [1, 8192]");
    if (k < 1 || k > n || k > 256) return set_err(WEFT_E_INVALID, "weft_topk: k must be in [1, min(n, 256)]");
    int np2 = 2;
    while (np2 < n) np2 <<= 1;
    if (st->n_tiles == 0) return WEFT_OK;
    WeftTopkArgs a{*st, d_table, d_out_idx, n, k, np2, 0};
    return launch(g_mod.f_topk, (unsigned)st->n_tiles, 256, (unsigned)(8 * np2), stream, &a);
}

int weft_refine(const WeftFitState* st, const WeftCands* cands, const int32_t* d_topk, const WeftRefineOpts* o,
                WeftCode* d_out_codes, WeftPrim* d_out_prims, float* d_out_params, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !cands || !d_topk || !o || !d_out_prims) return set_err(WEFT_E_INVALID, "weft_refine: null argument");
    if (cands->d_prims) return set_err(WEFT_E_INVALID, "weft_refine: drawn candidates only");
    if (!pow2_le64(o->k) || o->k > cands->n_cand)
        return set_err(WEFT_E_INVALID, "weft_refine: k must be a power of two <= min(64, n_cand)");
    if (o->n_iter < 0 || (o->n_iter > 0 && !o->d_sched)) return set_err(WEFT_E_INVALID, "weft_refine: bad schedule");
    int waves = 1;
    for (int c = 0; c < 4; ++c) {
        const int cpw = o->cpw_by_size[c];
        if (cpw < 0) {
            if (!(cpw == -2 || cpw == -4 || cpw == -8) || (waves > 1 && waves != -cpw))
                return set_err(WEFT_E_INVALID, "weft_refine: a negative cpw (waves per candidate) must be -2, -4 or "
                                               "-8, the same for every size");
            waves = -cpw;
        } else if (!refine_cpw_ok(cpw, o->k)) {
            return set_err(WEFT_E_INVALID, "weft_refine: cpw must be powers of two that divide k or are "
                                           "multiples of it by at most WEFT_WORK_MAX_TILES");
        }
    }
    if (o->n_work < 0 || (o->n_work > 0 && (!o->d_work_tile || !o->d_work_j0)))
        return set_err(WEFT_E_INVALID, "weft_refine: bad work list");
    if (o->n_work == 0) return WEFT_OK;
    WeftRefineArgs a{*st, *cands, *o, d_topk, d_out_codes, d_out_prims, d_out_params};
    return launch(waves > 1 ? g_mod.f_refine_mw : g_mod.f_refine, (unsigned)o->n_work, (unsigned)(64 * waves), 0,
                  stream, &a);
}

int weft_fit_score_split(const WeftFitState* st, const WeftCands* cands, const WeftSplitWork* w,
                         float* d_out_sse, float* d_out_mse, double* d_part_obj, int32_t* d_part_idx,
                         void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !cands || !w || !d_part_obj || !d_part_idx)
        return set_err(WEFT_E_INVALID, "weft_fit_score_split: null argument");
    if (cands->n_cand < 1) return set_err(WEFT_E_INVALID, "weft_fit_score_split: n_cand < 1");
    const int max_lanes[4] = {2, 8, 32, 32};
    for (int c = 0; c < 4; ++c) {
        const int cpw = w->cpw_by_size[c], tpw = w->tpw_by_size[c];
        if (!pow2_le64(cpw) || !pow2_le64(tpw) || tpw > WEFT_WORK_MAX_TILES || cpw * tpw > 64 ||
            64 / (cpw * tpw) > max_lanes[c])
            return set_err(WEFT_E_INVALID, "weft_fit_score_split: lanes per candidate, 64 / (cpw * tpw), must be "
                                           "a power of two <= 2, 8, 32, 32 for 8, 16, 32, 64-px tiles");
    }
    if (w->n_work < 0 || (w->n_work > 0 && (!w->d_work_tile || !w->d_work_c0)))
        return set_err(WEFT_E_INVALID, "weft_fit_score_split: bad work list");
    if (w->n_work == 0) return WEFT_OK;
    WeftSplitScoreArgs a{*st, *cands, w->d_work_tile, w->d_work_c0,
                         {w->cpw_by_size[0], w->cpw_by_size[1], w->cpw_by_size[2], w->cpw_by_size[3]},
                         {w->tpw_by_size[0], w->tpw_by_size[1], w->tpw_by_size[2], w->tpw_by_size[3]},
                         d_out_sse, d_out_mse, d_part_obj, d_part_idx};
    return launch(g_mod.f_split, (unsigned)w->n_work, 64, 0, stream, &a);
}

int64_t weft_rans_stage_bytes(int64_t raw_length) {
    // the encoder writes a rANS payload only while it stays shorter than the raw form,
    // so a slot never holds more than the raw form's 5 + n bytes
    const int64_t n = raw_length > 0 ? raw_length : 0;
    return (5 + n + 15) / 16 * 16;
}

int weft_rans_encode_chunks(const uint8_t* d_raw, WeftChunkDesc* d_descs, int n_chunks, uint8_t* d_stage,
                            const int64_t* d_stage_off, uint8_t* d_coded, int64_t* d_coded_total, int32_t* d_status,
                            void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n_chunks < 0) return set_err(WEFT_E_INVALID, "weft_rans_encode_chunks: n_chunks < 0");
    if (n_chunks == 0) return WEFT_OK;
    if (!d_raw || !d_descs || !d_stage || !d_stage_off || !d_coded || !d_coded_total || !d_status)
        return set_err(WEFT_E_INVALID, "weft_rans_encode_chunks: null argument");
    WeftRansEncodeArgs a{d_raw, d_descs, d_stage, d_stage_off, d_coded, d_coded_total, d_status, n_chunks, 0};
    rc = launch(g_mod.f_rans_enc, (unsigned)n_chunks, 256, 0, stream, &a);
    if (rc) return rc;
    rc = launch(g_mod.f_rans_scan, 1, 1024, 0, stream, &a);
    if (rc) return rc;
    return launch(g_mod.f_rans_pack, (unsigned)n_chunks, 256, 0, stream, &a);
}

int weft_rans_decode_chunks(const uint8_t* d_coded, const WeftChunkDesc* d_descs, int n_chunks, uint8_t* d_out_raw,
                            int32_t* d_status, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n_chunks < 0) return set_err(WEFT_E_INVALID, "weft_rans_decode_chunks: n_chunks < 0");
    if (n_chunks == 0) return WEFT_OK;
    if (!d_coded || !d_descs || !d_out_raw || !d_status)
        return set_err(WEFT_E_INVALID, "weft_rans_decode_chunks: null argument");
    WeftRansDecodeArgs a{d_coded, d_descs, d_out_raw, d_status, n_chunks, 0};
    return launch(g_mod.f_rans_dec, (unsigned)n_chunks, 64, 0, stream, &a);
}

int weft_batch_refine_fit(const WeftFitState* st, const WeftCands* cands, const WeftBatch* b,
                          const WeftRefineOpts* o, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!st || !cands || !b || !o || !b->d_tile_order || !b->d_queue || !b->d_win_idx || !b->d_win_obj ||
        !b->d_win_acc)
        return set_err(WEFT_E_INVALID, "weft_batch_refine_fit: null argument");
    if (cands->d_prims) return set_err(WEFT_E_INVALID, "weft_batch_refine_fit: drawn candidates only");
    if (cands->n_cand < 1 || cands->n_cand > 8192)
        return set_err(WEFT_E_INVALID, "weft_batch_refine_fit: n_cand must be in [1, 8192]");
    if (b->block < 64 || b->block > 256 || b->block % 64)
        return set_err(WEFT_E_INVALID, "weft_batch_refine_fit: block must be 64, 128, 192 or 256");
    if (b->grid < 1 || b->n_queue < 0 || b->n_steps < 0 || b->n_images < 0)
        return set_err(WEFT_E_INVALID, "weft_batch_refine_fit: bad sizes");
    if (b->n_images > 0 && (!b->d_tile_image || !b->d_image_tiles || !b->d_image_left || !b->d_image_done))
        return set_err(WEFT_E_INVALID, "weft_batch_refine_fit: image counters missing");
    if (!pow2_le64(o->k) || o->k > cands->n_cand)
        return set_err(WEFT_E_INVALID, "weft_batch_refine_fit: k must be a power of two <= min(64, n_cand)");
    if (o->n_iter < 0 || (o->n_iter > 0 && !o->d_sched))
        return set_err(WEFT_E_INVALID, "weft_batch_refine_fit: bad schedule");
    for (int c = 0; c < 4; ++c) {
        const int cpw = o->cpw_by_size[c];
        if (cpw < 0 ? !((cpw == -2 || cpw == -4) && 64 * -cpw <= b->block) : !refine_cpw_ok(cpw, o->k))
            return set_err(WEFT_E_INVALID, "weft_batch_refine_fit: cpw must be powers of two that divide k or are "
                                           "multiples of it by at most WEFT_WORK_MAX_TILES, or -2 / -4 waves per "
                                           "candidate within the workgroup");
    }
    if (b->d_group_first && (b->n_groups < 0 || b->group_max < 1 || b->group_max > 64 / o->k))
        return set_err(WEFT_E_INVALID, "weft_batch_refine_fit: groups need n_groups >= 0 and group_max in 1 .. 64 / k");
    WeftBatchRefineArgs a{};
    a.group_first = b->d_group_first;
    a.n_groups = b->d_group_first ? b->n_groups : 0;
    a.group_max = b->d_group_first ? b->group_max : 1;
    a.b = WeftBatchArgs{*st, *cands, b->d_tile_order, b->d_tile_image, b->d_image_tiles, b->d_image_left,
                        reinterpret_cast<unsigned long long*>(b->d_image_done),
                        reinterpret_cast<unsigned long long*>(b->d_t_start), b->d_queue, b->d_win_idx,
                        b->d_win_obj, b->d_win_acc, b->d_win_code, b->n_queue, b->n_images > 0 ? b->n_images : 0,
                        b->step0, b->n_steps};
    if (b->n_images == 0) a.b.image_left = nullptr;
    a.o = *o;
    // exact re-score: lanes per candidate, a power of two within the pairwise tree's leaf
    // count and small enough for all k + 1 candidates to fit the workgroup in one pass
    const int leaf_limit[4] = {2, 8, 32, 32};
    const int k1 = o->k + 1;
    for (int c = 0; c < 4; ++c) {
        int ls = 1;
        while (ls * 2 <= leaf_limit[c] && ls * 2 * k1 <= b->block) ls *= 2;
        a.split_lanes[c] = ls;
    }
    int np2 = 2;
    while (np2 < cands->n_cand) np2 <<= 1;
    a.n_pow2 = np2;
    const int reset_n = b->n_images > 1 ? b->n_images : 1;
    rc = launch(g_mod.f_batch_reset, (unsigned)((reset_n + 255) / 256), 256, 0, stream, &a.b);
    if (rc) return rc;
    const unsigned lds = (unsigned)(36 * 512 + 8 * np2);
    return launch(g_mod.f_batch_refine, (unsigned)b->grid, (unsigned)b->block, lds, stream, &a);
}

int weft_wallclock_rate_khz(int* khz) {
    int rc = require_init();
    if (rc) return rc;
    if (!khz) return set_err(WEFT_E_INVALID, "weft_wallclock_rate_khz: null");
    WEFT_HIP(hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, g_mod.device), "wall clock rate");
    return WEFT_OK;
}

int weft_render_tiles(const WeftTileDecodeDesc* d_descs, int n_tiles, const WeftPrim* d_prims,
                      float* d_out, int width, int height, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n_tiles < 0 || width < 0 || height < 0) return set_err(WEFT_E_INVALID, "weft_render_tiles: bad sizes");
    if (n_tiles == 0) return WEFT_OK;
    if (!d_descs || !d_prims || !d_out) return set_err(WEFT_E_INVALID, "weft_render_tiles: null pointer");
    WeftRenderArgs a{d_descs, d_prims, d_out, nullptr, n_tiles, width, height, 0, 0, 0};
    return launch(g_mod.f_render, (unsigned)n_tiles, 256, 0, stream, &a);
}

int weft_render_scene(const WeftTileDecodeDesc* d_descs, int n_tiles, const WeftPrim* d_prims,
                      const float* d_res1, int n_res1, int res1_grid, float* d_out, int width, int height,
                      void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n_tiles < 0 || width < 0 || height < 0 || n_res1 < 0 || n_res1 > n_tiles)
        return set_err(WEFT_E_INVALID, "weft_render_scene: bad sizes");
    if (n_res1 > 0 && (!d_res1 || res1_grid < 1 || res1_grid > 64))
        return set_err(WEFT_E_INVALID, "weft_render_scene: RES1 grids need a pointer and a grid size in 1..64");
    if (n_tiles == 0) return WEFT_OK;
    if (!d_descs || !d_prims || !d_out) return set_err(WEFT_E_INVALID, "weft_render_scene: null pointer");
    WeftRenderArgs a{d_descs, d_prims, d_out, n_res1 > 0 ? d_res1 : nullptr, n_tiles, width, height, 1,
                     n_res1, res1_grid};
    return launch(g_mod.f_render, (unsigned)n_tiles, 256, 0, stream, &a);
}

int weft_template_masks(int tile_size, uint32_t* d_mask, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (tile_size < 1 || tile_size > 64)
        return set_err(WEFT_E_UNSUPPORTED_TILE, "weft_template_masks: tile size must be in 1..64");
    if (!d_mask) return set_err(WEFT_E_INVALID, "weft_template_masks: null pointer");
    WeftTemplateMaskArgs a{d_mask, tile_size, 0};
    return launch(g_mod.f_tmpl_masks, 1, 256, 0, stream, &a);
}

int weft_template_fit(const float* d_image, int width, int height, int tile_size, const uint32_t* d_mask,
                      int32_t* d_model, float* d_c0, float* d_c1, float* d_mse, float* d_c0_all, float* d_c1_all,
                      void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (tile_size < 1 || tile_size > 64)
        return set_err(WEFT_E_UNSUPPORTED_TILE, "weft_template_fit: tile size must be in 1..64");
    if (width <= 0 || height <= 0) return set_err(WEFT_E_INVALID, "weft_template_fit: image dimensions must be > 0");
    if ((int64_t)width * height * 3 >= ((int64_t)1 << 62))
        return set_err(WEFT_E_TOO_LARGE, "weft_template_fit: image too large");
    const int64_t cols = (width + tile_size - 1) / tile_size;
    const int64_t rows = (height + tile_size - 1) / tile_size;
    if (cols * rows > 0x7fffffff) return set_err(WEFT_E_TOO_LARGE, "weft_template_fit: too many tiles");
    if (!d_image || !d_mask || !d_model || !d_c0 || !d_c1 || !d_mse || !d_c0_all || !d_c1_all)
        return set_err(WEFT_E_INVALID, "weft_template_fit: null pointer");
    WeftTemplateFitArgs a{d_image, d_mask, d_model, d_c0, d_c1, d_mse, d_c0_all, d_c1_all,
                          width, height, tile_size, (int32_t)cols, (int32_t)rows, 0};
    // 14 tiles of 18 templates per 256-thread workgroup, static LDS only
    const int64_t groups = (cols * rows + 13) / 14;
    return launch(g_mod.f_tmpl_fit, (unsigned)groups, 256, 0, stream, &a);
}

// ---- selection primitives: workspace layout shared by the size query and the launchers
namespace {
constexpr int64_t SEL_TILE_ELEMS = 2048;
constexpr int64_t SEL_RS_TILE = 2048;  // keys per workgroup of a radix pass
constexpr int64_t SEL_MAX_SORT = (int64_t)1 << 30;

int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }
int64_t sel_blocks(int64_t n) { return (n + SEL_TILE_ELEMS - 1) / SEL_TILE_ELEMS; }
int64_t sel_argmin_grid(int64_t n) {
    const int64_t g = sel_blocks(n);
    return g < 1024 ? (g < 1 ? 1 : g) : 1024;
}
int64_t sel_rs_blocks(int64_t n) { return (n + SEL_RS_TILE - 1) / SEL_RS_TILE; }

struct SelLayout {  // byte offsets into the workspace
    int64_t f64 = 0, i64 = 0, k0 = 0, k1 = 0, v0 = 0, v1 = 0, vb0 = 0, vb1 = 0, table = 0, prefix = 0, flags = 0,
            pos = 0, total = 0;
};

SelLayout sel_layout(int op, int64_t n) {
    SelLayout L;
    int64_t off = 0;
    if (op == WEFT_SELECT_ARGMIN) {
        const int64_t g = sel_argmin_grid(n);
        L.f64 = off;
        off += align256(8 * g);
        L.i64 = off;
        off += align256(8 * g);
    } else if (op == WEFT_SELECT_SCAN || op == WEFT_SELECT_COMPACT) {
        L.i64 = off;
        off += align256(8 * sel_blocks(n));
    } else {
        const int64_t nt = 256 * sel_rs_blocks(n);  // digit table entries
        L.k0 = off;
        off += align256(8 * n);
        L.k1 = off;
        off += align256(8 * n);
        L.v0 = off;
        off += align256(4 * n);
        L.v1 = off;
        off += align256(4 * n);
        L.vb0 = off;
        off += align256(4 * n);
        L.vb1 = off;
        off += align256(4 * n);
        L.table = off;
        off += align256(8 * nt);
        L.prefix = off;
        off += align256(8 * nt);
        L.i64 = off;
        off += align256(8 * (sel_blocks(nt) > sel_blocks(n) ? sel_blocks(nt) : sel_blocks(n)));
        L.flags = off;
        off += align256(n);
        L.pos = off;
        off += align256(8 * n);
    }
    L.total = off;
    return L;
}

// Three-pass scan of A.i64_in (or compaction of A.u8_in) over A.n elements.
int sel_scan_passes(WeftSelectArgs& a, void* stream) {
    a.n_blocks = sel_blocks(a.n);
    int rc = launch(g_mod.f_sel_tile_sums, (unsigned)a.n_blocks, 256, 0, stream, &a);
    if (rc) return rc;
    rc = launch(g_mod.f_sel_scan_tiles, 1, 256, 0, stream, &a);
    if (rc) return rc;
    return launch(g_mod.f_sel_scan_apply, (unsigned)a.n_blocks, 256, 0, stream, &a);
}

// One stable radix pass over bits [shift, shift + 8) of the keys: digit counts per
// workgroup, their digit-major exclusive scan, then the scatter.
int sel_radix_pass(const WeftSelectArgs& base, const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                   uint32_t* vals_out, int shift, char* w, const SelLayout& L, void* stream) {
    WeftSelectArgs a = base;
    a.rs_keys_in = keys_in;
    a.rs_vals_in = vals_in;
    a.rs_keys_out = keys_out;
    a.rs_vals_out = vals_out;
    a.rs_shift = shift;
    a.n_blocks = sel_rs_blocks(a.n);
    a.rs_table = (int64_t*)(w + L.table);
    a.rs_prefix = (const int64_t*)(w + L.prefix);
    int rc = launch(g_mod.f_sel_rs_hist, (unsigned)a.n_blocks, 256, 0, stream, &a);
    if (rc) return rc;
    WeftSelectArgs c{};
    c.i64_in = a.rs_table;
    c.n = 256 * a.n_blocks;
    c.out_i64 = (int64_t*)(w + L.prefix);
    c.work_i64 = (int64_t*)(w + L.i64);
    if ((rc = sel_scan_passes(c, stream))) return rc;
    return launch(g_mod.f_sel_rs_scatter, (unsigned)a.n_blocks, 256, 0, stream, &a);
}

int sel_check_work(int op, int64_t n, void* d_work, int64_t work_bytes, const char* who) {
    if (!d_work) return set_err(WEFT_E_INVALID, std::string(who) + ": null workspace");
    if (work_bytes < sel_layout(op, n).total)
        return set_err(WEFT_E_INVALID, std::string(who) + ": workspace smaller than weft_select_work_bytes");
    return WEFT_OK;
}
}  // namespace

int64_t weft_select_work_bytes(int op, int64_t n) {
    if (op < WEFT_SELECT_ARGMIN || op > WEFT_SELECT_TOPK || n < 0) return -1;
    return sel_layout(op, n).total;
}

int weft_argmin_f64(const double* d_values, int64_t n, int64_t* d_out_index, void* d_work, int64_t work_bytes,
                    void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n <= 0) return set_err(WEFT_E_INVALID, "weft_argmin_f64: n must be > 0");
    if (!d_values || !d_out_index) return set_err(WEFT_E_INVALID, "weft_argmin_f64: null pointer");
    if ((rc = sel_check_work(WEFT_SELECT_ARGMIN, n, d_work, work_bytes, "weft_argmin_f64"))) return rc;
    const SelLayout L = sel_layout(WEFT_SELECT_ARGMIN, n);
    WeftSelectArgs a{};
    a.f64_in = d_values;
    a.n = n;
    a.work_f64 = (double*)((char*)d_work + L.f64);
    a.work_i64 = (int64_t*)((char*)d_work + L.i64);
    a.out_i64 = d_out_index;
    const int64_t g = sel_argmin_grid(n);
    a.n_blocks = g;
    rc = launch(g_mod.f_sel_argmin_partial, (unsigned)g, 256, 0, stream, &a);
    if (rc) return rc;
    return launch(g_mod.f_sel_argmin_final, 1, 256, 0, stream, &a);
}

int weft_exclusive_scan_i64(const int64_t* d_in, int64_t n, int64_t* d_out, int64_t* d_total, void* d_work,
                            int64_t work_bytes, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n < 0) return set_err(WEFT_E_INVALID, "weft_exclusive_scan_i64: n < 0");
    if (n == 0) return WEFT_OK;
    if (!d_in || !d_out) return set_err(WEFT_E_INVALID, "weft_exclusive_scan_i64: null pointer");
    if ((rc = sel_check_work(WEFT_SELECT_SCAN, n, d_work, work_bytes, "weft_exclusive_scan_i64"))) return rc;
    WeftSelectArgs a{};
    a.i64_in = d_in;
    a.n = n;
    a.out_i64 = d_out;
    a.out_total = d_total;
    a.work_i64 = (int64_t*)((char*)d_work + sel_layout(WEFT_SELECT_SCAN, n).i64);
    return sel_scan_passes(a, stream);
}

int weft_compact_flags(const uint8_t* d_flags, int64_t n, int64_t* d_out_indices, int64_t* d_out_count,
                       void* d_work, int64_t work_bytes, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n < 0) return set_err(WEFT_E_INVALID, "weft_compact_flags: n < 0");
    if (n == 0) return WEFT_OK;
    if (!d_flags || !d_out_indices || !d_out_count) return set_err(WEFT_E_INVALID, "weft_compact_flags: null pointer");
    if ((rc = sel_check_work(WEFT_SELECT_COMPACT, n, d_work, work_bytes, "weft_compact_flags"))) return rc;
    WeftSelectArgs a{};
    a.u8_in = d_flags;
    a.n = n;
    a.out_i64 = d_out_indices;
    a.out_total = d_out_count;
    a.work_i64 = (int64_t*)((char*)d_work + sel_layout(WEFT_SELECT_COMPACT, n).i64);
    return sel_scan_passes(a, stream);
}

int weft_segmented_topk(const double* d_scores, const int64_t* d_segments, int64_t n, int64_t k, int largest,
                        int64_t* d_out_indices, int64_t* d_out_count, void* d_work, int64_t work_bytes,
                        void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n < 0 || k < 0) return set_err(WEFT_E_INVALID, "weft_segmented_topk: n and k must be >= 0");
    if (n > SEL_MAX_SORT) return set_err(WEFT_E_TOO_LARGE, "weft_segmented_topk: more than 2^30 scores");
    if (!d_out_count) return set_err(WEFT_E_INVALID, "weft_segmented_topk: null count pointer");
    if (n == 0 || k == 0) {  // count = 0, written by a kernel (the scan of no tiles)
        WeftSelectArgs z{};
        z.out_total = d_out_count;
        return launch(g_mod.f_sel_scan_tiles, 1, 256, 0, stream, &z);
    }
    if (!d_scores || !d_segments || !d_out_indices) return set_err(WEFT_E_INVALID, "weft_segmented_topk: null pointer");
    if ((rc = sel_check_work(WEFT_SELECT_TOPK, n, d_work, work_bytes, "weft_segmented_topk"))) return rc;
    const SelLayout L = sel_layout(WEFT_SELECT_TOPK, n);
    char* w = (char*)d_work;
    uint64_t* k0 = (uint64_t*)(w + L.k0);
    uint64_t* k1 = (uint64_t*)(w + L.k1);
    uint32_t* v0 = (uint32_t*)(w + L.v0);
    uint32_t* v1 = (uint32_t*)(w + L.v1);
    uint32_t* vb0 = (uint32_t*)(w + L.vb0);
    uint32_t* vb1 = (uint32_t*)(w + L.vb1);
    WeftSelectArgs a{};
    a.f64_in = d_scores;
    a.i64_in = d_segments;
    a.n = n;
    a.k = k;
    a.largest = largest ? 1 : 0;
    const unsigned elem_blocks = (unsigned)((n + 255) / 256);
    // A: score order; 8 passes end back in (k0, v0)
    a.rs_keys_out = k0;
    a.rs_vals_out = v0;
    if ((rc = launch(g_mod.f_sel_rs_score_keys, elem_blocks, 256, 0, stream, &a))) return rc;
    for (int pass = 0; pass < 8 && rc == 0; ++pass) {
        const bool even = (pass & 1) == 0;
        rc = sel_radix_pass(a, even ? k0 : k1, even ? v0 : v1, even ? k1 : k0, even ? v1 : v0, 8 * pass, w, L,
                            stream);
    }
    if (rc) return rc;
    // B: segments in A order (values = A position); 8 passes end back in (k1, vb0)
    a.rs_order = v0;
    a.rs_keys_out = k1;
    a.rs_vals_out = vb0;
    if ((rc = launch(g_mod.f_sel_rs_segment_keys, elem_blocks, 256, 0, stream, &a))) return rc;
    for (int pass = 0; pass < 8 && rc == 0; ++pass) {
        const bool even = (pass & 1) == 0;
        rc = sel_radix_pass(a, even ? k1 : k0, even ? vb0 : vb1, even ? k0 : k1, even ? vb1 : vb0, 8 * pass, w, L,
                            stream);
    }
    if (rc) return rc;
    a.rs_keys_in = k1;
    a.rs_vals_in = vb0;
    a.flags = (uint8_t*)(w + L.flags);
    if ((rc = launch(g_mod.f_sel_rs_keep, elem_blocks, 256, 0, stream, &a))) return rc;
    // kept A positions, in order, and their count
    WeftSelectArgs c{};
    c.u8_in = a.flags;
    c.n = n;
    c.out_i64 = (int64_t*)(w + L.pos);
    c.out_total = d_out_count;
    c.work_i64 = (int64_t*)(w + L.i64);
    if ((rc = sel_scan_passes(c, stream))) return rc;
    a.positions = c.out_i64;
    a.count_in = d_out_count;
    a.out_i64 = d_out_indices;
    return launch(g_mod.f_sel_rs_emit, elem_blocks, 256, 0, stream, &a);
}

int weft_gen_candidates(const float* d_pixels, const int64_t* d_tile_pix, const int32_t* d_tiles, int n_tiles,
                        int tile_size, int quality, WeftPrim* d_out, double* d_out_f64, int32_t* d_out_count,
                        int max_cands, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (tile_size != 8 && tile_size != 16 && tile_size != 32 && tile_size != 64)
        return set_err(WEFT_E_UNSUPPORTED_TILE, "weft_gen_candidates: tile size must be 8, 16, 32 or 64");
    if (n_tiles < 0) return set_err(WEFT_E_INVALID, "weft_gen_candidates: n_tiles < 0");
    if (max_cands < WEFT_GEN_MAX_CANDS)
        return set_err(WEFT_E_INVALID, "weft_gen_candidates: max_cands below WEFT_GEN_MAX_CANDS");
    if (n_tiles == 0) return WEFT_OK;
    if (!d_pixels || !d_tile_pix || !d_tiles || !d_out || !d_out_count)
        return set_err(WEFT_E_INVALID, "weft_gen_candidates: null pointer");
    WeftCandGenArgs a{};
    a.pixels = d_pixels;
    a.tile_pix = d_tile_pix;
    a.tiles = d_tiles;
    a.out = d_out;
    a.out_f64 = d_out_f64;
    a.out_count = d_out_count;
    a.n_tiles = n_tiles;
    a.tile_size = tile_size;
    a.quality = quality;
    a.max_cands = max_cands;
    // np.linspace(0, pi, n, endpoint=False)[i] = i * (pi / n); the host's libm cos / sin, as
    // the reference's math.cos / math.sin
    const int n_ang = quality >= 80 ? 16 : (quality >= 50 ? 12 : 8);
    const double pi = 3.141592653589793;
    for (int i = 0; i < n_ang; ++i) {
        const double ang = (double)i * (pi / (double)n_ang);
        a.cos_ang[i] = std::cos(ang);
        a.sin_ang[i] = std::sin(ang);
    }
    for (int i = 0; i < 6; ++i) {
        const double ang = (double)i * (pi / 6.0);
        a.cos6[i] = std::cos(ang);
        a.sin6[i] = std::sin(ang);
    }
    const unsigned lds = (unsigned)(32 * tile_size * tile_size);
    return launch(g_mod.f_gen_cands, (unsigned)n_tiles, 256, lds, stream, &a);
}

int weft_gen_residual_candidates(const float* d_pixels, const float* d_pred, int clip_pred, const int64_t* d_tile_pix,
                                 const int32_t* d_tiles, int n_tiles, int tile_size, WeftPrim* d_out,
                                 double* d_out_f64, int32_t* d_out_count, int max_cands, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (tile_size != 8 && tile_size != 16 && tile_size != 32 && tile_size != 64)
        return set_err(WEFT_E_UNSUPPORTED_TILE, "weft_gen_residual_candidates: tile size must be 8, 16, 32 or 64");
    if (n_tiles < 0) return set_err(WEFT_E_INVALID, "weft_gen_residual_candidates: n_tiles < 0");
    if (max_cands < WEFT_GEN_RES_MAX_CANDS)
        return set_err(WEFT_E_INVALID, "weft_gen_residual_candidates: max_cands below WEFT_GEN_RES_MAX_CANDS");
    if (n_tiles == 0) return WEFT_OK;
    if (!d_pixels || !d_pred || !d_tile_pix || !d_tiles || !d_out || !d_out_count)
        return set_err(WEFT_E_INVALID, "weft_gen_residual_candidates: null pointer");
    WeftResGenArgs a{d_pixels, d_pred, d_tile_pix, d_tiles, d_out, d_out_f64, d_out_count,
                     n_tiles, tile_size, clip_pred ? 1 : 0, max_cands};
    const unsigned lds = (unsigned)(28 * tile_size * tile_size);
    return launch(g_mod.f_gen_res, (unsigned)n_tiles, 256, lds, stream, &a);
}

int weft_gen_edge_candidates(const float* d_pixels, const int64_t* d_tile_pix, const int32_t* d_tiles, int n_tiles,
                             int tile_size, int quality, WeftPrim* d_out, double* d_out_f64, int32_t* d_out_count,
                             int max_cands, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (tile_size != 8 && tile_size != 16 && tile_size != 32 && tile_size != 64)
        return set_err(WEFT_E_UNSUPPORTED_TILE, "weft_gen_edge_candidates: tile size must be 8, 16, 32 or 64");
    if (n_tiles < 0) return set_err(WEFT_E_INVALID, "weft_gen_edge_candidates: n_tiles < 0");
    if (max_cands < WEFT_GEN_EDGE_MAX_CANDS)
        return set_err(WEFT_E_INVALID, "weft_gen_edge_candidates: max_cands below WEFT_GEN_EDGE_MAX_CANDS");
    if (n_tiles == 0) return WEFT_OK;
    if (!d_pixels || !d_tile_pix || !d_tiles || !d_out || !d_out_count)
        return set_err(WEFT_E_INVALID, "weft_gen_edge_candidates: null pointer");
    WeftEdgeGenArgs a{};
    a.pixels = d_pixels;
    a.tile_pix = d_tile_pix;
    a.tiles = d_tiles;
    a.out = d_out;
    a.out_f64 = d_out_f64;
    a.out_count = d_out_count;
    a.n_tiles = n_tiles;
    a.tile_size = tile_size;
    a.quality = quality;
    a.max_cands = max_cands;
    const int n = tile_size * tile_size;
    a.min_small = (int)std::floor((double)(n - 1) * 0.6) + 2;
    // np.linspace(0, pi, 13): i * (pi / 12), the last edge pi itself; the orientations as the
    // reference's Python floats, their cos / sin from the host's libm (math.cos / math.sin)
    const double pi = 3.141592653589793;
    const double step = pi / 12.0;
    for (int i = 0; i < 12; ++i) a.edges[i] = (double)i * step;
    a.edges[12] = pi;
    double theta[28];
    for (int i = 0; i < 12; ++i) theta[i] = (a.edges[i] + a.edges[i + 1]) / 2.0;
    theta[12] = 0.0;
    theta[13] = pi / 4.0;
    theta[14] = pi / 2.0;
    theta[15] = 3.0 * pi / 4.0;
    for (int i = 0; i < 12; ++i) theta[16 + i] = a.edges[i] + pi / 24.0;
    for (int i = 0; i < 28; ++i) {
        a.cos_o[i] = std::cos(theta[i]);
        a.sin_o[i] = std::sin(theta[i]);
    }
    const unsigned lds = (unsigned)(33 * n);
    return launch(g_mod.f_gen_edge, (unsigned)n_tiles, 256, lds, stream, &a);
}

int weft_merge_pools(const WeftPrim* d_a, const double* d_a_f64, const int32_t* d_a_count, int a_cap,
                     const WeftPrim* d_b, const double* d_b_f64, const int32_t* d_b_count, int b_cap, int n_pools,
                     WeftPrim* d_out, double* d_out_f64, int32_t* d_out_count, int out_cap, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (n_pools < 0 || a_cap < 0 || b_cap < 0) return set_err(WEFT_E_INVALID, "weft_merge_pools: negative size");
    if ((int64_t)a_cap + b_cap > WEFT_MERGE_MAX_CANDS)
        return set_err(WEFT_E_INVALID, "weft_merge_pools: a_cap + b_cap above WEFT_MERGE_MAX_CANDS");
    if (out_cap < a_cap + b_cap) return set_err(WEFT_E_INVALID, "weft_merge_pools: out_cap below a_cap + b_cap");
    if (n_pools == 0) return WEFT_OK;
    if (!d_a || !d_a_f64 || !d_a_count || !d_b || !d_b_f64 || !d_b_count || !d_out || !d_out_count)
        return set_err(WEFT_E_INVALID, "weft_merge_pools: null pointer");
    WeftMergePoolsArgs m{d_a, d_a_f64, d_a_count, d_b, d_b_f64, d_b_count, d_out, d_out_f64, d_out_count,
                         a_cap, b_cap, out_cap, n_pools};
    return launch(g_mod.f_merge_pools, (unsigned)n_pools, 256, 0, stream, &m);
}

}  // extern "C"
