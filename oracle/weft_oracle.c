/*
 * weft_oracle.c -- CPU restatement of the Weft tile-fit scoring path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (weft_amd/) links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg do, as the checker and as the timed CPU baseline.
 *
 * What it restates (reference = jarmak-personal/weft, paths relative to the repo):
 *   render.render_tile            src/weft/render.py:127-234
 *     _grid / _scaled_grid        src/weft/render.py:29-78
 *     _point_line_distance        src/weft/render.py:85-95
 *     _eval_curve_distance        src/weft/render.py:98-111
 *     _inside_triangle            src/weft/render.py:114-124
 *   encoder._tile_objective       src/weft/encoder.py:271-285  (np.mean pairwise tree)
 *   gpu_score.batch_mse           src/weft/gpu_score.py:45-57
 *   encoder._estimate_bits + LUT  src/weft/encoder.py:242-268
 *   encoder._greedy_add           src/weft/encoder.py:752-817
 *   cuda_backend.cccl_argmin      src/weft/cuda_backend.py:80-98 (first index on ties)
 *   entropy._build_model          src/weft/entropy.py:30-87   (order-0 frequency model)
 *   entropy._encode_rans / encode_bytes   src/weft/entropy.py:90-122, 190-198
 *   entropy._decode_rans / decode_bytes   src/weft/entropy.py:125-187, 201-217
 *
 * Float semantics.  NumPy 2 (NEP 50) evaluates `float32_array OP python_float` in
 * float32 after rounding the Python float to float32; Python-scalar sub-expressions
 * (triangle `den`, linear/line `vx`, `vy`, `den`, sigma) are float64.  Every float32
 * operation below is written out one rounding at a time and the file is compiled with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Kinds 6 (rotated ellipse), 7 (rotated rectangle) and 8 (triangle given by centre,
 * radii and angle; rendered exactly like kind 4 on its three vertices) have no
 * counterpart in the reference.  Their coverage tests are defined here and in the HIP
 * kernels identically; they are pinned by closed forms in tests/test_oracle.py.
 *
 * Parity status:
 *   kinds 0, 1, 4, 6, 7, 8 render, SSE, MSE, objective, argmin, greedy add:
 *     pinned bit-exact against fixtures written by the reference itself
 *     (tests/golden/) and against NumPy's own reductions.
 *   kinds 2, 3 (Gaussian line / curve): np.exp(float32) is NumPy's SIMD exp; this
 *     file calls libm expf, so those two kinds are pinned to 1e-6 relative only.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t kind;
    float geom[7];
    float color0[3];
    float color1[3];
    float alpha;
} orc_prim; /* 60 bytes: the reference's PrimData / _PRIM_DTYPE, gpu_render.py:241-247 */

/* Bytes per primitive kind, encoder.py:242-256 (unknown kinds 20), plus the 17-byte
 * parametric records this build adds for kinds 6, 7, 8. */
static int orc_prim_bytes(int kind) {
    switch (kind) {
    case 0: return 7;
    case 1: return 18;
    case 2: return 17;
    case 3: return 21;
    case 4: return 19;
    case 6: case 7: case 8: return 17;
    default: return 20;
    }
}

int orc_prim_bytes_c(int kind) { return orc_prim_bytes(kind); }

static inline float f32clip01(float v) {
    /* np.clip(x, 0.0, 1.0) == minimum(maximum(x, 0), 1) */
    float t = v < 0.0f ? 0.0f : v;
    return t > 1.0f ? 1.0f : t;
}

/* float32 alpha as render.py:185 builds it: np.float32(max(0.0, min(1.0, alpha))). */
static inline float alpha32(float alpha) {
    double a = (double)alpha;
    a = a < 1.0 ? a : 1.0;
    a = a > 0.0 ? a : 0.0;
    return (float)a;
}

/* ---------------------------------------------------------------------------
 * Coverage / per-pixel source for one primitive at sample (px, py).
 * Returns alpha (float32) and writes the source colour.
 * ------------------------------------------------------------------------- */

static float tri_mask(const float* g, float px, float py) {
    /* render.py:114-124.  geom holds Python floats equal to float32 values. */
    double x0 = g[0], y0 = g[1], x1 = g[2], y1 = g[3], x2 = g[4], y2 = g[5];
    double den = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2);
    if (fabs(den) < 1e-6) return 0.0f;
    float A = (float)(y1 - y2), B = (float)(x2 - x1);
    float C = (float)(y2 - y0), E = (float)(x0 - x2);
    float D = (float)den;
    float X2 = (float)x2, Y2 = (float)y2;
    float dx = px - X2;
    float dy = py - Y2;
    float ta = A * dx;
    float tb = B * dy;
    float na = ta + tb;
    float a = na / D;
    float tc = C * dx;
    float te = E * dy;
    float nb = tc + te;
    float b = nb / D;
    float om = 1.0f - a;
    float c = om - b;
    return (a >= 0.0f && b >= 0.0f && c >= 0.0f) ? 1.0f : 0.0f;
}

/* Kinds 6 and 7: geom = (cx, cy, rx, ry, cos t, sin t).  The sample is moved into the
 * primitive's frame scaled by the radii:
 *   p = (cos t / rx) * dx + (sin t / rx) * dy
 *   q = (cos t / ry) * dy - (sin t / ry) * dx
 * ellipse: p*p + q*q <= 1 ; rectangle: |p| <= 1 and |q| <= 1. */
static float ell_rect_mask(int kind, const float* g, float px, float py) {
    float rx = g[2], ry = g[3];
    if (!(rx > 0.0f) || !(ry > 0.0f)) return 0.0f;
    float ca = g[4] / rx, sa = g[5] / rx;
    float cb = g[4] / ry, sb = g[5] / ry;
    float dx = px - g[0];
    float dy = py - g[1];
    float p1 = ca * dx;
    float p2 = sa * dy;
    float p = p1 + p2;
    float q1 = cb * dy;
    float q2 = sb * dx;
    float q = q1 - q2;
    if (kind == 6) {
        float pp = p * p;
        float qq = q * q;
        float s = pp + qq;
        return s <= 1.0f ? 1.0f : 0.0f;
    }
    return (fabsf(p) <= 1.0f && fabsf(q) <= 1.0f) ? 1.0f : 0.0f;
}

static float seg_dist_f64params(float px, float py, double x0, double y0, double x1, double y1) {
    /* _point_line_distance with Python-float endpoints (kind 2). */
    double vx = x1 - x0, vy = y1 - y0;
    double den = vx * vx + vy * vy + 1e-8;
    float VX = (float)vx, VY = (float)vy, DN = (float)den;
    float X0 = (float)x0, Y0 = (float)y0;
    float u = px - X0;
    float w = py - Y0;
    float t1 = u * VX;
    float t2 = w * VY;
    float t = (t1 + t2) / DN;
    t = f32clip01(t);
    float cx = X0 + t * VX;
    float cy = Y0 + t * VY;
    float dx = px - cx;
    float dy = py - cy;
    float d2a = dx * dx;
    float d2b = dy * dy;
    return sqrtf(d2a + d2b);
}

static float seg_dist_f32params(float px, float py, float x0, float y0, float x1, float y1) {
    /* _point_line_distance with np.float32 endpoints (kind 3): scalar maths in float32. */
    float vx = x1 - x0, vy = y1 - y0;
    float vv = vx * vx;
    float ww = vy * vy;
    float den = (vv + ww) + (float)1e-8;
    float u = px - x0;
    float w = py - y0;
    float t1 = u * vx;
    float t2 = w * vy;
    float t = (t1 + t2) / den;
    t = f32clip01(t);
    float cx = x0 + t * vx;
    float cy = y0 + t * vy;
    float dx = px - cx;
    float dy = py - cy;
    float d2a = dx * dx;
    float d2b = dy * dy;
    return sqrtf(d2a + d2b);
}

static float gauss_alpha(float dist, double thickness, float a) {
    double sigma = thickness > 0.5 ? thickness : 0.5;
    double inv_2s2 = 1.0 / (2.0 * sigma * sigma);
    float d2 = dist * dist;
    float arg = (-d2) * (float)inv_2s2;
    float e = expf(arg);
    return e * a;
}

/* np.linspace(0, 1, 32, dtype=float32) */
static void curve_ts(float* ts) {
    double step = 1.0 / 31.0;
    for (int i = 0; i < 32; ++i) ts[i] = (float)((double)i * step + 0.0);
    ts[31] = 1.0f;
}

static float curve_dist(const float* g, float px, float py, const float* cts) {
    /* _eval_curve_distance: 32 float32 points, min over 31 segments. */
    float ptx[32], pty[32];
    for (int i = 0; i < 32; ++i) {
        float t = cts[i];
        float u = 1.0f - t;
        float uu = u * u, tu2 = (2.0f * u) * t, tt = t * t;
        float ax = uu * g[0], bx = tu2 * g[2], cx = tt * g[4];
        float ay = uu * g[1], by = tu2 * g[3], cy = tt * g[5];
        ptx[i] = (ax + bx) + cx;
        pty[i] = (ay + by) + cy;
    }
    float best = 1e9f;
    for (int i = 0; i < 31; ++i) {
        float d = seg_dist_f32params(px, py, ptx[i], pty[i], ptx[i + 1], pty[i + 1]);
        best = d < best ? d : best; /* np.minimum (no NaNs here) */
    }
    return best;
}

/* Composite one primitive over one pixel of the running (unclipped) canvas. */
static void composite_px(const orc_prim* P, float px, float py, float* rgb, const float* cts) {
    float a = alpha32(P->alpha);
    float s[3] = {P->color0[0], P->color0[1], P->color0[2]};
    float al;
    switch (P->kind) {
    case 0:
        al = a;
        break;
    case 1: {
        const float* g = P->geom;
        double x0 = g[0], y0 = g[1], x1 = g[2], y1 = g[3];
        double vx = x1 - x0, vy = y1 - y0;
        double den = vx * vx + vy * vy + 1e-8;
        float VX = (float)vx, VY = (float)vy, DN = (float)den;
        float u = px - (float)x0;
        float w = py - (float)y0;
        float t1 = u * VX;
        float t2 = w * VY;
        float t = f32clip01((t1 + t2) / DN);
        float omt = 1.0f - t;
        for (int c = 0; c < 3; ++c) {
            float k0 = P->color0[c] * omt;
            float k1 = P->color1[c] * t;
            s[c] = k0 + k1;
        }
        al = a;
        break;
    }
    case 2: {
        const float* g = P->geom;
        float d = seg_dist_f64params(px, py, g[0], g[1], g[2], g[3]);
        al = gauss_alpha(d, (double)g[4], a);
        break;
    }
    case 3: {
        float d = curve_dist(P->geom, px, py, cts);
        al = gauss_alpha(d, (double)P->geom[6], a);
        break;
    }
    case 4:
    case 8:
        al = tri_mask(P->geom, px, py) * a;
        break;
    case 6:
    case 7:
        al = ell_rect_mask(P->kind, P->geom, px, py) * a;
        break;
    default:
        return; /* render.py:225-226: unknown kinds are skipped */
    }
    float oma = 1.0f - al;
    for (int c = 0; c < 3; ++c) {
        float k0 = s[c] * al;
        float k1 = rgb[c] * oma;
        rgb[c] = k0 + k1;
    }
}

/* All-const fast path, render.py:175-182: alpha stays a Python float. */
static void composite_const_fast(const orc_prim* P, float* rgb) {
    double a = (double)P->alpha;
    a = a < 1.0 ? a : 1.0;
    a = a > 0.0 ? a : 0.0;
    float A = (float)a;
    float OM = (float)(1.0 - a);
    for (int c = 0; c < 3; ++c) {
        float k0 = P->color0[c] * A;
        float k1 = rgb[c] * OM;
        rgb[c] = k0 + k1;
    }
}

/* Sample grid: (arange(out, f32) + 0.5) * f32(src/out); equals _grid when src == out. */
static inline float grid_coord(int i, int src, int out) {
    float base = (float)i + 0.5f;
    if (src == out) return base;
    float step = (float)((double)src / (double)out);
    return base * step;
}

/* render_tile(prims, tile_size=ts_out, residual_rgb, source_size=ts_src) without the
 * final clip when do_clip == 0 (the running canvas the greedy search caches). */
int orc_render_tile(const orc_prim* prims, int n, int ts_out, int ts_src, const int* residual,
                    int do_clip, float* out) {
    float cts[32];
    curve_ts(cts);
    int all_const = 1;
    for (int i = 0; i < n; ++i) {
        if (prims[i].kind == 5) return -1; /* bicubic tiles: outside this path */
        if (prims[i].kind != 0) all_const = 0;
    }
    float rbias[3] = {0.0f, 0.0f, 0.0f};
    if (residual) {
        for (int c = 0; c < 3; ++c) rbias[c] = (float)residual[c] / 255.0f;
    }
    for (int y = 0; y < ts_out; ++y) {
        float py = grid_coord(y, ts_src, ts_out);
        for (int x = 0; x < ts_out; ++x) {
            float px = grid_coord(x, ts_src, ts_out);
            float rgb[3] = {0.0f, 0.0f, 0.0f};
            for (int i = 0; i < n; ++i) {
                if (all_const)
                    composite_const_fast(&prims[i], rgb);
                else
                    composite_px(&prims[i], px, py, rgb, cts);
            }
            float* o = out + 3 * ((size_t)y * ts_out + x);
            for (int c = 0; c < 3; ++c) {
                float v = rgb[c];
                if (residual) v = v + rbias[c];
                o[c] = do_clip ? f32clip01(v) : v;
            }
        }
    }
    return 0;
}

/* Composite one primitive over an existing unclipped canvas (ts x ts, native grid). */
void orc_composite_canvas(const orc_prim* P, int ts, float* canvas) {
    float cts[32];
    curve_ts(cts);
    for (int y = 0; y < ts; ++y) {
        float py = (float)y + 0.5f;
        for (int x = 0; x < ts; ++x) {
            float px = (float)x + 0.5f;
            composite_px(P, px, py, canvas + 3 * ((size_t)y * ts + x), cts);
        }
    }
}

/* ---------------------------------------------------------------------------
 * NumPy float32 pairwise summation (np.add.reduce over a contiguous buffer):
 * blocks of <= 128 elements use 8 interleaved accumulators; longer runs split at
 * n/2 rounded down to a multiple of 8; the iterator hands the loop at most 8192
 * elements at a time and the chunk sums are added left to right.
 * ------------------------------------------------------------------------- */
static float pw_sum(const float* a, long n) {
    if (n < 8) {
        float r = -0.0f;
        for (long i = 0; i < n; ++i) r = r + a[i];
        return r;
    }
    if (n <= 128) {
        float r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        long i;
        for (i = 8; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] = r[j] + a[i + j];
        float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res = res + a[i];
        return res;
    }
    long n2 = n / 2;
    n2 -= n2 % 8;
    return pw_sum(a, n2) + pw_sum(a + n2, n - n2);
}

float orc_np_sum_f32(const float* a, long n) {
    float res = 0.0f;
    for (long s = 0; s < n; s += 8192) {
        long m = n - s < 8192 ? n - s : 8192;
        res = res + pw_sum(a + s, m);
    }
    return res;
}

/* float(np.mean((src - pred) ** 2)) as float32 SSE and MSE; encoder.py:282. */
void orc_sse_mse(const float* src, const float* pred, long n, float* sse, float* mse) {
    float* e = (float*)malloc(sizeof(float) * (size_t)n);
    for (long i = 0; i < n; ++i) {
        float d = src[i] - pred[i];
        e[i] = d * d;
    }
    float s = orc_np_sum_f32(e, n);
    free(e);
    *sse = s;
    *mse = s / (float)n;
}

static int est_bits(const orc_prim* sel, int n_sel, int extra_kind, int include_residual) {
    int bits = 8;
    for (int i = 0; i < n_sel; ++i) bits += orc_prim_bytes(sel[i].kind) * 8;
    if (extra_kind >= 0) bits += orc_prim_bytes(extra_kind) * 8;
    if (include_residual) bits += 24;
    return bits;
}

int orc_estimate_bits(const orc_prim* sel, int n_sel, int include_residual) {
    return est_bits(sel, n_sel, -1, include_residual);
}

double orc_tile_lambda(int quality) {
    int q = quality < 0 ? 0 : (quality > 100 ? 100 : quality);
    double base = ((101 - q) / 100.0);
    base = base * base; /* Python ** 2 on a float is an exact square of the double */
    return q >= 90 ? base * 0.0002 : base * 0.001;
}

/* _tile_objective with residual (0,0,0) and no RES1 map. */
double orc_tile_objective(const float* src, int ts, const orc_prim* sel, int n_sel, double lam,
                          float* pred_out) {
    orc_prim none;
    if (n_sel == 0 || !sel) { memset(&none, 0, sizeof(none)); sel = &none; n_sel = 0; }
    long n = (long)ts * ts * 3;
    float* pred = pred_out ? pred_out : (float*)malloc(sizeof(float) * (size_t)n);
    orc_render_tile(sel, n_sel, ts, ts, NULL, 1, pred);
    float sse, mse;
    orc_sse_mse(src, pred, n, &sse, &mse);
    if (!pred_out) free(pred);
    return (double)mse + lam * (double)est_bits(sel, n_sel, -1, 0);
}

/* _greedy_add CPU branch (GPU scorer returning None).  Returns the number of
 * primitives selected; sel_idx receives indices into the original pool. */
int orc_greedy_add(const float* src, int ts, const orc_prim* pool, int n_pool, int max_prims,
                   double lam, int* sel_idx, double* best_obj_out) {
    long n = (long)ts * ts * 3;
    orc_prim* sel = (orc_prim*)calloc((size_t)(max_prims + 1), sizeof(orc_prim));
    int* alive = (int*)malloc(sizeof(int) * (size_t)(n_pool > 0 ? n_pool : 1));
    int* order = (int*)malloc(sizeof(int) * (size_t)(n_pool > 0 ? n_pool : 1));
    double* objs = (double*)malloc(sizeof(double) * (size_t)(n_pool > 0 ? n_pool : 1));
    float* pred = (float*)malloc(sizeof(float) * (size_t)n);
    int n_sel = 0, n_alive = n_pool;
    for (int i = 0; i < n_pool; ++i) order[i] = i;
    double best_obj = orc_tile_objective(src, ts, sel, 0, lam, NULL);
    while (n_alive > 0 && n_sel < max_prims) {
        for (int k = 0; k < n_alive; ++k) {
            sel[n_sel] = pool[order[k]];
            orc_render_tile(sel, n_sel + 1, ts, ts, NULL, 1, pred);
            float sse, mse;
            orc_sse_mse(src, pred, n, &sse, &mse); /* batch_mse == np.mean per row */
            int rate = est_bits(sel, n_sel, pool[order[k]].kind, 0);
            objs[k] = (double)mse + lam * (double)rate;
        }
        int mi = 0;
        for (int k = 1; k < n_alive; ++k)
            if (objs[k] < objs[mi]) mi = k;
        if (!(objs[mi] < best_obj)) break;
        sel[n_sel] = pool[order[mi]];
        sel_idx[n_sel] = order[mi];
        n_sel++;
        best_obj = objs[mi];
        for (int k = mi; k < n_alive - 1; ++k) order[k] = order[k + 1]; /* pool.pop */
        n_alive--;
    }
    (void)alive;
    *best_obj_out = best_obj;
    free(sel);
    free(alive);
    free(order);
    free(objs);
    free(pred);
    return n_sel;
}

/* ---------------------------------------------------------------------------
 * Candidate codes (the synthetic configurations' random candidates).
 * A candidate is drawn from two 64-bit hashes of (seed, step, tile, index);
 * fields are integer codes on fixed grids, so float values are exact by
 * construction and identical on host and device.
 *   family  : 0 ellipse (kind 6), 1 rectangle (kind 7), 2 triangle (kind 8);
 *             stratified: index group g = idx / 64 of a tile takes family (r + g) mod 3,
 *             with r = ((base >> 40) & 0xFFFF) * 3 >> 16 from the tile's base hash
 *   cx, cy  : u16, value = code / 1024 (tile-local pixels), confined to the tile
 *   rx, ry  : u16 code from a log-uniform table over [4, 64) px, value = code / 1024
 *   theta   : 12-bit code, angle = code * pi / 4096, cos/sin from a table
 *   r, g, b : u8, value = code / 255
 *   alpha   : u16 in [19661, 65535], value = code / 65535  (alpha in [0.3, 1])
 * ------------------------------------------------------------------------- */
typedef struct {
    uint16_t family, cx, cy, rx, ry, theta, r, g, b, alpha;
} orc_code;

static uint64_t smix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

#define ORC_ALPHA_LO 19661u
static uint16_t g_logr[4096];
static float g_cos[4096], g_sin[4096];
static int g_tables_ready = 0;

static void orc_tables(void) {
    if (g_tables_ready) return;
    for (int k = 0; k < 4096; ++k) {
        double r = 4.0 * pow(16.0, ((double)k + 0.5) / 4096.0);
        g_logr[k] = (uint16_t)llround(r * 1024.0);
        double th = (double)k * 3.14159265358979323846 / 4096.0;
        g_cos[k] = (float)cos(th);
        g_sin[k] = (float)sin(th);
    }
    g_tables_ready = 1;
}

void orc_get_tables(uint16_t* logr, float* cosv, float* sinv) {
    orc_tables();
    memcpy(logr, g_logr, sizeof(g_logr));
    memcpy(cosv, g_cos, sizeof(g_cos));
    memcpy(sinv, g_sin, sizeof(g_sin));
}

void orc_draw_code(uint64_t seed, uint32_t step, uint32_t tile, uint32_t idx, int ts, orc_code* c) {
    uint64_t base = smix(smix(smix(seed) ^ (uint64_t)step) ^ (uint64_t)tile);
    uint64_t h1 = smix(base ^ (2ull * idx));
    uint64_t h2 = smix(base ^ (2ull * idx + 1ull));
    orc_tables();
    c->family = (uint16_t)(((uint32_t)((((base >> 40) & 0xFFFFull) * 3ull) >> 16) + (idx >> 6)) % 3u);
    uint64_t span = (uint64_t)ts * 1024ull;
    c->cx = (uint16_t)((((h1 >> 16) & 0xFFFFull) * span) >> 16);
    c->cy = (uint16_t)((((h1 >> 32) & 0xFFFFull) * span) >> 16);
    c->rx = g_logr[(h1 >> 48) & 0xFFFull];
    c->ry = g_logr[h2 & 0xFFFull];
    c->theta = (uint16_t)((h2 >> 12) & 0xFFFull);
    c->r = (uint16_t)((h2 >> 24) & 0xFFull);
    c->g = (uint16_t)((h2 >> 32) & 0xFFull);
    c->b = (uint16_t)((h2 >> 40) & 0xFFull);
    c->alpha = (uint16_t)(ORC_ALPHA_LO + ((((h2 >> 48) & 0xFFFFull) * (65536ull - ORC_ALPHA_LO)) >> 16));
}

/* Code -> float record.  Triangle vertices sit on the rotated ellipse at
 * parameter angles -90, 30 and 150 degrees. */
void orc_code_to_prim(const orc_code* c, orc_prim* P) {
    orc_tables();
    memset(P, 0, sizeof(*P));
    float cx = (float)c->cx * (1.0f / 1024.0f);
    float cy = (float)c->cy * (1.0f / 1024.0f);
    float rx = (float)c->rx * (1.0f / 1024.0f);
    float ry = (float)c->ry * (1.0f / 1024.0f);
    float ct = g_cos[c->theta & 0xFFF], st = g_sin[c->theta & 0xFFF];
    P->color0[0] = (float)((double)c->r / 255.0);
    P->color0[1] = (float)((double)c->g / 255.0);
    P->color0[2] = (float)((double)c->b / 255.0);
    P->alpha = (float)((double)c->alpha / 65535.0);
    if (c->family == 2) {
        static const float ux[3] = {0.0f, 0.8660254037844386f, -0.8660254037844386f};
        static const float uy[3] = {-1.0f, 0.5f, 0.5f};
        P->kind = 8;
        for (int k = 0; k < 3; ++k) {
            float lx = rx * ux[k];
            float ly = ry * uy[k];
            float ax = ct * lx, bx = st * ly;
            float ay = st * lx, by = ct * ly;
            P->geom[2 * k] = cx + (ax - bx);
            P->geom[2 * k + 1] = cy + (ay + by);
        }
    } else {
        P->kind = c->family == 0 ? 6 : 7;
        P->geom[0] = cx;
        P->geom[1] = cy;
        P->geom[2] = rx;
        P->geom[3] = ry;
        P->geom[4] = ct;
        P->geom[5] = st;
    }
}

/* ---------------------------------------------------------------------------
 * Plain tile fit: every step draws n_cand fresh candidates per tile, scores
 * canvas+candidate, keeps the first minimum of the objective and accepts it when
 * it beats the tile's current objective (encoder.py:752-817 semantics, float64
 * objective, rate = _estimate_bits of the grown stack).
 *
 * target/canvas: tile-major packed (tile k at float offset pix_off[k]*3, size ts[k]).
 * out_sse (optional): [n_steps][cand_total] float32 SSE table where cand_total =
 * n_tiles * n_cand (tile-major); entries of tiles that are full stay untouched.
 * win_idx/win_obj/win_acc: [n_steps][n_tiles].
 * n_sel, cur_obj: per-tile state, updated in place (initialise n_sel = 0 and
 * cur_obj < 0 to have it computed).
 * ------------------------------------------------------------------------- */
/* Objective of a tile: mode 0 is _greedy_add's float64 ``float(mse) + lam * bits``
 * (encoder.py:794-797); mode 1 the batched round's float32 ``mses + lam * rates`` with
 * float32 arrays (encoder.py:2793-2800: lam rounded to float32, product and sum rounded to
 * float32), widened to float64 for the comparison with the tile's objective (2802). */
double orc_objective(float mse, int bits, double lam, int mode) {
    if (mode == 0) return (double)mse + lam * (double)bits;
    float prod = (float)lam * (float)bits;
    float o = mse + prod;
    return (double)o;
}

int orc_fit_plain_mode(const float* target, float* canvas, const int64_t* pix_off,
                       const int32_t* tsz, int n_tiles, int n_cand, int depth, uint64_t seed,
                       int step0, int n_steps, double lam, int* n_sel, int* sel_bits,
                       double* cur_obj, float* out_sse, int32_t* win_idx, double* win_obj,
                       int32_t* win_acc, orc_code* win_code, int n_threads,
                       const int32_t* tile_ids, int obj_mode);

int orc_fit_plain_ids(const float* target, float* canvas, const int64_t* pix_off,
                      const int32_t* tsz, int n_tiles, int n_cand, int depth, uint64_t seed,
                      int step0, int n_steps, double lam, int* n_sel, int* sel_bits,
                      double* cur_obj, float* out_sse, int32_t* win_idx, double* win_obj,
                      int32_t* win_acc, orc_code* win_code, int n_threads,
                      const int32_t* tile_ids) {
    return orc_fit_plain_mode(target, canvas, pix_off, tsz, n_tiles, n_cand, depth, seed, step0, n_steps,
                              lam, n_sel, sel_bits, cur_obj, out_sse, win_idx, win_obj, win_acc, win_code,
                              n_threads, tile_ids, 0);
}

int orc_fit_plain_mode(const float* target, float* canvas, const int64_t* pix_off,
                       const int32_t* tsz, int n_tiles, int n_cand, int depth, uint64_t seed,
                       int step0, int n_steps, double lam, int* n_sel, int* sel_bits,
                       double* cur_obj, float* out_sse, int32_t* win_idx, double* win_obj,
                       int32_t* win_acc, orc_code* win_code, int n_threads,
                       const int32_t* tile_ids, int obj_mode) {
    orc_tables();
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    for (int s = 0; s < n_steps; ++s) {
        int step = step0 + s;
#pragma omp parallel for schedule(dynamic, 1)
        for (int t = 0; t < n_tiles; ++t) {
            int ts = tsz[t];
            int gid = tile_ids ? tile_ids[t] : t; /* the tile's index in the whole fit */
            long n = (long)ts * ts * 3;
            const float* src = target + pix_off[t] * 3;
            float* cv = canvas + pix_off[t] * 3;
            float* pred = (float*)malloc(sizeof(float) * (size_t)n);
            float* e = (float*)malloc(sizeof(float) * (size_t)n);
            if (cur_obj[t] < 0.0) {
                for (long i = 0; i < n; ++i) {
                    float d = src[i] - f32clip01(cv[i]);
                    e[i] = d * d;
                }
                float sse = orc_np_sum_f32(e, n);
                float mse = sse / (float)n;
                cur_obj[t] = orc_objective(mse, sel_bits[t], lam, obj_mode);
            }
            int32_t best = -1;
            double best_obj = 0.0;
            if (n_sel[t] < depth) {
                for (int k = 0; k < n_cand; ++k) {
                    orc_code code;
                    orc_prim P;
                    orc_draw_code(seed, (uint32_t)step, (uint32_t)gid, (uint32_t)k, ts, &code);
                    orc_code_to_prim(&code, &P);
                    memcpy(pred, cv, sizeof(float) * (size_t)n);
                    orc_composite_canvas(&P, ts, pred);
                    for (long i = 0; i < n; ++i) {
                        float d = src[i] - f32clip01(pred[i]);
                        e[i] = d * d;
                    }
                    float sse = orc_np_sum_f32(e, n);
                    float mse = sse / (float)n;
                    if (out_sse) out_sse[((size_t)s * n_tiles + t) * n_cand + k] = sse;
                    double obj = orc_objective(mse, sel_bits[t] + orc_prim_bytes(P.kind) * 8, lam, obj_mode);
                    if (best < 0 || obj < best_obj) {
                        best = k;
                        best_obj = obj;
                    }
                }
            }
            int acc = 0;
            if (best >= 0 && best_obj < cur_obj[t]) {
                orc_code code;
                orc_prim P;
                orc_draw_code(seed, (uint32_t)step, (uint32_t)gid, (uint32_t)best, ts, &code);
                orc_code_to_prim(&code, &P);
                orc_composite_canvas(&P, ts, cv);
                n_sel[t] += 1;
                sel_bits[t] += orc_prim_bytes(P.kind) * 8;
                cur_obj[t] = best_obj;
                acc = 1;
                if (win_code) win_code[(size_t)s * n_tiles + t] = code;
            }
            win_idx[(size_t)s * n_tiles + t] = best;
            win_obj[(size_t)s * n_tiles + t] = best_obj;
            win_acc[(size_t)s * n_tiles + t] = acc;
            free(pred);
            free(e);
        }
    }
    return 0;
}

int orc_fit_plain(const float* target, float* canvas, const int64_t* pix_off, const int32_t* tsz,
                  int n_tiles, int n_cand, int depth, uint64_t seed, int step0, int n_steps,
                  double lam, int* n_sel, int* sel_bits, double* cur_obj, float* out_sse,
                  int32_t* win_idx, double* win_obj, int32_t* win_acc, orc_code* win_code,
                  int n_threads) {
    return orc_fit_plain_ids(target, canvas, pix_off, tsz, n_tiles, n_cand, depth, seed, step0,
                             n_steps, lam, n_sel, sel_bits, cur_obj, out_sse, win_idx, win_obj,
                             win_acc, win_code, n_threads, NULL);
}

/* Score an explicit candidate list over one tile's cached canvas (no state change). */
void orc_score_candidates(const float* src, const float* canvas, int ts, const orc_prim* cands,
                          int n_cand, float* out_sse, float* out_mse) {
    long n = (long)ts * ts * 3;
    float* pred = (float*)malloc(sizeof(float) * (size_t)n);
    float* e = (float*)malloc(sizeof(float) * (size_t)n);
    for (int k = 0; k < n_cand; ++k) {
        memcpy(pred, canvas, sizeof(float) * (size_t)n);
        orc_composite_canvas(&cands[k], ts, pred);
        for (long i = 0; i < n; ++i) {
            float d = src[i] - f32clip01(pred[i]);
            e[i] = d * d;
        }
        float sse = orc_np_sum_f32(e, n);
        out_sse[k] = sse;
        out_mse[k] = sse / (float)n;
    }
    free(pred);
    free(e);
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- Static order-0 rANS (src/weft/entropy.py) ----------------------------------
 * 12-bit probabilities, 32-bit state in [2^23, 2^31), bytes emitted low first while
 * encoding backwards; the decoder reads the emitted bytes from the end. */
#define ORC_PROB_BITS 12
#define ORC_TOTFREQ (1 << ORC_PROB_BITS)
#define ORC_RANS_L (1u << 23)

/* _build_model: counts -> frequencies summing to 4096 (entropy.py:30-87). */
void orc_rans_model(const uint8_t* data, int64_t n, int32_t* freqs /* 256 */) {
    int64_t counts[256] = {0};
    for (int64_t i = 0; i < n; ++i) counts[data[i]] += 1;
    for (int s = 0; s < 256; ++s) freqs[s] = 0;
    if (n <= 0) {
        freqs[0] = ORC_TOTFREQ;
        return;
    }
    int64_t cur = 0;
    for (int s = 0; s < 256; ++s) {
        if (counts[s] > 0) {
            /* max(1, int(round((c / total) * TOTFREQ))): Python rounds half to even */
            double r = nearbyint(((double)counts[s] / (double)n) * (double)ORC_TOTFREQ);
            freqs[s] = r < 1.0 ? 1 : (int32_t)r;
        }
        cur += freqs[s];
    }
    if (cur == 0) {
        freqs[0] = ORC_TOTFREQ;
        cur = ORC_TOTFREQ;
    }
    int64_t diff = ORC_TOTFREQ - cur;
    /* sorted(range(256), key=counts, reverse=True): stable, ties keep ascending symbol */
    int order[256];
    for (int s = 0; s < 256; ++s) order[s] = s;
    for (int i = 1; i < 256; ++i) {  /* insertion sort, stable */
        int v = order[i], j = i - 1;
        while (j >= 0 && counts[order[j]] < counts[v]) {
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = v;
    }
    int64_t k = 0;
    while (diff != 0 && k < 100000) {
        int idx = order[k % 256];
        if (counts[idx] == 0) idx = 0;
        if (diff > 0) {
            freqs[idx] += 1;
            diff -= 1;
        } else if (freqs[idx] > 1) {
            freqs[idx] -= 1;
            diff += 1;
        }
        k += 1;
    }
    int64_t sum = 0;
    for (int s = 0; s < 256; ++s) sum += freqs[s];
    if (sum != ORC_TOTFREQ) freqs[0] += (int32_t)(ORC_TOTFREQ - sum);
}

static void put_u32(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24);
}
static uint32_t get_u32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

/* encode_bytes (entropy.py:190-198): rANS when strictly shorter than the raw form.
 * Returns the payload length, or -1 if `cap` is too small / a frequency is zero. */
int64_t orc_rans_encode(const uint8_t* data, int64_t n, uint8_t* out, int64_t cap) {
    const int64_t raw_len = 5 + n;
    if (n > 0) {
        int32_t freqs[256];
        int32_t cum[257];
        orc_rans_model(data, n, freqs);
        int32_t run = 0;
        for (int s = 0; s < 256; ++s) {
            cum[s] = run;
            run += freqs[s];
        }
        cum[256] = run;
        int n_pairs = 0;
        for (int s = 0; s < 256; ++s) n_pairs += freqs[s] > 0;
        const int64_t hdr = 1 + 4 + 2 + 3 * (int64_t)n_pairs + 4 + 4;
        uint8_t* em = (uint8_t*)malloc((size_t)(2 * n + 16));
        if (!em) return -1;
        int64_t ne = 0;
        uint64_t state = ORC_RANS_L;
        for (int64_t i = n - 1; i >= 0; --i) {
            const int sym = data[i];
            const int64_t freq = freqs[sym];
            if (freq <= 0) {
                free(em);
                return -1;
            }
            const uint64_t x_max = ((uint64_t)(ORC_RANS_L >> ORC_PROB_BITS) << 8) * (uint64_t)freq;
            while (state >= x_max) {
                em[ne++] = (uint8_t)(state & 0xFF);
                state >>= 8;
            }
            state = ((state / (uint64_t)freq) << ORC_PROB_BITS) + (state % (uint64_t)freq) + (uint64_t)cum[sym];
        }
        const int64_t rans_len = hdr + ne;
        if (rans_len < raw_len) {
            if (rans_len > cap) {
                free(em);
                return -1;
            }
            uint8_t* p = out;
            *p++ = 1;
            put_u32(p, (uint32_t)n); p += 4;
            p[0] = (uint8_t)n_pairs; p[1] = (uint8_t)(n_pairs >> 8); p += 2;
            for (int s = 0; s < 256; ++s) {
                if (freqs[s] > 0) {
                    p[0] = (uint8_t)s; p[1] = (uint8_t)freqs[s]; p[2] = (uint8_t)(freqs[s] >> 8);
                    p += 3;
                }
            }
            put_u32(p, (uint32_t)state); p += 4;
            put_u32(p, (uint32_t)ne); p += 4;
            for (int64_t i = 0; i < ne; ++i) p[i] = em[i];
            free(em);
            return rans_len;
        }
        free(em);
    }
    if (raw_len > cap) return -1;
    out[0] = 0;
    put_u32(out + 1, (uint32_t)n);
    for (int64_t i = 0; i < n; ++i) out[5 + i] = data[i];
    return raw_len;
}

/* decode_bytes (entropy.py:201-217, 125-187).  Returns the decoded length, or a
 * negative code: -2 truncated, -3 raw length mismatch, -4 unknown mode, -5 truncated
 * model, -6 frequency sum, -7 emitted length mismatch, -8 output capacity. */
int64_t orc_rans_decode(const uint8_t* pl, int64_t n, uint8_t* out, int64_t cap) {
    if (n == 0) return 0;
    const int mode = pl[0];
    if (mode == 0) {
        if (n < 5) return -2;
        const int64_t len = get_u32(pl + 1);
        if (5 + len != n) return -3;
        if (len > cap) return -8;
        for (int64_t i = 0; i < len; ++i) out[i] = pl[5 + i];
        return len;
    }
    if (mode != 1) return -4;
    if (n < 7) return -2;
    const int64_t size = get_u32(pl + 1);
    const int n_syms = pl[5] | (pl[6] << 8);
    int64_t off = 7;
    if (off + 3 * (int64_t)n_syms + 8 > n) return -5;
    int32_t freqs[256] = {0};
    for (int i = 0; i < n_syms; ++i) {
        freqs[pl[off]] = pl[off + 1] | (pl[off + 2] << 8);
        off += 3;
    }
    int32_t cum[257];
    int32_t run = 0;
    for (int s = 0; s < 256; ++s) {
        cum[s] = run;
        run += freqs[s];
    }
    cum[256] = run;
    if (run != ORC_TOTFREQ) return -6;
    uint8_t lookup[ORC_TOTFREQ];
    for (int s = 0; s < 256; ++s)
        for (int32_t q = cum[s]; q < cum[s] + freqs[s]; ++q) lookup[q] = (uint8_t)s;
    uint32_t state = get_u32(pl + off);
    off += 4;
    const int64_t ne = get_u32(pl + off);
    off += 4;
    if (off + ne != n) return -7;
    if (size > cap) return -8;
    const uint8_t* em = pl + off;
    int64_t ptr = ne - 1;
    for (int64_t i = 0; i < size; ++i) {
        const uint32_t slot = state & (ORC_TOTFREQ - 1);
        const int sym = lookup[slot];
        out[i] = (uint8_t)sym;
        state = (uint32_t)freqs[sym] * (state >> ORC_PROB_BITS) + (slot - (uint32_t)cum[sym]);
        while (state < ORC_RANS_L && ptr >= 0) state = (state << 8) | em[ptr--];
    }
    return size;
}

/* ------------------------------------------------------------------------------------
 * Stage-3 template fit: gpu_encoder.weft_model_bucket / weft_tile_fit_candidates
 * (src/weft/gpu_encoder.py:42-251) and the result assembly of fit_tiles_gpu_constant
 * (gpu_encoder.py:411-562, without the final np.clip, which the caller applies).
 *
 * The reference accumulates the bucket sums and squared errors with float32 atomics in
 * an unspecified order.  This restatement adds them in pixel order in float64 and rounds
 * once; everything else is the reference's float32 arithmetic, one rounding at a time.
 * ------------------------------------------------------------------------------------ */
#define ORC_TMPL_MODELS 18

static float orc_minf(float a, float b) { return a < b ? a : b; }
static float orc_maxf(float a, float b) { return a > b ? a : b; }

/* gpu_encoder.py:48-60 */
static float orc_tmpl_seg(float px, float py, float x0, float y0, float x1, float y1) {
    float vx = x1 - x0, vy = y1 - y0;
    float vv = vx * vx;
    vv = vv + vy * vy;
    float den = vv + 1e-8f;
    float num = (px - x0) * vx;
    num = num + (py - y0) * vy;
    float t = num / den;
    t = orc_maxf(0.0f, orc_minf(1.0f, t));
    float qx = x0 + t * vx, qy = y0 + t * vy;
    float dx = px - qx, dy = py - qy;
    float d2 = dx * dx;
    d2 = d2 + dy * dy;
    return sqrtf(d2);
}

/* gpu_encoder.py:62-69 */
static int orc_tmpl_tri(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2) {
    float den = (y1 - y2) * (x0 - x2);
    den = den + (x2 - x1) * (y0 - y2);
    if (fabsf(den) < 1e-6f) return 0;
    float a = (y1 - y2) * (px - x2);
    a = (a + (x2 - x1) * (py - y2)) / den;
    float b = (y2 - y0) * (px - x2);
    b = (b + (x0 - x2) * (py - y2)) / den;
    float c = 1.0f - a;
    c = c - b;
    return a >= 0.0f && b >= 0.0f && c >= 0.0f;
}

/* gpu_encoder.py:71-87 */
static float orc_tmpl_curve(float px, float py, float x0, float y0, float cx, float cy, float x1, float y1) {
    float best = 1e9f, ax = x0, ay = y0;
    for (int i = 1; i <= 12; ++i) {
        float t = (float)i / 12.0f;
        float u = 1.0f - t;
        float uu = u * u, ut2 = (2.0f * u) * t, tt = t * t;
        float qx = uu * x0;
        qx = qx + ut2 * cx;
        qx = qx + tt * x1;
        float qy = uu * y0;
        qy = qy + ut2 * cy;
        qy = qy + tt * y1;
        best = orc_minf(best, orc_tmpl_seg(px, py, ax, ay, qx, qy));
        ax = qx;
        ay = qy;
    }
    return best;
}

/* gpu_encoder.py:89-129; cos / sin of the float32 angle rounded once to float32 */
int orc_tmpl_bucket(int m, int lx, int ly, int tile_size) {
    float x = (float)lx, y = (float)ly, ts = (float)(tile_size - 1);
    float cx = ts * 0.5f, cy = ts * 0.5f;
    if (m == 0) return 0;
    if (m <= 8) {
        float angle = (float)(m - 1) * (3.14159265358979323846f / 8.0f);
        float nx = (float)cos((double)angle), ny = (float)sin((double)angle);
        float proj = (x - cx) * nx;
        proj = proj + (y - cy) * ny;
        return proj < 0.0f ? 0 : 1;
    }
    if (m == 9) return fabsf(x - cx) <= 1.5f ? 0 : 1;
    if (m == 10) return fabsf(y - cy) <= 1.5f ? 0 : 1;
    if (m == 11) return fabsf(y - x) <= 1.5f ? 0 : 1;
    if (m == 12) return orc_tmpl_tri(x, y, 0.0f, 0.0f, ts, 0.0f, 0.0f, ts) ? 0 : 1;
    if (m == 13) return orc_tmpl_tri(x, y, ts, 0.0f, ts, ts, 0.0f, 0.0f) ? 0 : 1;
    if (m == 14) return orc_tmpl_tri(x, y, 0.0f, ts, ts, ts, 0.0f, 0.0f) ? 0 : 1;
    float d;
    if (m == 15) d = orc_tmpl_curve(x, y, 0.0f, cy, cx, 0.0f, ts, cy);
    else if (m == 16) d = orc_tmpl_curve(x, y, 0.0f, cy, cx, ts, ts, cy);
    else d = orc_tmpl_curve(x, y, cx, 0.0f, 0.0f, cy, cx, ts);
    return d <= 1.25f ? 0 : 1;
}

/* gpu_encoder.py:132-251 for every tile of an H x W x 3 float32 image */
void orc_template_fit(const float* img, int w, int h, int ts, int32_t* model, float* c0, float* c1, float* mse,
                      float* c0_all, float* c1_all) {
    const int cols = (w + ts - 1) / ts, rows = (h + ts - 1) / ts, n = ts * ts;
    unsigned char* bk = (unsigned char*)malloc((size_t)n * ORC_TMPL_MODELS);
    for (int m = 0; m < ORC_TMPL_MODELS; ++m)
        for (int i = 0; i < n; ++i) bk[m * n + i] = (unsigned char)orc_tmpl_bucket(m, i % ts, i / ts, ts);
    for (int t = 0; t < cols * rows; ++t) {
        const int x0 = (t % cols) * ts, y0 = (t / cols) * ts;
        float col[ORC_TMPL_MODELS][2][3];
        float e32[ORC_TMPL_MODELS];
        for (int m = 0; m < ORC_TMPL_MODELS; ++m) {
            double s[2][3] = {{0, 0, 0}, {0, 0, 0}};
            int cnt[2] = {0, 0};
            for (int i = 0; i < n; ++i) {
                int x = x0 + i % ts, y = y0 + i / ts;
                if (x >= w) x = w - 1;
                if (y >= h) y = h - 1;
                const float* p = img + ((long)y * w + x) * 3;
                const int b = bk[m * n + i];
                cnt[b] += 1;
                for (int c = 0; c < 3; ++c) s[b][c] += (double)p[c];
            }
            for (int c = 0; c < 3; ++c) {
                float inv0 = 1.0f / (float)(cnt[0] > 1 ? cnt[0] : 1);
                float inv1 = 1.0f / (float)(cnt[1] > 1 ? cnt[1] : 1);
                col[m][0][c] = (float)s[0][c] * inv0;
                col[m][1][c] = m == 0 ? col[m][0][c] : (float)s[1][c] * inv1;
            }
            double err = 0.0;
            for (int i = 0; i < n; ++i) {
                int x = x0 + i % ts, y = y0 + i / ts;
                if (x >= w) x = w - 1;
                if (y >= h) y = h - 1;
                const float* p = img + ((long)y * w + x) * 3;
                const float* q = col[m][bk[m * n + i]];
                float dr = p[0] - q[0], dg = p[1] - q[1], db = p[2] - q[2];
                float e = dr * dr;
                e = e + dg * dg;
                e = e + db * db;
                err += (double)e;
            }
            e32[m] = (float)err * (1.0f / (float)n);
        }
        int best = 0;
        for (int m = 0; m < ORC_TMPL_MODELS; ++m)
            if (e32[m] < e32[best]) best = m;
        model[t] = best;
        for (int m = 0; m < ORC_TMPL_MODELS; ++m) {
            mse[(long)t * ORC_TMPL_MODELS + m] = e32[m];
            for (int c = 0; c < 3; ++c) {
                c0_all[((long)t * ORC_TMPL_MODELS + m) * 3 + c] = col[m][0][c];
                c1_all[((long)t * ORC_TMPL_MODELS + m) * 3 + c] = col[m][1][c];
            }
        }
        for (int c = 0; c < 3; ++c) {
            c0[(long)t * 3 + c] = col[best][0][c];
            c1[(long)t * 3 + c] = col[best][1][c];
        }
    }
    free(bk);
}
