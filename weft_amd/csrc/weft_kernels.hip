// weft_kernels.hip -- gfx950 kernels for the Weft tile-fit scoring path and the
// primitive-walk renderer.  Built offline into a code object (hipcc --genco) that
// libweft_hip.so embeds and loads with hipModuleLoadData.
//
// Exactness contract (reference = jarmak-personal/weft, src/weft/render.py:127-234,
// src/weft/encoder.py:271-285): every float32 operation is the one NumPy performs for
// render_tile / np.mean, in the same order, with no fused multiply-add (the file is
// compiled with -ffp-contract=off) and IEEE division.  Per-candidate SSE is summed in
// NumPy's pairwise order: 8 interleaved accumulators over leaf blocks of 96 elements
// (128 for 64-px tiles), a balanced binary tree over the leaves, and 8192-element
// chunks added left to right.  Kinds 2/3 use the device expf, which is not NumPy's
// SIMD exp: those two kinds agree to ~1e-7, every other kind is bit-exact.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "weft_abi.h"

// Largest tile size whose per-column candidate terms are kept in registers (2 * TS
// VGPRs); larger tiles form them per pixel.
#ifndef WEFT_COLPRE_MAX
#define WEFT_COLPRE_MAX 8
#endif

// Candidate tables, filled by weft_init through hipModuleGetGlobal.
__device__ uint16_t weft_tab_logr[WEFT_TABLE_N];
__device__ float weft_tab_cos[WEFT_TABLE_N];
__device__ float weft_tab_sin[WEFT_TABLE_N];
__device__ float weft_tab_u8[256];  // RN32(c / 255)

// Scratch of the allocation-free fit step (weft_fit_step_static): class orders and
// partial minima.  It lives as long as the loaded code object (one copy per device);
// weft_init takes its address with hipModuleGetGlobal and the host passes that address
// to the kernels.  External linkage so that the module lookup by name finds it.
__device__ __align__(256) unsigned char weft_scratch[WEFT_STATIC_SCRATCH_BYTES];

namespace {

enum : int { CL_CONST = 0, CL_ELL = 1, CL_RECT = 2, CL_TRI = 3, CL_GEN = 4, CL_NONE = 5 };

__device__ __forceinline__ float clip01(float v) { return __builtin_amdgcn_fmed3f(v, 0.0f, 1.0f); }

__device__ __forceinline__ int prim_bytes(int kind) {
    // encoder.py:242-256 (unknown kinds cost 20 bytes) + 17-byte records for 6/7/8
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

// ---------------------------------------------------------------------------
// Generic per-pixel composite of one primitive (all kinds), render.py:184-229.
// rgb holds the running unclipped canvas; (px, py) the sample position.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float tri_alpha(const float* g, float px, float py) {
    double x0 = g[0], y0 = g[1], x1 = g[2], y1 = g[3], x2 = g[4], y2 = g[5];
    double den = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2);
    if (fabs(den) < 1e-6) return 0.0f;
    float A = (float)(y1 - y2), B = (float)(x2 - x1);
    float C = (float)(y2 - y0), E = (float)(x0 - x2);
    float D = (float)den;
    float dx = px - (float)x2;
    float dy = py - (float)y2;
    float na = A * dx + B * dy;
    float a = na / D;
    float nb = C * dx + E * dy;
    float b = nb / D;
    float c = (1.0f - a) - b;
    return (a >= 0.0f && b >= 0.0f && c >= 0.0f) ? 1.0f : 0.0f;
}

__device__ __forceinline__ float ell_rect_alpha(int kind, const float* g, float px, float py) {
    float rx = g[2], ry = g[3];
    if (!(rx > 0.0f) || !(ry > 0.0f)) return 0.0f;
    float ca = g[4] / rx, sa = g[5] / rx, cb = g[4] / ry, sb = g[5] / ry;
    float dx = px - g[0];
    float dy = py - g[1];
    float p = ca * dx + sa * dy;
    float q = cb * dy - sb * dx;
    if (kind == 6) return (p * p + q * q) <= 1.0f ? 1.0f : 0.0f;
    return (fabsf(p) <= 1.0f && fabsf(q) <= 1.0f) ? 1.0f : 0.0f;
}

__device__ __forceinline__ float seg_dist_f32(float px, float py, float x0, float y0, float x1,
                                              float y1) {
    float vx = x1 - x0, vy = y1 - y0;
    float den = (vx * vx + vy * vy) + 1e-8f;
    float t = ((px - x0) * vx + (py - y0) * vy) / den;
    t = clip01(t);
    float cx = x0 + t * vx;
    float cy = y0 + t * vy;
    float dx = px - cx, dy = py - cy;
    return sqrtf(dx * dx + dy * dy);
}

__device__ __forceinline__ float gauss_alpha(float dist, float thickness, float a) {
    double sigma = (double)thickness > 0.5 ? (double)thickness : 0.5;
    double inv_2s2 = 1.0 / (2.0 * sigma * sigma);
    float arg = (-(dist * dist)) * (float)inv_2s2;
    return expf(arg) * a;
}

__device__ void composite_px(const WeftPrim& P, float px, float py, float& r, float& g, float& b) {
    const float a = clip01(P.alpha);
    float s0 = P.color0[0], s1 = P.color0[1], s2 = P.color0[2];
    float al;
    switch (P.kind) {
    case 0:
        al = a;
        break;
    case 1: {
        double x0 = P.geom[0], y0 = P.geom[1], x1 = P.geom[2], y1 = P.geom[3];
        double vx = x1 - x0, vy = y1 - y0;
        double den = vx * vx + vy * vy + 1e-8;
        float VX = (float)vx, VY = (float)vy, DN = (float)den;
        float t = clip01(((px - (float)x0) * VX + (py - (float)y0) * VY) / DN);
        float omt = 1.0f - t;
        s0 = P.color0[0] * omt + P.color1[0] * t;
        s1 = P.color0[1] * omt + P.color1[1] * t;
        s2 = P.color0[2] * omt + P.color1[2] * t;
        al = a;
        break;
    }
    case 2: {
        double x0 = P.geom[0], y0 = P.geom[1], x1 = P.geom[2], y1 = P.geom[3];
        double vx = x1 - x0, vy = y1 - y0;
        double den = vx * vx + vy * vy + 1e-8;
        float VX = (float)vx, VY = (float)vy, DN = (float)den;
        float X0 = (float)x0, Y0 = (float)y0;
        float t = clip01(((px - X0) * VX + (py - Y0) * VY) / DN);
        float cx = X0 + t * VX, cy = Y0 + t * VY;
        float dx = px - cx, dy = py - cy;
        al = gauss_alpha(sqrtf(dx * dx + dy * dy), P.geom[4], a);
        break;
    }
    case 3: {
        // _eval_curve_distance: 32 float32 points (np.linspace in float64, cast), 31 segments
        float best = 1e9f;
        float prevx = 0.0f, prevy = 0.0f;
        for (int i = 0; i < 32; ++i) {
            float t = (i == 31) ? 1.0f : (float)((double)i * (1.0 / 31.0));
            float u = 1.0f - t;
            float uu = u * u, tu2 = (2.0f * u) * t, tt = t * t;
            float x = (uu * P.geom[0] + tu2 * P.geom[2]) + tt * P.geom[4];
            float y = (uu * P.geom[1] + tu2 * P.geom[3]) + tt * P.geom[5];
            if (i > 0) best = fminf(best, seg_dist_f32(px, py, prevx, prevy, x, y));
            prevx = x;
            prevy = y;
        }
        al = gauss_alpha(best, P.geom[6], a);
        break;
    }
    case 4:
    case 8:
        al = tri_alpha(P.geom, px, py) * a;
        break;
    case 6:
    case 7:
        al = ell_rect_alpha(P.kind, P.geom, px, py) * a;
        break;
    default:
        return;  // render.py:225-226 skips unknown kinds
    }
    const float oma = 1.0f - al;
    r = s0 * al + r * oma;
    g = s1 * al + g * oma;
    b = s2 * al + b * oma;
}

// composite_px restricted to the constant and hard-edge kinds (0, 4, 6, 7, 8): the same
// operations, without the code of the Gaussian and gradient kinds; other kinds are
// skipped as composite_px skips unknown ones.
__device__ __forceinline__ void composite_hard_px(const WeftPrim& P, float px, float py, float& r, float& g,
                                                  float& b) {
    const float a = clip01(P.alpha);
    float al;
    switch (P.kind) {
    case 0: al = a; break;
    case 4:
    case 8: al = tri_alpha(P.geom, px, py) * a; break;
    case 6:
    case 7: al = ell_rect_alpha(P.kind, P.geom, px, py) * a; break;
    default: return;
    }
    const float oma = 1.0f - al;
    r = P.color0[0] * al + r * oma;
    g = P.color0[1] * al + g * oma;
    b = P.color0[2] * al + b * oma;
}

// composite_px over 8 consecutive pixels of one row (x0 + i + 0.5, py), i = 0..7: the same
// float operations per pixel, with every per-primitive and per-row term formed once.
__device__ __forceinline__ void composite_row8(const WeftPrim& P, float x0, float py, float (&r)[8], float (&g)[8],
                                               float (&b)[8]) {
    const float a = clip01(P.alpha);
    float al[8];
    float s0[8], s1[8], s2[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        s0[i] = P.color0[0];
        s1[i] = P.color0[1];
        s2[i] = P.color0[2];
    }
    switch (P.kind) {
    case 0:
#pragma unroll
        for (int i = 0; i < 8; ++i) al[i] = a;
        break;
    case 1:
    case 2: {
        const double gx0 = P.geom[0], gy0 = P.geom[1], gx1 = P.geom[2], gy1 = P.geom[3];
        const double vx = gx1 - gx0, vy = gy1 - gy0;
        const double den = vx * vx + vy * vy + 1e-8;
        const float VX = (float)vx, VY = (float)vy, DN = (float)den;
        const float X0 = (float)gx0, Y0 = (float)gy0;
        const float ty = (py - Y0) * VY;
        if (P.kind == 1) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const float px = (x0 + (float)i) + 0.5f;
                const float t = clip01(((px - X0) * VX + ty) / DN);
                const float omt = 1.0f - t;
                s0[i] = P.color0[0] * omt + P.color1[0] * t;
                s1[i] = P.color0[1] * omt + P.color1[1] * t;
                s2[i] = P.color0[2] * omt + P.color1[2] * t;
                al[i] = a;
            }
        } else {
            const double sigma = (double)P.geom[4] > 0.5 ? (double)P.geom[4] : 0.5;
            const float k2 = (float)(1.0 / (2.0 * sigma * sigma));
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const float px = (x0 + (float)i) + 0.5f;
                const float t = clip01(((px - X0) * VX + ty) / DN);
                const float cx = X0 + t * VX, cy = Y0 + t * VY;
                const float dx = px - cx, dy = py - cy;
                const float dist = sqrtf(dx * dx + dy * dy);
                al[i] = expf((-(dist * dist)) * k2) * a;
            }
        }
        break;
    }
    case 3: {
        float best[8];  // squared distances; 1e18 > RN(1e9)^2
#pragma unroll
        for (int i = 0; i < 8; ++i) best[i] = 1e18f;
        float prevx = 0.0f, prevy = 0.0f;
        for (int k = 0; k < 32; ++k) {
            const float t = (k == 31) ? 1.0f : (float)((double)k * (1.0 / 31.0));
            const float u = 1.0f - t;
            const float uu = u * u, tu2 = (2.0f * u) * t, tt = t * t;
            const float x = (uu * P.geom[0] + tu2 * P.geom[2]) + tt * P.geom[4];
            const float y = (uu * P.geom[1] + tu2 * P.geom[3]) + tt * P.geom[5];
            if (k > 0) {  // seg_dist_f32(px, py, prevx, prevy, x, y) for the 8 pixels
                const float vx = x - prevx, vy = y - prevy;
                const float den = (vx * vx + vy * vy) + 1e-8f;
                const float ty = (py - prevy) * vy;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float px = (x0 + (float)i) + 0.5f;
                    const float tt2 = clip01(((px - prevx) * vx + ty) / den);
                    const float qx = prevx + tt2 * vx, qy = prevy + tt2 * vy;
                    const float dx = px - qx, dy = py - qy;
                    best[i] = fminf(best[i], dx * dx + dy * dy);
                }
            }
            prevx = x;
            prevy = y;
        }
        const double sigma = (double)P.geom[6] > 0.5 ? (double)P.geom[6] : 0.5;
        const float k2 = (float)(1.0 / (2.0 * sigma * sigma));
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float d = sqrtf(best[i]);  // = the minimum of the segments' sqrtf (monotone)
            al[i] = expf((-(d * d)) * k2) * a;
        }
        break;
    }
    case 4:
    case 8: {
        const float* gm = P.geom;
        const double gx0 = gm[0], gy0 = gm[1], gx1 = gm[2], gy1 = gm[3], gx2 = gm[4], gy2 = gm[5];
        const double den = (gy1 - gy2) * (gx0 - gx2) + (gx2 - gx1) * (gy0 - gy2);
        if (fabs(den) < 1e-6) {
#pragma unroll
            for (int i = 0; i < 8; ++i) al[i] = 0.0f * a;
            break;
        }
        const float A = (float)(gy1 - gy2), B = (float)(gx2 - gx1);
        const float C = (float)(gy2 - gy0), E = (float)(gx0 - gx2);
        const float D = (float)den;
        const float dy = py - (float)gy2;
        const float bdy = B * dy, edy = E * dy;
        const float X2 = (float)gx2;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float dx = ((x0 + (float)i) + 0.5f) - X2;
            const float ta = (A * dx + bdy) / D;
            const float tb = (C * dx + edy) / D;
            const float tc = (1.0f - ta) - tb;
            al[i] = ((ta >= 0.0f && tb >= 0.0f && tc >= 0.0f) ? 1.0f : 0.0f) * a;
        }
        break;
    }
    case 6:
    case 7: {
        const float* gm = P.geom;
        const float rx = gm[2], ry = gm[3];
        if (!(rx > 0.0f) || !(ry > 0.0f)) {
#pragma unroll
            for (int i = 0; i < 8; ++i) al[i] = 0.0f * a;
            break;
        }
        const float ca = gm[4] / rx, sa = gm[5] / rx, cb = gm[4] / ry, sb = gm[5] / ry;
        const float dy = py - gm[1];
        const float sady = sa * dy, cbdy = cb * dy;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float dx = ((x0 + (float)i) + 0.5f) - gm[0];
            const float p = ca * dx + sady;
            const float q = cbdy - sb * dx;
            const bool in = P.kind == 6 ? (p * p + q * q) <= 1.0f : (fabsf(p) <= 1.0f && fabsf(q) <= 1.0f);
            al[i] = (in ? 1.0f : 0.0f) * a;
        }
        break;
    }
    default:
        return;  // render.py:225-226 skips unknown kinds
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float oma = 1.0f - al[i];
        r[i] = s0[i] * al[i] + r[i] * oma;
        g[i] = s1[i] * al[i] + g[i] * oma;
        b[i] = s2[i] * al[i] + b[i] * oma;
    }
}

// Out-of-line copy for the per-pixel loops of the generic paths: keeps the unrolled
// pixel bodies small and the fast classes' register allocation unaffected.
__device__ __noinline__ void composite_px_call(const WeftPrim& P, float px, float py, float& r,
                                               float& g, float& b) {
    composite_px(P, px, py, r, g, b);
}

// ---------------------------------------------------------------------------
// NumPy pairwise float32 summation, streamed in element order.
// ---------------------------------------------------------------------------
// Leaf size of NumPy's pairwise recursion for n = 3*TS*TS elements (halving at
// multiples of 8 until <= 128): 96 for 8/16/32-px tiles, 128 for 64-px tiles.
template <int TS> struct PwCfg { static constexpr int BLOCK = (TS == 64) ? 128 : 96; };

struct PwState {
    float r[8];
    float s[8];

    // Push a finished leaf (block j, 0-based) into the binary counter.  j is
    // wave-uniform, so the branches are scalar.
    __device__ __forceinline__ void push(float v, int j) {
        if (j & 1) {
            v = s[0] + v;
            if (j & 2) {
                v = s[1] + v;
                if (j & 4) {
                    v = s[2] + v;
                    if (j & 8) {
                        v = s[3] + v;
                        if (j & 16) {
                            v = s[4] + v;
                            if (j & 32) {
                                v = s[5] + v;
                                if (j & 64) { v = s[6] + v; s[7] = v; }
                                else s[6] = v;
                            } else s[5] = v;
                        } else s[4] = v;
                    } else s[3] = v;
                } else s[2] = v;
            } else s[1] = v;
        } else {
            s[0] = v;
        }
    }

    // Element `off` (0..23, compile-time) of the 24-element group starting at element
    // `base` (runtime, wave-uniform, a multiple of 24).  Accumulators start each leaf at
    // +0, and 0 + e == e exactly, so this equals NumPy's "first eight initialise".
    template <int BLOCK>
    __device__ __forceinline__ void feed(int off, float e, int base) {
        r[off & 7] = r[off & 7] + e;
        // A leaf can only end after an element whose index is 7 mod 8; when the leaf size
        // is a multiple of 24 it ends on a group boundary (index 23 mod 24).
        if ((off & 7) == 7 && (BLOCK % 24 != 0 || off % 24 == 23)) {
            const int end = base + off + 1;
            if (end % BLOCK == 0) {
                const float v = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
                push(v, end / BLOCK - 1);
#pragma unroll
                for (int j = 0; j < 8; ++j) r[j] = 0.0f;
            }
        }
    }

    template <int TS>
    __device__ __forceinline__ float result() const {
        if constexpr (TS == 8) return s[1];
        else if constexpr (TS == 16) return s[3];
        else if constexpr (TS == 32) return s[5];
        else return s[6] + s[5];  // 8192-element chunk, then the 4096-element rest
    }
};

// One group of 8 pixels (24 elements from element `base`) of row y, columns gx..gx+7.
template <int TS, class Px>
__device__ __forceinline__ void pw_group(Px& px, PwState& st, int y, int gx, float gxf, int base) {
    constexpr int BLOCK = PwCfg<TS>::BLOCK;
    if constexpr (Px::kGroup) px.group(gx, gxf);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        float e0, e1, e2;
        px.pixel(y, gx + i, i, gxf + (float)i, e0, e1, e2);
        st.template feed<BLOCK>(3 * i + 0, e0, base);
        st.template feed<BLOCK>(3 * i + 1, e1, base);
        st.template feed<BLOCK>(3 * i + 2, e2, base);
    }
}

// Pairwise SSE over rows [y0, y1) of a tile, pixels in row-major order, 8 pixels
// (24 elements) per unrolled group; the state carries over between calls so a tile can
// be processed in row bands.  Px provides row(y) and pixel(y, x, i, pxf, e0, e1, e2) (the
// three squared errors of pixel (y, x); i = x mod 8, a compile-time constant, and pxf =
// x + 0.5).
template <int TS, class Px>
__device__ __forceinline__ void pairwise_rows(Px& px, PwState& st, int y0, int y1) {
    int base = 3 * TS * y0;
#pragma unroll 1
    for (int y = y0; y < y1; ++y) {
        px.row(y);
        if constexpr (Px::kUnrollRow) {
            // whole row unrolled: column index x is a compile-time constant
#pragma unroll
            for (int g = 0; g < TS / 8; ++g) {
                pw_group<TS>(px, st, y, 8 * g, 8.0f * (float)g + 0.5f, base);
                base += 24;
            }
        } else {
            float gxf = 0.5f;  // sample x of the group's first pixel (exact)
#pragma unroll 1
            for (int gx = 0; gx < TS; gx += 8) {
                pw_group<TS>(px, st, y, gx, gxf, base);
                base += 24;
                gxf += 8.0f;
            }
        }
    }
}

__device__ __forceinline__ void pw_init(PwState& st) {
#pragma unroll
    for (int j = 0; j < 8; ++j) st.r[j] = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) st.s[j] = 0.0f;
}

template <int TS, class Px>
__device__ __forceinline__ float pairwise_tile(Px& px) {
    PwState st;
    pw_init(st);
    pairwise_rows<TS>(px, st, 0, TS);
    return st.template result<TS>();
}

// ---------------------------------------------------------------------------
// Pixel sources.
// ---------------------------------------------------------------------------

// Reference-semantics stack walk over a zero canvas; target read from global memory.
struct PxStack {
    static constexpr bool kUnrollRow = false;
    static constexpr bool kGroup = false;
    const float* tgt;
    const WeftPrim* prims;
    int n;
    int ts;
    float py;
    __device__ __forceinline__ void row(int y) { py = (float)y + 0.5f; }
    __device__ __forceinline__ void pixel(int y, int x, int, float px, float& e0, float& e1, float& e2) {
        float r = 0.0f, g = 0.0f, b = 0.0f;
        for (int i = 0; i < n; ++i) composite_px_call(prims[i], px, py, r, g, b);
        const float* t = tgt + 3 * (y * ts + x);
        float d0 = t[0] - clip01(r), d1 = t[1] - clip01(g), d2 = t[2] - clip01(b);
        e0 = d0 * d0;
        e1 = d1 * d1;
        e2 = d2 * d2;
    }
};


// LDS band of rows [y0, y0 + ROWS): per pixel two float4 and one float,
//   A = (c_r, c_g, c_b, t_r), B = (k_r, k_g, k_b, t_g), T = t_b,
// with c the unclipped canvas, k = clip(c) (the prediction where the candidate does not
// cover the pixel) and t the target.
struct LdsBand {
    const float4* A;  // [2 * pixels]: A, B interleaved
    const float* T;   // [pixels]
};

template <int TS>
struct PxLdsBase {
    LdsBand L;
    int y0;
    float ks0, ks1, ks2, om;
    __device__ __forceinline__ void finish(int y, int x, bool in, float& e0, float& e1, float& e2) {
        const int p = (y - y0) * TS + x;
        const float4 a = L.A[2 * p], b = L.A[2 * p + 1];
        const float tb = L.T[p];
        // clip(s*a + c*(1 - a)) where covered (the clip folds into the add), clip(c) elsewhere
        const float w0 = clip01(ks0 + a.x * om);
        const float w1 = clip01(ks1 + a.y * om);
        const float w2 = clip01(ks2 + a.z * om);
        const float v0 = in ? w0 : b.x;
        const float v1 = in ? w1 : b.y;
        const float v2 = in ? w2 : b.z;
        const float d0 = a.w - v0, d1 = b.w - v1, d2 = tb - v2;
        e0 = d0 * d0;
        e1 = d1 * d1;
        e2 = d2 * d2;
    }
};

template <int TS>
struct PxConst : PxLdsBase<TS> {
    static constexpr bool kUnrollRow = false;
    static constexpr bool kGroup = false;
    __device__ __forceinline__ void row(int) {}
    __device__ __forceinline__ void pixel(int y, int x, int, float, float& e0, float& e1, float& e2) {
        this->finish(y, x, true, e0, e1, e2);
    }
};

// Ellipse / rectangle: p = ca*dx + sa*dy, q = cb*dy - sb*dx.  The dy terms are formed
// once per row; for tiles up to 32 px the dx terms (which do not depend on the row) are
// formed once per candidate and kept in registers, with every row fully unrolled.
template <int TS, bool ELL>
struct PxEllRect : PxLdsBase<TS> {
    static constexpr bool kUnrollRow = TS <= WEFT_COLPRE_MAX;
    static constexpr bool kGroup = false;
    static constexpr int NCOL = kUnrollRow ? TS : 1;
    float cx, cy, ca, sa, cb, sb;
    float p2, q1;
    float P1[NCOL], Q2[NCOL];
    __device__ __forceinline__ void init() {
        if constexpr (kUnrollRow) {
#pragma unroll
            for (int x = 0; x < TS; ++x) {
                const float dx = ((float)x + 0.5f) - cx;
                P1[x] = ca * dx;
                Q2[x] = sb * dx;
            }
        }
    }
    __device__ __forceinline__ void row(int y) {
        const float dy = ((float)y + 0.5f) - cy;
        p2 = sa * dy;
        q1 = cb * dy;
    }
    __device__ __forceinline__ void pixel(int y, int x, int i, float pxf, float& e0, float& e1, float& e2) {
        float p, q;
        if constexpr (kUnrollRow) {
            p = P1[x] + p2;
            q = q1 - Q2[x];
        } else {
            const float dx = pxf - cx;
            p = ca * dx + p2;
            q = q1 - sb * dx;
        }
        bool in;
        if constexpr (ELL) in = (p * p + q * q) <= 1.0f;
        else in = (fabsf(p) <= 1.0f) && (fabsf(q) <= 1.0f);
        this->finish(y, x, in, e0, e1, e2);
    }
};

// Triangle (render.py:114-124): a = RN(na / D), b = RN(nb / D) with
// na = A dx + B dy, nb = C dx + E dy.  The float32 quotient is formed as
// RN32(na * RN64(1 / D)): that float64 product is within 2^-52 (relative) of na / D while
// a float32 quotient is never within 2^-48 of a rounding midpoint, so both round to the
// same float -- the IEEE division's result, at the cost of one float64 multiply.
template <int TS>
struct PxTri : PxLdsBase<TS> {
    static constexpr bool kGroup = true;
    static constexpr bool kUnrollRow = TS <= WEFT_COLPRE_MAX;
    static constexpr int NCOL = kUnrollRow ? TS : 1;
    float X2, Y2, A, B, C, E;
    double RD;  // RN64(1 / D)
    float RF;   // RN32(1 / D)
    float bdy, edy;
    float AD[NCOL], CD[NCOL];
    __device__ __forceinline__ void init() {
        if constexpr (kUnrollRow) {
#pragma unroll
            for (int x = 0; x < TS; ++x) {
                const float dx = ((float)x + 0.5f) - X2;
                AD[x] = A * dx;
                CD[x] = C * dx;
            }
        }
    }
    __device__ __forceinline__ void row(int y) {
        const float dy = ((float)y + 0.5f) - Y2;
        bdy = B * dy;
        edy = E * dy;
    }
    // Numerators of pixel x (column group start gx, sample x pxf) of the current row.
    __device__ __forceinline__ void numer(int x, int i, float pxf, float& na, float& nb) const {
        if constexpr (kUnrollRow) {
            na = AD[x] + bdy;
            nb = CD[x] + edy;
        } else {
            const float dx = pxf - X2;
            na = A * dx + bdy;
            nb = C * dx + edy;
        }
    }
    // Coverage of the 8 pixels from column gx (sample x gxf) on: the inside test on the
    // float32 quotients a = RN(na / D), b = RN(nb / D) and c = RN(RN(1 - a) - b)
    // (render.py:118-123), from the approximations a1 = RN(na * RF), b1, c1 (RF = RN(1 / D))
    // and m = min(a1, b1, c1).  a1 and b1 are within 2^-22 (relative) of a and b, so
    // m >= 2^-19 puts all three quotients clearly inside (then a, b <= 1 and c1 is within
    // 2^-20 of c); m <= -2^-19 is outside for the quotients too: a1 or b1 that negative fixes
    // the sign of a or b, and a c1 that negative with a1, b1 > -2^-19 means a + b exceeds 1
    // by more than c's error, which grows only like 2^-22 (|a| + |b|).  Lanes with |m| below
    // 2^-19 take the exact quotients RN(na * RN64(1 / D)) for the group, and so does every
    // pixel of a candidate with RF = NaN (coefficients that are not finite or beyond 2^60,
    // see cand_consts): then m is NaN.
    bool inb[8];
    __device__ __forceinline__ void group(int gx, float gxf) {
        bool anynear = false;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            float na, nb;
            numer(gx + i, i, gxf + (float)i, na, nb);
            const float a1 = na * RF;
            const float b1 = nb * RF;
            const float c1 = (1.0f - a1) - b1;
            const float m = fminf(fminf(a1, b1), c1);
            inb[i] = m >= 0.0f;
            anynear |= !(fabsf(m) >= 0x1p-19f);
        }
        if (anynear) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                float na, nb;
                numer(gx + i, i, gxf + (float)i, na, nb);
                const float a = (float)((double)na * RD);
                const float b = (float)((double)nb * RD);
                const float c = (1.0f - a) - b;
                inb[i] = (a >= 0.0f) && (b >= 0.0f) && (c >= 0.0f);
            }
        }
    }
    __device__ __forceinline__ void pixel(int y, int x, int i, float, float& e0, float& e1, float& e2) {
        this->finish(y, x, inb[i], e0, e1, e2);
    }
};

// Any kind over the LDS canvas through the generic composite, 8 pixels of a row at a time
// (composite_row8: the per-primitive and per-row terms, a curve's polyline among them,
// formed once per group).
template <int TS>
struct PxGeneral {
    static constexpr bool kUnrollRow = false;
    static constexpr bool kGroup = true;
    LdsBand L;
    int y0, yr;
    WeftPrim P;
    float py;
    float cr[8], cg[8], cb[8];
    __device__ __forceinline__ void row(int y) {
        yr = y;
        py = (float)y + 0.5f;
    }
    __device__ __forceinline__ void group(int gx, float) {
        const int p0 = (yr - y0) * TS + gx;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float4 a = L.A[2 * (p0 + i)];
            cr[i] = a.x;
            cg[i] = a.y;
            cb[i] = a.z;
        }
        composite_row8(P, (float)gx, py, cr, cg, cb);
    }
    __device__ __forceinline__ void pixel(int y, int x, int i, float, float& e0, float& e1, float& e2) {
        const int p = (y - y0) * TS + x;
        const float4 a = L.A[2 * p], b = L.A[2 * p + 1];
        const float d0 = a.w - clip01(cr[i]), d1 = b.w - clip01(cg[i]), d2 = L.T[p] - clip01(cb[i]);
        e0 = d0 * d0;
        e1 = d1 * d1;
        e2 = d2 * d2;
    }
};

template <int TS>
__device__ __noinline__ void score_general(LdsBand L, WeftPrim P, PwState& st, int y0, int y1) {
    PxGeneral<TS> px;
    px.L = L;
    px.y0 = y0;
    px.P = P;
    pairwise_rows<TS>(px, st, y0, y1);
}

// ---------------------------------------------------------------------------
// Candidate codes, drawn from (seed, step, tile, index).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t smix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ uint64_t draw_base(uint64_t seed, uint32_t step, uint32_t tile) {
    return smix(smix(smix(seed) ^ (uint64_t)step) ^ (uint64_t)tile);
}

// Family of candidate idx, stratified in groups of 64 consecutive indices (one wavefront
// once the scorers have sorted a tile's candidates by class): group g of a tile takes
// family (r + g) mod 3, r drawn from the tile's base hash, so every family gets a third of
// the groups on average and no wavefront holds candidates of two families.
__device__ __forceinline__ uint32_t draw_rot(uint64_t base) {
    return (uint32_t)((((base >> 40) & 0xFFFFull) * 3ull) >> 16);
}

__device__ __forceinline__ int draw_family(uint64_t base, uint32_t idx) {
    return (int)((draw_rot(base) + (idx >> 6)) % 3u);
}

// Index at position `slot` (< n) of a tile's n drawn candidates stably sorted by class.
// A drawn candidate's class is its family's (ellipse < rectangle < triangle; drawn radii
// are at least 4 px, so none folds to the constant class), so the order lists family 0's
// groups of 64, then family 1's, then 2's, each in index order; only the last group
// (G - 1) can be partial, and it is the last of its family.
__device__ __forceinline__ int drawn_sorted_index(uint64_t base, int n, int slot) {
    const int r = (int)draw_rot(base);
    const int G = (n + 63) >> 6;
    int start = 0;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const int g0 = (f - r + 3) % 3;
        const int ng = g0 < G ? (G - 1 - g0) / 3 + 1 : 0;
        int cnt = ng * 64;
        if (ng > 0 && (G - 1 - g0) % 3 == 0) cnt -= 64 * G - n;
        if (slot < start + cnt) {
            const int local = slot - start;
            return ((g0 + 3 * (local >> 6)) << 6) + (local & 63);
        }
        start += cnt;
    }
    return -1;
}

__device__ __forceinline__ WeftCode draw_code(uint64_t base, uint32_t idx, int ts) {
    const uint64_t h1 = smix(base ^ (2ull * idx));
    const uint64_t h2 = smix(base ^ (2ull * idx + 1ull));
    const uint64_t span = (uint64_t)ts * 1024ull;
    WeftCode c;
    c.family = (uint16_t)draw_family(base, idx);
    c.cx = (uint16_t)((((h1 >> 16) & 0xFFFFull) * span) >> 16);
    c.cy = (uint16_t)((((h1 >> 32) & 0xFFFFull) * span) >> 16);
    c.rx = weft_tab_logr[(h1 >> 48) & 0xFFFull];
    c.ry = weft_tab_logr[h2 & 0xFFFull];
    c.theta = (uint16_t)((h2 >> 12) & 0xFFFull);
    c.r = (uint16_t)((h2 >> 24) & 0xFFull);
    c.g = (uint16_t)((h2 >> 32) & 0xFFull);
    c.b = (uint16_t)((h2 >> 40) & 0xFFull);
    c.alpha = (uint16_t)(19661u + ((((h2 >> 48) & 0xFFFFull) * (65536ull - 19661ull)) >> 16));
    return c;
}

__device__ __forceinline__ WeftPrim code_to_prim(const WeftCode& c) {
    WeftPrim P;
    const float cx = (float)c.cx * (1.0f / 1024.0f);
    const float cy = (float)c.cy * (1.0f / 1024.0f);
    const float rx = (float)c.rx * (1.0f / 1024.0f);
    const float ry = (float)c.ry * (1.0f / 1024.0f);
    const float ct = weft_tab_cos[c.theta & 0xFFF], st = weft_tab_sin[c.theta & 0xFFF];
    // RN32(c / 255) from a table, and RN32(a / 65535) as a float32 division: each equals the
    // float64 quotient rounded to float32 (no quotient of these codes is within 2^-41
    // relative of a float32 rounding midpoint, so the double rounding never differs)
    P.color0[0] = weft_tab_u8[c.r & 0xFF];
    P.color0[1] = weft_tab_u8[c.g & 0xFF];
    P.color0[2] = weft_tab_u8[c.b & 0xFF];
    P.color1[0] = P.color1[1] = P.color1[2] = 0.0f;
    P.alpha = (float)c.alpha / 65535.0f;
    P.geom[6] = 0.0f;
    if (c.family == 2) {
        const float ux[3] = {0.0f, 0.8660254037844386f, -0.8660254037844386f};
        const float uy[3] = {-1.0f, 0.5f, 0.5f};
        P.kind = 8;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float lx = rx * ux[k];
            const float ly = ry * uy[k];
            P.geom[2 * k] = cx + (ct * lx - st * ly);
            P.geom[2 * k + 1] = cy + (st * lx + ct * ly);
        }
    } else {
        P.kind = c.family == 0 ? 6 : 7;
        P.geom[0] = cx;
        P.geom[1] = cy;
        P.geom[2] = rx;
        P.geom[3] = ry;
        P.geom[4] = ct;
        P.geom[5] = st;
    }
    return P;
}

__device__ __forceinline__ WeftPrim load_prim(const WeftPrim* p) {
    WeftPrim P;
    P.kind = p->kind;
#pragma unroll
    for (int i = 0; i < 7; ++i) P.geom[i] = p->geom[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        P.color0[i] = p->color0[i];
        P.color1[i] = p->color1[i];
    }
    P.alpha = p->alpha;
    return P;
}

__device__ __forceinline__ int class_of_kind(int kind) {
    switch (kind) {
    case 0: return CL_CONST;
    case 4: case 8: return CL_TRI;
    case 6: return CL_ELL;
    case 7: return CL_RECT;
    case 1: case 2: case 3: return CL_GEN;
    default: return CL_CONST;  // skipped kinds leave the canvas unchanged
    }
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const unsigned lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

__device__ __forceinline__ void argmin_pair(double& o, int& i, double o2, int i2) {
    if (o2 < o || (o2 == o && i2 < i)) {
        o = o2;
        i = i2;
    }
}

__device__ __forceinline__ double objective(float mse, int bits, double lam, int mode) {
    if (mode == 0) return (double)mse + lam * (double)bits;
    const float r32 = (float)bits;
    const float o = mse + (float)lam * r32;
    return (double)o;
}

// ---------------------------------------------------------------------------
// Plain fit: score one work item (one tile, `blockDim.x` candidates).
// ---------------------------------------------------------------------------
// Rows of a tile held in LDS at a time (36 bytes per pixel): at most 512 pixels, i.e.
// whole 8- and 16-px tiles, 16-row bands of 32-px tiles and 8-row bands of 64-px tiles
// (<= 18 KiB), so eight 256-thread workgroups fit on a CU.
template <int TS> struct Band { static constexpr int ROWS = TS * TS <= 512 ? TS : 512 / TS; };

// Per-candidate constants of the hard-edge classes.
struct CandConsts {
    float ks0, ks1, ks2, om;               // s * a per channel, 1 - a
    float g0, g1, k11, k12, k21, k22;      // class-specific coverage constants
    double rd;                             // triangles: RN64(1 / D)
    float rf;                              // triangles: RN32(RN64(1 / D)), or NaN: exact quotients only
};

// Constants of candidate P (class `cls`) for the hard-edge scorers; returns the class
// with the empty-coverage cases (degenerate triangle, non-positive radius, skipped
// kind) folded into the no-op constant, whose canvas passes through unchanged.
__device__ __forceinline__ int cand_consts(const WeftPrim& P, int cls, CandConsts& cc) {
    const float a = clip01(P.alpha);
    cc.ks0 = P.color0[0] * a;
    cc.ks1 = P.color0[1] * a;
    cc.ks2 = P.color0[2] * a;
    cc.om = 1.0f - a;
    cc.g0 = cc.g1 = cc.k11 = cc.k12 = cc.k21 = cc.k22 = 0.0f;
    cc.rd = 0.0;
    cc.rf = 0.0f;
    if (cls == CL_TRI) {
        const float* g = P.geom;
        double x0 = g[0], y0 = g[1], x1 = g[2], y1 = g[3], x2 = g[4], y2 = g[5];
        double den = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2);
        if (fabs(den) < 1e-6) {
            cls = CL_CONST;
        } else {
            cc.k11 = (float)(y1 - y2); cc.k12 = (float)(x2 - x1);
            cc.k21 = (float)(y2 - y0); cc.k22 = (float)(x0 - x2);
            cc.rd = 1.0 / (double)(float)den;
            cc.g0 = (float)x2; cc.g1 = (float)y2;
            // the filtered inside test needs products that cannot overflow and a normal
            // reciprocal; otherwise every pixel takes the exact quotients
            const float lim = 0x1p60f;
            const float rf = (float)cc.rd;
            const bool fast = fabsf(cc.k11) <= lim && fabsf(cc.k12) <= lim && fabsf(cc.k21) <= lim &&
                              fabsf(cc.k22) <= lim && fabsf(cc.g0) <= lim && fabsf(cc.g1) <= lim &&
                              fabsf(rf) >= 0x1p-120f && fabsf(rf) <= lim;
            cc.rf = fast ? rf : __builtin_nanf("");
        }
    }
    if (cls == CL_ELL || cls == CL_RECT) {
        if (!(P.geom[2] > 0.0f && P.geom[3] > 0.0f)) {
            cls = CL_CONST;
        } else {
            cc.g0 = P.geom[0]; cc.g1 = P.geom[1];
            cc.k11 = P.geom[4] / P.geom[2]; cc.k12 = P.geom[5] / P.geom[2];
            cc.k21 = P.geom[4] / P.geom[3]; cc.k22 = P.geom[5] / P.geom[3];
        }
    }
    if (cls == CL_CONST && P.kind != 0) {
        cc.ks0 = cc.ks1 = cc.ks2 = 0.0f;
        cc.om = 1.0f;
    }
    return cls;
}

template <int TS, bool ELL>
__device__ __forceinline__ void score_ellrect(const CandConsts& c, LdsBand L, PwState& pw, int y0, int y1) {
    PxEllRect<TS, ELL> px;
    px.L = L; px.y0 = y0; px.ks0 = c.ks0; px.ks1 = c.ks1; px.ks2 = c.ks2; px.om = c.om;
    px.cx = c.g0; px.cy = c.g1; px.ca = c.k11; px.sa = c.k12; px.cb = c.k21; px.sb = c.k22;
    px.init();
    pairwise_rows<TS>(px, pw, y0, y1);
}

template <int TS>
__device__ __forceinline__ void score_tri(const CandConsts& c, LdsBand L, PwState& pw, int y0, int y1) {
    PxTri<TS> px;
    px.L = L; px.y0 = y0; px.ks0 = c.ks0; px.ks1 = c.ks1; px.ks2 = c.ks2; px.om = c.om;
    px.X2 = c.g0; px.Y2 = c.g1; px.A = c.k11; px.B = c.k12; px.C = c.k21; px.E = c.k22; px.RD = c.rd;
    px.RF = c.rf;
    px.init();
    pairwise_rows<TS>(px, pw, y0, y1);
}

template <int TS, bool GEN>
__device__ __forceinline__ void score_band(int mycls, const CandConsts& c, const WeftPrim& P,
                                           LdsBand L, PwState& pw, int y0, int y1) {
    if (mycls == CL_ELL) {
        score_ellrect<TS, true>(c, L, pw, y0, y1);
    } else if (mycls == CL_RECT) {
        score_ellrect<TS, false>(c, L, pw, y0, y1);
    } else if (mycls == CL_TRI) {
        score_tri<TS>(c, L, pw, y0, y1);
    } else if (mycls == CL_CONST) {
        PxConst<TS> px;
        px.L = L; px.y0 = y0; px.ks0 = c.ks0; px.ks1 = c.ks1; px.ks2 = c.ks2; px.om = c.om;
        pairwise_rows<TS>(px, pw, y0, y1);
    } else if (mycls == CL_GEN) {
        // kinds 1-3 need the generic composite; the GEN=false variants are only
        // launched when the host has checked that no such candidate is present
        if constexpr (GEN) score_general<TS>(L, P, pw, y0, y1);
        else pw.s[0] = pw.s[1] = pw.s[3] = pw.s[5] = pw.s[6] = __builtin_nanf("");
    }
}

// Class of candidate k of `tile` (drawn or explicit), with the empty-coverage cases
// (degenerate triangle, non-positive radius) folded into the no-op constant.
__device__ __forceinline__ int candidate_class(const WeftCands& cd, uint64_t dbase, int64_t cand_base,
                                               int k) {
    if (cd.d_prims == nullptr) {
        const int fam = draw_family(dbase, (uint32_t)k);
        return fam == 0 ? CL_ELL : (fam == 1 ? CL_RECT : CL_TRI);
    }
    const WeftPrim* pp = cd.d_prims + cand_base + k;
    int cls = class_of_kind(pp->kind);
    if ((cls == CL_ELL || cls == CL_RECT) && !(pp->geom[2] > 0.0f && pp->geom[3] > 0.0f))
        cls = CL_CONST;
    if (cls == CL_TRI) {
        const float* g = pp->geom;
        double x0 = g[0], y0 = g[1], x1 = g[2], y1 = g[3], x2 = g[4], y2 = g[5];
        double den = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2);
        if (fabs(den) < 1e-6) cls = CL_CONST;
    }
    return cls;
}

// Live candidates of a tile's pool, and whether the tile takes part in this step.
__device__ __forceinline__ int tile_ncand(const WeftCands& cd, int t) {
    return cd.d_tile_ncand ? cd.d_tile_ncand[t] : cd.n_cand;
}
__device__ __forceinline__ bool tile_skipped(const WeftFitState& st, const WeftCands& cd, int t) {
    return st.d_tile_nsel[t] >= st.depth || (cd.d_tile_active != nullptr && cd.d_tile_active[t] == 0);
}

// ---------------------------------------------------------------------------
// Plain fit: score one work item (one tile, `blockDim.x` consecutive slots of the
// tile's class-sorted candidate order, so waves are class-uniform except at the at
// most few class boundaries of a tile).
// ---------------------------------------------------------------------------
template <int TS, bool GEN>
__device__ void fit_score_body(const WeftFitScoreArgs& A) {
    extern __shared__ float4 weft_smem[];
    constexpr int ROWS = Band<TS>::ROWS;
    constexpr int BPX = ROWS * TS;  // pixels per LDS band
    float4* LA = weft_smem;                                        // [2 * BPX]
    float* LT = reinterpret_cast<float*>(LA + 2 * BPX);            // [BPX]
    double* red_o = reinterpret_cast<double*>(LT + BPX);           // [16]
    int* red_i = reinterpret_cast<int*>(red_o + 16);               // [16]

    const int w = blockIdx.x;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int wave = tid >> 6;
    const int nwaves = nthr >> 6;
    const int tile = A.wk.d_work_tile[w];
    const int c0 = A.wk.d_work_cand0[w];
    const WeftFitState& st = A.st;

    const int ncand = tile_ncand(A.cd, tile);
    if (tile_skipped(st, A.cd, tile) || c0 >= ncand) {
        if (tid == 0) {
            A.part_obj[w] = __builtin_huge_val();
            A.part_idx[w] = -1;
        }
        return;
    }

    // This lane's candidate: slot c0 + tid of the tile's class-sorted order.
    const bool drawn = A.cd.d_prims == nullptr;
    const uint64_t dbase = draw_base(A.cd.seed, A.cd.step, (uint32_t)tile);
    const int64_t cand_base = drawn ? 0
                              : (A.cd.d_tile_cand ? A.cd.d_tile_cand[tile]
                                                  : (int64_t)tile * A.cd.n_cand);
    const int slot = c0 + tid;
    const int ks = slot >= ncand ? -1
                   : (drawn ? drawn_sorted_index(dbase, A.cd.n_cand, slot)
                            : A.order[(int64_t)tile * A.cd.n_cand + slot]);
    WeftPrim P;
    P.kind = -1;
    P.alpha = 0.0f;
    P.color0[0] = P.color0[1] = P.color0[2] = 0.0f;
    int mycls = CL_NONE;
    if (ks >= 0) {
        if (drawn) P = code_to_prim(draw_code(dbase, (uint32_t)ks, TS));
        else P = load_prim(A.cd.d_prims + cand_base + ks);
        mycls = class_of_kind(P.kind);
    }
    CandConsts cc;
    mycls = cand_consts(P, mycls, cc);

    // Stage the tile band by band (coalesced loads) and score it.
    const int64_t base = st.d_tile_pix[tile] * 3;
    const float* T = st.d_target + base;
    const float* Cv = st.d_canvas + base;
    float* LAf = reinterpret_cast<float*>(LA);
    const LdsBand L{LA, LT};
    // A split item of a 64-px tile scores rows [32 half, 32 half + 32) only: leaves 0-47 or
    // 48-95 of the tile's 96 pairwise leaves (see the combine kernel).
    int y_lo = 0, y_hi = TS, split = -1;
    if constexpr (TS == 64) {
        if (A.wk.d_work_split) {
            split = A.wk.d_work_split[w];
            if (split >= 0) {
                y_lo = 32 * (split & 1);
                y_hi = y_lo + 32;
            }
        }
    }
    PwState pw;
    pw_init(pw);
#pragma unroll 1
    for (int y0 = y_lo; y0 < y_hi; y0 += ROWS) {
        if (y0 > y_lo) __syncthreads();  // every wave is done with the previous band
        const int e0 = 3 * y0 * TS;
        for (int i = tid; i < 3 * BPX; i += nthr) {
            const int p = i / 3, ch = i - 3 * (i / 3);
            const float cv = Cv[e0 + i];
            const float tv = T[e0 + i];
            LAf[8 * p + ch] = cv;
            LAf[8 * p + 4 + ch] = clip01(cv);
            if (ch < 2) LAf[8 * p + 3 + 4 * ch] = tv;
            else LT[p] = tv;
        }
        __syncthreads();
        score_band<TS, GEN>(mycls, cc, P, L, pw, y0, y0 + ROWS);
    }
    if constexpr (TS == 64) {
        if (split >= 0) {
            // rows 0-31: subtrees [0..31] (stack level 5) and [32..47] (level 4); rows 32-63,
            // from a zero stack: [48..63] (level 6) and the 32-leaf rest chunk (level 5)
            const int half = split & 1, pair = split >> 1;
            const float2 v = half == 0 ? make_float2(pw.s[5], pw.s[4]) : make_float2(pw.s[6], pw.s[5]);
            reinterpret_cast<float2*>(A.wk.d_split_sse)[((int64_t)pair * 2 + half) * nthr + tid] = v;
            if (tid == 0) {
                A.part_obj[w] = __builtin_huge_val();
                A.part_idx[w] = -1;
            }
            return;
        }
    }
    const float sse = pw.template result<TS>();

    // Tables, objective and the work item's first minimum.
    double obj = __builtin_huge_val();
    int idx = 0x7fffffff;
    if (ks >= 0) {
        const float mse = sse / (float)(3 * TS * TS);
        const int64_t ti = (int64_t)tile * A.cd.n_cand + ks;
        if (A.out_sse) A.out_sse[ti] = sse;
        if (A.out_mse) A.out_mse[ti] = mse;
        obj = objective(mse, st.d_tile_bits[tile] + prim_bytes(P.kind) * 8, st.lam, st.obj_mode);
        idx = ks;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o2 = __shfl_xor(obj, off, 64);
        const int i2 = __shfl_xor(idx, off, 64);
        argmin_pair(obj, idx, o2, i2);
    }
    if ((tid & 63) == 0) {
        red_o[wave] = obj;
        red_i[wave] = idx;
    }
    __syncthreads();
    if (tid == 0) {
        double bo = red_o[0];
        int bi = red_i[0];
        for (int ww = 1; ww < nwaves; ++ww) argmin_pair(bo, bi, red_o[ww], red_i[ww]);
        A.part_obj[w] = bo;
        A.part_idx[w] = bi == 0x7fffffff ? -1 : bi;
    }
}

// Split pairs of 64-px tiles (one workgroup of the scoring block size per pair): each
// candidate's SSE from the two halves' partial sums, exactly as the single pass forms it --
// chunk A = [0..31] + ([32..47] + [48..63]), then chunk A + the rest chunk (NumPy's 8192-element
// buffer, then the remaining 4096) -- then the tables, the objective and the pair's first
// minimum, written to the first half's partial slot (the second half's holds +inf).
__device__ void fit_score_combine_body(const WeftFitScoreArgs& A) {
    __shared__ double red_o[16];
    __shared__ int red_i[16];
    constexpr int TS = 64;
    const int sp = blockIdx.x;
    const int tid = threadIdx.x, nthr = blockDim.x, wave = tid >> 6, nwaves = nthr >> 6;
    const int w0 = A.wk.d_split_pairs[2 * sp];
    const int tile = A.wk.d_work_tile[w0];
    const int c0 = A.wk.d_work_cand0[w0];
    const WeftFitState& st = A.st;
    const int ncand = tile_ncand(A.cd, tile);
    if (tile_skipped(st, A.cd, tile) || c0 >= ncand) return;  // both halves wrote +inf
    const bool drawn = A.cd.d_prims == nullptr;
    const uint64_t dbase = draw_base(A.cd.seed, A.cd.step, (uint32_t)tile);
    const int64_t cand_base = drawn ? 0
                              : (A.cd.d_tile_cand ? A.cd.d_tile_cand[tile]
                                                  : (int64_t)tile * A.cd.n_cand);
    const int slot = c0 + tid;
    const int ks = slot >= ncand ? -1
                   : (drawn ? drawn_sorted_index(dbase, A.cd.n_cand, slot)
                            : A.order[(int64_t)tile * A.cd.n_cand + slot]);
    double obj = __builtin_huge_val();
    int idx = 0x7fffffff;
    if (ks >= 0) {
        const int kind = drawn ? code_to_prim(draw_code(dbase, (uint32_t)ks, TS)).kind
                               : A.cd.d_prims[cand_base + ks].kind;
        const float2* part = reinterpret_cast<const float2*>(A.wk.d_split_sse);
        const float2 h0 = part[((int64_t)sp * 2 + 0) * nthr + tid];
        const float2 h1 = part[((int64_t)sp * 2 + 1) * nthr + tid];
        const float chunk_a = h0.x + (h0.y + h1.x);
        const float sse = chunk_a + h1.y;
        const float mse = sse / (float)(3 * TS * TS);
        const int64_t ti = (int64_t)tile * A.cd.n_cand + ks;
        if (A.out_sse) A.out_sse[ti] = sse;
        if (A.out_mse) A.out_mse[ti] = mse;
        obj = objective(mse, st.d_tile_bits[tile] + prim_bytes(kind) * 8, st.lam, st.obj_mode);
        idx = ks;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o2 = __shfl_xor(obj, off, 64);
        const int i2 = __shfl_xor(idx, off, 64);
        argmin_pair(obj, idx, o2, i2);
    }
    if ((tid & 63) == 0) {
        red_o[wave] = obj;
        red_i[wave] = idx;
    }
    __syncthreads();
    if (tid == 0) {
        double bo = red_o[0];
        int bi = red_i[0];
        for (int ww = 1; ww < nwaves; ++ww) argmin_pair(bo, bi, red_o[ww], red_i[ww]);
        A.part_obj[w0] = bo;
        A.part_idx[w0] = bi == 0x7fffffff ? -1 : bi;
    }
}

// Class-sorted candidate order of one tile (stable counting sort by class).
__device__ void fit_order_body(const WeftFitOrderArgs& A) {
    __shared__ int s_cnt[8];
    __shared__ int s_run[8];
    __shared__ int s_wcnt[16 * 8];
    __shared__ int s_woff[16 * 8];
    const int tile = blockIdx.x;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int wave = tid >> 6;
    const int nwaves = nthr >> 6;
    const WeftFitState& st = A.st;
    if (tile_skipped(st, A.cd, tile)) return;
    const int n = tile_ncand(A.cd, tile);
    const bool drawn = A.cd.d_prims == nullptr;
    const uint64_t dbase = draw_base(A.cd.seed, A.cd.step, (uint32_t)tile);
    const int64_t cand_base = drawn ? 0
                              : (A.cd.d_tile_cand ? A.cd.d_tile_cand[tile]
                                                  : (int64_t)tile * A.cd.n_cand);
    if (tid < 8) {
        s_cnt[tid] = 0;
        s_run[tid] = 0;
    }
    __syncthreads();
    // class counts, one LDS atomic per wave and class (drawn families come in runs of 64)
    for (int j0 = 0; j0 < n; j0 += nthr) {
        const int k = j0 + tid;
        const int cls = k < n ? candidate_class(A.cd, dbase, cand_base, k) : CL_NONE;
#pragma unroll
        for (int c = 0; c < WEFT_N_CLASSES; ++c) {
            const uint64_t bal = __ballot(cls == c);
            if ((tid & 63) == 0 && bal && c != CL_NONE) atomicAdd(&s_cnt[c], (int)__popcll(bal));
        }
    }
    __syncthreads();
    if (tid == 0) {
        int off = 0;
        for (int c = 0; c < WEFT_N_CLASSES; ++c) {
            const int v = s_cnt[c];
            s_cnt[c] = off;  // class base
            off += v;
        }
    }
    __syncthreads();
    int32_t* out = A.order + (int64_t)tile * A.cd.n_cand;
    for (int j0 = 0; j0 < n; j0 += nthr) {
        const int k = j0 + tid;
        const int cls = k < n ? candidate_class(A.cd, dbase, cand_base, k) : CL_NONE;
        uint64_t mine = 0;
#pragma unroll
        for (int c = 0; c < WEFT_N_CLASSES; ++c) {
            const uint64_t bal = __ballot(cls == c);
            if (cls == c) mine = bal;
            if ((tid & 63) == 0) s_wcnt[wave * 8 + c] = __popcll(bal);
        }
        __syncthreads();
        if (tid < WEFT_N_CLASSES) {
            int off = s_cnt[tid] + s_run[tid];
            for (int ww = 0; ww < nwaves; ++ww) {
                s_woff[ww * 8 + tid] = off;
                off += s_wcnt[ww * 8 + tid];
            }
            s_run[tid] = off - s_cnt[tid];
        }
        __syncthreads();
        if (k < n) out[s_woff[wave * 8 + cls] + __popcll(mine & lanemask_lt())] = k;
        __syncthreads();
    }
}


template <int TS>
__device__ float stack_sse(const float* tgt, const WeftPrim* prims, int n) {
    PxStack px{tgt, prims, n, TS, 0.0f};
    return pairwise_tile<TS>(px);
}

// ---------------------------------------------------------------------------
// Persistent batched fit: workgroups take tiles from a queue and run each tile's whole
// greedy fit (n_steps rounds of: score the tile's candidates in class order, first
// minimum, accept, composite) before taking the next tile.  Candidates are drawn
// from (seed, step0 + s, tile, index) exactly as the stepwise kernels draw them, so the
// result equals that of weft_fit_order/score/select run n_steps times.
// ---------------------------------------------------------------------------
struct BatchShared {
    double red_o[16];
    int red_i[16];
    int tile;
    int acc;
    int best;
    double best_obj;
};

template <int TS>
__device__ void batch_fit_tile(const WeftBatchArgs& A, int t, float4* LA, float* LT,
                               BatchShared& sh) {
    constexpr int ROWS = Band<TS>::ROWS;
    constexpr int NB = TS / ROWS;
    constexpr int BPX = ROWS * TS;
    const WeftFitState& st = A.st;
    const int tid = threadIdx.x, nthr = blockDim.x, wave = tid >> 6, nwaves = nthr >> 6;
    const int64_t base = st.d_tile_pix[t] * 3;
    float* cvp = st.d_canvas + base;
    const float* T = st.d_target + base;
    const int n = A.cd.n_cand;
    float* LAf = reinterpret_cast<float*>(LA);
    const LdsBand L{LA, LT};
    const int64_t T_all = st.n_tiles;
#pragma unroll 1
    for (int s = 0; s < A.n_steps; ++s) {
        const int64_t rec = (int64_t)s * T_all + t;
        if (st.d_tile_nsel[t] >= st.depth) {  // uniform: written by this workgroup only
            if (tid == 0) {
                A.win_idx[rec] = -1;
                A.win_obj[rec] = __builtin_huge_val();
                A.win_acc[rec] = 0;
            }
            continue;
        }
        const uint32_t step = A.step0 + (uint32_t)s;
        const uint64_t dbase = draw_base(A.cd.seed, step, (uint32_t)t);
        double bo = __builtin_huge_val();
        int bi = 0x7fffffff;
        bool staged = false;
#pragma unroll 1
        for (int c0 = 0; c0 < n; c0 += nthr) {
            const int slot = c0 + tid;
            const int ks = slot < n ? drawn_sorted_index(dbase, n, slot) : -1;
            WeftPrim P;
            P.kind = -1;
            P.alpha = 0.0f;
            P.color0[0] = P.color0[1] = P.color0[2] = 0.0f;
            int mycls = CL_NONE;
            if (ks >= 0) {
                P = code_to_prim(draw_code(dbase, (uint32_t)ks, TS));
                mycls = class_of_kind(P.kind);
            }
            CandConsts cc;
            mycls = cand_consts(P, mycls, cc);
            PwState pw;
            pw_init(pw);
#pragma unroll 1
            for (int y0 = 0; y0 < TS; y0 += ROWS) {
                if (NB > 1 || !staged) {
                    __syncthreads();
                    const int e0 = 3 * y0 * TS;
                    for (int i = tid; i < 3 * BPX; i += nthr) {
                        const int p = i / 3, ch = i - 3 * (i / 3);
                        const float cv = cvp[e0 + i];
                        const float tv = T[e0 + i];
                        LAf[8 * p + ch] = cv;
                        LAf[8 * p + 4 + ch] = clip01(cv);
                        if (ch < 2) LAf[8 * p + 3 + 4 * ch] = tv;
                        else LT[p] = tv;
                    }
                    __syncthreads();
                    staged = true;
                }
                score_band<TS, false>(mycls, cc, P, L, pw, y0, y0 + ROWS);
            }
            if (ks >= 0) {
                const float sse = pw.template result<TS>();
                const float mse = sse / (float)(3 * TS * TS);
                argmin_pair(bo, bi, objective(mse, st.d_tile_bits[t] + prim_bytes(P.kind) * 8, st.lam,
                                              st.obj_mode), ks);
            }
        }
        // first minimum over the workgroup
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double o2 = __shfl_xor(bo, off, 64);
            const int i2 = __shfl_xor(bi, off, 64);
            argmin_pair(bo, bi, o2, i2);
        }
        if ((tid & 63) == 0) {
            sh.red_o[wave] = bo;
            sh.red_i[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double o = sh.red_o[0];
            int i = sh.red_i[0];
            for (int ww = 1; ww < nwaves; ++ww) argmin_pair(o, i, sh.red_o[ww], sh.red_i[ww]);
            const bool found = i != 0x7fffffff;
            sh.best = found ? i : -1;
            sh.best_obj = o;
            sh.acc = found && (o < st.d_tile_obj[t]);
        }
        __syncthreads();
        const int win = sh.best;
        const int acc = sh.acc;
        if (acc) {
            const WeftCode code = draw_code(dbase, (uint32_t)win, TS);
            const WeftPrim Pw = code_to_prim(code);
            for (int p = tid; p < TS * TS; p += nthr) {
                const int y = p / TS, x = p - y * TS;
                float r = cvp[3 * p], g = cvp[3 * p + 1], b = cvp[3 * p + 2];
                composite_px(Pw, (float)x + 0.5f, (float)y + 0.5f, r, g, b);
                cvp[3 * p] = r;
                cvp[3 * p + 1] = g;
                cvp[3 * p + 2] = b;
            }
            if (tid == 0) {
                if (A.win_code) A.win_code[rec] = code;
                st.d_tile_nsel[t] += 1;
                st.d_tile_bits[t] += prim_bytes(Pw.kind) * 8;
                st.d_tile_obj[t] = sh.best_obj;
            }
        }
        if (tid == 0) {
            A.win_idx[rec] = win;
            A.win_obj[rec] = sh.best_obj;
            A.win_acc[rec] = acc;
        }
        __threadfence_block();
        __syncthreads();  // canvas and state are settled before the next round reads them
    }
}

__device__ void batch_fit_body(const WeftBatchArgs& A) {
    extern __shared__ float4 weft_smem[];
    __shared__ BatchShared sh;
    float4* LA = weft_smem;                                   // [2 * 512]
    float* LT = reinterpret_cast<float*>(LA + 2 * 512);       // [512]
    const int tid = threadIdx.x;
    if (tid == 0 && A.t_start) {
        atomicMin(A.t_start, (unsigned long long)__builtin_amdgcn_s_memrealtime());
    }
#pragma unroll 1
    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(A.queue, 1);
            sh.tile = q < A.n_queue ? A.tile_order[q] : -1;
        }
        __syncthreads();
        const int t = sh.tile;
        __syncthreads();  // every thread has read sh.tile before thread 0 may overwrite it
        if (t < 0) break;
        switch (A.st.d_tile_ts[t]) {
        case 8: batch_fit_tile<8>(A, t, LA, LT, sh); break;
        case 16: batch_fit_tile<16>(A, t, LA, LT, sh); break;
        case 32: batch_fit_tile<32>(A, t, LA, LT, sh); break;
        case 64: batch_fit_tile<64>(A, t, LA, LT, sh); break;
        default: break;
        }
        if (tid == 0 && A.image_left) {
            const int img = A.tile_image[t];
            if (atomicSub(&A.image_left[img], 1) == 1)
                A.image_done[img] = (unsigned long long)__builtin_amdgcn_s_memrealtime();
        }
    }
}

#include "weft_soft.h"
#include "weft_split.h"
#include "weft_rans.h"
#include "weft_batch_refine.h"
#include "weft_template.h"
#include "weft_select.h"
#include "weft_candgen.h"
#include "weft_edgegen.h"

}  // namespace

// ===========================================================================
// Kernels
// ===========================================================================

// One lane per evaluation (reference batch_tile_objective with CPU semantics).
extern "C" __global__ __launch_bounds__(256) void weft_score_evals_kernel(WeftScoreEvalsArgs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n_evals) return;
    const WeftEval ev = A.evals[i];
    const float* tgt = A.tiles + ev.tile_offset;
    const WeftPrim* pr = A.prims + ev.prim_offset;
    float sse;
    switch (ev.tile_size) {
    case 8: sse = stack_sse<8>(tgt, pr, ev.prim_count); break;
    case 16: sse = stack_sse<16>(tgt, pr, ev.prim_count); break;
    case 32: sse = stack_sse<32>(tgt, pr, ev.prim_count); break;
    case 64: sse = stack_sse<64>(tgt, pr, ev.prim_count); break;
    default: sse = __builtin_nanf(""); break;
    }
    const float n = (float)(3 * ev.tile_size * ev.tile_size);
    A.out_mse[i] = sse / n;
    if (A.out_sse) A.out_sse[i] = sse;
}

// One 64-lane workgroup per tile: objective of the current canvas, the pairwise tree of
// the clipped canvas's squared errors with every lane busy.  Lane = 8 * leaf lane + a:
// accumulator a of each of its leaf lane's leaves adds the leaf's elements a, a + 8, ...
// in order (reading them directly, as no candidate is composited); the 8 accumulators of a
// leaf meet in the ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7)) tree through xor
// shuffles 1, 2, 4 (a + b == b + a), and the leaves in the pairwise tree as in
// weft_split.h.  Equal to the sequential sum bit for bit.
// EMPTY: the canvas is zero (clip(0) - t = -t exactly), so it is not read.
template <int LS, bool EMPTY = false>
__device__ __forceinline__ float init_chunk(const float* tg, const float* cv, int c0, int nleaves) {
    const int lane = threadIdx.x & 63;
    const int a = lane & 7, ll = lane >> 3;
    const int L = nleaves < 8 ? nleaves : 8;  // leaf lanes in use (a power of two)
    const int m = nleaves / L;                 // leaves per leaf lane
    SplitTree T;
    T.n = 0;
#pragma unroll
    for (int q = 0; q < 7; ++q) T.lvl[q] = 0.0f;
#pragma unroll 1
    for (int li = 0; li < m; ++li) {
        float r = 0.0f;
        if (ll < L) {
            const int e0 = c0 + (ll * m + li) * LS + a;
#pragma unroll
            for (int k = 0; k < LS / 8; ++k) {
                const float d = (EMPTY ? 0.0f : clip01(cv[e0 + 8 * k])) - tg[e0 + 8 * k];
                r = r + d * d;
            }
        }
        r += __shfl_xor(r, 1, 64);
        r += __shfl_xor(r, 2, 64);
        r += __shfl_xor(r, 4, 64);
        split_tree_push(T, r);
    }
    float v = 0.0f;
#pragma unroll
    for (int q = 0; q < 7; ++q)
        if ((1 << q) == m) v = T.lvl[q];
    for (int off = 8; off < 8 * L; off <<= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int TS, bool EMPTY = false>
__device__ __forceinline__ float init_tile_sse(const float* tg, const float* cv) {
    if constexpr (TS == 64) {
        const float a = init_chunk<128, EMPTY>(tg, cv, 0, 64);
        const float b = init_chunk<128, EMPTY>(tg, cv, 8192, 32);
        return (0.0f + a) + b;
    } else {
        return init_chunk<96, EMPTY>(tg, cv, 0, (3 * TS * TS) / 96);
    }
}
extern "C" __global__ __launch_bounds__(256) void weft_fit_order_kernel(WeftFitOrderArgs A) { fit_order_body(A); }

// Fast variants: candidate kinds 0, 4, 6, 7, 8 (hard-edge coverage over the cached canvas).
extern "C" __global__ __launch_bounds__(1024) void weft_fit_score_ts8(WeftFitScoreArgs A) { fit_score_body<8, false>(A); }
extern "C" __global__ __launch_bounds__(1024) void weft_fit_score_ts16(WeftFitScoreArgs A) { fit_score_body<16, false>(A); }
extern "C" __global__ __launch_bounds__(1024) void weft_fit_score_ts32(WeftFitScoreArgs A) { fit_score_body<32, false>(A); }
extern "C" __global__ __launch_bounds__(1024) void weft_fit_score_ts64(WeftFitScoreArgs A) { fit_score_body<64, false>(A); }
// All tile sizes in one launch: each work item's size comes from its tile.
template <bool GEN>
__device__ __forceinline__ void fit_score_mixed(const WeftFitScoreArgs& A) {
    switch (A.st.d_tile_ts[A.wk.d_work_tile[blockIdx.x]]) {
    case 8: fit_score_body<8, GEN>(A); break;
    case 16: fit_score_body<16, GEN>(A); break;
    case 32: fit_score_body<32, GEN>(A); break;
    case 64: fit_score_body<64, GEN>(A); break;
    default: break;
    }
}
extern "C" __global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void weft_fit_score_mixed(WeftFitScoreArgs A) { fit_score_mixed<false>(A); }
extern "C" __global__ __launch_bounds__(1024) void weft_fit_score_gen_mixed(WeftFitScoreArgs A) { fit_score_mixed<true>(A); }
extern "C" __global__ __launch_bounds__(1024) void weft_fit_score_combine_kernel(WeftFitScoreArgs A) {
    fit_score_combine_body(A);
}

// Variants that also accept kinds 1-3 (linear gradient, Gaussian line, Gaussian curve).
extern "C" __global__ __launch_bounds__(1024) void weft_fit_score_gen_ts8(WeftFitScoreArgs A) { fit_score_body<8, true>(A); }
extern "C" __global__ __launch_bounds__(1024) void weft_fit_score_gen_ts16(WeftFitScoreArgs A) { fit_score_body<16, true>(A); }
extern "C" __global__ __launch_bounds__(1024) void weft_fit_score_gen_ts32(WeftFitScoreArgs A) { fit_score_body<32, true>(A); }
extern "C" __global__ __launch_bounds__(1024) void weft_fit_score_gen_ts64(WeftFitScoreArgs A) { fit_score_body<64, true>(A); }

// One workgroup per tile: pick the winner, accept, composite, update the state.
extern "C" __global__ __launch_bounds__(256) void weft_fit_select_kernel(WeftFitSelectArgs A) {
    __shared__ double s_obj;
    __shared__ int s_idx;
    __shared__ int s_acc;
    const int t = blockIdx.x;
    const int tid = threadIdx.x;
    const WeftFitState& st = A.st;
    if (tile_skipped(st, A.cd, t)) {
        if (tid == 0) {
            A.win_idx[t] = -1;
            A.win_obj[t] = __builtin_huge_val();
            A.win_acc[t] = 0;
        }
        return;
    }
    if (tid == 0) {
        double bo = __builtin_huge_val();
        int bi = 0x7fffffff;
        const int w0 = A.tile_work0[t], nw = A.tile_nwork[t];
        const int ws = A.tile_wstride ? A.tile_wstride[t] : 1;
        for (int j = 0, w = w0; j < nw; ++j, w += ws) {
            const int i2 = A.part_idx[w];
            if (i2 >= 0) argmin_pair(bo, bi, A.part_obj[w], i2);
        }
        const bool found = bi != 0x7fffffff;
        s_idx = found ? bi : -1;
        s_obj = bo;
        s_acc = found && ((A.cd.flags & WEFT_CANDS_FORCE_ACCEPT) != 0 || bo < st.d_tile_obj[t]);
    }
    __syncthreads();
    const int bi = s_idx;
    const int acc = s_acc;
    const int ts = st.d_tile_ts[t];
    WeftPrim P;
    P.kind = -1;
    if (acc) {
        if (A.cd.d_prims == nullptr) {
            const WeftCode code = draw_code(draw_base(A.cd.seed, A.cd.step, (uint32_t)t), (uint32_t)bi, ts);
            P = code_to_prim(code);
            if (tid == 0 && A.win_code) A.win_code[t] = code;
        } else {
            const int64_t cb = A.cd.d_tile_cand ? A.cd.d_tile_cand[t] : (int64_t)t * A.cd.n_cand;
            P = load_prim(A.cd.d_prims + cb + bi);
            if (A.cd.flags & WEFT_CANDS_REMOVE_WINNER) {
                // the winner leaves the pool: records bi+1 .. n-1 move down one place, in
                // chunks of one 32-bit word per thread (every thread has read the winner and
                // each chunk is read in full before any of it is written)
                __syncthreads();
                const int n = A.cd.d_tile_ncand[t];
                uint32_t* pool = reinterpret_cast<uint32_t*>(const_cast<WeftPrim*>(A.cd.d_prims) + cb);
                constexpr int RW = (int)(sizeof(WeftPrim) / 4);
                const int w0 = bi * RW, w1 = (n - 1) * RW;
                for (int c = w0; c < w1; c += blockDim.x) {
                    const int i = c + tid;
                    const uint32_t v = i < w1 ? pool[i + RW] : 0u;
                    __syncthreads();
                    if (i < w1) pool[i] = v;
                    __syncthreads();
                }
                if (tid == 0) {
                    A.cd.d_tile_ncand[t] = n - 1;
                    // the accepted stack grows down from the top of the tile's pool region
                    const_cast<WeftPrim*>(A.cd.d_prims)[cb + A.cd.n_cand - 1 - st.d_tile_nsel[t]] = P;
                }
            }
        }
        float* cv = st.d_canvas + st.d_tile_pix[t] * 3;
        for (int p = tid; p < ts * ts; p += blockDim.x) {
            const int y = p / ts, x = p - y * ts;
            float r = cv[3 * p], g = cv[3 * p + 1], b = cv[3 * p + 2];
            composite_px(P, (float)x + 0.5f, (float)y + 0.5f, r, g, b);
            cv[3 * p] = r;
            cv[3 * p + 1] = g;
            cv[3 * p + 2] = b;
        }
    }
    if (tid == 0) {
        A.win_idx[t] = bi;
        A.win_obj[t] = s_obj;
        A.win_acc[t] = acc;
        if (acc) {
            st.d_tile_nsel[t] += 1;
            st.d_tile_bits[t] += prim_bytes(P.kind) * 8;
            st.d_tile_obj[t] = s_obj;
        }
    }
}

// Round control of the reference's batched greedy loop: one workgroup over all tiles.
extern "C" __global__ __launch_bounds__(1024) void weft_adaptive_gate_kernel(WeftAdaptGateArgs A) {
    __shared__ int s_sum[2][16];
    const WeftAdaptive& ad = A.ad;
    const WeftFitState& st = A.st;
    int n_act = 0, n_acc = 0;
    for (int t = threadIdx.x; t < ad.n_tiles; t += blockDim.x) {
        int cx = ad.d_complex[t];
        const int ncand = ad.d_tile_ncand[t];
        const bool room = st.d_tile_nsel[t] < ad.d_max_prims[t];
        if (A.mode == WEFT_ADAPT_START) cx = st.d_tile_obj[t] > ad.threshold && ncand > 0;
        if (A.mode == WEFT_ADAPT_KEEP) cx = ad.d_win_acc[t] != 0;
        if (A.mode == WEFT_ADAPT_COUNT) {
            n_acc += ad.d_win_acc[t] != 0;
            continue;
        }
        ad.d_complex[t] = cx;
        const int act = A.mode == WEFT_ADAPT_RESEED ? (cx && room) : (cx && ncand > 0 && room);
        ad.d_active[t] = act;
        n_act += act;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        n_act += __shfl_xor(n_act, off, 64);
        n_acc += __shfl_xor(n_acc, off, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_sum[0][wave] = n_act;
        s_sum[1][wave] = n_acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0, b = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            a += s_sum[0][w];
            b += s_sum[1][w];
        }
        if (A.mode == WEFT_ADAPT_COUNT) ad.d_counts[1] = b;
        else ad.d_counts[0] = a;
        if (A.mode == WEFT_ADAPT_RESEED) ad.d_counts[1] = 0;  // records the append drops
    }
}

// Re-seed append: one workgroup per list entry, its new records behind the tile's live ones.
extern "C" __global__ __launch_bounds__(256) void weft_adaptive_append_kernel(WeftAdaptAppendArgs A) {
    const int i = blockIdx.x;
    if (i >= A.n) return;
    const int t = A.tiles[i];
    if (!A.ad.d_active[t]) return;
    const int n0 = A.ad.d_tile_ncand[t];
    const int want = A.fresh_count[i];
    const int room = A.pool_cap - n0 - A.st.d_tile_nsel[t];  // the accepted stack holds the top
    const int k = max(0, want < room ? want : room);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(A.fresh + (int64_t)i * A.max_new);
    uint32_t* dst = reinterpret_cast<uint32_t*>(A.pools + (int64_t)t * A.pool_cap + n0);
    constexpr int RW = (int)(sizeof(WeftPrim) / 4);
    for (int w = threadIdx.x; w < k * RW; w += blockDim.x) dst[w] = src[w];
    if (threadIdx.x == 0) {
        A.ad.d_tile_ncand[t] = n0 + k;
        if (want > k) atomicAdd(&A.ad.d_counts[1], want - k);
    }
}

extern "C" __global__ __launch_bounds__(256) void weft_batch_fit_kernel(WeftBatchArgs A) { batch_fit_body(A); }

// Resets the queue counter, start time and per-image counters of a batched fit.
extern "C" __global__ __launch_bounds__(256) void weft_batch_reset_kernel(WeftBatchArgs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        A.queue[0] = 0;
        if (A.t_start) A.t_start[0] = ~0ull;
    }
    if (A.image_left && i < A.n_images) {
        A.image_left[i] = A.image_tiles[i];
        A.image_done[i] = 0ull;
    }
}

// Candidate codes / records as the fit kernels draw them.
extern "C" __global__ __launch_bounds__(256) void weft_draw_kernel(WeftDrawArgs A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)A.n_tiles * A.n_cand) return;
    const int j = (int)(i / A.n_cand);
    const int k = (int)(i - (int64_t)j * A.n_cand);
    const WeftCode c = draw_code(draw_base(A.seed, A.step, (uint32_t)A.tile_ids[j]), (uint32_t)k, A.tile_ts[j]);
    if (A.codes) A.codes[i] = c;
    if (A.prims) {
        const WeftPrim P = code_to_prim(c);
        WeftPrim* o = A.prims + i;
        o->kind = P.kind;
        for (int q = 0; q < 7; ++q) o->geom[q] = P.geom[q];
        for (int q = 0; q < 3; ++q) {
            o->color0[q] = P.color0[q];
            o->color1[q] = P.color1[q];
        }
        o->alpha = P.alpha;
    }
}

// RES1: bilinear weights of the g x g residual grid along one axis of a tile rendered at
// `up` pixels (upsample_residual_map, src/weft/render.py:474-515).  The sample position
// is float32 (pixel centre times the float32 grid step, minus 0.5, clamped to the grid);
// its fraction is float64, as NumPy promotes float32 minus int32.
struct Res1Axis {
    int i0, i1;
    double f;
};

__device__ __forceinline__ Res1Axis res1_axis(int i, float step, int g) {
    const float raw = ((float)i + 0.5f) * step - 0.5f;
    const float cl = fminf(fmaxf(raw, 0.0f), (float)(g - 1));
    const int i0 = (int)floorf(cl);
    return Res1Axis{i0, min(i0 + 1, g - 1), (double)cl - (double)i0};
}

// Decode: one workgroup per tile, threads over the tile's output pixels.  Plain mode
// (weft_render_tiles) samples the source square with per-axis steps source / target_w and
// source / target_h, as the reference decode kernel does; scene mode (weft_render_scene)
// renders a square of desc._pad pixels with step source / _pad on both axes and keeps its
// top-left target_w x target_h, as render_scene_adaptive renders then crops
// (src/weft/render.py:322-345), then adds the tile's RES1 grid when it has one.
extern "C" __global__ __launch_bounds__(256) void weft_render_tiles_kernel(WeftRenderArgs A) {
    const int t = blockIdx.x;
    if (t >= A.n_tiles) return;
    const WeftTileDecodeDesc d = A.descs[t];
    const WeftPrim* pr = A.prims + d.prim_offset;
    const int sx = A.scene ? d._pad : d.target_w;
    const int sy = A.scene ? d._pad : d.target_h;
    const float stepx = d.source_size == sx ? 1.0f : (float)((double)d.source_size / (double)sx);
    const float stepy = d.source_size == sy ? 1.0f : (float)((double)d.source_size / (double)sy);
    const bool res1 = A.res1 != nullptr && t < A.n_res1;
    const int g = A.res1_grid;
    const float* map = res1 ? A.res1 + (int64_t)t * g * g * 3 : nullptr;
    const float ustep = res1 ? (float)((double)g / (double)d.target_h) : 0.0f;
    const int npx = d.target_w * d.target_h;
    for (int p = threadIdx.x; p < npx; p += blockDim.x) {
        const int ly = p / d.target_w, lx = p - ly * d.target_w;
        const float fx = ((float)lx + 0.5f) * stepx;
        const float fy = ((float)ly + 0.5f) * stepy;
        float r = 0.0f, g3 = 0.0f, b = 0.0f;
        for (int i = 0; i < d.prim_count; ++i) composite_px(pr[i], fx, fy, r, g3, b);
        float v[3] = {clip01(r + d.res0_r), clip01(g3 + d.res0_g), clip01(b + d.res0_b)};
        if (res1) {
            const Res1Axis ax = res1_axis(lx, ustep, g), ay = res1_axis(ly, ustep, g);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double c00 = map[(ay.i0 * g + ax.i0) * 3 + c], c01 = map[(ay.i0 * g + ax.i1) * 3 + c];
                const double c10 = map[(ay.i1 * g + ax.i0) * 3 + c], c11 = map[(ay.i1 * g + ax.i1) * 3 + c];
                const double top = c00 * (1.0 - ax.f) + c01 * ax.f;
                const double bot = c10 * (1.0 - ax.f) + c11 * ax.f;
                const double up = top * (1.0 - ay.f) + bot * ay.f;
                const double s = (double)v[c] + up;
                v[c] = (float)(s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s));  // np.clip keeps -0.0
            }
        }
        const int gx = d.target_x0 + lx, gy = d.target_y0 + ly;
        if (gx >= 0 && gx < A.width && gy >= 0 && gy < A.height) {
            const int64_t o = ((int64_t)gy * A.width + gx) * 3;
            A.out[o] = v[0];
            A.out[o + 1] = v[1];
            A.out[o + 2] = v[2];
        }
    }
}

// Soft-edge SSE and its gradient for groups of k candidates (one workgroup each: 64 lanes, or
// 64 W lanes for one candidate over W waves).
#ifndef WEFT_SOFT_MIN_WAVES
#define WEFT_SOFT_MIN_WAVES 1
#endif
extern "C" __global__ __launch_bounds__(512, WEFT_SOFT_MIN_WAVES) void weft_soft_eval_kernel(WeftSoftArgs A) {
    soft_eval_body(A);
}

// Per-tile top-k indices of a value table (one workgroup per tile).
extern "C" __global__ __launch_bounds__(256) void weft_topk_kernel(WeftTopkArgs A) { topk_body(A); }

// Adam refinement of each tile's top-k candidates (one 64-lane workgroup per work item), and
// the form for a work split that spreads a size's candidates over W waves (one workgroup of
// 64 W lanes per item).
extern "C" __global__ __launch_bounds__(64, WEFT_SOFT_MIN_WAVES) void weft_refine_kernel(WeftRefineArgs A) {
    refine_body<false>(A);
}
extern "C" __global__ __launch_bounds__(512, WEFT_SOFT_MIN_WAVES) void weft_refine_mw_kernel(WeftRefineArgs A) {
    refine_body<true>(A);
}

// Exact scores of a few candidates per tile, each candidate's pixels split over lanes.
extern "C" __global__ __launch_bounds__(64) void weft_fit_score_split_kernel(WeftSplitScoreArgs A) { split_score_body(A); }

// Chunked PRIM entropy coding.
extern "C" __global__ __launch_bounds__(256) void weft_rans_encode_kernel(WeftRansEncodeArgs A) { rans_encode_body(A); }
extern "C" __global__ __launch_bounds__(1024) void weft_rans_scan_kernel(WeftRansEncodeArgs A) { rans_scan_body(A); }
extern "C" __global__ __launch_bounds__(256) void weft_rans_pack_kernel(WeftRansEncodeArgs A) { rans_pack_body(A); }
extern "C" __global__ __launch_bounds__(64) void weft_rans_decode_kernel(WeftRansDecodeArgs A) { rans_decode_body(A); }

// Persistent refined fit (tiles of many images, every round refined and re-scored).
#ifndef WEFT_BATCH_REFINE_MIN_WAVES
// Waves per SIMD the register allocation must allow: 4 caps the kernel at 128 VGPRs (the
// refinement phase spills a little to scratch), which runs the whole fit faster than the
// uncapped 201 VGPRs at 2 waves per SIMD.
#define WEFT_BATCH_REFINE_MIN_WAVES 4
#endif
extern "C" __global__ __launch_bounds__(256, WEFT_BATCH_REFINE_MIN_WAVES) void weft_batch_refine_kernel(
    WeftBatchRefineArgs A) {
    batch_refine_body(A);
}

// One wave per tile: the objective of its canvas.  With A.reset the tile first goes back to
// an empty canvas and stack (canvas zero, nsel 0, bits A.bits0) and the objective is that of
// the empty canvas, which is not read back.
template <bool RESET>
__device__ __forceinline__ void fit_init_body(const WeftFitInitArgs& A) {
    const int t = blockIdx.x;
    const WeftFitState& st = A.st;
    const int64_t base = st.d_tile_pix[t] * 3;
    const float* tg = st.d_target + base;
    float* cv = st.d_canvas + base;
    const int ts = st.d_tile_ts[t];
    if (RESET)
        for (int i = threadIdx.x; i < 3 * ts * ts; i += 64) cv[i] = 0.0f;
    float sse;
    switch (ts) {
    case 8: sse = init_tile_sse<8, RESET>(tg, cv); break;
    case 16: sse = init_tile_sse<16, RESET>(tg, cv); break;
    case 32: sse = init_tile_sse<32, RESET>(tg, cv); break;
    case 64: sse = init_tile_sse<64, RESET>(tg, cv); break;
    default: sse = __builtin_nanf(""); break;
    }
    if (threadIdx.x == 0) {
        const int bits = RESET ? A.bits0 : st.d_tile_bits[t];
        if (RESET) {
            st.d_tile_nsel[t] = 0;
            st.d_tile_bits[t] = bits;
        }
        const float mse = sse / (float)(3 * ts * ts);
        st.d_tile_obj[t] = objective(mse, bits, st.lam, st.obj_mode);
    }
}

extern "C" __global__ __launch_bounds__(64) void weft_fit_init_kernel(WeftFitInitArgs A) {
    if (A.reset) fit_init_body<true>(A);
    else fit_init_body<false>(A);
}

// Drop-in stack scoring, each eval's pixels split over lanes (one tile size per launch).
extern "C" __global__ __launch_bounds__(64) void weft_score_evals_split_kernel(WeftScoreEvalsSplitArgs A) {
    switch (A.tile_size) {
    case 8: score_evals_split_item<8>(A); break;
    case 16: score_evals_split_item<16>(A); break;
    case 32: score_evals_split_item<32>(A); break;
    case 64: score_evals_split_item<64>(A); break;
    default: break;
    }
}

// Stage-3 template fit: bucket table of one tile size, then one workgroup per tile.
extern "C" __global__ __launch_bounds__(256) void weft_template_masks_kernel(WeftTemplateMaskArgs A) {
    tmpl_masks_body(A);
}
extern "C" __global__ __launch_bounds__(256) void weft_template_fit_kernel(WeftTemplateFitArgs A) { tmpl_fit_body(A); }

// Selection primitives (argmin, exclusive scan, compaction, segmented top-k).
extern "C" __global__ __launch_bounds__(256) void weft_sel_argmin_partial_kernel(WeftSelectArgs A) {
    sel_argmin_partial_body(A);
}
extern "C" __global__ __launch_bounds__(256) void weft_sel_argmin_final_kernel(WeftSelectArgs A) {
    sel_argmin_final_body(A);
}
extern "C" __global__ __launch_bounds__(256) void weft_sel_tile_sums_kernel(WeftSelectArgs A) { sel_tile_sums_body(A); }
extern "C" __global__ __launch_bounds__(256) void weft_sel_scan_tiles_kernel(WeftSelectArgs A) {
    sel_scan_tiles_body(A);
}
extern "C" __global__ __launch_bounds__(256) void weft_sel_scan_apply_kernel(WeftSelectArgs A) {
    sel_scan_apply_body(A);
}
extern "C" __global__ __launch_bounds__(256) void weft_sel_rs_score_keys_kernel(WeftSelectArgs A) {
    sel_rs_score_keys_body(A);
}
extern "C" __global__ __launch_bounds__(256) void weft_sel_rs_hist_kernel(WeftSelectArgs A) { sel_rs_hist_body(A); }
extern "C" __global__ __launch_bounds__(256) void weft_sel_rs_scatter_kernel(WeftSelectArgs A) {
    sel_rs_scatter_body(A);
}
extern "C" __global__ __launch_bounds__(256) void weft_sel_rs_segment_keys_kernel(WeftSelectArgs A) {
    sel_rs_segment_keys_body(A);
}
extern "C" __global__ __launch_bounds__(256) void weft_sel_rs_keep_kernel(WeftSelectArgs A) { sel_rs_keep_body(A); }
extern "C" __global__ __launch_bounds__(256) void weft_sel_rs_emit_kernel(WeftSelectArgs A) { sel_rs_emit_body(A); }

// Template candidate pools, one workgroup per tile of one size.
extern "C" __global__ __launch_bounds__(256) void weft_gen_candidates_kernel(WeftCandGenArgs A) { cg_body(A); }
extern "C" __global__ __launch_bounds__(256) void weft_gen_residual_candidates_kernel(WeftResGenArgs A) {
    cr_body(A);
}
// Edge-driven candidate pools, one workgroup per tile of one size.
extern "C" __global__ __launch_bounds__(256) void weft_gen_edge_candidates_kernel(WeftEdgeGenArgs A) { ed_body(A); }
extern "C" __global__ __launch_bounds__(256) void weft_merge_pools_kernel(WeftMergePoolsArgs A) { mg_body(A); }
