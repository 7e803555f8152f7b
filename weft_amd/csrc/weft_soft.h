// Soft-edge surrogate of the three drawn candidate families, its hand-written
// gradient, and the per-tile Adam refinement built on it.  Included by
// weft_kernels.hip inside its anonymous namespace (uses clip01, draw_code,
// code_to_prim and the candidate tables defined there).
//
// Parameters (9 floats): cx, cy, rx, ry, theta, r, g, b, a.  The hard-edge kinds:
//   ellipse (6):   p = (c dx + s dy) / rx, q = (c dy - s dx) / ry, inside iff p^2+q^2 <= 1
//   rectangle (7): inside iff |p| <= 1 and |q| <= 1
//   triangle (8):  vertices C + Rot(theta) (rx ux_j, ry uy_j), (ux, uy) at -90, 30, 150 deg
// with (dx, dy) = pixel centre - (cx, cy) and c, s = cos, sin theta.  The soft
// coverage m replaces the indicator:
//   ellipse:   m = sig((1 - rho) sqrt(rx ry) / w), rho = sqrt(p^2 + q^2)
//              ((1 - rho) sqrt(rx ry) is the signed distance for a circle)
//   polygons:  m = prod_i sig(d_i / w), d_i the signed distance to edge i
//              (rectangle: rx -+ u, ry -+ v in the rotated frame; triangle: its 3 edges)
// and the pixel is out_k = s_k A + canvas_k (1 - A), A = a m, pred = clip(out, 0, 1),
// SSE = sum_k,px (pred_k - target_k)^2.
//
// The backward pass accumulates per lane a few sums over the pixels (of dSSE/dd times
// 1, dx, dy for each edge; of the ellipse's radial terms) and applies the chain rule
// to the parameters once per candidate.

#ifndef WEFT_SOFT_UNROLL
// Pixels in flight per lane (memory-level parallelism vs registers).  Measured on the
// 4096-tile refinement, one is the fastest setting (20.0 ms against 21.5 ms with 2 and
// 24.4 ms with 4).
#define WEFT_SOFT_UNROLL 1
#endif

constexpr int SOFT_UNROLL = WEFT_SOFT_UNROLL;
constexpr int SOFT_NP = 9;
constexpr float SOFT_PI = 3.14159265358979323846f;

struct SoftCand {
    int kind;
    float p[SOFT_NP];
};

struct SoftPre {
    float c, s;
    float R, irx, iry;                 // ellipse
    float nx[4], ny[4], e[4];          // polygon edges: d = nx dx + ny dy + e
    float wx[3], wy[3];                // triangle vertex offsets from the centre
};

struct SoftAcc {
    float sse, gc0, gc1, gc2, ga;
    float q[12];
};

__device__ __forceinline__ float soft_sig(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

__device__ __forceinline__ void soft_pre(const SoftCand& C, SoftPre& P) {
    const float rx = C.p[2], ry = C.p[3];
    __sincosf(C.p[4], &P.s, &P.c);  // theta in [0, pi]: the hardware sine and cosine
    const float c = P.c, s = P.s;
    P.R = sqrtf(rx * ry);
    P.irx = 1.0f / rx;
    P.iry = 1.0f / ry;
    // Polygons use four edges; the triangle's fourth is always inside (coverage factor
    // exactly 1, gradient 0), so every polygon runs the same code.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        P.nx[i] = P.ny[i] = 0.0f;
        P.e[i] = 1e6f;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) P.wx[j] = P.wy[j] = 0.0f;
    if (C.kind == 7) {
        P.nx[0] = -c; P.ny[0] = -s; P.e[0] = rx;   // rx - u
        P.nx[1] = c;  P.ny[1] = s;  P.e[1] = rx;   // rx + u
        P.nx[2] = s;  P.ny[2] = -c; P.e[2] = ry;   // ry - v
        P.nx[3] = -s; P.ny[3] = c;  P.e[3] = ry;   // ry + v
    } else if (C.kind == 8) {
        const float ux[3] = {0.0f, 0.8660254037844386f, -0.8660254037844386f};
        const float uy[3] = {-1.0f, 0.5f, 0.5f};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float lx = rx * ux[j], ly = ry * uy[j];
            P.wx[j] = c * lx - s * ly;
            P.wy[j] = s * lx + c * ly;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int b = i == 2 ? 0 : i + 1;
            const float ex = P.wx[b] - P.wx[i], ey = P.wy[b] - P.wy[i];
            const float il = 1.0f / sqrtf(ex * ex + ey * ey);
            P.nx[i] = -ey * il;
            P.ny[i] = ex * il;
            P.e[i] = -(P.nx[i] * P.wx[i] + P.ny[i] * P.wy[i]);
        }
    }
}

__device__ __forceinline__ void soft_acc_zero(SoftAcc& A) {
    A.sse = A.gc0 = A.gc1 = A.gc2 = A.ga = 0.0f;
#pragma unroll
    for (int i = 0; i < 12; ++i) A.q[i] = 0.0f;
}

// One pixel: (dx, dy) from the candidate centre, canvas cv, target tg.  The surrogate is
// this build's own definition (no reference arithmetic to match), so its multiply-adds are
// fused explicitly with fmaf: the same instructions wherever it is inlined.
template <bool GRAD>
__device__ __forceinline__ void soft_px(const SoftCand& C, const SoftPre& P, float invw, float dx,
                                        float dy, float cv0, float cv1, float cv2, float tg0,
                                        float tg1, float tg2, SoftAcc& A) {
    float m, rho = 0.0f, pp = 0.0f, qq = 0.0f;
    float sg[4];
    if (C.kind == 6) {
        pp = fmaf(P.c, dx, P.s * dy) * P.irx;
        qq = fmaf(P.c, dy, -(P.s * dx)) * P.iry;
        rho = sqrtf(fmaf(pp, pp, qq * qq));
        m = soft_sig((1.0f - rho) * P.R * invw);
        sg[0] = m;
        sg[1] = sg[2] = sg[3] = 1.0f;
    } else {
        // three edges for every polygon; the rectangle's fourth only for rectangles (the
        // triangle's fourth edge is always inside: factor 1, gradient 0)
#pragma unroll
        for (int i = 0; i < 3; ++i) sg[i] = soft_sig(fmaf(P.nx[i], dx, fmaf(P.ny[i], dy, P.e[i])) * invw);
        m = (sg[0] * sg[1]) * sg[2];
        sg[3] = 1.0f;
        if (C.kind == 7) {
            sg[3] = soft_sig(fmaf(P.nx[3], dx, fmaf(P.ny[3], dy, P.e[3])) * invw);
            m = m * sg[3];
        }
    }
    const float a = C.p[8];
    const float Aa = a * m;
    const float om = 1.0f - Aa;
    const float o0 = fmaf(C.p[5], Aa, cv0 * om);
    const float o1 = fmaf(C.p[6], Aa, cv1 * om);
    const float o2 = fmaf(C.p[7], Aa, cv2 * om);
    const float e0 = clip01(o0) - tg0, e1 = clip01(o1) - tg1, e2 = clip01(o2) - tg2;
    A.sse = fmaf(e0, e0, fmaf(e1, e1, fmaf(e2, e2, A.sse)));
    if constexpr (GRAD) {
        const float g0 = (o0 >= 0.0f && o0 <= 1.0f) ? 2.0f * e0 : 0.0f;
        const float g1 = (o1 >= 0.0f && o1 <= 1.0f) ? 2.0f * e1 : 0.0f;
        const float g2 = (o2 >= 0.0f && o2 <= 1.0f) ? 2.0f * e2 : 0.0f;
        A.gc0 = fmaf(g0, Aa, A.gc0);
        A.gc1 = fmaf(g1, Aa, A.gc1);
        A.gc2 = fmaf(g2, Aa, A.gc2);
        const float G = fmaf(g0, C.p[5] - cv0, fmaf(g1, C.p[6] - cv1, g2 * (C.p[7] - cv2)));
        A.ga = fmaf(G, m, A.ga);
        const float H = G * a;  // d SSE / d m
        // the pixel's terms first, then one add per sum (stores in the two branches
        // would be merged by index and push the sums out of registers)
        float c[12];
        if (C.kind == 6) {
            const float h = H * m * (1.0f - m) * invw;    // d SSE / d (signed distance)
            const float K = -h * P.R / fmaxf(rho, 1e-30f);
            const float Kp = K * pp, Kq = K * qq;
            c[0] = Kp;
            c[1] = Kq;
            c[2] = Kp * pp;
            c[3] = Kq * qq;
            c[4] = Kp * qq;
            c[5] = h * (1.0f - rho);
#pragma unroll
            for (int i = 6; i < 12; ++i) c[i] = 0.0f;
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float h = H * m * (1.0f - sg[i]) * invw;  // d SSE / d d_i
                c[3 * i] = h;
                c[3 * i + 1] = h * dx;
                c[3 * i + 2] = h * dy;
            }
        }
#pragma unroll
        for (int i = 0; i < 12; ++i) A.q[i] += c[i];
    }
}

// Chain rule from the accumulated sums to the 9 parameters.
__device__ __forceinline__ void soft_finish(const SoftCand& C, const SoftPre& P, const SoftAcc& A,
                                            float* g) {
    const float c = P.c, s = P.s, rx = C.p[2], ry = C.p[3];
    g[5] = A.gc0;
    g[6] = A.gc1;
    g[7] = A.gc2;
    g[8] = A.ga;
    if (C.kind == 6) {
        g[0] = -c * P.irx * A.q[0] + s * P.iry * A.q[1];
        g[1] = -s * P.irx * A.q[0] - c * P.iry * A.q[1];
        g[2] = -A.q[2] * P.irx + A.q[5] * P.R * 0.5f * P.irx;
        g[3] = -A.q[3] * P.iry + A.q[5] * P.R * 0.5f * P.iry;
        g[4] = A.q[4] * (ry * P.irx - rx * P.iry);
        return;
    }
    float gx = 0.0f, gy = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        gx -= P.nx[i] * A.q[3 * i];
        gy -= P.ny[i] * A.q[3 * i];
    }
    g[0] = gx;
    g[1] = gy;
    if (C.kind == 7) {
        g[2] = A.q[0] + A.q[3];
        g[3] = A.q[6] + A.q[9];
        // d d_i / d theta = (d n_i / d theta) . (dx, dy)
        g[4] = (s * A.q[1] - c * A.q[2]) + (-s * A.q[4] + c * A.q[5]) + (c * A.q[7] + s * A.q[8]) +
               (-c * A.q[10] - s * A.q[11]);
        return;
    }
    // triangle: edges (i -> i+1) with n = perp(E)/|E|, e = -n . W_i
    float gwx[3] = {0.0f, 0.0f, 0.0f}, gwy[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int b = i == 2 ? 0 : i + 1;
        const float ex = P.wx[b] - P.wx[i], ey = P.wy[b] - P.wy[i];
        const float l2 = ex * ex + ey * ey;
        const float il = 1.0f / sqrtf(l2);
        const float S = A.q[3 * i];
        const float px = A.q[3 * i + 1] - S * P.wx[i];   // d SSE / d n
        const float py = A.q[3 * i + 2] - S * P.wy[i];
        const float pperp = -px * ey + py * ex;           // P . perp(E)
        const float il3 = il * il * il;
        const float qx = py * il - pperp * ex * il3;      // d SSE / d E
        const float qy = -px * il - pperp * ey * il3;
        gwx[b] += qx;
        gwy[b] += qy;
        gwx[i] -= qx + S * P.nx[i];
        gwy[i] -= qy + S * P.ny[i];
    }
    const float ux[3] = {0.0f, 0.8660254037844386f, -0.8660254037844386f};
    const float uy[3] = {-1.0f, 0.5f, 0.5f};
    float grx = 0.0f, gry = 0.0f, gth = 0.0f;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        grx += ux[j] * (gwx[j] * c + gwy[j] * s);
        gry += uy[j] * (-gwx[j] * s + gwy[j] * c);
        gth += -gwx[j] * P.wy[j] + gwy[j] * P.wx[j];
    }
    g[2] = grx;
    g[3] = gry;
    g[4] = gth;
}

// One level of the butterfly sum: v + (v of lane i ^ OFF), the levels taken in the order
// OFF = 1, 2, 4, ...  After the levels below OFF, every lane of an aligned group of OFF
// lanes holds the same value, so any partner in the other half of the aligned 2·OFF group
// gives the same sum, bit for bit.  The quad permutations (OFF 1, 2), the row's half
// mirror (i <-> 7 - i, OFF 4) and its mirror (i <-> 15 - i, OFF 8) are such partners and
// run as DPP operand modifiers; OFF 16 and 32 go through the LDS crossbar.
template <int OFF>
__device__ __forceinline__ float bfly_add(float v) {
    if constexpr (OFF <= 8) {
        constexpr int ctrl = OFF == 1 ? 0xB1 : (OFF == 2 ? 0x4E : (OFF == 4 ? 0x141 : 0x140));
        return v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, false));
    } else {
        return v + __shfl_xor(v, OFF, 64);
    }
}

// v summed over its aligned group of L lanes (L a power of two, uniform over the wave) by
// the butterfly of bfly_add; every lane of the group ends with the same total.
__device__ __forceinline__ float lane_group_sum(float v, int L) {
    if (L > 1) v = bfly_add<1>(v);
    if (L > 2) v = bfly_add<2>(v);
    if (L > 4) v = bfly_add<4>(v);
    if (L > 8) v = bfly_add<8>(v);
    if (L > 16) v = bfly_add<16>(v);
    if (L > 32) v = bfly_add<32>(v);
    return v;
}

template <int OFF>
__device__ __forceinline__ void soft_acc_level(SoftAcc& A) {
    A.sse = bfly_add<OFF>(A.sse);
    A.gc0 = bfly_add<OFF>(A.gc0);
    A.gc1 = bfly_add<OFF>(A.gc1);
    A.gc2 = bfly_add<OFF>(A.gc2);
    A.ga = bfly_add<OFF>(A.ga);
#pragma unroll
    for (int i = 0; i < 12; ++i) A.q[i] = bfly_add<OFF>(A.q[i]);
}

// Sums over the lanes of one candidate (the L lanes i with i / L equal, L a power of two,
// uniform over the wave); every lane ends with the same totals.  Level by level over all
// seventeen sums, so the crossbar latencies of a level overlap.
__device__ __forceinline__ void soft_acc_reduce(SoftAcc& A, int L) {
    if (L > 1) soft_acc_level<1>(A);
    if (L > 2) soft_acc_level<2>(A);
    if (L > 4) soft_acc_level<4>(A);
    if (L > 8) soft_acc_level<8>(A);
    if (L > 16) soft_acc_level<16>(A);
    if (L > 32) soft_acc_level<32>(A);
}

// Waves per candidate at most (a candidate's lanes span W = 2, 4 or 8 waves of one workgroup).
constexpr int SOFT_MAX_WAVES = 8;

// The butterfly's levels above one wave, for a candidate whose L = 64 W lanes are the W waves
// of the workgroup: after soft_acc_reduce every lane of wave w holds its wave's totals s_w;
// level 64 forms s_w + s_(w^1), level 128 adds the partner pair's sum, level 256 the partner
// quad's, so every wave ends with the same bits (float addition commutes).  All W waves must
// call it; `parity` alternates between calls (two barriers per call already order the buffers'
// reuse; the alternation keeps consecutive calls' writes apart).
__device__ __forceinline__ void soft_acc_waves(SoftAcc& A, int W, int parity) {
    constexpr int NS = 17;
    __shared__ float s_part[2][SOFT_MAX_WAVES][NS];
    __shared__ float s_tot[2][SOFT_MAX_WAVES][NS];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (lane == 0) {
        float* mine = s_part[parity][wave];
        mine[0] = A.sse;
        mine[1] = A.gc0;
        mine[2] = A.gc1;
        mine[3] = A.gc2;
        mine[4] = A.ga;
#pragma unroll
        for (int i = 0; i < 12; ++i) mine[5 + i] = A.q[i];
    }
    __syncthreads();
    // lane i < 17 of each wave forms sum i over the aligned group of W waves holding the
    // candidate, level by level
    if (lane < NS) {
        const int g0 = wave & ~(W - 1);
        float lv[SOFT_MAX_WAVES];
#pragma unroll
        for (int w = 0; w < SOFT_MAX_WAVES; ++w) lv[w] = w < W ? s_part[parity][g0 + w][lane] : 0.0f;
#pragma unroll
        for (int off = 1; off < SOFT_MAX_WAVES; off <<= 1)
#pragma unroll
            for (int w = 0; w + off < SOFT_MAX_WAVES; w += 2 * off)
                if (off < W) lv[w] = lv[w] + lv[w + off];
        s_tot[parity][wave][lane] = lv[0];
    }
    __syncthreads();
    const float* t = s_tot[parity][wave];
    A.sse = t[0];
    A.gc0 = t[1];
    A.gc1 = t[2];
    A.gc2 = t[3];
    A.ga = t[4];
#pragma unroll
    for (int i = 0; i < 12; ++i) A.q[i] = t[5 + i];
}

// SSE (and sums for the gradient) of candidate C over tile pixels slice, slice + L, ...
// (ts and L powers of two: every lane takes ts * ts / L pixels, or none when L > ts * ts and
// slice >= ts * ts).
// With PREFETCH the next pixel's canvas and target are loaded before the current pixel is
// computed, so a lane keeps one pixel's loads in flight behind its arithmetic (six more
// registers; the persistent kernel, short of registers, goes without).  UNROLL pixels per
// loop trip.  The sums are accumulated in the same order in every variant.
template <bool GRAD, int UNROLL = SOFT_UNROLL, bool PREFETCH = true>
__device__ __forceinline__ void soft_tile(const SoftCand& C, const SoftPre& P, float invw,
                                          const float* tg, const float* cv, int ts, int slice, int L,
                                          SoftAcc& A) {
    soft_acc_zero(A);
    const int npx = ts * ts;
    const int lg = __builtin_ctz((unsigned)ts);
    if constexpr (PREFETCH) {
        const int p0 = slice < npx ? slice : 0;  // a lane past the tile (L > ts * ts) takes no pixel
        float n0 = cv[3 * p0], n1 = cv[3 * p0 + 1], n2 = cv[3 * p0 + 2];
        float m0 = tg[3 * p0], m1 = tg[3 * p0 + 1], m2 = tg[3 * p0 + 2];
#pragma unroll UNROLL
        for (int px = slice; px < npx; px += L) {
            const float c0 = n0, c1 = n1, c2 = n2, t0 = m0, t1 = m1, t2 = m2;
            const int pn = px + L < npx ? px + L : px;  // the last trip reloads its own pixel
            n0 = cv[3 * pn];
            n1 = cv[3 * pn + 1];
            n2 = cv[3 * pn + 2];
            m0 = tg[3 * pn];
            m1 = tg[3 * pn + 1];
            m2 = tg[3 * pn + 2];
            const int y = px >> lg, x = px & (ts - 1);
            const float dx = ((float)x + 0.5f) - C.p[0];
            const float dy = ((float)y + 0.5f) - C.p[1];
            soft_px<GRAD>(C, P, invw, dx, dy, c0, c1, c2, t0, t1, t2, A);
        }
    } else {
#pragma unroll UNROLL
        for (int px = slice; px < npx; px += L) {
            const int y = px >> lg, x = px & (ts - 1);
            const float dx = ((float)x + 0.5f) - C.p[0];
            const float dy = ((float)y + 0.5f) - C.p[1];
            soft_px<GRAD>(C, P, invw, dx, dy, cv[3 * px], cv[3 * px + 1], cv[3 * px + 2], tg[3 * px],
                          tg[3 * px + 1], tg[3 * px + 2], A);
        }
    }
    soft_acc_reduce(A, L);
}

// ---- conversions between codes and continuous parameters ------------------------
__device__ __forceinline__ SoftCand soft_from_code(const WeftCode& c) {
    SoftCand C;
    C.kind = 6 + (int)c.family;
    C.p[0] = (float)c.cx * (1.0f / 1024.0f);
    C.p[1] = (float)c.cy * (1.0f / 1024.0f);
    C.p[2] = (float)c.rx * (1.0f / 1024.0f);
    C.p[3] = (float)c.ry * (1.0f / 1024.0f);
    C.p[4] = (float)c.theta * (SOFT_PI / 4096.0f);
    C.p[5] = (float)((double)c.r / 255.0);
    C.p[6] = (float)((double)c.g / 255.0);
    C.p[7] = (float)((double)c.b / 255.0);
    C.p[8] = (float)((double)c.alpha / 65535.0);
    return C;
}

constexpr float SOFT_R_MIN = 1.0f;
constexpr float SOFT_R_MAX = 65535.0f / 1024.0f;
constexpr float SOFT_A_MIN = 19661.0f / 65535.0f;
constexpr float SOFT_TH_MAX = SOFT_PI * (4095.0f / 4096.0f);

// Keep the parameters inside what a code can hold (theta periodic for the centrally
// symmetric ellipse and rectangle, clamped for the triangle).
__device__ __forceinline__ void soft_project(SoftCand& C, int ts) {
    const float span = (float)ts - (1.0f / 1024.0f);
    C.p[0] = fminf(fmaxf(C.p[0], 0.0f), span);
    C.p[1] = fminf(fmaxf(C.p[1], 0.0f), span);
    C.p[2] = fminf(fmaxf(C.p[2], SOFT_R_MIN), SOFT_R_MAX);
    C.p[3] = fminf(fmaxf(C.p[3], SOFT_R_MIN), SOFT_R_MAX);
    if (C.kind == 8) {
        C.p[4] = fminf(fmaxf(C.p[4], 0.0f), SOFT_TH_MAX);
    } else {
        if (C.p[4] < 0.0f) C.p[4] = C.p[4] + SOFT_PI;
        if (C.p[4] >= SOFT_PI) C.p[4] = C.p[4] - SOFT_PI;
        C.p[4] = fminf(fmaxf(C.p[4], 0.0f), SOFT_PI);
    }
#pragma unroll
    for (int k = 5; k < 8; ++k) C.p[k] = clip01(C.p[k]);
    C.p[8] = fminf(fmaxf(C.p[8], SOFT_A_MIN), 1.0f);
}

__device__ __forceinline__ uint16_t soft_q(float v, float scale, float lo, float hi) {
    return (uint16_t)fminf(fmaxf(rintf(v * scale), lo), hi);
}

__device__ __forceinline__ WeftCode soft_to_code(const SoftCand& C, int ts) {
    WeftCode c;
    c.family = (uint16_t)(C.kind - 6);
    const float cmax = (float)(ts * 1024 - 1);
    c.cx = soft_q(C.p[0], 1024.0f, 0.0f, cmax);
    c.cy = soft_q(C.p[1], 1024.0f, 0.0f, cmax);
    c.rx = soft_q(C.p[2], 1024.0f, 1024.0f, 65535.0f);
    c.ry = soft_q(C.p[3], 1024.0f, 1024.0f, 65535.0f);
    const float th = rintf(C.p[4] * (4096.0f / SOFT_PI));
    c.theta = (uint16_t)(C.kind == 8 ? fminf(fmaxf(th, 0.0f), 4095.0f)
                                     : (th >= 4096.0f ? 0.0f : fmaxf(th, 0.0f)));
    c.r = soft_q(C.p[5], 255.0f, 0.0f, 255.0f);
    c.g = soft_q(C.p[6], 255.0f, 0.0f, 255.0f);
    c.b = soft_q(C.p[7], 255.0f, 0.0f, 255.0f);
    c.alpha = soft_q(C.p[8], 65535.0f, 19661.0f, 65535.0f);
    return c;
}

// ---- kernels' bodies -------------------------------------------------------------
// One workgroup per group of k candidates of one tile; L = 64 / k lanes per candidate split
// the tile's pixels, or with A.b.waves = W > 1 (k = 1) the 64 W lanes of W waves.
__device__ void soft_eval_body(const WeftSoftArgs& A) {
    const int g = blockIdx.x;
    const int k = A.b.k;
    const int W = A.b.waves > 1 ? A.b.waves : 1;
    const int L = W > 1 ? 64 * W : 64 / k;
    const int lane = threadIdx.x;
    const int j = lane / L, slice = lane - j * L;
    const int t = A.b.d_group_tile[g];
    const int ts = A.st.d_tile_ts[t];
    const int64_t base = A.st.d_tile_pix[t] * 3;
    const int64_t ci = (int64_t)g * k + j;
    SoftCand C;
    C.kind = A.b.d_kinds[ci];
#pragma unroll
    for (int i = 0; i < SOFT_NP; ++i) C.p[i] = A.b.d_params[ci * SOFT_NP + i];
    SoftPre P;
    soft_pre(C, P);
    SoftAcc S;
    const float invw = 1.0f / A.b.w;
    if (A.b.d_out_grad) {
        soft_tile<true>(C, P, invw, A.st.d_target + base, A.st.d_canvas + base, ts, slice, L, S);
        if (W > 1) soft_acc_waves(S, W, 0);
        if (slice == 0) {
            float gr[SOFT_NP];
            soft_finish(C, P, S, gr);
#pragma unroll
            for (int i = 0; i < SOFT_NP; ++i) A.b.d_out_grad[ci * SOFT_NP + i] = gr[i];
        }
    } else {
        soft_tile<false>(C, P, invw, A.st.d_target + base, A.st.d_canvas + base, ts, slice, L, S);
        if (W > 1) soft_acc_waves(S, W, 0);
    }
    if (slice == 0) A.b.d_out_sse[ci] = S.sse;
}

// (value, index) key of a score: order-preserving map of the float bits (-0 as +0, NaN
// last) above the index, so keys are distinct and ordered as a stable ascending sort.
__device__ __forceinline__ unsigned long long topk_key(float v, int i) {
    const unsigned int b = v == 0.0f ? 0u : __float_as_uint(v);
    const unsigned int hi = v != v ? 0xFFFFFFFEu : ((b & 0x80000000u) ? ~b : (b | 0x80000000u));
    return ((unsigned long long)hi << 32) | (unsigned int)i;
}

// Segmented top-k: for small k, k rounds of a workgroup-wide minimum over the keys above
// the last one taken; otherwise a bitonic sort of the keys in LDS.
__device__ void topk_body(const WeftTopkArgs& A) {
    extern __shared__ float4 weft_smem[];
    unsigned long long* key = reinterpret_cast<unsigned long long*>(weft_smem);
    const int t = blockIdx.x;
    const int tid = threadIdx.x, nthr = blockDim.x;
    if (A.st.d_tile_nsel[t] >= A.st.depth) return;
    const int n = A.n, N = A.n_pow2;
    const float* row = A.table + (int64_t)t * n;
    if (A.k <= 8) {
        __shared__ unsigned long long s_wmin[WEFT_MAX_BLOCK / 64];
        __shared__ unsigned long long s_last;
        unsigned long long last = 0ull;
        for (int r = 0; r < A.k; ++r) {
            unsigned long long m = ~0ull;
            for (int i = tid; i < n; i += nthr) {
                const unsigned long long kk = topk_key(row[i], i);
                if ((r == 0 || kk > last) && kk < m) m = kk;
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const unsigned long long o = __shfl_xor(m, off, 64);
                m = o < m ? o : m;
            }
            if ((tid & 63) == 0) s_wmin[tid >> 6] = m;
            __syncthreads();
            if (tid == 0) {
                unsigned long long mm = s_wmin[0];
                for (int w = 1; w < (nthr >> 6); ++w) mm = s_wmin[w] < mm ? s_wmin[w] : mm;
                A.out_idx[(int64_t)t * A.k + r] = (int32_t)(mm & 0xFFFFFFFFull);
                s_last = mm;
            }
            __syncthreads();
            last = s_last;
        }
        return;
    }
    for (int i = tid; i < N; i += nthr) key[i] = i < n ? topk_key(row[i], i) : ~0ull;
    __syncthreads();
    for (int size = 2; size <= N; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < (N >> 1); i += nthr) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const unsigned long long a = key[lo], b = key[hi];
                if ((a > b) == asc) {
                    key[lo] = b;
                    key[hi] = a;
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < A.k; i += nthr) A.out_idx[(int64_t)t * A.k + i] = (int32_t)(key[i] & 0xFFFFFFFFull);
}

// n_iter Adam steps on candidate C's soft SSE over its tile, L lanes per candidate
// (this lane takes pixels slice, slice + L, ...), projected onto the code ranges after
// every step.  Every lane of the candidate ends with the same parameters.  MW: L may exceed
// 64 (the candidate's lanes span L / 64 waves, every one of which must call this); without
// it L <= 64 and the code keeps no cross-wave step (fewer registers).
template <int UNROLL = SOFT_UNROLL, bool PREFETCH = true, bool MW = true>
__device__ __forceinline__ void refine_candidate(const WeftRefineOpts& o, const float* tg, const float* cv,
                                                 int ts, int slice, int L, SoftCand& C) {
    float m1[SOFT_NP], m2[SOFT_NP];
#pragma unroll
    for (int i = 0; i < SOFT_NP; ++i) m1[i] = m2[i] = 0.0f;
#pragma unroll 1
    for (int it = 0; it < o.n_iter; ++it) {
        // edge width and the reciprocal bias corrections of this step
        const float w = o.d_sched[4 * it], ibc1 = o.d_sched[4 * it + 1], ibc2 = o.d_sched[4 * it + 2];
        SoftPre P;
        soft_pre(C, P);
        SoftAcc S;
        soft_tile<true, UNROLL, PREFETCH>(C, P, 1.0f / w, tg, cv, ts, slice, L, S);
        if constexpr (MW) {
            if (L > 64) soft_acc_waves(S, L >> 6, it & 1);
        }
        float gr[SOFT_NP];
        soft_finish(C, P, S, gr);
#pragma unroll
        for (int i = 0; i < SOFT_NP; ++i) {
            m1[i] = o.beta1 * m1[i] + o.one_minus_beta1 * gr[i];
            m2[i] = o.beta2 * m2[i] + o.one_minus_beta2 * (gr[i] * gr[i]);
            const float mh = m1[i] * ibc1;
            const float vh = m2[i] * ibc2;
            C.p[i] = C.p[i] - o.lr[i] * (mh / (sqrtf(vh) + o.eps));
        }
        soft_project(C, ts);
    }
}

// Adam on the soft SSE of each tile's top-k drawn candidates; one workgroup per work item:
// cpw candidates, 64 / cpw lanes of the first wave each, of one tile (cpw <= k) or of
// cpw / k tiles (cpw > k, all k candidates of each); a size with cpw = -W spreads one
// candidate over W waves (64 W lanes).  The workgroup has 64 times the largest such W lanes;
// an item of one wave leaves the others idle.  Every lane of a candidate holds the same
// parameters and moments (the butterfly sums give all of them the same totals).
template <bool MW>
__device__ void refine_body(const WeftRefineArgs& A) {
    const int item = blockIdx.x;
    const WeftFitState& st = A.st;
    const int k = A.o.k;
    const int ts = st.d_tile_ts[A.o.d_work_tile[(int64_t)item * WEFT_WORK_MAX_TILES]];
    const int cls = ts == 8 ? 0 : (ts == 16 ? 1 : (ts == 32 ? 2 : 3));
    const int cpw_raw = A.o.cpw_by_size[cls];
    const int W = MW && cpw_raw < 0 ? -cpw_raw : 1;
    if ((int)(threadIdx.x >> 6) >= W) return;  // the whole wave: no barrier waits for it
    const int cpw = cpw_raw < 0 ? 1 : cpw_raw;
    const int cpt = cpw < k ? cpw : k;  // candidates per tile in this item
    const int L = 64 * W / cpw;
    const int lane = threadIdx.x;
    const int q = lane / L, slice = lane - q * L;
    const int g = q / cpt;  // tile slot: a candidate's lanes share it, so they leave together
    const int t = A.o.d_work_tile[(int64_t)item * WEFT_WORK_MAX_TILES + g];
    if (t < 0 || st.d_tile_nsel[t] >= st.depth) return;
    const int j0 = A.o.d_work_j0[item];
    const int j = j0 + (q - g * cpt);
    const int64_t base = st.d_tile_pix[t] * 3;
    const float* tg = st.d_target + base;
    const float* cv = st.d_canvas + base;
    const uint64_t dbase = draw_base(A.cd.seed, A.cd.step, (uint32_t)t);
    const int64_t slot0 = (int64_t)t * (k + 1);
    const int idx = A.topk[(int64_t)t * k + j];
    const WeftCode code0 = draw_code(dbase, (uint32_t)idx, ts);
    if (j == 0 && slice == 0) {
        // slot 0: the best drawn candidate (top-1) as drawn
        const WeftCode best = draw_code(dbase, (uint32_t)A.topk[(int64_t)t * k], ts);
        if (A.out_codes) A.out_codes[slot0] = best;
        A.out_prims[slot0] = code_to_prim(best);
    }
    SoftCand C = soft_from_code(code0);
    refine_candidate<SOFT_UNROLL, true, MW>(A.o, tg, cv, ts, slice, L, C);
    if (slice == 0) {
        const WeftCode qc = soft_to_code(C, ts);
        if (A.out_codes) A.out_codes[slot0 + 1 + j] = qc;
        A.out_prims[slot0 + 1 + j] = code_to_prim(qc);
        if (A.out_params) {
#pragma unroll
            for (int i = 0; i < SOFT_NP; ++i) A.out_params[((int64_t)t * k + j) * SOFT_NP + i] = C.p[i];
        }
    }
}
