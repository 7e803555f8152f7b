// Exact scoring of a few candidates per tile with the tile's pixels split over lanes.
// Included by weft_kernels.hip inside its anonymous namespace.
//
// The SSE of a candidate is NumPy's float32 pairwise sum over the tile's elements in
// (y, x, channel) order: chunks of 8192 elements added left to right, each chunk a
// binary tree (halves rounded down to multiples of 8) over leaves of at most 128
// elements, a leaf summed with 8 interleaved accumulators.  For the tile sizes here
// the trees are perfect: 8, 16, 32 px -> 2, 8, 32 leaves of 96 elements; 64 px ->
// chunks of 64 and 32 leaves of 128 elements.  So L lanes (a power of two dividing
// the leaf counts) can each reduce a contiguous run of leaves to its subtree's value,
// and a butterfly over the L lanes adds the subtrees in tree order -- the same value
// as the sequential sum, bit for bit (a + b == b + a in IEEE arithmetic).

struct SplitTree {
    float lvl[7];  // up to 64 leaves per lane
    int n;
};

__device__ __forceinline__ void split_tree_push(SplitTree& T, float v) {
    int j = 0;
    int n = T.n;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        if (n & 1) {
            v = T.lvl[q] + v;  // left subtree (earlier leaves) + right
            n >>= 1;
            j = q + 1;
        } else {
            break;
        }
    }
    // store at level j (static indices only)
#pragma unroll
    for (int q = 0; q < 7; ++q)
        if (q == j) T.lvl[q] = v;
    T.n += 1;
}

// Per-pixel squared errors (pixel p = (y, x) of a TS-px tile):
// a candidate composited over the tile's canvas (HARD: constant or hard-edge kinds only,
// a smaller composite) ...
template <bool HARD>
struct PxOverCanvas {
    static constexpr bool kRow8 = false;
    const WeftPrim& P;
    const float* cv;
    const float* tg;
    __device__ __forceinline__ void operator()(int p, int y, int x, float& q0, float& q1, float& q2) const {
        float o0 = cv[3 * p], o1 = cv[3 * p + 1], o2 = cv[3 * p + 2];
        if constexpr (HARD) composite_hard_px(P, (float)x + 0.5f, (float)y + 0.5f, o0, o1, o2);
        else composite_px(P, (float)x + 0.5f, (float)y + 0.5f, o0, o1, o2);
        const float d0 = clip01(o0) - tg[3 * p];
        const float d1 = clip01(o1) - tg[3 * p + 1];
        const float d2 = clip01(o2) - tg[3 * p + 2];
        q0 = d0 * d0;
        q1 = d1 * d1;
        q2 = d2 * d2;
    }
};

// ... or a whole stack rendered over black (render_tile, then the clipped error), eight
// pixels of a row at a time so that each primitive's constants are formed once per group.
template <int TS>
struct PxStackWalk {
    static constexpr bool kRow8 = true;
    const WeftPrim* prims;
    int n;
    const float* tg;
    // squared errors of pixels pg .. pg + 7 (one row: TS is a multiple of 8)
    __device__ __forceinline__ void row8(int pg, float (&q)[24]) const {
        const int y = pg / TS, x0 = pg - y * TS;
        float r[8], g[8], b[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = g[i] = b[i] = 0.0f;
        const float py = (float)y + 0.5f;
        for (int k = 0; k < n; ++k) composite_row8(prims[k], (float)x0, py, r, g, b);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float* t = tg + 3 * (pg + i);
            const float d0 = clip01(r[i]) - t[0];
            const float d1 = clip01(g[i]) - t[1];
            const float d2 = clip01(b[i]) - t[2];
            q[3 * i] = d0 * d0;
            q[3 * i + 1] = d1 * d1;
            q[3 * i + 2] = d2 * d2;
        }
    }
};

// Sum of elements [e0, e0 + LS) of the squared errors (one leaf).
template <int TS, int LS, class F>
__device__ __forceinline__ float split_leaf(const F& f, int e0) {
    constexpr int NPX = TS * TS;
    float r[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = 0.0f;
    const int p0 = e0 / 3;
    const int p1 = (e0 + LS + 2) / 3;  // one past the last pixel touching the leaf
#pragma unroll 1
    for (int pg = p0 & ~7; pg < p1; pg += 8) {
        float qg[24];
        if constexpr (F::kRow8) f.row8(pg, qg);  // pg + 7 < NPX: NPX and pg are multiples of 8
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int p = pg + i;
            float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
            if constexpr (F::kRow8) {
                q0 = qg[3 * i];
                q1 = qg[3 * i + 1];
                q2 = qg[3 * i + 2];
            } else if (p >= p0 && p < p1 && p < NPX) {
                const int y = p / TS, x = p - y * TS;
                f(p, y, x, q0, q1, q2);
            }
            const int e = 3 * p;
            // element 3p + c goes to accumulator (3i + c) % 8 (pg is a multiple of 8)
            r[(3 * i) % 8] += (e >= e0 && e < e0 + LS) ? q0 : 0.0f;
            r[(3 * i + 1) % 8] += (e + 1 >= e0 && e + 1 < e0 + LS) ? q1 : 0.0f;
            r[(3 * i + 2) % 8] += (e + 2 >= e0 && e + 2 < e0 + LS) ? q2 : 0.0f;
        }
    }
    return ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
}

// Perfect tree over the nleaves leaves of one chunk, L lanes; every lane returns the sum.
template <int TS, int LS, class F>
__device__ __forceinline__ float split_chunk(const F& f, int c0, int nleaves, int L, int sl) {
    const int m = nleaves / L;
    SplitTree T;
    T.n = 0;
#pragma unroll
    for (int q = 0; q < 7; ++q) T.lvl[q] = 0.0f;
#pragma unroll 1
    for (int li = 0; li < m; ++li) split_tree_push(T, split_leaf<TS, LS>(f, c0 + (sl * m + li) * LS));
    // m is a power of two: the root is the single level left
    float v = 0.0f;
#pragma unroll
    for (int q = 0; q < 7; ++q)
        if ((1 << q) == m) v = T.lvl[q];
    return lane_group_sum(v, L);
}

template <int TS, class F>
__device__ __forceinline__ float split_sse_f(const F& f, int L, int sl) {
    if constexpr (TS == 64) {
        const float a = split_chunk<TS, 128>(f, 0, 64, L, sl);
        const float b = split_chunk<TS, 128>(f, 8192, 32, L, sl);
        return (0.0f + a) + b;
    } else {
        return split_chunk<TS, 96>(f, 0, (3 * TS * TS) / 96, L, sl);
    }
}

template <int TS, bool HARD = false>
__device__ float split_sse(const WeftPrim& P, const float* cv, const float* tg, int L, int sl) {
    return split_sse_f<TS>(PxOverCanvas<HARD>{P, cv, tg}, L, sl);
}

// Drop-in stack scoring with lanes split: evals d_order[0 .. n) of one tile size, L lanes
// each (L = the pairwise tree's leaf limit of the size), 64 / L evals per workgroup.
template <int TS>
__device__ void score_evals_split_item(const WeftScoreEvalsSplitArgs& A) {
    constexpr int L = TS == 8 ? 2 : (TS == 16 ? 8 : 32);
    const int lane = threadIdx.x;
    const int j = blockIdx.x * (64 / L) + lane / L, sl = lane & (L - 1);
    const bool live = j < A.n;
    const int ei = live ? A.order[j] : 0;
    const WeftEval ev = A.evals[ei];
    const bool ok = live && ev.tile_size == TS;
    const PxStackWalk<TS> f{A.prims + ev.prim_offset, ok ? ev.prim_count : 0, A.tiles + ev.tile_offset};
    const float sse = split_sse_f<TS>(f, L, sl);
    if (live && sl == 0) {
        const float v = ok ? sse : __builtin_nanf("");
        A.out_mse[ei] = v / (float)(3 * TS * TS);
        if (A.out_sse) A.out_sse[ei] = v;
    }
}

template <int TS>
__device__ void split_score_item(const WeftSplitScoreArgs& A, int item, int cpw, int tpw) {
    const WeftFitState& st = A.st;
    const int G = 64 / tpw;  // lanes per tile slot
    const int L = G / cpw;   // lanes per candidate
    const int lane = threadIdx.x;
    const int g = lane / G, gl = lane - g * G;
    const int jl = gl / L, sl = gl - jl * L;
    const int64_t slot = (int64_t)item * WEFT_WORK_MAX_TILES + g;
    const int t = A.work_tile[slot];
    const int tt = t >= 0 ? t : A.work_tile[(int64_t)item * WEFT_WORK_MAX_TILES];  // pixels read by an unused slot
    const bool tile_live = t >= 0 && st.d_tile_nsel[t] < st.depth;
    const int n = A.cd.n_cand;
    const int c = A.work_c0[item] + jl;
    const bool live = tile_live && c < n;
    WeftPrim P;
    P.kind = -1;
    P.alpha = 0.0f;
    P.color0[0] = P.color0[1] = P.color0[2] = 0.0f;
    if (live) {
        if (A.cd.d_prims == nullptr) {
            P = code_to_prim(draw_code(draw_base(A.cd.seed, A.cd.step, (uint32_t)t), (uint32_t)c, TS));
        } else {
            const int64_t cb = A.cd.d_tile_cand ? A.cd.d_tile_cand[t] : (int64_t)t * n;
            P = load_prim(A.cd.d_prims + cb + c);
        }
    }
    const int64_t base = st.d_tile_pix[tt] * 3;
    const float sse = split_sse<TS>(P, st.d_canvas + base, st.d_target + base, L, sl);
    double obj = __builtin_huge_val();
    int idx = 0x7fffffff;
    if (live) {
        const float mse = sse / (float)(3 * TS * TS);
        if (sl == 0) {
            const int64_t ti = (int64_t)t * n + c;
            if (A.out_sse) A.out_sse[ti] = sse;
            if (A.out_mse) A.out_mse[ti] = mse;
        }
        obj = objective(mse, st.d_tile_bits[t] + prim_bytes(P.kind) * 8, st.lam, st.obj_mode);
        idx = c;
    }
    // first minimum over the lanes of this tile slot (xor offsets below G stay in it)
    for (int off = G >> 1; off >= 1; off >>= 1) {
        const double o2 = __shfl_xor(obj, off, 64);
        const int i2 = __shfl_xor(idx, off, 64);
        argmin_pair(obj, idx, o2, i2);
    }
    if (gl == 0 && t >= 0) {
        A.part_obj[slot] = obj;
        A.part_idx[slot] = idx == 0x7fffffff ? -1 : idx;
    }
}

__device__ void split_score_body(const WeftSplitScoreArgs& A) {
    const int item = blockIdx.x;
    const int t0 = A.work_tile[(int64_t)item * WEFT_WORK_MAX_TILES];
    switch (A.st.d_tile_ts[t0]) {
    case 8: split_score_item<8>(A, item, A.cpw[0], A.tpw[0]); break;
    case 16: split_score_item<16>(A, item, A.cpw[1], A.tpw[1]); break;
    case 32: split_score_item<32>(A, item, A.cpw[2], A.tpw[2]); break;
    case 64: split_score_item<64>(A, item, A.cpw[3], A.tpw[3]); break;
    default: break;
    }
}
