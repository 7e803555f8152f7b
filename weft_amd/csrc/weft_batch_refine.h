// Persistent refined fit: the batched fit of weft_batch_fit with, in every round, the
// tiles' top-k drawn candidates refined by Adam on the soft SSE and the refined codes
// re-scored exactly before anything is accepted.  Included by weft_kernels.hip inside
// its anonymous namespace, after the batched fit, the soft surrogate and the split scorer.
//
// The queue hands out groups of tiles of one size (one tile per group when the host gives
// no group table).  Per group and round (one workgroup): each tile in turn has its n_cand
// drawn candidates class-sorted and scored (as weft_batch_fit) and the (MSE bits, index)
// keys of its candidates ranked for its top k (k workgroup-wide minima for k <= 8, else a
// bitonic sort); then the top k of all the group's tiles are refined together, with the
// work split of weft_refine_kernel (cpw candidates per wave, 64 / cpw lanes each; small
// tiles share a wave when cpw > k); then the k refined codes and the unrefined best of every
// tile are scored exactly together (each candidate's pixels split over lanes along the
// pairwise tree), and each tile accepts its first minimum if it beats its objective.  The
// result equals that of the stepwise refined round (weft_fit_score + weft_topk +
// weft_refine + weft_fit_score_split + weft_fit_select) with the same work split.

#ifndef WEFT_BATCH_REFINE_UNROLL
#define WEFT_BATCH_REFINE_UNROLL 1  // pixels in flight in the refinement (registers are shared with scoring)
#endif

constexpr int BR_GMAX = 64;  // tiles per group, and at most 64 / k of them

struct RefineShared {
    int topk[64];                 // [tile of the group][k]
    WeftCode codes[BR_GMAX + 64];  // [tile][k + 1]: the unrefined best, then the refined
    double obj[BR_GMAX + 64];     // exact objectives of codes
    int gtile[BR_GMAX];           // the group's tiles
    int live[BR_GMAX];            // the tile's stack is not full this round
    int win[BR_GMAX];             // accepted slot (-1 none) of each tile this round
    double win_obj[BR_GMAX];
    unsigned long long wmin[16];  // per-wave minima of the top-k rounds (up to 1024 threads)
    unsigned long long last;
};

// Adam on the top k of the group's G tiles, with the work split of weft_refine_kernel: work
// item i holds cpw candidates (64 / cpw lanes each) -- cpw of one tile when cpw <= k, or all
// k of each of cpw / k tiles -- and the workgroup's waves take the items in turn; with
// cpw = -W each aligned group of W waves takes one candidate at a time.  Writes
// rs.codes[g * (k + 1) + 1 + j].
template <int TS>
__device__ __forceinline__ void refine_group(const WeftBatchRefineArgs& R, uint32_t step, int G, RefineShared& rs) {
    constexpr int CLS = TS == 8 ? 0 : (TS == 16 ? 1 : (TS == 32 ? 2 : 3));
    const WeftFitState& st = R.b.st;
    const int tid = threadIdx.x, wave = tid >> 6, nwaves = blockDim.x >> 6, lane = tid & 63;
    const int k = R.o.k;
    const int K1 = k + 1;
    const int cpw_raw = R.o.cpw_by_size[CLS];
    // cpw = -W: every wave runs every round's n_iter steps, each with a workgroup barrier, so
    // the barriers line up; a wave group without a live candidate in a round (or the spare
    // waves when W does not divide the waves) refines a stand-in and keeps nothing.
    const bool mw = cpw_raw < 0;
    const int W = mw ? -cpw_raw : 1;
    const int cpw = mw ? 1 : cpw_raw;
    const int cpt = cpw < k ? cpw : k;           // candidates of one tile in an item
    const int tpw = cpw > k ? cpw / k : 1;       // tiles per item
    const int ipt = cpw < k ? k / cpw : 1;       // items per tile
    const int items = cpw > k ? (G + tpw - 1) / tpw : G * ipt;
    const int L = 64 * W / cpw;
    const int per = nwaves / W;                  // items in flight (aligned groups of W waves)
    const int gi = wave / W;
    const int rounds = (items + per - 1) / per;
    const int q = mw ? 0 : lane / L;
    const int slice = mw ? (wave - gi * W) * 64 + lane : lane - q * L;
#pragma unroll 1
    for (int r = 0; mw ? r < rounds : r * per + gi < items; ++r) {
        const int it = r * per + gi;
        const int g = cpw > k ? it * tpw + q / cpt : it / ipt;
        const int j = cpw > k ? q - (q / cpt) * cpt : (it - (it / ipt) * ipt) * cpw + q;
        const bool keep = gi < per && it < items && g < G && rs.live[g];
        if (!mw && !keep) continue;  // a candidate's lanes leave together
        const int t = rs.gtile[keep ? g : 0];
        const int64_t base = st.d_tile_pix[t] * 3;
        const uint64_t dbase = draw_base(R.b.cd.seed, step, (uint32_t)t);
        SoftCand C = soft_from_code(draw_code(dbase, keep ? (uint32_t)rs.topk[g * k + j] : 0u, TS));
        refine_candidate<WEFT_BATCH_REFINE_UNROLL, false>(R.o, st.d_target + base, st.d_canvas + base, TS, slice, L,
                                                          C);
        if (keep && slice == 0) rs.codes[g * K1 + 1 + j] = soft_to_code(C, TS);
    }
}

// Exact objectives of the group's codes (G tiles x (k + 1)), each candidate's pixels split
// over lanes along the pairwise tree, into rs.obj.
template <int TS>
__device__ __forceinline__ void rescore_group(const WeftBatchRefineArgs& R, int G, RefineShared& rs) {
    constexpr int CLS = TS == 8 ? 0 : (TS == 16 ? 1 : (TS == 32 ? 2 : 3));
    const WeftFitState& st = R.b.st;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int K1 = R.o.k + 1;
    const int total = G * K1;
    const int Ls = R.split_lanes[CLS];
    const int per = nthr / Ls;
#pragma unroll 1
    for (int c0 = 0; c0 < total; c0 += per) {
        const int c = c0 + tid / Ls, sl = tid - (tid / Ls) * Ls;
        const int g = c < total ? c / K1 : 0;
        const bool live = c < total && rs.live[g];
        const int t = rs.gtile[g];
        const int64_t base = st.d_tile_pix[t] * 3;
        WeftPrim P;
        P.kind = -1;
        P.alpha = 0.0f;
        P.color0[0] = P.color0[1] = P.color0[2] = 0.0f;
        if (live) P = code_to_prim(rs.codes[c]);
        const float sse = split_sse<TS, true>(P, st.d_canvas + base, st.d_target + base, Ls, sl);
        if (live && sl == 0) {
            const float mse = sse / (float)(3 * TS * TS);
            rs.obj[c] = objective(mse, st.d_tile_bits[t] + prim_bytes(P.kind) * 8, st.lam, st.obj_mode);
        }
    }
}

// Inlined into the kernel on purpose: a call would pass the kernel arguments by address,
// which copies the whole argument block into scratch memory.
//
// Scores tile t's drawn candidates of this round (class-sorted, one lane per candidate,
// canvas and target through LDS bands) and leaves its top k in rs.topk[g * k ..] and its
// unrefined best in rs.codes[g * (k + 1)].
template <int TS>
__device__ __forceinline__ void score_topk(const WeftBatchRefineArgs& R, int t, int g, uint32_t step, float4* LA,
                                           float* LT, unsigned long long* keys, RefineShared& rs) {
    const WeftBatchArgs& A = R.b;
    constexpr int ROWS = Band<TS>::ROWS;
    constexpr int NB = TS / ROWS;
    constexpr int BPX = ROWS * TS;
    const WeftFitState& st = A.st;
    const int tid = threadIdx.x, nthr = blockDim.x, wave = tid >> 6, nwaves = nthr >> 6, lane = tid & 63;
    const int64_t base = st.d_tile_pix[t] * 3;
    const float* cvp = st.d_canvas + base;
    const float* T = st.d_target + base;
    const int n = A.cd.n_cand;
    const int k = R.o.k;
    const int np2 = R.n_pow2;
    float* LAf = reinterpret_cast<float*>(LA);
    const LdsBand L{LA, LT};
    const uint64_t dbase = draw_base(A.cd.seed, step, (uint32_t)t);
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
            const float mse = pw.template result<TS>() / (float)(3 * TS * TS);
            keys[ks] = ((unsigned long long)__float_as_uint(mse) << 32) | (unsigned int)ks;
        }
    }
    for (int i = n + tid; i < np2; i += nthr) keys[i] = ~0ull;
    __syncthreads();
    // ---- top k by (MSE, index) (keys are distinct): for k <= 8, k workgroup-wide
    // minima, each above the last; above that, a bitonic sort of the keys ----
    if (k <= 8) {
        unsigned long long last = 0ull;
        for (int r = 0; r < k; ++r) {
            unsigned long long m = ~0ull;
            for (int i = tid; i < n; i += nthr) {
                const unsigned long long kk = keys[i];
                if ((r == 0 || kk > last) && kk < m) m = kk;
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const unsigned long long o = __shfl_xor(m, off, 64);
                m = o < m ? o : m;
            }
            if (lane == 0) rs.wmin[wave] = m;
            __syncthreads();
            if (tid == 0) {
                unsigned long long mm = rs.wmin[0];
                for (int ww = 1; ww < nwaves; ++ww) mm = rs.wmin[ww] < mm ? rs.wmin[ww] : mm;
                rs.topk[g * k + r] = (int)(mm & 0xFFFFFFFFull);
                rs.last = mm;
            }
            __syncthreads();
            last = rs.last;
        }
    } else {
        for (int size = 2; size <= np2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = tid; i < (np2 >> 1); i += nthr) {
                    const int lo = 2 * i - (i & (stride - 1));
                    const int hi = lo + stride;
                    const bool asc = (lo & size) == 0;
                    const unsigned long long a = keys[lo], b = keys[hi];
                    if ((a > b) == asc) {
                        keys[lo] = b;
                        keys[hi] = a;
                    }
                }
                __syncthreads();
            }
        }
        if (tid < k) rs.topk[g * k + tid] = (int)(keys[tid] & 0xFFFFFFFFull);
        __syncthreads();
    }
    if (tid == 0) rs.codes[g * (k + 1)] = draw_code(dbase, (uint32_t)rs.topk[g * k], TS);
}

// The fit of a group of G tiles of TS px (rs.gtile[0 .. G - 1]), all rounds.
template <int TS>
__device__ __forceinline__ void batch_refine_group(const WeftBatchRefineArgs& R, int G, float4* LA, float* LT,
                                                   unsigned long long* keys, RefineShared& rs) {
    const WeftBatchArgs& A = R.b;
    const WeftFitState& st = A.st;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int k = R.o.k;
    const int K1 = k + 1;
    const int64_t T_all = st.n_tiles;
#pragma unroll 1
    for (int s = 0; s < A.n_steps; ++s) {
        const uint32_t step = A.step0 + (uint32_t)s;
        // ---- exact scores of each live tile's drawn candidates, its top k ----
        int any = 0;
#pragma unroll 1
        for (int g = 0; g < G; ++g) {
            const int t = rs.gtile[g];
            const int live = st.d_tile_nsel[t] < st.depth;  // uniform: written by this workgroup only
            if (tid == 0) rs.live[g] = live;
            if (!live) {
                if (tid == 0) {
                    const int64_t rec = (int64_t)s * T_all + t;
                    A.win_idx[rec] = -1;
                    A.win_obj[rec] = __builtin_huge_val();
                    A.win_acc[rec] = 0;
                }
                continue;
            }
            any = 1;
            score_topk<TS>(R, t, g, step, LA, LT, keys, rs);
        }
        __syncthreads();
        if (!any) continue;
        // ---- Adam on the group's top k; exact scores of the best drawn and the refined ----
        refine_group<TS>(R, step, G, rs);
        __syncthreads();
        rescore_group<TS>(R, G, rs);
        __syncthreads();
        // ---- per tile: the first minimum of (objective, slot), accepted if it beats the tile's ----
        for (int g = tid; g < G; g += nthr) {
            int win = -1;
            double wo = __builtin_huge_val();
            if (rs.live[g]) {
                const int t = rs.gtile[g];
                double bo = __builtin_huge_val();
                int bi = 0x7fffffff;
                for (int c = 0; c < K1; ++c) argmin_pair(bo, bi, rs.obj[g * K1 + c], c);
                const bool found = bi != 0x7fffffff;
                const bool acc = found && (bo < st.d_tile_obj[t]);
                const int64_t rec = (int64_t)s * T_all + t;
                A.win_idx[rec] = found ? bi : -1;
                A.win_obj[rec] = bo;
                A.win_acc[rec] = acc;
                if (acc) {
                    win = bi;
                    wo = bo;
                }
            }
            rs.win[g] = win;
            rs.win_obj[g] = wo;
        }
        __syncthreads();
#pragma unroll 1
        for (int g = 0; g < G; ++g) {
            const int win = rs.win[g];
            if (win < 0) continue;  // uniform
            const int t = rs.gtile[g];
            float* cvp = st.d_canvas + st.d_tile_pix[t] * 3;
            const WeftCode code = rs.codes[g * K1 + win];
            const WeftPrim Pw = code_to_prim(code);
            for (int p = tid; p < TS * TS; p += nthr) {
                const int y = p / TS, x = p - y * TS;
                float r = cvp[3 * p], gg = cvp[3 * p + 1], b = cvp[3 * p + 2];
                composite_px(Pw, (float)x + 0.5f, (float)y + 0.5f, r, gg, b);
                cvp[3 * p] = r;
                cvp[3 * p + 1] = gg;
                cvp[3 * p + 2] = b;
            }
            if (tid == 0) {
                if (A.win_code) A.win_code[(int64_t)s * T_all + t] = code;
                st.d_tile_nsel[t] += 1;
                st.d_tile_bits[t] += prim_bytes(Pw.kind) * 8;
                st.d_tile_obj[t] = rs.win_obj[g];
            }
        }
        __threadfence_block();
        __syncthreads();  // canvases and state are settled before the next round reads them
    }
}

__device__ __forceinline__ void batch_refine_body(const WeftBatchRefineArgs& R) {
    extern __shared__ float4 weft_smem[];
    __shared__ BatchShared sh;
    __shared__ RefineShared rs;
    const WeftBatchArgs& A = R.b;
    float4* LA = weft_smem;                                             // [2 * 512]
    float* LT = reinterpret_cast<float*>(LA + 2 * 512);                 // [512]
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(LT + 512);  // [n_pow2]
    const int tid = threadIdx.x;
    // queue entries: groups of tiles of one size (one tile each without a group table),
    // taken gmax tiles at a time
    const int n_entries = R.group_first ? R.n_groups : A.n_queue;
    const int gmax = R.group_first ? R.group_max : 1;
    if (tid == 0 && A.t_start) atomicMin(A.t_start, (unsigned long long)__builtin_amdgcn_s_memrealtime());
#pragma unroll 1
    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(A.queue, 1);
            int first = -1, last = -1;
            if (q < n_entries) {
                if (R.group_first) {
                    first = max(0, min(R.group_first[q], A.n_queue));
                    last = max(first, min(R.group_first[q + 1], A.n_queue));
                } else {
                    first = q;
                    last = q + 1;
                }
            }
            sh.tile = first;
            sh.best = last;
        }
        __syncthreads();
        const int first = sh.tile, last = sh.best;
        __syncthreads();
        if (first < 0) break;
#pragma unroll 1
        for (int g0 = first; g0 < last;) {
            // the next run of at most gmax tiles of the first one's size
            if (tid == 0) {
                const int lim = min(gmax, last - g0);
                const int ts0 = A.st.d_tile_ts[A.tile_order[g0]];
                int G = 0;
                while (G < lim) {
                    const int t = A.tile_order[g0 + G];
                    if (A.st.d_tile_ts[t] != ts0) break;
                    rs.gtile[G++] = t;
                }
                sh.acc = G;
            }
            __syncthreads();
            const int G = sh.acc;
            switch (A.st.d_tile_ts[rs.gtile[0]]) {
            case 8: batch_refine_group<8>(R, G, LA, LT, keys, rs); break;
            case 16: batch_refine_group<16>(R, G, LA, LT, keys, rs); break;
            case 32: batch_refine_group<32>(R, G, LA, LT, keys, rs); break;
            case 64: batch_refine_group<64>(R, G, LA, LT, keys, rs); break;
            default: break;
            }
            if (tid == 0 && A.image_left) {
                for (int g = 0; g < G; ++g) {
                    const int img = A.tile_image[rs.gtile[g]];
                    if (atomicSub(&A.image_left[img], 1) == 1)
                        A.image_done[img] = (unsigned long long)__builtin_amdgcn_s_memrealtime();
                }
            }
            g0 += G;
            __syncthreads();
        }
    }
}
