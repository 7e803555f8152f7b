// Template candidate pools on the device: encoder._generate_candidates
// (src/weft/encoder.py:535-745) for every tile of one size, one 256-thread workgroup per
// tile.  Included by weft_kernels.hip inside its anonymous namespace.
//
// The pool is built in the reference's order and with its arithmetic: float32 image
// statistics reduced in NumPy's orders (pairwise for flat means, row by row for means over
// leading axes, float32 percentile interpolation), float64 geometry written as the
// reference's Python expressions, and the reference's de-duplication keys (geometry on a
// 1/4 grid, colours 1/64, alpha 1/32).  Three choices the reference leaves to NumPy
// internals are fixed: a stable luminance sort for the k-means palette, row-major side sums
// for the linear patches, and ascending (luminance, index) order for the six brightest and
// darkest pixels that pair into triangles.

constexpr int CG_MAXC = 352;        // pre-de-duplication pool bound (341 at quality >= 80)
constexpr int CG_KEY = 12;          // longest de-duplication key
constexpr int CG_MAX_LINES = 16 + 36;
constexpr int CG_THREADS = 256;

__device__ __forceinline__ double cg_ct(double v) {  // encoder._ct on one channel
    const double m = v < 1.0 ? v : 1.0;
    return m > 0.0 ? m : 0.0;
}

__device__ __forceinline__ uint32_t cg_fkey(float v) {  // float order as unsigned order
    if (v == 0.0f) v = 0.0f;
    const uint32_t u = __float_as_uint(v);
    return (u >> 31) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float cg_funkey(uint32_t k) {
    return __uint_as_float((k >> 31) ? (k & 0x7fffffffu) : ~k);
}

// NumPy's pairwise float32 sum of a[0 .. n) (n a power of two >= 8): 128-element leaves
// with 8 accumulators, combined as a perfect binary tree; the reduction adds it to 0.
__device__ float cg_pairwise(const float* a, int n) {
    const int leaf = n < 128 ? n : 128;
    const int nl = n / leaf;  // a power of two <= 32
    SplitTree T;
    T.n = 0;
#pragma unroll
    for (int q = 0; q < 7; ++q) T.lvl[q] = 0.0f;
    for (int l = 0; l < nl; ++l) {
        const float* b = a + l * leaf;
        float r[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] = b[j];
        for (int i = 8; i < leaf; i += 8)
#pragma unroll
            for (int j = 0; j < 8; ++j) r[j] += b[i + j];
        split_tree_push(T, ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7])));
    }
    float v = 0.0f;
#pragma unroll
    for (int q = 0; q < 7; ++q)
        if ((1 << q) == nl) v = T.lvl[q];
    return 0.0f + v;
}

// float32 percentile q (0.25f or 0.75f) of n sorted keys, NumPy's "linear" method.
__device__ float cg_percentile(const uint32_t* keys, int n, float q) {
    const float virt = (float)(n - 1) * q;
    const float prev = floorf(virt);
    if (virt >= (float)(n - 1)) return cg_funkey(keys[n - 1]);
    const int i = (int)prev;
    const float a = cg_funkey(keys[i]), b = cg_funkey(keys[i + 1]);
    const float gamma = virt - prev;
    const float diff = b - a;
    float r = a + diff * gamma;
    if (gamma >= 0.5f) r = b - diff * (1.0f - gamma);
    return r;
}

struct CgShared {
    int keys[CG_MAXC][CG_KEY];
    int keep[CG_MAXC];
    uint32_t hash[CG_MAXC];
    float line_col[CG_MAX_LINES][2][3];
    double line_geom[CG_MAX_LINES][4];
    float mean_all[3], quad[4][3], sub[16][3], km[9][3], ext[3][3];
    int ext_n[3];
    float avg_mag, gmx, gmy;
    double pal[13][3];
    int n_pal, grad_lines, n_tri_x, n_kept;
    double base_alpha, gline[4];
    int pk[6], vl[6];
    // 3-means regions: pixel count, x and y sums, min of x n + p, max of x n + (n - 1 - p)
    int km_cnt[3], km_sx[3], km_sy[3], km_lo[3], km_hi[3];
};

// Sequential float32 sum of plane values over the pixels (x, y) in [x0, x1) x [y0, y1), row-major.
__device__ __forceinline__ float cg_block_sum(const float* plane, int ts, int x0, int x1, int y0, int y1) {
    float s = 0.0f;
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) s += plane[y * ts + x];
    return s;
}

// s + a[0] + a[1] + ... + a[n - 1], left to right in float32 (loads issued ahead of the adds).
__device__ __forceinline__ float cg_seq_add(float s, const float* a, int n) {
    int i = 0;
    for (; i + 8 <= n; i += 8) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = a[i + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) s += v[j];
    }
    for (; i < n; ++i) s += a[i];
    return s;
}

// The same sum over the pixels of [x0, x0 + w) x [y0, y0 + h), row by row.
__device__ __forceinline__ float cg_rows_add(const float* plane, int ts, int x0, int w, int y0, int h) {
    float s = 0.0f;
    for (int y = y0; y < y0 + h; ++y) s = cg_seq_add(s, plane + y * ts + x0, w);
    return s;
}

__device__ __forceinline__ int cg_wave_sum(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ int cg_wave_min(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int cg_wave_max(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int cg_q(double v, double s) {  // int(v * s + 0.5), truncated
    const double t = v * s + 0.5;
    return t >= 2147483647.0 ? 2147483647 : (t <= -2147483647.0 ? -2147483647 : (int)t);
}

// Candidate j of the pool -> its record and de-duplication key.
__device__ void cg_emit(int (*keys)[CG_KEY], WeftPrim* out, double* f64, int j, int kind, const double* g, int ng,
                        const double* c0, const double* c1, double alpha) {
    WeftPrim P;
    P.kind = kind;
#pragma unroll
    for (int i = 0; i < 7; ++i) P.geom[i] = i < ng ? (float)g[i] : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        P.color0[c] = (float)c0[c];
        P.color1[c] = c1 ? (float)c1[c] : 0.0f;
    }
    P.alpha = (float)alpha;
    out[j] = P;
    if (f64) {  // the same values before rounding to float32: geometry, colours, alpha
        double* d = f64 + (int64_t)j * 14;
        for (int i = 0; i < 7; ++i) d[i] = i < ng ? g[i] : 0.0;
        for (int c = 0; c < 3; ++c) {
            d[7 + c] = c0[c];
            d[10 + c] = c1 ? c1[c] : 0.0;
        }
        d[13] = alpha;
    }
    int* k = keys[j];
    int n = 0;
    k[n++] = kind;
    for (int i = 0; i < ng; ++i) k[n++] = cg_q(g[i], 4.0);
    for (int c = 0; c < 3; ++c) k[n++] = cg_q(c0[c], 64.0);
    if (c1)
        for (int c = 0; c < 3; ++c) k[n++] = cg_q(c1[c], 64.0);
    k[n++] = cg_q(alpha, 32.0);
    while (n < CG_KEY) k[n++] = -2147483647 - 1;
}

// De-duplication (keep the first candidate of every key, _dedup_candidates), compaction in
// place (every kept record is read before any is written), and records of no kind in the
// slots past the pool.
__device__ __forceinline__ uint32_t cg_key_hash(const int* k) {
    uint32_t h = 0x811c9dc5u;
#pragma unroll
    for (int t = 0; t < CG_KEY; ++t) h = (h ^ (uint32_t)k[t]) * 0x01000193u;
    return h ^ (h >> 15);
}

// keep[j] = 1 for the first candidate of every key (hashes compared first, keys on a match).
__device__ void cg_first_of_key(int (*keys)[CG_KEY], uint32_t* hash, int* keep, int n_total) {
    const int tid = threadIdx.x;
    for (int j = tid; j < n_total; j += CG_THREADS) hash[j] = cg_key_hash(keys[j]);
    __syncthreads();
    for (int j = tid; j < n_total; j += CG_THREADS) {
        const uint32_t hj = hash[j];
        int k = 1;
        int i = 0;
        // four hashes per trip (their loads in flight together); a trip with a match looks at
        // its four in order, keys compared in full
        for (; i + 4 <= j && k; i += 4) {
            const uint32_t h0 = hash[i], h1 = hash[i + 1], h2 = hash[i + 2], h3 = hash[i + 3];
            if (h0 == hj || h1 == hj || h2 == hj || h3 == hj) {
                for (int u = i; u < i + 4 && k; ++u)
                    if (hash[u] == hj) {
                        bool same = true;
                        for (int t = 0; t < CG_KEY && same; ++t) same = keys[u][t] == keys[j][t];
                        if (same) k = 0;
                    }
            }
        }
        for (; i < j && k; ++i)
            if (hash[i] == hj) {
                bool same = true;
                for (int t = 0; t < CG_KEY && same; ++t) same = keys[i][t] == keys[j][t];
                if (same) k = 0;
            }
        keep[j] = k;
    }
}

// keep[j] (1 to keep, 0 to drop, j < n <= 2 * CG_THREADS) becomes the candidate's rank among
// the kept ones, or -1: ballots over chunks of 64 candidates, then the chunks' counts in
// order.  Returns the number kept (to every thread); ends with a barrier.
__device__ int cg_keep_ranks(int* keep, int n) {
    __shared__ int s_chunk[2 * CG_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63;
    int kq[2], rank[2];
    for (int q = 0; q < 2; ++q) {
        const int j = tid + q * CG_THREADS;
        kq[q] = j < n ? keep[j] : 0;
        const unsigned long long m = __ballot(kq[q] != 0);
        rank[q] = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) s_chunk[(tid >> 6) + q * (CG_THREADS / 64)] = __popcll(m);
    }
    __syncthreads();
    int off[2] = {0, 0}, total = 0;
    for (int c = 0; c < 2 * CG_THREADS / 64; ++c) {
        const int v = s_chunk[c];
        if (c < (tid >> 6)) off[0] += v;
        if (c < (tid >> 6) + CG_THREADS / 64) off[1] += v;
        total += v;
    }
    for (int q = 0; q < 2; ++q) {
        const int j = tid + q * CG_THREADS;
        if (j < n) keep[j] = kq[q] ? off[q] + rank[q] : -1;
    }
    __syncthreads();
    return total;
}

__device__ void cg_finish(int (*keys)[CG_KEY], uint32_t* hash, int* keep, int* n_kept_lds, WeftPrim* out,
                          double* f64, int n_total, int max_cands, int32_t* out_count) {
    const int tid = threadIdx.x;
    cg_first_of_key(keys, hash, keep, n_total);
    __syncthreads();
    const int total = cg_keep_ranks(keep, n_total);
    if (tid == 0) {
        *out_count = total;
        *n_kept_lds = total;
    }
    WeftPrim mine[2];
    int dst[2] = {-1, -1};
    for (int q = 0; q < 2; ++q) {
        const int j = tid + q * CG_THREADS;
        if (j < n_total && keep[j] >= 0) {
            mine[q] = out[j];
            dst[q] = keep[j];
        }
    }
    __syncthreads();
    for (int q = 0; q < 2; ++q)
        if (dst[q] >= 0) out[dst[q]] = mine[q];
    const int n_kept = total;
    for (int j = n_kept + tid; j < max_cands; j += CG_THREADS) {
        WeftPrim P;
        P.kind = -1;
        for (int i = 0; i < 7; ++i) P.geom[i] = 0.0f;
        for (int c = 0; c < 3; ++c) P.color0[c] = P.color1[c] = 0.0f;
        P.alpha = 0.0f;
        out[j] = P;
    }
    if (f64) {
        double v[2][14];
        for (int q = 0; q < 2; ++q)
            if (dst[q] >= 0)
                for (int i = 0; i < 14; ++i) v[q][i] = f64[(int64_t)(tid + q * CG_THREADS) * 14 + i];
        __syncthreads();
        for (int q = 0; q < 2; ++q)
            if (dst[q] >= 0)
                for (int i = 0; i < 14; ++i) f64[(int64_t)dst[q] * 14 + i] = v[q][i];
    }
}

// Work is split so that each independent in-order sum runs on lanes of its own wavefront
// (lanes of one wavefront running different loops would take turns).
__device__ void cg_body(const WeftCandGenArgs& A) {
    __shared__ CgShared S;
    extern __shared__ float4 weft_smem[];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ts = A.tile_size, n = ts * ts;
    const int slot = blockIdx.x;
    const int tile = A.tiles[slot];
    const double hf = (double)(ts - 1), wf = (double)(ts - 1);
    float* R = reinterpret_cast<float*>(weft_smem);
    float* G = R + n;
    float* B = G + n;
    float* LUM = B + n;
    float* MAG = LUM + n;
    float* GX = MAG + n;              // gx, gy: then the luminance sort keys (2n words)
    float* GY = GX + n;
    uint64_t* SORT = reinterpret_cast<uint64_t*>(GX);
    uint32_t* MSORT = reinterpret_cast<uint32_t*>(GY + n);
    WeftPrim* out = A.out + (int64_t)slot * A.max_cands;
    double* f64 = A.out_f64 ? A.out_f64 + (int64_t)slot * A.max_cands * 14 : nullptr;

    // pixels and luminance (float32, left to right as NumPy evaluates it)
    const float* src = A.pixels + A.tile_pix[tile] * 3;
    for (int p = tid; p < n; p += CG_THREADS) {
        const float r = src[3 * p], g = src[3 * p + 1], b = src[3 * p + 2];
        R[p] = r;
        G[p] = g;
        B[p] = b;
        LUM[p] = (0.2126f * r + 0.7152f * g) + 0.0722f * b;
    }
    __syncthreads();
    // np.gradient (central differences inside, one-sided at the borders) and the magnitude
    for (int p = tid; p < n; p += CG_THREADS) {
        const int y = p / ts, x = p - y * ts;
        const float gx = x == 0 ? LUM[p + 1] - LUM[p] : (x == ts - 1 ? LUM[p] - LUM[p - 1] : (LUM[p + 1] - LUM[p - 1]) / 2.0f);
        const float gy = y == 0 ? LUM[p + ts] - LUM[p] : (y == ts - 1 ? LUM[p] - LUM[p - ts] : (LUM[p + ts] - LUM[p - ts]) / 2.0f);
        GX[p] = gx;
        GY[p] = gy;
        MAG[p] = sqrtf(gx * gx + gy * gy);
    }
    __syncthreads();
    // flat means (pairwise) and block means (row by row), one kind per wavefront
    if (wave == 0) {
        if (lane < 3) {
            const float v = cg_pairwise(lane == 0 ? MAG : (lane == 1 ? GX : GY), n) / (float)n;
            if (lane == 0) S.avg_mag = v;
            else if (lane == 1) S.gmx = v;
            else S.gmy = v;
        }
    } else if (wave == 1) {
        if (lane < 3) S.mean_all[lane] = cg_seq_add(0.0f, R + lane * n, n) / (float)n;
    } else if (wave == 2) {
        if (lane < 12) {
            const int q = lane / 3, c = lane % 3, qh = ts / 2;
            S.quad[q][c] = cg_rows_add(R + c * n, ts, (q & 1) * qh, qh, (q >> 1) * qh, qh) / (float)(qh * qh);
        }
    } else {
        if (lane < 48) {
            const int q = lane / 3, c = lane % 3, sh = ts / 4;
            S.sub[q][c] = cg_rows_add(R + c * n, ts, (q & 3) * sh, sh, (q >> 2) * sh, sh) / (float)(sh * sh);
        }
    }
    if (tid < 3) {
        S.km_cnt[tid] = S.km_sx[tid] = S.km_sy[tid] = 0;
        S.km_lo[tid] = 2147483647;
        S.km_hi[tid] = -1;
    }
    __syncthreads();
    // sort keys: (luminance, index) for the palette and the extreme pixels; magnitude values
    for (int p = tid; p < n; p += CG_THREADS) {
        SORT[p] = ((uint64_t)cg_fkey(LUM[p]) << 32) | (uint32_t)p;
        MSORT[p] = cg_fkey(MAG[p]);
    }
    __syncthreads();
    for (int k = 2; k <= n; k <<= 1)
        for (int j = k >> 1; j >= 1; j >>= 1) {
            for (int t = tid; t < n / 2; t += CG_THREADS) {
                const int lo = (t / j) * 2 * j + (t % j), hi = lo + j;
                const bool up = (lo & k) == 0;
                const uint64_t a = SORT[lo], b = SORT[hi];
                if ((b < a) == up) {
                    SORT[lo] = b;
                    SORT[hi] = a;
                }
                const uint32_t c = MSORT[lo], d = MSORT[hi];
                if ((d < c) == up) {
                    MSORT[lo] = d;
                    MSORT[hi] = c;
                }
            }
            __syncthreads();
        }
    // extreme pixels and the linear-patch geometry
    const int n_ang = A.quality >= 80 ? 16 : (A.quality >= 50 ? 12 : 8);
    const int n_lines = n_ang + 36;
    if (tid < 6) {
        S.vl[tid] = (int)(SORT[tid] & 0xffffffffu);
        S.pk[tid] = (int)(SORT[n - 6 + tid] & 0xffffffffu);
    } else if (tid >= 32 && tid < 32 + n_lines) {
        const int l = tid - 32;
        double vx, vy, cx, cy, f;
        if (l < n_ang) {
            vx = A.cos_ang[l];
            vy = A.sin_ang[l];
            cx = wf * 0.5;
            cy = hf * 0.5;
            f = 0.6;
        } else {
            const int o = (l - n_ang) / 6, a = (l - n_ang) % 6;
            const double ox[6] = {0.25, 0.75, 0.25, 0.75, 0.5, 0.5}, oy[6] = {0.25, 0.25, 0.75, 0.75, 0.25, 0.75};
            vx = A.cos6[a];
            vy = A.sin6[a];
            cx = wf * ox[o];
            cy = hf * oy[o];
            f = 0.4;
        }
        S.line_geom[l][0] = cx - vx * wf * f;
        S.line_geom[l][1] = cy - vy * hf * f;
        S.line_geom[l][2] = cx + vx * wf * f;
        S.line_geom[l][3] = cy + vy * hf * f;
    }
    __syncthreads();
    // percentiles (every thread reads the few sorted keys it needs); then the masked means and
    // the linear patches' side means in one loop over the pixels in order (waves 0-2: lanes
    // 0-8 the bright / dark / edge masks, lanes 9 .. 9 + 3 n_lines - 1 the lines), the k-means
    // groups on wave 3
    float p25, p75, gthr;
    {
        auto lum_at = [&](int i) { return cg_funkey((uint32_t)(SORT[i] >> 32)); };
        const float virt25 = (float)(n - 1) * 0.25f, virt75 = (float)(n - 1) * 0.75f;
        auto pct = [&](float virt) {
            const float prev = floorf(virt);
            const int i = (int)prev;
            const float a = lum_at(i), b = lum_at(i + 1);
            const float gamma = virt - prev, diff = b - a;
            float r = a + diff * gamma;
            if (gamma >= 0.5f) r = b - diff * (1.0f - gamma);
            return r;
        };
        p25 = pct(virt25);
        p75 = pct(virt75);
        gthr = cg_percentile(MSORT, n, 0.75f);
    }
    if (tid < 9 + 3 * n_lines) {
        const bool side_lane = tid >= 9;
        const int e = tid / 3, l = (tid - 9) / 3, c = side_lane ? (tid - 9) % 3 : tid % 3;
        float x0 = 0.0f, y0 = 0.0f, dyl = 0.0f, dxl = 0.0f;
        if (side_lane) {
            x0 = (float)S.line_geom[l][0];
            y0 = (float)S.line_geom[l][1];
            const float x1 = (float)S.line_geom[l][2], y1 = (float)S.line_geom[l][3];
            dyl = y1 - y0;
            dxl = x1 - x0;
        }
        const float* plane = R + c * n;
        float sa = 0.0f, sb = 0.0f;
        int na = 0;
        for (int y = 0; y < ts; ++y) {
            const float ty = ((float)y - y0) * dxl;
            for (int x = 0; x < ts; x += 8) {
                const int p = y * ts + x;
                float v[8], lm[8], mg[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    v[j] = plane[p + j];
                    lm[j] = LUM[p + j];
                    mg[j] = MAG[p + j];
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float side = ((float)(x + j) - x0) * dyl - ty;
                    const bool m = side_lane ? side >= 0.0f
                                             : (e == 0 ? lm[j] > p75 : (e == 1 ? lm[j] < p25 : mg[j] >= gthr));
                    if (m) {
                        sa += v[j];
                        ++na;
                    } else {
                        sb += v[j];
                    }
                }
            }
        }
        if (!side_lane) {
            S.ext[e][c] = na ? sa / (float)na : 0.0f;
            if (c == 0) S.ext_n[e] = na;
        } else {
            const int nb = n - na;
            S.line_col[l][0][c] = na ? sa / (float)na : S.mean_all[c];
            S.line_col[l][1][c] = nb ? sb / (float)nb : S.mean_all[c];
        }
    } else if (wave == 3 && lane < 27) {
        const int g = lane / 3, c = lane % 3;
        const int k = g < 2 ? 2 : (g < 5 ? 3 : 4);
        const int jj = g < 2 ? g : (g < 5 ? g - 2 : g - 5);
        const double step = (double)n / (double)k;
        const int a = (int)((double)jj * step), b = (int)((double)(jj + 1) * step);
        const float* plane = R + c * n;
        float v;
        if (b <= a) {
            v = plane[(int)(SORT[a < n - 1 ? a : n - 1] & 0xffffffffu)];
        } else {
            float s = 0.0f;
            int i = a;
            for (; i + 8 <= b; i += 8) {
                float w[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) w[q] = plane[(int)(SORT[i + q] & 0xffffffffu)];
#pragma unroll
                for (int q = 0; q < 8; ++q) s += w[q];
            }
            for (; i < b; ++i) s += plane[(int)(SORT[i] & 0xffffffffu)];
            v = s / (float)(b - a);
        }
        S.km[g][c] = v;
    }
    __syncthreads();
    // k-means labels (quality >= 80; first of the smallest float32 distances) and their regions'
    // integer statistics, reduced per wavefront
    if (A.quality >= 80) {
        int cnt[3] = {0, 0, 0}, sx[3] = {0, 0, 0}, sy[3] = {0, 0, 0};
        int lo[3] = {2147483647, 2147483647, 2147483647}, hi[3] = {-1, -1, -1};
        for (int p = tid; p < n; p += CG_THREADS) {
            int best = 0;
            float bd = 0.0f;
            for (int k = 0; k < 3; ++k) {
                float d = 0.0f;
                for (int c = 0; c < 3; ++c) {
                    const float e = (R + c * n)[p] - S.km[2 + k][c];
                    d = d + e * e;
                }
                if (k == 0 || d < bd) {
                    bd = d;
                    best = k;
                }
            }
            const int y = p / ts, x = p - y * ts;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (best == k) {
                    ++cnt[k];
                    sx[k] += x;
                    sy[k] += y;
                    lo[k] = min(lo[k], x * n + p);
                    hi[k] = max(hi[k], x * n + (n - 1 - p));
                }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int c = cg_wave_sum(cnt[k]), a = cg_wave_sum(sx[k]), b = cg_wave_sum(sy[k]);
            const int l = cg_wave_min(lo[k]), h = cg_wave_max(hi[k]);
            if (lane == 0 && c) {
                atomicAdd(&S.km_cnt[k], c);
                atomicAdd(&S.km_sx[k], a);
                atomicAdd(&S.km_sy[k], b);
                atomicMin(&S.km_lo[k], l);
                atomicMax(&S.km_hi[k], h);
            }
        }
    }
    // palette, alphas, gradient line, section sizes
    if (tid == 0) {
        auto near = [&](const double* c) {
            for (int i = 0; i < S.n_pal; ++i) {
                double d = c[0] - S.pal[i][0];
                double s = d * d;
                d = c[1] - S.pal[i][1];
                s = s + d * d;
                d = c[2] - S.pal[i][2];
                s = s + d * d;
                if (s < 0.002) return true;
            }
            return false;
        };
        S.n_pal = 0;
        for (int c = 0; c < 3; ++c) S.pal[0][c] = cg_ct((double)S.mean_all[c]);
        S.n_pal = 1;
        for (int g = 0; g < 9; ++g) {
            double c[3];
            for (int i = 0; i < 3; ++i) c[i] = cg_ct((double)S.km[g][i]);
            if (!near(c)) {
                for (int i = 0; i < 3; ++i) S.pal[S.n_pal][i] = c[i];
                ++S.n_pal;
            }
        }
        for (int e = 0; e < 3; ++e) {  // bright, dark, edge (the tile mean when the mask is empty)
            double c[3];
            for (int i = 0; i < 3; ++i) c[i] = S.ext_n[e] ? cg_ct((double)S.ext[e][i]) : S.pal[0][i];
            for (int i = 0; i < 3; ++i) S.ext[e][i] = (float)c[i];  // clamped copies for the lines
            if (!near(c)) {
                for (int i = 0; i < 3; ++i) S.pal[S.n_pal][i] = c[i];
                ++S.n_pal;
            }
        }
        const double ba = 0.25 + (double)S.avg_mag * 3.0;
        S.base_alpha = ba < 1.0 ? ba : 1.0;
        const double gmx = S.gmx, gmy = S.gmy;
        const double gnorm = sqrt(gmx * gmx + gmy * gmy);
        S.grad_lines = gnorm > 1e-4 ? 10 : 0;
        if (S.grad_lines) {
            const double lx = -gmy / (gnorm + 1e-8), ly = gmx / (gnorm + 1e-8);
            const double cx = wf * 0.5, cy = hf * 0.5;
            S.gline[0] = cx - lx * wf * 0.7;
            S.gline[1] = cy - ly * hf * 0.7;
            S.gline[2] = cx + lx * wf * 0.7;
            S.gline[3] = cy + ly * hf * 0.7;
        }
        S.n_tri_x = 8;  // 4 brightest-pixel and 4 darkest-pixel triangles (6 extreme pixels each)
    }
    __syncthreads();
    // the pool, in the reference's order
    int n_km_tri = 0;
    if (A.quality >= 80)
        for (int k = 0; k < 3; ++k) n_km_tri += S.km_cnt[k] >= 3 ? 2 : 0;
    const int pc_total = S.n_pal < 3 ? S.n_pal : 3;
    const int n_total =
        4 * S.n_pal + 8 + 16 + 3 * n_ang + 36 + 80 + S.grad_lines + 45 + 8 * pc_total + 8 + S.n_tri_x + n_km_tri;
    for (int j = tid; j < n_total; j += CG_THREADS) {
        double g[7] = {0, 0, 0, 0, 0, 0, 0};
        double c0[3], c1[3];
        int r = j;
        const double ext0[3] = {S.ext[2][0], S.ext[2][1], S.ext[2][2]};  // edge colour
        auto col = [&](double* d, const float* s) {
            for (int c = 0; c < 3; ++c) d[c] = (double)s[c];
        };
        auto colct = [&](double* d, const float* s) {
            for (int c = 0; c < 3; ++c) d[c] = cg_ct((double)s[c]);
        };
        if (r < 4 * S.n_pal) {  // constant patches of the palette
            const double al[4] = {1.0, 0.7, 0.4, 0.2};
            cg_emit(S.keys, out, f64, j, 0, g, 0, S.pal[r / 4], nullptr, al[r % 4]);
            continue;
        }
        r -= 4 * S.n_pal;
        if (r < 8) {  // quadrant means
            colct(c0, S.quad[r / 2]);
            cg_emit(S.keys, out, f64, j, 0, g, 0, c0, nullptr, (r & 1) ? 0.3 : 0.5);
            continue;
        }
        r -= 8;
        if (r < 16) {  // 4 x 4 block means
            colct(c0, S.sub[r]);
            cg_emit(S.keys, out, f64, j, 0, g, 0, c0, nullptr, 0.25);
            continue;
        }
        r -= 16;
        if (r < 3 * n_ang + 36) {  // linear patches
            const int l = r < 3 * n_ang ? r / 3 : n_ang + (r - 3 * n_ang);
            const double al[3] = {1.0, 0.7, 0.4};
            const double a = r < 3 * n_ang ? al[r % 3] : 0.5;
            for (int i = 0; i < 4; ++i) g[i] = S.line_geom[l][i];
            col(c0, S.line_col[l][0]);
            col(c1, S.line_col[l][1]);
            cg_emit(S.keys, out, f64, j, 1, g, 4, c0, c1, a);
            continue;
        }
        r -= 3 * n_ang + 36;
        const double thick[5] = {0.4, 0.8, 1.2, 2.0, 3.0};
        if (r < 80) {  // lines
            const int t = r / 16, q = r % 16;
            g[4] = thick[t];
            if (q < 10) {
                const double fr[5] = {0.0, 0.25, 0.5, 0.75, 1.0};
                const double f = fr[q / 2];
                if ((q & 1) == 0) {
                    g[0] = wf * f; g[1] = 0.0; g[2] = wf * f; g[3] = hf;
                } else {
                    g[0] = 0.0; g[1] = hf * f; g[2] = wf; g[3] = hf * f;
                }
                cg_emit(S.keys, out, f64, j, 2, g, 5, ext0, nullptr, S.base_alpha);
            } else {
                const int i = q - 10;
                const int e = i / 2 == 0 ? 2 : (i / 2 == 1 ? 0 : 1);  // edge, bright, dark
                col(c0, S.ext[e]);
                if ((i & 1) == 0) {
                    g[0] = 0.0; g[1] = 0.0; g[2] = wf; g[3] = hf;
                } else {
                    g[0] = 0.0; g[1] = hf; g[2] = wf; g[3] = 0.0;
                }
                cg_emit(S.keys, out, f64, j, 2, g, 5, c0, nullptr, S.base_alpha);
            }
            continue;
        }
        r -= 80;
        if (r < S.grad_lines) {  // lines across the mean gradient
            for (int i = 0; i < 4; ++i) g[i] = S.gline[i];
            g[4] = thick[r / 2];
            col(c0, S.ext[(r & 1) ? 0 : 2]);
            cg_emit(S.keys, out, f64, j, 2, g, 5, c0, nullptr, S.base_alpha);
            continue;
        }
        r -= S.grad_lines;
        if (r < 45) {  // quadratic curves
            const double cf[3] = {0.25, 0.5, 0.75};
            if (r < 27) {
                const double tt[3] = {0.5, 1.0, 2.0};
                g[0] = 0.0; g[1] = hf * 0.5; g[2] = wf * cf[(r / 3) % 3]; g[3] = hf * cf[r / 9];
                g[4] = wf; g[5] = hf * 0.5; g[6] = tt[r % 3];
            } else {
                const int q = r - 27;
                g[0] = wf * 0.5; g[1] = 0.0; g[2] = wf * cf[q / 6]; g[3] = hf * cf[(q / 2) % 3];
                g[4] = wf * 0.5; g[5] = hf; g[6] = (q & 1) ? 1.0 : 0.5;
            }
            cg_emit(S.keys, out, f64, j, 3, g, 7, ext0, nullptr, S.base_alpha);
            continue;
        }
        r -= 45;
        const int pc = S.n_pal < 3 ? S.n_pal : 3;
        if (r < 8 * pc) {  // corner-centre triangles
            const double cx[4] = {0.0, wf, wf, 0.0}, cy[4] = {0.0, 0.0, hf, hf};
            const int i = r / (2 * pc), q = r % (2 * pc);
            g[0] = cx[i]; g[1] = cy[i]; g[2] = cx[(i + 1) & 3]; g[3] = cy[(i + 1) & 3];
            g[4] = wf * 0.5; g[5] = hf * 0.5;
            cg_emit(S.keys, out, f64, j, 4, g, 6, S.pal[q / 2], nullptr, (q & 1) ? 0.3 : 0.6);
            continue;
        }
        r -= 8 * pc;
        if (r < 8) {  // half-tile triangles
            const double hv[4][6] = {{0, 0, wf, 0, wf, hf}, {0, 0, wf, hf, 0, hf}, {0, 0, wf, 0, 0, hf},
                                     {wf, 0, wf, hf, 0, hf}};
            for (int i = 0; i < 6; ++i) g[i] = hv[r / 2][i];
            cg_emit(S.keys, out, f64, j, 4, g, 6, S.pal[0], nullptr, (r & 1) ? 0.3 : 0.5);
            continue;
        }
        r -= 8;
        if (r < S.n_tri_x) {  // triangles on the brightest, then the darkest, pixels
            const int* ix = r < 4 ? S.pk : S.vl;
            const int i = r & 3;
            float cs[3];
            for (int c = 0; c < 3; ++c) {
                float s = 0.0f;
                for (int v = 0; v < 3; ++v) s += (R + c * n)[ix[i + v]];
                cs[c] = s / 3.0f;
            }
            for (int v = 0; v < 3; ++v) {
                g[2 * v] = (double)(ix[i + v] % ts);
                g[2 * v + 1] = (double)(ix[i + v] / ts);
            }
            colct(c0, cs);
            cg_emit(S.keys, out, f64, j, 4, g, 6, c0, nullptr, 0.5);
            continue;
        }
        r -= S.n_tri_x;
        {  // k-means region triangles (quality >= 80)
            int lab = 0, q = r;
            while (lab < 3 && (S.km_cnt[lab] < 3 || q >= 2)) {
                if (S.km_cnt[lab] >= 3) q -= 2;
                ++lab;
            }
            const int a = S.km_lo[lab] % n, b = (n - 1) - S.km_hi[lab] % n;
            const double cnt = (double)S.km_cnt[lab];
            g[0] = (double)(a % ts); g[1] = (double)(a / ts);
            g[2] = (double)(b % ts); g[3] = (double)(b / ts);
            g[4] = (double)S.km_sx[lab] / cnt; g[5] = (double)S.km_sy[lab] / cnt;
            colct(c0, S.km[2 + lab]);
            cg_emit(S.keys, out, f64, j, 4, g, 6, c0, nullptr, q ? 0.3 : 0.5);
        }
    }
    __syncthreads();
    cg_finish(S.keys, S.hash, S.keep, &S.n_kept, out, f64, n_total, A.max_cands, A.out_count + slot);
}

// ---------------------------------------------------------------------------------------
// Residual candidate pools: encoder._generate_residual_candidates (src/weft/encoder.py:
// 1255-1366), candidates aimed at the error between a tile and its current prediction.
// The error pixels are ranked by x with a stable order (the reference's np.argsort does not
// promise one for equal x); it picks the curves' end and middle points.
constexpr int CR_MAXC = 128;  // 117 before de-duplication at most

struct CrShared {
    int keys[CR_MAXC][CG_KEY];
    int keep[CR_MAXC];
    uint32_t hash[CR_MAXC];
    float mt[3], mp[3], me[3], mall[3], quad[4][3];
    int quad_n[4];
    float line_col[4][2][3];
    double line_geom[4][4];
    int line_na[4];
    float emx, emy;
    int n_total, n_kept;
    // error pixels: count, coordinate sums, min of x n + p, max of x n + (n - 1 - p), first and
    // last pixel, pixels per column
    int n_hi, sx, sy, key_lo, key_hi, p_first, p_last;
    int colc[64];
    int xmin, y_at_xmin, xmax, y_at_xmax, ymin, ymax;  // extent of the error pixels
    int o_first, o_mid, o_last;                       // pixel indices in the stable x-order
    double hi_col[3], corr[3];
};

__device__ __forceinline__ float cr_clip(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

// Work is split so that each independent in-order sum runs on lanes of its own wavefront;
// the prediction is staged in LDS once the gradients and sort keys are consumed.
__device__ void cr_body(const WeftResGenArgs& A) {
    __shared__ CrShared S;
    extern __shared__ float4 weft_smem[];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ts = A.tile_size, n = ts * ts;
    const int slot = blockIdx.x;
    const int tile = A.tiles[slot];
    const double hf = (double)(ts - 1), wf = (double)(ts - 1);
    float* T = reinterpret_cast<float*>(weft_smem);  // target planes
    float* EL = T + 3 * n;                            // mean absolute error, then the error mask
    float* GX = EL + n;
    float* GY = GX + n;
    uint32_t* SK = reinterpret_cast<uint32_t*>(GY + n);
    float* P = GX;                                    // prediction planes, over GX, GY and SK
    int* HI = reinterpret_cast<int*>(EL);
    WeftPrim* out = A.out + (int64_t)slot * A.max_cands;
    double* f64 = A.out_f64 ? A.out_f64 + (int64_t)slot * A.max_cands * 14 : nullptr;

    const int64_t base = A.tile_pix[tile] * 3;
    auto pred_at = [&](int p, int c) {
        const float q = A.pred[base + 3 * p + c];
        return A.clip_pred ? clip01(q) : q;
    };
    for (int p = tid; p < n; p += CG_THREADS) {
        float e[3];
        for (int c = 0; c < 3; ++c) {
            const float t = A.pixels[base + 3 * p + c];
            T[c * n + p] = t;
            e[c] = t - pred_at(p, c);
        }
        EL[p] = ((0.0f + fabsf(e[0]) + fabsf(e[1])) + fabsf(e[2])) / 3.0f;
    }
    if (tid < 64) S.colc[tid] = 0;
    if (tid == 0) {
        S.n_hi = S.sx = S.sy = 0;
        S.key_lo = S.p_first = 2147483647;
        S.key_hi = S.p_last = -1;
    }
    __syncthreads();
    for (int p = tid; p < n; p += CG_THREADS) {
        const int y = p / ts, x = p - y * ts;
        GX[p] = x == 0 ? EL[p + 1] - EL[p] : (x == ts - 1 ? EL[p] - EL[p - 1] : (EL[p + 1] - EL[p - 1]) / 2.0f);
        GY[p] = y == 0 ? EL[p + ts] - EL[p] : (y == ts - 1 ? EL[p] - EL[p - ts] : (EL[p + ts] - EL[p - ts]) / 2.0f);
        SK[p] = cg_fkey(EL[p]);
    }
    __syncthreads();
    if (tid < 2) {  // mean error gradient (pairwise), while the others start sorting
        const float v = cg_pairwise(tid == 0 ? GX : GY, n) / (float)n;
        if (tid == 0) S.emx = v;
        else S.emy = v;
    }
    for (int k = 2; k <= n; k <<= 1)
        for (int j = k >> 1; j >= 1; j >>= 1) {
            for (int t = tid; t < n / 2; t += CG_THREADS) {
                const int lo = (t / j) * 2 * j + (t % j), hi = lo + j;
                const bool up = (lo & k) == 0;
                const uint32_t a = SK[lo], b = SK[hi];
                if ((b < a) == up) {
                    SK[lo] = b;
                    SK[hi] = a;
                }
            }
            __syncthreads();
        }
    const float thresh = cg_percentile(SK, n, 60.0f / 100.0f);
    __syncthreads();
    // error mask, prediction planes, and the error pixels' positions reduced per wavefront
    {
        int cnt = 0, sx = 0, sy = 0, lo = 2147483647, hi = -1, pf = 2147483647, pl = -1;
        for (int p = tid; p < n; p += CG_THREADS) {
            const int h = EL[p] >= thresh ? 1 : 0;
            HI[p] = h;
            for (int c = 0; c < 3; ++c) P[c * n + p] = pred_at(p, c);
            if (h) {
                const int y = p / ts, x = p - y * ts;
                ++cnt;
                sx += x;
                sy += y;
                lo = min(lo, x * n + p);
                hi = max(hi, x * n + (n - 1 - p));
                pf = min(pf, p);
                pl = max(pl, p);
                atomicAdd(&S.colc[x], 1);
            }
        }
        cnt = cg_wave_sum(cnt);
        sx = cg_wave_sum(sx);
        sy = cg_wave_sum(sy);
        lo = cg_wave_min(lo);
        hi = cg_wave_max(hi);
        pf = cg_wave_min(pf);
        pl = cg_wave_max(pl);
        if (lane == 0 && cnt) {
            atomicAdd(&S.n_hi, cnt);
            atomicAdd(&S.sx, sx);
            atomicAdd(&S.sy, sy);
            atomicMin(&S.key_lo, lo);
            atomicMax(&S.key_hi, hi);
            atomicMin(&S.p_first, pf);
            atomicMax(&S.p_last, pl);
        }
    }
    __syncthreads();
    // means over the error pixels (row-major), quadrant means, the tile mean, line side means
    if (wave == 0) {
        if (lane < 9) {
            const int w = lane / 3, c = lane % 3;
            const float* tp = T + c * n;
            const float* pp = P + c * n;
            float s = 0.0f;
            int cnt = 0;
            for (int p0 = 0; p0 < n; p0 += 8) {
                int h[8];
                float t[8], q[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    h[j] = HI[p0 + j];
                    t[j] = tp[p0 + j];
                    q[j] = pp[p0 + j];
                }
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (h[j]) {
                        s += w == 0 ? t[j] : (w == 1 ? q[j] : t[j] - q[j]);
                        ++cnt;
                    }
            }
            const float m = cnt ? s / (float)cnt : 0.0f;
            if (w == 0) S.mt[c] = m;
            else if (w == 1) S.mp[c] = m;
            else S.me[c] = m;
        }
    } else if (wave == 1) {
        if (lane < 12) {
            const int q = lane / 3, c = lane % 3, qh = ts / 2;
            const int x0 = (q & 1) * qh, y0 = (q >> 1) * qh;
            float s = 0.0f;
            int cnt = 0;
            for (int y = y0; y < y0 + qh; ++y)
                for (int x = x0; x < x0 + qh; x += 4) {
                    int h[4];
                    float v[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        h[j] = HI[y * ts + x + j];
                        v[j] = T[c * n + y * ts + x + j];
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (h[j]) {
                            s += v[j];
                            ++cnt;
                        }
                }
            S.quad[q][c] = cnt ? s / (float)cnt : 0.0f;
            if (c == 0) S.quad_n[q] = cnt;
        }
    } else if (wave == 2) {
        if (lane < 3) S.mall[lane] = cg_seq_add(0.0f, T + lane * n, n) / (float)n;
    } else if (lane < 12) {
        // the four lines: along the mean error gradient at 0.5, 0.3, 0.7 of the tile and across
        // it at the centre (_split_by_line: the Python-float line enters as float32)
        const int l = lane / 3, c = lane % 3;
        const double emx = S.emx, emy = S.emy;
        const double en = sqrt(emx * emx + emy * emy) + 1e-8;
        const double vx = emx / en, vy = emy / en;
        double g[4];
        if (l < 3) {
            const double offs[3] = {0.5, 0.3, 0.7};
            const double cx = wf * offs[l], cy = hf * offs[l];
            g[0] = cx - vx * wf * 0.6;
            g[1] = cy - vy * hf * 0.6;
            g[2] = cx + vx * wf * 0.6;
            g[3] = cy + vy * hf * 0.6;
        } else {
            const double pvx = -vy, pvy = vx, cx = wf * 0.5, cy = hf * 0.5;
            g[0] = cx - pvx * wf * 0.6;
            g[1] = cy - pvy * hf * 0.6;
            g[2] = cx + pvx * wf * 0.6;
            g[3] = cy + pvy * hf * 0.6;
        }
        if (c == 0)
            for (int i = 0; i < 4; ++i) S.line_geom[l][i] = g[i];
        const float x0 = (float)g[0], y0 = (float)g[1];
        const float dy = (float)(g[3] - g[1]), dx = (float)(g[2] - g[0]);
        const float* plane = T + c * n;
        float sa = 0.0f, sb = 0.0f;
        int na = 0;
        for (int y = 0; y < ts; ++y) {
            const float ty = ((float)y - y0) * dx;
            for (int x = 0; x < ts; x += 8) {
                float v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = plane[y * ts + x + j];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float side = ((float)(x + j) - x0) * dy - ty;
                    if (side >= 0.0f) {
                        sa += v[j];
                        ++na;
                    } else {
                        sb += v[j];
                    }
                }
            }
        }
        const int nb = n - na;
        S.line_col[l][0][c] = na ? sa / (float)na : 0.0f;  // an empty side takes the tile mean below
        S.line_col[l][1][c] = nb ? sb / (float)nb : 0.0f;
        if (c == 0) S.line_na[l] = na;
    }
    __syncthreads();
    if (tid == 0) {
        const int cnt = S.n_hi;
        S.xmin = cnt ? S.key_lo / n : 0;
        S.y_at_xmin = cnt ? (S.key_lo % n) / ts : 0;
        S.xmax = cnt ? S.key_hi / n : 0;
        S.y_at_xmax = cnt ? ((n - 1) - S.key_hi % n) / ts : 0;
        S.ymin = cnt ? S.p_first / ts : 0;
        S.ymax = cnt ? S.p_last / ts : 0;
        // stable sort by x: rank r lies in the first column whose running count passes r, at
        // the (r - count before) -th error pixel of that column from the top
        int ranks[3] = {0, cnt / 2, cnt - 1};
        int res[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) {
            int r = ranks[k], x = 0;
            while (x < ts && r >= S.colc[x]) {
                r -= S.colc[x];
                ++x;
            }
            int y = 0;
            if (x < ts)
                for (; y < ts; ++y)
                    if (HI[y * ts + x]) {
                        if (r == 0) break;
                        --r;
                    }
            res[k] = y * ts + x;
        }
        S.o_first = res[0];
        S.o_mid = res[1];
        S.o_last = res[2];
        for (int c = 0; c < 3; ++c) {
            S.hi_col[c] = cg_ct((double)cr_clip(S.mt[c]));
            S.corr[c] = cg_ct((double)cr_clip(S.mp[c] + S.me[c]));
        }
        for (int l = 0; l < 4; ++l)
            for (int c = 0; c < 3; ++c) {
                if (S.line_na[l] == 0) S.line_col[l][0][c] = S.mall[c];
                if (S.line_na[l] == n) S.line_col[l][1][c] = S.mall[c];
            }
        int nq = 0;
        for (int q = 0; q < 4; ++q) nq += S.quad_n[q] ? 3 : 0;
        S.n_total = S.n_hi == 0 ? 0 : 15 + nq + 11 + 64 + (S.n_hi >= 6 ? 3 : 0) + (S.n_hi >= 3 ? 12 : 0);
    }
    __syncthreads();

    const int n_total = S.n_total;
    const double ecx = S.n_hi ? (double)S.sx / (double)S.n_hi : 0.0;
    const double ecy = S.n_hi ? (double)S.sy / (double)S.n_hi : 0.0;
    for (int j = tid; j < n_total; j += CG_THREADS) {
        double g[7] = {0, 0, 0, 0, 0, 0, 0};
        double c0[3], c1[3];
        int r = j;
        if (r < 15) {  // constant patches: error, corrected and target colour
            const double al[5] = {0.6, 0.4, 0.25, 0.15, 0.08};
            cg_emit(S.keys, out, f64, j, 0, g, 0, r / 5 == 1 ? S.corr : S.hi_col, nullptr, al[r % 5]);
            continue;
        }
        r -= 15;
        int q = 0, rq = r;
        for (; q < 4; ++q) {
            if (!S.quad_n[q]) continue;
            if (rq < 3) break;
            rq -= 3;
        }
        if (q < 4) {  // quadrant means of the error pixels
            const double al[3] = {0.35, 0.2, 0.1};
            for (int c = 0; c < 3; ++c) c0[c] = cg_ct((double)cr_clip(S.quad[q][c]));
            cg_emit(S.keys, out, f64, j, 0, g, 0, c0, nullptr, al[rq]);
            continue;
        }
        r = rq;
        if (r < 11) {  // linear patches along the error gradient and across it
            const int l = r < 9 ? r / 3 : 3;
            const double al9[3] = {0.5, 0.3, 0.15};
            const double a = r < 9 ? al9[r % 3] : (r == 9 ? 0.4 : 0.2);
            for (int i = 0; i < 4; ++i) g[i] = S.line_geom[l][i];
            for (int c = 0; c < 3; ++c) {
                c0[c] = (double)S.line_col[l][0][c];
                c1[c] = (double)S.line_col[l][1][c];
            }
            cg_emit(S.keys, out, f64, j, 1, g, 4, c0, c1, a);
            continue;
        }
        r -= 11;
        if (r < 60) {  // lines through the error centroid
            const double tt[5] = {0.3, 0.6, 1.0, 1.5, 2.5}, al[3] = {0.5, 0.3, 0.15};
            const int t = r / 12, rem = r % 12;
            const int a = (rem % 6) / 2;
            g[4] = tt[t];
            if ((rem & 1) == 0) {
                g[0] = ecx; g[1] = 0.0; g[2] = ecx; g[3] = hf;
            } else {
                g[0] = 0.0; g[1] = ecy; g[2] = wf; g[3] = ecy;
            }
            cg_emit(S.keys, out, f64, j, 2, g, 5, S.hi_col, nullptr, al[a]);  // target colour = error colour
            continue;
        }
        r -= 60;
        if (r < 4) {  // diagonals
            g[4] = r < 2 ? 0.5 : 1.0;
            if ((r & 1) == 0) {
                g[0] = 0.0; g[1] = 0.0; g[2] = wf; g[3] = hf;
            } else {
                g[0] = 0.0; g[1] = hf; g[2] = wf; g[3] = 0.0;
            }
            cg_emit(S.keys, out, f64, j, 2, g, 5, S.hi_col, nullptr, 0.25);
            continue;
        }
        r -= 4;
        if (S.n_hi >= 6) {
            if (r < 3) {  // a curve through the error pixels' x-order ends and middle
                const double p0x = S.o_first % ts, p0y = S.o_first / ts;
                const double pmx = S.o_mid % ts, pmy = S.o_mid / ts;
                const double p1x = S.o_last % ts, p1y = S.o_last / ts;
                double cx = 2 * pmx - (p0x + p1x) / 2, cy = 2 * pmy - (p0y + p1y) / 2;
                cx = cx < wf ? cx : wf;
                cx = cx > 0.0 ? cx : 0.0;
                cy = cy < hf ? cy : hf;
                cy = cy > 0.0 ? cy : 0.0;
                g[0] = p0x; g[1] = p0y; g[2] = cx; g[3] = cy; g[4] = p1x; g[5] = p1y;
                g[6] = r == 0 ? 0.5 : (r == 1 ? 1.0 : 2.0);
                cg_emit(S.keys, out, f64, j, 3, g, 7, S.hi_col, nullptr, 0.3);
                continue;
            }
            r -= 3;
        }
        {  // triangles over the error pixels' extent (error colour, then the equal target colour)
            const double al[3] = {0.4, 0.2, 0.1};
            const int gi = r / 6, a = (r % 6) / 2;
            g[0] = (double)S.xmin;
            g[1] = (double)(gi == 0 ? S.y_at_xmin : S.ymin);
            g[2] = (double)S.xmax;
            g[3] = (double)(gi == 0 ? S.y_at_xmax : S.ymax);
            g[4] = ecx;
            g[5] = ecy;
            cg_emit(S.keys, out, f64, j, 4, g, 6, S.hi_col, nullptr, al[a]);
        }
    }
    __syncthreads();
    cg_finish(S.keys, S.hash, S.keep, &S.n_kept, out, f64, n_total, A.max_cands, A.out_count + slot);
}
