// Edge-driven candidate pools on the device: edge_analysis.generate_edge_driven_candidates
// (src/weft/edge_analysis.py:139-315) for every tile of one size, one 256-thread workgroup
// per tile.  Included by weft_kernels.hip inside its anonymous namespace, after the
// template pools (it uses their statistics helpers, emission and de-duplication).
//
// Per tile: the Sobel gradients as the reference's slices form them (every tap reads
// lum[y-1][x-1], taps in row-major order, float32), the magnitude, the angle modulo pi
// (float32 arctangent: atan of min / max rounded to float32, then pi/2 - t, pi - t and the
// sign in float32, atan(1) = 0.7853981; then NumPy's float32 remainder), the 12-bin
// magnitude histogram with NumPy's pairwise float32 sums over each bin's pixels in order,
// the dominant orientations; then the palette and quadrant patches, two-colour linear
// patches across each orientation, and triangles over the 3-means regions, in the
// reference's order and de-duplicated by its keys.  The line and curve sections need edge
// points, which the reference finds only where the 60th percentile of the magnitudes
// reaches 1e-6; the Sobel taps cancel inside the tile (to zero or a rounding residue), so
// that needs values far outside [0, 1].  Such a tile gets the count -1 instead of a pool.

constexpr int ED_MAXC = 128;  // 4 x 10 palette + 8 quadrant + 5 x 12 linear + 6 triangle = 114
constexpr int ED_BINS = 12;

struct EdShared {
    int keys[ED_MAXC][CG_KEY];
    int keep[ED_MAXC];
    uint32_t hash[ED_MAXC];
    float mean_all[3], quad[4][3], km[9][3];
    float hist[ED_BINS];
    int bin_cnt[ED_BINS], bin_off[ED_BINS];
    int orient[ED_BINS];
    double line_geom[ED_BINS][4];
    double line_col[ED_BINS][2][3];
    double pal[10][3];
    int n_orient, n_pal, n_kept, small_cnt;
    // 3-means regions: pixel count, x and y sums, min of x n + p, max of x n + (n - 1 - p)
    int km_cnt[3], km_sx[3], km_sy[3], km_lo[3], km_hi[3];
};

// NumPy's float32 pairwise sum of a[0 .. n) as np.add.reduce forms it: 0 + pw(a, n), where
// pw adds fewer than 8 values in order, up to 128 values with 8 interleaved accumulators (the
// remainder after the last full group in order), and more by halves rounded down to a
// multiple of 8.  One thread; an explicit stack instead of recursion.
__device__ float ed_leaf(const float* a, int n) {
    if (n < 8) {
        float res = 0.0f;
        for (int i = 0; i < n; ++i) res += a[i];
        return res;
    }
    float r[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    int i = 8;
    for (; i < n - (n % 8); i += 8)
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
}

__device__ float ed_pairwise(const float* a, int n) {
    int st_off[16], st_len[16], st_state[16];
    float st_left[16];
    int sp = 0, off = 0, len = n;
    float v = 0.0f;
    for (;;) {
        while (len > 128) {  // pending right half, descend into the left one
            int n2 = len / 2;
            n2 -= n2 % 8;
            st_off[sp] = off + n2;
            st_len[sp] = len - n2;
            st_state[sp] = 0;
            ++sp;
            len = n2;
        }
        v = ed_leaf(a + off, len);
        bool descend = false;
        while (sp > 0) {
            const int f = sp - 1;
            if (st_state[f] == 0) {  // v is the left value: now the right half
                st_left[f] = v;
                st_state[f] = 1;
                off = st_off[f];
                len = st_len[f];
                descend = true;
                break;
            }
            v = st_left[f] + v;
            --sp;
        }
        if (!descend) break;
    }
    return 0.0f + v;
}

// The reference's Sobel gradients at pixel (x, y) of a ts x ts tile: tap (dy, dx) covers
// rows [1, ts-1), [1, ts), [2, ts) for dy = 0, 1, 2 (columns likewise with dx) and adds
// k[dy][dx] * lum[y-1][x-1] (edge_analysis.py:29-36).
__device__ __forceinline__ void ed_sobel(const float* LUM, int ts, int x, int y, float& gx, float& gy) {
    constexpr float KX[3][3] = {{-1.0f, 0.0f, 1.0f}, {-2.0f, 0.0f, 2.0f}, {-1.0f, 0.0f, 1.0f}};
    constexpr float KY[3][3] = {{-1.0f, -2.0f, -1.0f}, {0.0f, 0.0f, 0.0f}, {1.0f, 2.0f, 1.0f}};
    const int lo[3] = {1, 1, 2};
    const int hi[3] = {ts - 1, ts, ts};
    gx = 0.0f;
    gy = 0.0f;
    if (x < 1 || y < 1) return;
    const float L = LUM[(y - 1) * ts + (x - 1)];
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx)
            if (y >= lo[dy] && y < hi[dy] && x >= lo[dx] && x < hi[dx]) {
                gx = gx + KX[dy][dx] * L;
                gy = gy + KY[dy][dx] * L;
            }
}

// The float32 angle modulo pi of (gx, gy), as the module comment states.
__device__ __forceinline__ float ed_angle(float gx, float gy) {
    const float ay = fabsf(gy), ax = fabsf(gx);
    const float big = fmaxf(ay, ax), small = fminf(ay, ax);
    float t = big > 0.0f ? (float)atan((double)small / (double)big) : 0.0f;
    if (ay == ax && ax > 0.0f) t = 0.7853981f;
    if (ay > ax) t = 1.5707964f - t;
    if (signbit(gx)) t = 3.1415927f - t;
    if (signbit(gy)) t = -t;
    float mod = fmodf(t, 3.1415927f);  // NumPy's divmod remainder with a positive divisor
    if (mod != 0.0f) {
        if (mod < 0.0f) mod += 3.1415927f;
    } else {
        mod = 0.0f;
    }
    return mod;
}

// The dominant orientations from the histogram (table indices into cos_o / sin_o); returns
// their number.  np.max and np.argmax: NaN propagates and is the first maximum.
__device__ int ed_orientations(const float* hist, int* orient) {
    float hmax = hist[0];
    int amax = 0;
    for (int i = 1; i < ED_BINS && hmax == hmax; ++i)
        if (hist[i] > hmax || hist[i] != hist[i]) {
            hmax = hist[i];
            amax = i;
        }
    int m = 0;
    if (hmax < 1e-6f) {  // no edges: 0, pi/4, pi/2, 3pi/4 (table entries 12 .. 15)
        for (int i = 0; i < 4; ++i) orient[m++] = ED_BINS + i;
    } else {
        const float thr = hmax * 0.2f;
        for (int i = 0; i < ED_BINS; ++i)
            if (hist[i] >= thr) orient[m++] = i;
        if (m == 0) orient[m++] = ED_BINS + 4 + amax;  // edges[argmax] + pi / 24 (entries 16 .. 27)
    }
    return m;
}

// Work is split so that each independent in-order sum runs on lanes of its own wavefront
// (lanes of one wavefront running different loops would take turns).
__device__ void ed_body(const WeftEdgeGenArgs& A) {
    __shared__ EdShared S;
    extern __shared__ float4 weft_smem[];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ts = A.tile_size, n = ts * ts;
    const int slot = blockIdx.x;
    const int tile = A.tiles[slot];
    const double hf = (double)(ts - 1), wf = (double)(ts - 1);
    float* R = reinterpret_cast<float*>(weft_smem);              // R, G, B planes
    float* LUM = R + 3 * n;
    float* MAG = LUM + n;
    float* MS = MAG + n;                                         // each bin's magnitudes, in order
    uint64_t* SORT = reinterpret_cast<uint64_t*>(MS + n);        // (luminance, index) keys
    uint8_t* BIN = reinterpret_cast<uint8_t*>(SORT + n);         // histogram bin (ED_BINS: none)
    WeftPrim* out = A.out + (int64_t)slot * A.max_cands;
    double* f64 = A.out_f64 ? A.out_f64 + (int64_t)slot * A.max_cands * 14 : nullptr;

    const float* src = A.pixels + A.tile_pix[tile] * 3;
    for (int p = tid; p < n; p += CG_THREADS) {
        const float r = src[3 * p], g = src[3 * p + 1], b = src[3 * p + 2];
        R[p] = r;
        R[n + p] = g;
        R[2 * n + p] = b;
        LUM[p] = (0.2126f * r + 0.7152f * g) + 0.0722f * b;
    }
    if (tid == 0) S.small_cnt = 0;
    if (tid < 3) {
        S.km_cnt[tid] = S.km_sx[tid] = S.km_sy[tid] = 0;
        S.km_lo[tid] = 2147483647;
        S.km_hi[tid] = -1;
    }
    __syncthreads();
    // gradients, magnitude, histogram bin, sort key; magnitudes below 9.99e-7 counted for the
    // edge-point test (a float32 interpolation of two of them stays below the reference's 1e-6)
    int small = 0;
    for (int p = tid; p < n; p += CG_THREADS) {
        const int y = p / ts, x = p - y * ts;
        float gx, gy;
        ed_sobel(LUM, ts, x, y, gx, gy);
        const float mag = sqrtf(gx * gx + gy * gy);
        MAG[p] = mag;
        small += (double)mag < 9.99e-7;
        const double a = (double)ed_angle(gx, gy);
        int bin = ED_BINS;
        for (int i = 0; i < ED_BINS; ++i)
            if (a >= A.edges[i] && a < A.edges[i + 1]) bin = i;
        BIN[p] = (uint8_t)bin;
        SORT[p] = ((uint64_t)cg_fkey(LUM[p]) << 32) | (uint32_t)p;
    }
    small = cg_wave_sum(small);
    if (lane == 0) atomicAdd(&S.small_cnt, small);
    __syncthreads();
    if (wave == 0) {
        // each bin's magnitudes gathered in pixel order: counts by ballots, then each lane's
        // slot = its bin's running offset + the earlier lanes of that bin (n is a multiple of 64)
        int cnt[ED_BINS];
#pragma unroll
        for (int b = 0; b < ED_BINS; ++b) cnt[b] = 0;
        for (int p0 = 0; p0 < n; p0 += 64) {
            const int bin = BIN[p0 + lane];
#pragma unroll
            for (int b = 0; b < ED_BINS; ++b) cnt[b] += __popcll(__ballot(bin == b));
        }
        int off[ED_BINS];
        int run = 0;
#pragma unroll
        for (int b = 0; b < ED_BINS; ++b) {
            off[b] = run;
            run += cnt[b];
        }
        if (lane < ED_BINS) {
#pragma unroll
            for (int b = 0; b < ED_BINS; ++b)
                if (lane == b) {
                    S.bin_cnt[b] = cnt[b];
                    S.bin_off[b] = off[b];
                }
        }
        for (int p0 = 0; p0 < n; p0 += 64) {
            const int p = p0 + lane;
            const int bin = BIN[p];
            const float mag = MAG[p];
#pragma unroll
            for (int b = 0; b < ED_BINS; ++b) {
                const uint64_t m = __ballot(bin == b);
                if (bin == b) {
                    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                    MS[off[b] + below] = mag;
                }
                off[b] += __popcll(m);
            }
        }
    } else if (wave == 1) {
        if (lane < 3) S.mean_all[lane] = cg_seq_add(0.0f, R + lane * n, n) / (float)n;  // row by row = in order
    } else if (wave == 2) {
        if (lane < 12) {
            const int q = lane / 3, c = lane % 3, qh = ts / 2;
            const int x0 = (q & 1) * qh, y0 = (q >> 1) * qh;
            S.quad[q][c] = cg_rows_add(R + c * n, ts, x0, qh, y0, qh) / (float)(qh * qh);
        }
    }
    __syncthreads();
    if (tid < ED_BINS) S.hist[tid] = ed_pairwise(MS + S.bin_off[tid], S.bin_cnt[tid]);
    // luminance sort for the 2-, 3- and 4-means palettes (stable: keys carry the index)
    for (int k = 2; k <= n; k <<= 1)
        for (int j = k >> 1; j >= 1; j >>= 1) {
            __syncthreads();
            for (int t = tid; t < n / 2; t += CG_THREADS) {
                const int lo = (t / j) * 2 * j + (t % j), hi = lo + j;
                const bool up = (lo & k) == 0;
                const uint64_t a = SORT[lo], b = SORT[hi];
                if ((b < a) == up) {
                    SORT[lo] = b;
                    SORT[hi] = a;
                }
            }
        }
    __syncthreads();
    if (wave == 0) {
        if (lane == 0) S.n_orient = ed_orientations(S.hist, S.orient);
    } else if (wave == 1) {
        // k-means groups: equal-count runs of the luminance-sorted pixels, summed in order
        if (lane < 27) {
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
    } else if (wave == 2) {
        // two-colour split of every orientation's line: float32 side test, sums in pixel order
        int orient[ED_BINS];
        const int n_or = ed_orientations(S.hist, orient);
        const int l = lane / 3, c = lane % 3;
        if (l < n_or) {
            const int o = orient[l];
            const double vx = A.cos_o[o], vy = A.sin_o[o];
            const double nx = -vy, ny = vx;
            const double cx = wf * 0.5, cy = hf * 0.5;
            const double lx0 = cx - nx * wf * 0.6, ly0 = cy - ny * hf * 0.6;
            const double lx1 = cx + nx * wf * 0.6, ly1 = cy + ny * hf * 0.6;
            if (c == 0) {
                S.line_geom[l][0] = lx0;
                S.line_geom[l][1] = ly0;
                S.line_geom[l][2] = lx1;
                S.line_geom[l][3] = ly1;
            }
            const float x0 = (float)lx0, y0 = (float)ly0;
            const float dyl = (float)(ly1 - ly0), dxl = (float)(lx1 - lx0);
            const float* plane = R + c * n;
            float sa = 0.0f, sb = 0.0f;
            int na = 0;
            for (int y = 0; y < ts; ++y) {
                const float ty = ((float)y - y0) * dxl;
                const float* row = plane + y * ts;
                for (int x = 0; x < ts; x += 8) {
                    float v[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[j] = row[x + j];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const float side = ((float)(x + j) - x0) * dyl - ty;
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
            const double avg = cg_ct((double)S.mean_all[c]);
            S.line_col[l][0][c] = na ? cg_ct((double)(sa / (float)na)) : avg;
            S.line_col[l][1][c] = nb ? cg_ct((double)(sb / (float)nb)) : avg;
        }
    }
    __syncthreads();
    // 3-means labels (first of the smallest float32 distances) and their regions' integer
    // statistics, reduced per wavefront; the palette (one thread)
    if (A.quality >= 60) {
        int cnt[3] = {0, 0, 0}, sx[3] = {0, 0, 0}, sy[3] = {0, 0, 0};
        int lo[3] = {2147483647, 2147483647, 2147483647}, hi[3] = {-1, -1, -1};
        for (int p = tid; p < n; p += CG_THREADS) {
            int best = 0;
            float bd = 0.0f;
            for (int k = 0; k < 3; ++k) {
                float d = 0.0f;
                for (int c = 0; c < 3; ++c) {
                    const float e = R[c * n + p] - S.km[2 + k][c];
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
    }
    __syncthreads();

    // the pool, in the reference's order
    int n_tri = 0;
    if (A.quality >= 60)
        for (int k = 0; k < 3; ++k) n_tri += S.km_cnt[k] >= 3 ? 2 : 0;
    const int n_pal = S.n_pal, n_orient = S.n_orient;
    const int n_total = 4 * n_pal + 8 + 5 * n_orient + n_tri;
    for (int j = tid; j < n_total; j += CG_THREADS) {
        double g[7] = {0, 0, 0, 0, 0, 0, 0};
        double c0[3];
        int r = j;
        if (r < 4 * n_pal) {  // constant patches of the palette
            const double al[4] = {1.0, 0.7, 0.4, 0.2};
            cg_emit(S.keys, out, f64, j, 0, g, 0, S.pal[r / 4], nullptr, al[r % 4]);
            continue;
        }
        r -= 4 * n_pal;
        if (r < 8) {  // quadrant means
            for (int c = 0; c < 3; ++c) c0[c] = cg_ct((double)S.quad[r / 2][c]);
            cg_emit(S.keys, out, f64, j, 0, g, 0, c0, nullptr, (r & 1) ? 0.3 : 0.5);
            continue;
        }
        r -= 8;
        if (r < 5 * n_orient) {  // linear patches across each orientation
            const int l = r / 5, q = r % 5;
            const double* lg = S.line_geom[l];
            double a = 0.5;
            if (q < 3) {
                const double al[3] = {1.0, 0.7, 0.4};
                a = al[q];
                for (int i = 0; i < 4; ++i) g[i] = lg[i];
            } else {
                const int o = S.orient[l];
                const double nx = -A.sin_o[o], ny = A.cos_o[o];
                const double cx = wf * 0.5, cy = hf * 0.5;
                const double off = q == 3 ? 0.25 : 0.75;
                const double ox = cx + nx * wf * (off - 0.5), oy = cy + ny * hf * (off - 0.5);
                g[0] = ox - nx * wf * 0.4;
                g[1] = oy - ny * hf * 0.4;
                g[2] = ox + nx * wf * 0.4;
                g[3] = oy + ny * hf * 0.4;
            }
            cg_emit(S.keys, out, f64, j, 1, g, 4, S.line_col[l][0], S.line_col[l][1], a);
            continue;
        }
        r -= 5 * n_orient;
        {  // triangles over the 3-means regions of 3 pixels or more: leftmost, rightmost, centroid
            int lab = 0, q = r;
            while (lab < 3 && (S.km_cnt[lab] < 3 || q >= 2)) {
                if (S.km_cnt[lab] >= 3) q -= 2;
                ++lab;
            }
            const int a = S.km_lo[lab] % n, b = (n - 1) - S.km_hi[lab] % n;
            const double cnt = (double)S.km_cnt[lab];
            auto clampx = [&](double v) { return v > wf ? wf : (v < 0.0 ? 0.0 : v); };
            auto clampy = [&](double v) { return v > hf ? hf : (v < 0.0 ? 0.0 : v); };
            g[0] = clampx((double)(a % ts));
            g[1] = clampy((double)(a / ts));
            g[2] = clampx((double)(b % ts));
            g[3] = clampy((double)(b / ts));
            g[4] = clampx((double)S.km_sx[lab] / cnt);
            g[5] = clampy((double)S.km_sy[lab] / cnt);
            for (int c = 0; c < 3; ++c) c0[c] = cg_ct((double)S.km[2 + lab][c]);
            cg_emit(S.keys, out, f64, j, 4, g, 6, c0, nullptr, q ? 0.3 : 0.6);
        }
    }
    __syncthreads();
    cg_finish(S.keys, S.hash, S.keep, &S.n_kept, out, f64, n_total, A.max_cands, A.out_count + slot);
    // Edge points (the 60th percentile of the magnitudes at or above 1e-6) would need the line
    // and curve sections, which are not built: unless sorted magnitude floor(0.6 (n - 1)) + 1
    // is below 9.99e-7, the tile gets no pool
    if (tid == 0 && S.small_cnt < A.min_small) A.out_count[slot] = -1;
}

// Merged pools: _dedup_candidates(a + b) per tile (encoder.py:1383-1385, 2513-2515, the edge-
// driven pool followed by the template pool).  Keys are formed from the float64 values, as
// the generators form them; a count outside [0, cap] on either side gives the count -1.
constexpr int MG_MAXC = 2 * CG_THREADS;

__device__ __forceinline__ void mg_key(int* k, int kind, const double* d) {
    const int ng = kind == 0 ? 0 : (kind == 1 ? 4 : (kind == 2 ? 5 : (kind == 4 ? 6 : 7)));
    int n = 0;
    k[n++] = kind;
    for (int i = 0; i < ng; ++i) k[n++] = cg_q(d[i], 4.0);
    for (int c = 0; c < 3; ++c) k[n++] = cg_q(d[7 + c], 64.0);
    if (kind == 1)
        for (int c = 0; c < 3; ++c) k[n++] = cg_q(d[10 + c], 64.0);
    k[n++] = cg_q(d[13], 32.0);
    while (n < CG_KEY) k[n++] = -2147483647 - 1;
}

__device__ void mg_body(const WeftMergePoolsArgs& A) {
    __shared__ int keys[MG_MAXC][CG_KEY];
    __shared__ int keep[MG_MAXC];
    __shared__ uint32_t hash[MG_MAXC];
    const int tid = threadIdx.x, slot = blockIdx.x;
    const int na = A.a_count[slot], nb = A.b_count[slot];
    const bool ok = na >= 0 && na <= A.a_cap && nb >= 0 && nb <= A.b_cap;
    const int n = ok ? na + nb : 0;
    const WeftPrim* a = A.a + (int64_t)slot * A.a_cap;
    const WeftPrim* b = A.b + (int64_t)slot * A.b_cap;
    const double* af = A.a_f64 + (int64_t)slot * A.a_cap * 14;
    const double* bf = A.b_f64 + (int64_t)slot * A.b_cap * 14;
    WeftPrim* out = A.out + (int64_t)slot * A.out_cap;
    for (int j = tid; j < n; j += CG_THREADS) {
        const bool in_a = j < na;
        const int i = in_a ? j : j - na;
        mg_key(keys[j], (in_a ? a : b)[i].kind, (in_a ? af : bf) + (int64_t)i * 14);
    }
    __syncthreads();
    cg_first_of_key(keys, hash, keep, n);
    __syncthreads();
    const int run = cg_keep_ranks(keep, n);
    if (tid == 0) A.out_count[slot] = ok ? run : -1;
    for (int j = tid; j < n; j += CG_THREADS) {
        const int d = keep[j];
        if (d < 0) continue;
        const bool in_a = j < na;
        const int i = in_a ? j : j - na;
        out[d] = (in_a ? a : b)[i];
        if (A.out_f64) {
            const double* s = (in_a ? af : bf) + (int64_t)i * 14;
            double* o = A.out_f64 + ((int64_t)slot * A.out_cap + d) * 14;
            for (int q = 0; q < 14; ++q) o[q] = s[q];
        }
    }
    for (int j = run + tid; j < A.out_cap; j += CG_THREADS) {
        WeftPrim P;
        P.kind = -1;
        for (int i = 0; i < 7; ++i) P.geom[i] = 0.0f;
        for (int c = 0; c < 3; ++c) P.color0[c] = P.color1[c] = 0.0f;
        P.alpha = 0.0f;
        out[j] = P;
    }
}
