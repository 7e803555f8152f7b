// Device-wide selection primitives behind the reference's CCCL helpers
// (src/weft/cuda_backend.py:80-181): first-index argmin of float64 values, exclusive scan
// of int64 values, compaction of flags to indices, and per-segment top-k.
// Included by weft_kernels.hip inside its anonymous namespace.
//
// Every result is deterministic.  Orders follow NumPy: NaN sorts after +inf and compares
// as the minimum in argmin (np.argmin returns the first NaN), -0.0 equals +0.0, and equal
// keys keep index order.

constexpr int SEL_BLOCK = 256;
constexpr int SEL_TILE = 2048;       // elements per workgroup in scan / count passes (8 per thread)

// ---------------------------------------------------------------- argmin (first index)
// (value, index) pairs ordered as np.argmin sees them: NaN first, then by value, then index.
__device__ __forceinline__ bool sel_before(double a, int64_t ia, double b, int64_t ib) {
    const bool na = a != a, nb = b != b;
    if (na || nb) return na && (!nb || ia < ib);
    return a < b || (a == b && ia < ib);
}

__device__ __forceinline__ void sel_argmin_wave(double& v, int64_t& i) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double v2 = __shfl_xor(v, off, 64);
        const int64_t i2 = __shfl_xor(i, off, 64);
        if (sel_before(v2, i2, v, i)) {
            v = v2;
            i = i2;
        }
    }
}

// One workgroup reduces values [lo, hi) into (*out_v, *out_i); an empty range gives index -1.
__device__ void sel_argmin_range(const double* vals, const int64_t* idx_in, int64_t lo, int64_t hi, double* out_v,
                                 int64_t* out_i) {
    __shared__ double wv[SEL_BLOCK / 64];
    __shared__ int64_t wi[SEL_BLOCK / 64];
    double v = __builtin_huge_val();
    int64_t i = -1;
    for (int64_t p = lo + threadIdx.x; p < hi; p += blockDim.x) {
        const double x = vals[p];
        const int64_t ix = idx_in ? idx_in[p] : p;
        if (ix >= 0 && (i < 0 || sel_before(x, ix, v, i))) {
            v = x;
            i = ix;
        }
    }
    // empty lanes (index -1) must lose every comparison
    if (i < 0) {
        v = __builtin_huge_val();
        i = INT64_MAX;
    }
    sel_argmin_wave(v, i);
    if ((threadIdx.x & 63) == 0) {
        wv[threadIdx.x >> 6] = v;
        wi[threadIdx.x >> 6] = i;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            if (sel_before(wv[w], wi[w], v, i)) {
                v = wv[w];
                i = wi[w];
            }
        *out_v = v;
        *out_i = i == INT64_MAX ? -1 : i;
    }
}

__device__ void sel_argmin_partial_body(const WeftSelectArgs& A) {
    const int64_t per = (A.n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * per;
    const int64_t hi = lo + per < A.n ? lo + per : A.n;
    sel_argmin_range(A.f64_in, nullptr, lo, hi, A.work_f64 + blockIdx.x, A.work_i64 + blockIdx.x);
}

__device__ void sel_argmin_final_body(const WeftSelectArgs& A) {
    double v;
    int64_t i;
    sel_argmin_range(A.work_f64, A.work_i64, 0, A.n_blocks, &v, &i);
    if (threadIdx.x == 0) A.out_i64[0] = i;
}

// ---------------------------------------------------------------- exclusive scan (int64)
// Inclusive scan of one value per thread across the workgroup (Hillis-Steele in LDS).
__device__ __forceinline__ int64_t sel_block_inclusive(int64_t x, int64_t* buf) {
    const int t = threadIdx.x;
    buf[t] = x;
    __syncthreads();
    for (int off = 1; off < SEL_BLOCK; off <<= 1) {
        const int64_t y = t >= off ? buf[t - off] : 0;
        __syncthreads();
        buf[t] += y;
        __syncthreads();
    }
    return buf[t];
}

// Value of element p for the scan: int64 input, or 0/1 flags (compaction).
__device__ __forceinline__ int64_t sel_load(const WeftSelectArgs& A, int64_t p) {
    if (p >= A.n) return 0;
    return A.i64_in ? A.i64_in[p] : (int64_t)(A.u8_in[p] != 0);
}

// Pass 1: sum of every tile of SEL_TILE elements -> work_i64[tile].
__device__ void sel_tile_sums_body(const WeftSelectArgs& A) {
    __shared__ int64_t buf[SEL_BLOCK];
    const int64_t base = (int64_t)blockIdx.x * SEL_TILE + threadIdx.x * (SEL_TILE / SEL_BLOCK);
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SEL_TILE / SEL_BLOCK; ++k) s += sel_load(A, base + k);
    const int64_t tot = sel_block_inclusive(s, buf);
    if (threadIdx.x == SEL_BLOCK - 1) A.work_i64[blockIdx.x] = tot;
}

// Pass 2 (one workgroup): exclusive scan of the n_blocks tile sums in place; the grand
// total goes to *out_total.
__device__ void sel_scan_tiles_body(const WeftSelectArgs& A) {
    __shared__ int64_t buf[SEL_BLOCK];
    int64_t carry = 0;
    for (int64_t c0 = 0; c0 < A.n_blocks; c0 += SEL_BLOCK) {
        const int64_t p = c0 + threadIdx.x;
        const int64_t x = p < A.n_blocks ? A.work_i64[p] : 0;
        const int64_t inc = sel_block_inclusive(x, buf);
        if (p < A.n_blocks) A.work_i64[p] = carry + inc - x;
        carry += buf[SEL_BLOCK - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0 && A.out_total) *A.out_total = carry;
}

// Pass 3: each tile rescans its elements from its offset.  Scan: out[p] = exclusive
// prefix.  Compaction (u8_in): out[prefix] = p for every set flag.
__device__ void sel_scan_apply_body(const WeftSelectArgs& A) {
    __shared__ int64_t buf[SEL_BLOCK];
    constexpr int PER = SEL_TILE / SEL_BLOCK;
    const int64_t base = (int64_t)blockIdx.x * SEL_TILE + threadIdx.x * PER;
    int64_t x[PER];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        x[k] = sel_load(A, base + k);
        s += x[k];
    }
    int64_t run = A.work_i64[blockIdx.x] + sel_block_inclusive(s, buf) - s;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int64_t p = base + k;
        if (p < A.n) {
            if (A.i64_in) A.out_i64[p] = run;
            else if (x[k]) A.out_i64[run] = p;
        }
        run += x[k];
    }
}

// ---------------------------------------------------------------- segmented top-k
// Two stable LSD radix sorts (8-bit digits, 8 passes each) of 64-bit keys with 32-bit
// values:
//   A: score keys in NumPy's ascending order (`largest` sorts the complemented keys of the
//      elements taken in reverse index order, which is NumPy's stable ascending order read
//      backwards), values = element index p; position q in this order is the output order.
//   B: segment keys in A order, values = q; stability keeps every segment's run in A order.
// An element of B is kept when fewer than k earlier elements of its run exist (the element
// k places back belongs to another segment); its flag is set at q, and compacting the flags
// lists the kept q in output order.

constexpr int RS_ITEMS = 8;                      // keys per thread per pass
constexpr int RS_TILE = SEL_BLOCK * RS_ITEMS;    // keys per workgroup
constexpr int RS_WAVES = SEL_BLOCK / 64;
static_assert(SEL_BLOCK == 256, "one thread per radix digit");

// float64 -> uint64 whose unsigned order is NumPy's ascending order (-0.0 = +0.0, NaN last).
__device__ __forceinline__ uint64_t sel_f64_key(double v) {
    if (v != v) return ~0ull;
    if (v == 0.0) v = 0.0;
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}

__device__ __forceinline__ uint64_t sel_lanemask_lt() {
    const unsigned lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

__device__ __forceinline__ uint32_t rs_digit(uint64_t key, int shift) { return (uint32_t)(key >> shift) & 255u; }

// A's input: keys of the scores, values = index.
__device__ void sel_rs_score_keys_body(const WeftSelectArgs& A) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    const int64_t p = A.largest ? A.n - 1 - j : j;
    const uint64_t key = sel_f64_key(A.f64_in[p]);
    A.rs_keys_out[j] = A.largest ? ~key : key;
    A.rs_vals_out[j] = (uint32_t)p;
}

// Per workgroup, the count of each digit among its RS_TILE keys, stored digit-major:
// table[digit * n_blocks + block].
__device__ void sel_rs_hist_body(const WeftSelectArgs& A) {
    __shared__ int s_h[RS_WAVES][256];
    const int t = threadIdx.x, w = t >> 6;
    for (int i = t; i < RS_WAVES * 256; i += SEL_BLOCK) (&s_h[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int64_t j = base + (int64_t)r * SEL_BLOCK + t;
        if (j < A.n) atomicAdd(&s_h[w][rs_digit(A.rs_keys_in[j], A.rs_shift)], 1);
    }
    __syncthreads();
    int c = 0;
#pragma unroll
    for (int ww = 0; ww < RS_WAVES; ++ww) c += s_h[ww][t];
    A.rs_table[(int64_t)t * A.n_blocks + blockIdx.x] = c;
}

// Stable scatter of one pass.  The workgroup walks its keys in index order, SEL_BLOCK at a
// time; a key's place is the digit's global start for this workgroup (exclusive scan of the
// digit-major table) + the keys of that digit earlier in the workgroup: earlier rounds,
// earlier waves of this round, and lower lanes of its wave (found by matching the 8 digit
// bits across the wave with ballots).
__device__ void sel_rs_scatter_body(const WeftSelectArgs& A) {
    __shared__ int64_t s_base[256];
    __shared__ int s_run[256];
    __shared__ int s_cnt[RS_WAVES][256];
    __shared__ int s_off[RS_WAVES][256];
    const int t = threadIdx.x, w = t >> 6;
    s_base[t] = A.rs_prefix[(int64_t)t * A.n_blocks + blockIdx.x];
    s_run[t] = 0;
    const uint64_t lt = sel_lanemask_lt();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 1
    for (int r = 0; r < RS_ITEMS; ++r) {
        for (int i = t; i < RS_WAVES * 256; i += SEL_BLOCK) (&s_cnt[0][0])[i] = 0;
        __syncthreads();
        const int64_t j = base + (int64_t)r * SEL_BLOCK + t;
        const bool valid = j < A.n;
        const uint64_t key = valid ? A.rs_keys_in[j] : 0ull;
        const uint32_t val = valid ? A.rs_vals_in[j] : 0u;
        const uint32_t d = rs_digit(key, A.rs_shift);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) s_cnt[w][d] = __popcll(peers);
        __syncthreads();
        int run = s_run[t];  // thread t: digit t
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ++ww) {
            s_off[ww][t] = run;
            run += s_cnt[ww][t];
        }
        s_run[t] = run;
        __syncthreads();
        if (valid) {
            const int64_t pos = s_base[d] + s_off[w][d] + rank;
            A.rs_keys_out[pos] = key;
            A.rs_vals_out[pos] = val;
        }
    }
}

// B's input: the segment key of each element in A order, values = position in A order.
__device__ void sel_rs_segment_keys_body(const WeftSelectArgs& A) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= A.n) return;
    A.rs_keys_out[q] = (uint64_t)A.i64_in[A.rs_order[q]] ^ (1ull << 63);
    A.rs_vals_out[q] = (uint32_t)q;
}

// Flag (at its A position) every element of B among the first k of its segment's run.
__device__ void sel_rs_keep_body(const WeftSelectArgs& A) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    const bool keep = j < A.k || A.rs_keys_in[j - A.k] != A.rs_keys_in[j];
    A.flags[A.rs_vals_in[j]] = keep ? 1 : 0;
}

// out[i] = index of the i-th kept element in A order.
__device__ void sel_rs_emit_body(const WeftSelectArgs& A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *A.count_in) return;
    A.out_i64[i] = (int64_t)A.rs_order[A.positions[i]];
}
