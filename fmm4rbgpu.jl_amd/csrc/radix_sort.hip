// radix_sort.hip -- stable LSD radix sort of 64-bit or 32-bit keys (with optional
// 32-bit values) for gfx950, 8-bit digits, three kernels per pass:
//   hist    : per-tile digit histogram (LDS atomics), digit-major table
//   scan    : per-digit exclusive scan over tiles + digit totals
//   scatter : stable in-tile ranking with wave64 ballots into LDS (the tile
//             ordered by digit), then each digit's run written contiguously
// Several independent arrays of the same length sort in one launch
// (blockIdx.y selects the array).  Used for the tree's per-axis presort
// (particle positions) and for grouping interaction pairs by target.
#include "fmm4rb_internal.hpp"

namespace fmm4rb {

constexpr int RS_THREADS = 256;  // 4 waves of 64
// keys per thread: a 6144-key tile writes each digit's run ~24 keys long (with 2048-key
// tiles, ~8 keys: partial HBM lines); the tree's presort at N = 25.6M 27.2 -> 25.5 ms
// (2048 -> 6144; 8192-key tiles hold a workgroup per CU and were slower)
constexpr int RS_ITEMS = 24;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_BINS = 256;

static inline int64_t rs_nblk(int64_t n) { return (n + RS_TILE - 1) / RS_TILE; }

size_t radix_scratch_words(int64_t n, int nseg) {
    return (size_t)nseg * RS_BINS * (size_t)rs_nblk(n) + (size_t)nseg * RS_BINS;
}

template <typename KT>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const KT *__restrict__ keys, int64_t n,
                                                        int shift, uint32_t *__restrict__ hist,
                                                        int nblk) {
    __shared__ uint32_t h[RS_BINS];
    const int seg = blockIdx.y;
    const KT *k = keys + (int64_t)seg * n;
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        int64_t i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(k[i] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[((int64_t)seg * RS_BINS + threadIdx.x) * nblk + blockIdx.x] = h[threadIdx.x];
}

// Block-wide inclusive scan of one value per thread (256 threads).
__device__ inline uint32_t block_incl_scan256(uint32_t v, uint32_t *s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < RS_THREADS; off <<= 1) {
        uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t r = s[threadIdx.x];
    __syncthreads();
    return r;
}

// Exclusive scan of each row (len entries) of rows[gridDim.x*gridDim.y][len];
// the row total goes to tot[row].  Thread t loads SC_I consecutive entries at once
// (independent loads), and one block-wide scan of the threads' sums covers ST * SC_I
// entries: a row of 12,500 tile counts (the tree's partition at N = 25.6M) takes one
// step instead of 49 barrier-bound 256-entry steps.
constexpr int SC_I = 16;
template <int ST>
__global__ __launch_bounds__(ST) void k_scan_rows(uint32_t *__restrict__ rows, int len,
                                                  uint32_t *__restrict__ tot) {
    __shared__ uint32_t ws[ST / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t row = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    uint32_t *r = rows + row * len;
    uint32_t carry = 0;
    for (int base = 0; base < len; base += ST * SC_I) {
        const int b0 = base + tid * SC_I;
        uint32_t v[SC_I], s = 0;
#pragma unroll
        for (int i = 0; i < SC_I; ++i) {
            v[i] = b0 + i < len ? r[b0 + i] : 0u;
            s += v[i];
        }
        uint32_t x = s;  // inclusive scan over the wave, then the waves' totals
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) ws[wave] = x;
        __syncthreads();
        uint32_t pre = 0, total = 0;
#pragma unroll
        for (int w = 0; w < ST / 64; ++w) {
            const uint32_t t = ws[w];
            pre += w < wave ? t : 0u;
            total += t;
        }
        uint32_t run = carry + pre + (x - s);
#pragma unroll
        for (int i = 0; i < SC_I; ++i) {
            if (b0 + i < len) r[b0 + i] = run;
            run += v[i];
        }
        carry += total;
        __syncthreads();  // ws is rewritten by the next step
    }
    if (tid == 0) tot[row] = carry;
}

// Stable in-tile ranking: each wave owns RS_ITEMS x 64 consecutive keys of the tile
// and ranks them 64 at a time in order, with per-wave digit counters in LDS (the
// wave's own LDS accesses are in order, so no block barrier is needed between its
// steps); one barrier and a per-digit prefix over the waves then give every key's
// slot in the digit-ordered tile.
template <typename KT, bool HASVAL>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const KT *__restrict__ kin, const uint32_t *__restrict__ vin, KT *__restrict__ kout,
    uint32_t *__restrict__ vout, int64_t n, int shift, const uint32_t *__restrict__ hist, int nblk,
    const uint32_t *__restrict__ dtot) {
    constexpr int RS_WAVES = RS_THREADS / 64;
    __shared__ uint32_t doff[RS_BINS];  // global position of this tile's first key of each digit
    __shared__ uint32_t lst[RS_BINS];   // local position of the same key within the tile
    __shared__ uint32_t wcnt[RS_WAVES][RS_BINS];  // per-wave counts, then offsets over waves
    __shared__ uint32_t s[RS_THREADS];
    __shared__ KT sk[RS_TILE];
    __shared__ uint32_t sv[HASVAL ? RS_TILE : 1];
    const int seg = blockIdx.y;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t so = (int64_t)seg * n;

    // digit base = exclusive scan of digit totals + this tile's prefix; the tile's
    // own count of each digit = next tile's prefix (or the total) - this prefix
    const uint32_t tv = dtot[seg * RS_BINS + tid];
    const uint32_t incl = block_incl_scan256(tv, s);
    const int64_t hrow = ((int64_t)seg * RS_BINS + tid) * nblk;
    const uint32_t pre = hist[hrow + blockIdx.x];
    const uint32_t cnt = (blockIdx.x + 1 < (unsigned)nblk ? hist[hrow + blockIdx.x + 1] : tv) - pre;
    doff[tid] = incl - tv + pre;
    const uint32_t lincl = block_incl_scan256(cnt, s);
    lst[tid] = lincl - cnt;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) wcnt[w][tid] = 0;
    __syncthreads();

    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const int64_t wbase = base + (int64_t)wave * (RS_ITEMS * 64);
    const int tn = (int)min<int64_t>(RS_TILE, n - base);
    KT key[RS_ITEMS];
    uint32_t val[RS_ITEMS], rnk[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int64_t i = wbase + r * 64 + lane;
        const bool valid = i < n;
        key[r] = valid ? kin[so + i] : (KT)0;
        val[r] = (HASVAL && valid) ? vin[so + i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const bool valid = wbase + r * 64 + lane < n;
        const uint32_t dg = (uint32_t)(key[r] >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dg >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        // this wave's keys of digit dg before this step, then the lower lanes of this step
        const uint32_t prior = valid ? wcnt[wave][dg] : 0u;
        rnk[r] = prior + (uint32_t)__popcll(peers & lt);
        if (valid && (peers & lt) == 0ull) wcnt[wave][dg] = prior + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {  // per digit: exclusive prefix of the wave counts
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) {
            const uint32_t c = wcnt[w][tid];
            wcnt[w][tid] = acc;
            acc += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        if (wbase + r * 64 + lane < n) {
            const uint32_t dg = (uint32_t)(key[r] >> shift) & 0xFFu;
            const uint32_t lp = lst[dg] + wcnt[wave][dg] + rnk[r];  // < tn
            sk[lp] = key[r];
            if (HASVAL) sv[lp] = val[r];
        }
    }
    __syncthreads();
    // the tile in digit order: consecutive slots of one digit go to consecutive positions
    for (int j = tid; j < tn; j += RS_THREADS) {
        const KT k = sk[j];
        const uint32_t dg = (uint32_t)(k >> shift) & 0xFFu;
        const uint32_t pos = doff[dg] + (uint32_t)j - lst[dg];
        kout[so + pos] = k;
        if (HASVAL) vout[so + pos] = sv[j];
    }
}

// rows[nrows] of len words (a grid of gx x gy rows when gy > 1)
static void scan_rows_grid(uint32_t *rows, dim3 grid, int len, uint32_t *tot, hipStream_t st) {
    if (len > 256 * SC_I)
        hipLaunchKernelGGL(k_scan_rows<1024>, grid, dim3(1024), 0, st, rows, len, tot);
    else
        hipLaunchKernelGGL(k_scan_rows<256>, grid, dim3(256), 0, st, rows, len, tot);
    FMM_LAUNCH_CHECK();
}

void scan_rows(uint32_t *rows, int nrows, int len, uint32_t *tot, hipStream_t st) {
    scan_rows_grid(rows, dim3(nrows), len, tot, st);
}

template <typename KT>
static bool radix_sort_t(KT *keys, KT *keys_alt, uint32_t *vals, uint32_t *vals_alt, int64_t n,
                         int nseg, int bit0, int bit1, uint32_t *scratch, hipStream_t st) {
    if (n <= 0 || bit1 <= bit0) return false;
    const int nblk = (int)rs_nblk(n);
    uint32_t *hist = scratch;
    uint32_t *dtot = scratch + (size_t)nseg * RS_BINS * nblk;
    bool alt = false;
    for (int shift = bit0; shift < bit1; shift += 8) {
        KT *ki = alt ? keys_alt : keys, *ko = alt ? keys : keys_alt;
        uint32_t *vi = alt ? vals_alt : vals, *vo = alt ? vals : vals_alt;
        dim3 grid(nblk, nseg);
        hipLaunchKernelGGL(k_rs_hist<KT>, grid, dim3(RS_THREADS), 0, st, ki, n, shift, hist, nblk);
        FMM_LAUNCH_CHECK();
        scan_rows_grid(hist, dim3(RS_BINS, nseg), nblk, dtot, st);
        if (vals)
            hipLaunchKernelGGL((k_rs_scatter<KT, true>), grid, dim3(RS_THREADS), 0, st, ki, vi, ko,
                               vo, n, shift, hist, nblk, dtot);
        else
            hipLaunchKernelGGL((k_rs_scatter<KT, false>), grid, dim3(RS_THREADS), 0, st, ki,
                               (const uint32_t *)nullptr, ko, (uint32_t *)nullptr, n, shift, hist,
                               nblk, dtot);
        FMM_LAUNCH_CHECK();
        alt = !alt;
    }
    return alt;
}

bool radix_sort(uint64_t *keys, uint64_t *keys_alt, uint32_t *vals, uint32_t *vals_alt, int64_t n,
                int nseg, int bit0, int bit1, uint32_t *scratch, hipStream_t st) {
    return radix_sort_t(keys, keys_alt, vals, vals_alt, n, nseg, bit0, bit1, scratch, st);
}

bool radix_sort32(uint32_t *keys, uint32_t *keys_alt, uint32_t *vals, uint32_t *vals_alt,
                  int64_t n, int nseg, int bit0, int bit1, uint32_t *scratch, hipStream_t st) {
    return radix_sort_t(keys, keys_alt, vals, vals_alt, n, nseg, bit0, bit1, scratch, st);
}

}  // namespace fmm4rb
