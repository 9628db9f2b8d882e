// tree.hip -- mean momentum and the array-based binary cluster tree on gfx950.
//
// Reference: ClusterTree / fillclusters! / BBox (src/cluster_tree/cluster_tree.jl:12-18,
// clusters.jl:21-70), kpermute! (src/utils/partition.jl:29-43), and the
// p_avg / stretch of the GPU driver (src/update_particles_field.jl:77-79).
//
// The reference splits each cluster recursively: bbox, split axis
// argmax(stretch .* (bmax - bmin)), and a quickselect that sends the
// ceil(c/2) smallest coordinates left.  Here every level is processed at once:
//   1. three stable radix sorts give, per axis d, the particle ids ordered by
//      (coordinate_d, id) -> perm_d, and every particle's rank r_d in that order;
//   2. three lists, one per axis, hold the particles of each cluster as rank
//      triples (r_x, r_y, r_z), list d ordered by r_d within every cluster;
//   3. at each level a cluster's bbox comes from the first/last entry of each
//      list, and its median split along axis a sends left exactly the
//      particles with r_a below the pivot rank (list a's entry at ceil(c/2)),
//      so every list entry decides its side from its own triple, with no
//      gather by particle id;
//   4. a stable segmented partition (wave ballots + a global prefix sum) moves
//      the three lists to the next level.
// Ties in coordinate are broken by particle id.  Each leaf's ids are finally
// sorted ascending (the canonical within-leaf order).
#include <cmath>

#include "fmm4rb_internal.hpp"

#pragma clang fp contract(off)

namespace fmm4rb {

void scan_rows(uint32_t *rows, int nrows, int len, uint32_t *tot, hipStream_t st);

constexpr int PAVG_THREADS = 256;
constexpr int PAVG_MAXBLK = 1024;
constexpr int PT_THREADS = 256, PT_ITEMS = 8, PT_TILE = PT_THREADS * PT_ITEMS;
constexpr int PT_WAVES = PT_THREADS / 64;

// ----------------------------------------------------------------- p_avg
// Deterministic fixed-order sum of the momenta and a finiteness check of all
// inputs.  update_particles_field.jl:77 (reduce(+, momenta)/N).  The same pass
// forms a 64-bit digest of the inputs, the wrapping sum over all 6N input words of
// a hash of (bits, index): independent of the summation order, so deterministic,
// and it tells whether a later call sees the inputs of an earlier one.
__device__ inline uint64_t word_hash(uint64_t x, uint64_t idx) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull * (idx + 1ull);  // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The same pass also forms each axis' minimum and maximum position (the range of
// the tree's quantised presort keys).
__global__ __launch_bounds__(PAVG_THREADS) void k_pavg_partial(
    const double *__restrict__ pos, const double *__restrict__ mom, int64_t N,
    double *__restrict__ part, unsigned long long *__restrict__ hpart, double *__restrict__ mmp,
    int *__restrict__ bad) {
    __shared__ double s[3][PAVG_THREADS];
    __shared__ double smm[6][PAVG_THREADS];
    __shared__ unsigned long long sh[PAVG_THREADS];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    double mn[3] = {__builtin_huge_val(), __builtin_huge_val(), __builtin_huge_val()};
    double mx[3] = {-__builtin_huge_val(), -__builtin_huge_val(), -__builtin_huge_val()};
    uint64_t h = 0;
    int b = 0;
    for (int64_t i = (int64_t)blockIdx.x * PAVG_THREADS + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * PAVG_THREADS) {
        double m0 = mom[3 * i], m1 = mom[3 * i + 1], m2 = mom[3 * i + 2];
        double x0 = pos[3 * i], x1 = pos[3 * i + 1], x2 = pos[3 * i + 2];
        a0 += m0;
        a1 += m1;
        a2 += m2;
        mn[0] = fmin(mn[0], x0);
        mn[1] = fmin(mn[1], x1);
        mn[2] = fmin(mn[2], x2);
        mx[0] = fmax(mx[0], x0);
        mx[1] = fmax(mx[1], x1);
        mx[2] = fmax(mx[2], x2);
        if (!(isfinite(m0) && isfinite(m1) && isfinite(m2) && isfinite(x0) && isfinite(x1) &&
              isfinite(x2)))
            b = 1;
        const double v[6] = {x0, x1, x2, m0, m1, m2};
        for (int e = 0; e < 6; ++e)
            h += word_hash((uint64_t)__double_as_longlong(v[e]), 6 * (uint64_t)i + e);
    }
    s[0][threadIdx.x] = a0;
    s[1][threadIdx.x] = a1;
    s[2][threadIdx.x] = a2;
    for (int d = 0; d < 3; ++d) {
        smm[d][threadIdx.x] = mn[d];
        smm[3 + d][threadIdx.x] = mx[d];
    }
    sh[threadIdx.x] = h;
    for (int off = PAVG_THREADS / 2; off > 0; off >>= 1) {
        __syncthreads();
        if (threadIdx.x < (unsigned)off) {
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] += s[d][threadIdx.x + off];
                smm[d][threadIdx.x] = fmin(smm[d][threadIdx.x], smm[d][threadIdx.x + off]);
                smm[3 + d][threadIdx.x] = fmax(smm[3 + d][threadIdx.x], smm[3 + d][threadIdx.x + off]);
            }
            sh[threadIdx.x] += sh[threadIdx.x + off];
        }
    }
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; ++d) part[3 * blockIdx.x + d] = s[d][0];
        for (int d = 0; d < 6; ++d) mmp[6 * blockIdx.x + d] = smm[d][0];
        hpart[blockIdx.x] = sh[0];
    }
    if (b) atomicOr(bad, 1);
}

__global__ __launch_bounds__(PAVG_THREADS) void k_pavg_final(
    const double *__restrict__ part, const unsigned long long *__restrict__ hpart,
    const double *__restrict__ mmp, int nblk, int64_t N, double *__restrict__ pavg4,
    unsigned long long *__restrict__ digest) {
    __shared__ double s[3][PAVG_THREADS];
    __shared__ double smm[6][PAVG_THREADS];
    __shared__ unsigned long long sh[PAVG_THREADS];
    double a[3] = {0.0, 0.0, 0.0};
    double m[6] = {__builtin_huge_val(), __builtin_huge_val(), __builtin_huge_val(),
                   -__builtin_huge_val(), -__builtin_huge_val(), -__builtin_huge_val()};
    uint64_t h = 0;
    for (int i = threadIdx.x; i < nblk; i += PAVG_THREADS) {
        for (int d = 0; d < 3; ++d) a[d] += part[3 * i + d];
        for (int d = 0; d < 3; ++d) {
            m[d] = fmin(m[d], mmp[6 * i + d]);
            m[3 + d] = fmax(m[3 + d], mmp[6 * i + 3 + d]);
        }
        h += hpart[i];
    }
    for (int d = 0; d < 3; ++d) s[d][threadIdx.x] = a[d];
    for (int d = 0; d < 6; ++d) smm[d][threadIdx.x] = m[d];
    sh[threadIdx.x] = h;
    for (int off = PAVG_THREADS / 2; off > 0; off >>= 1) {
        __syncthreads();
        if (threadIdx.x < (unsigned)off) {
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] += s[d][threadIdx.x + off];
                smm[d][threadIdx.x] = fmin(smm[d][threadIdx.x], smm[d][threadIdx.x + off]);
                smm[3 + d][threadIdx.x] = fmax(smm[3 + d][threadIdx.x], smm[3 + d][threadIdx.x + off]);
            }
            sh[threadIdx.x] += sh[threadIdx.x + off];
        }
    }
    if (threadIdx.x == 0) {
        double p0 = s[0][0] / (double)N, p1 = s[1][0] / (double)N, p2 = s[2][0] / (double)N;
        pavg4[0] = p0;
        pavg4[1] = p1;
        pavg4[2] = p2;
        pavg4[3] = sqrt(1.0 + ((p0 * p0 + p1 * p1) + p2 * p2));  // update_particles_field.jl:78
        digest[0] = sh[0];
        // position range per axis: pavg4[8 + d] min, pavg4[11 + d] max
        for (int d = 0; d < 6; ++d) pavg4[8 + d] = smm[d][0];
    }
}

// part: 3 PAVG_MAXBLK doubles, PAVG_MAXBLK digest words, then 6 PAVG_MAXBLK doubles
// (per-block position ranges); pavg4: p_avg, g_avg, the digest word, the flag word,
// two spare words, then the position range (min x y z, max x y z)
void launch_pavg(const double *pos, const double *mom, int64_t N, double *part, double *pavg4,
                 int *bad, hipStream_t st) {
    int nblk = (int)std::min<int64_t>(PAVG_MAXBLK, (N + PAVG_THREADS - 1) / PAVG_THREADS);
    unsigned long long *hpart = reinterpret_cast<unsigned long long *>(part + 3 * PAVG_MAXBLK);
    double *mmp = part + 4 * PAVG_MAXBLK;
    unsigned long long *digest = reinterpret_cast<unsigned long long *>(pavg4 + 4);
    hipLaunchKernelGGL(k_pavg_partial, dim3(nblk), dim3(PAVG_THREADS), 0, st, pos, mom, N, part,
                       hpart, mmp, bad);
    FMM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pavg_final, dim3(1), dim3(PAVG_THREADS), 0, st, part, hpart, mmp, nblk,
                       N, pavg4, digest);
    FMM_LAUNCH_CHECK();
}

// ------------------------------------------------------------ tree kernels
__device__ inline uint64_t ordered_key(double x) {
    x = x + 0.0;  // -0.0 and +0.0 share one key
    uint64_t u = (uint64_t)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ inline int node_level(int32_t k) { return 31 - __clz(k + 1); }

__device__ inline bool splitting(int32_t k, int l, const uint8_t *flags) {
    return node_level(k) == l && (flags[k] & 3) == 1;
}

__global__ void k_make_keys(const double *__restrict__ pos, int64_t N, uint64_t *__restrict__ keys,
                            uint32_t *__restrict__ vals, int32_t *__restrict__ seg,
                            uint32_t *__restrict__ flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (i == 0) *flag = 0u;
    for (int d = 0; d < 3; ++d) {
        keys[d * N + i] = ordered_key(pos[3 * i + d]);
        vals[d * N + i] = (uint32_t)i;
    }
    seg[i] = 0;
}

// Quantised presort: key = trunc((x - min_d) s_d) with s_d = QMAX / (max_d - min_d),
// non-decreasing in x, so a stable sort of (key, id) from id order gives the (x, id)
// order everywhere except inside runs of equal keys that hold different coordinates
// (positions closer than the range / 2^32); k_presort_mark / k_presort_fix order
// those runs.  4 radix passes over 4-byte keys instead of 8 over 8-byte ones.  (24-bit
// keys, 3 passes, were slower at N = 2^22: a Gaussian's central keys then hold ~1
// particle each, and the run fix-up grew from 45 to 307 us, more than the pass saved.)
constexpr int PRESORT_BITS = 32;
constexpr double PRESORT_QMAX = 4294967040.0;  // 2^32 - 256: the product's rounding stays below 2^32
constexpr int PRESORT_RUN_CAP = 256;           // longest run the fix-up orders

__global__ void k_make_keys32(const double *__restrict__ pos, int64_t N,
                              const double *__restrict__ pavg4, uint32_t *__restrict__ keys,
                              uint32_t *__restrict__ vals, int32_t *__restrict__ seg,
                              uint8_t *__restrict__ need, uint32_t *__restrict__ flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (i == 0) *flag = 0u;
    for (int d = 0; d < 3; ++d) {
        const double mn = pavg4[8 + d], w = pavg4[11 + d] - mn;
        const double q = PRESORT_QMAX / w;
        const double sc = (w > 0.0 && q < 1e300) ? q : 0.0;
        keys[d * N + i] = (uint32_t)((pos[3 * i + d] - mn) * sc);
        vals[d * N + i] = (uint32_t)i;
        need[d * N + i] = 0;
    }
    seg[i] = 0;
}

// A run of equal keys is out of (x, id) order exactly where a coordinate decreases
// (ids ascend within the run): mark the run's first position.  A run longer than
// PRESORT_RUN_CAP raises the flag instead (the caller then sorts the exact keys).
__global__ void k_presort_mark(int64_t N, const uint32_t *__restrict__ keys,
                               const uint32_t *__restrict__ perm, const double *__restrict__ pos,
                               uint8_t *__restrict__ need, uint32_t *__restrict__ flag) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * N) return;
    const int d = (int)(r / N);
    const int64_t i = r - (int64_t)d * N;
    const uint32_t *K = keys + (int64_t)d * N;
    const uint32_t *P = perm + (int64_t)d * N;
    if (i == 0 || K[i] != K[i - 1]) return;
    const double xa = pos[3 * (int64_t)P[i - 1] + d] + 0.0, xb = pos[3 * (int64_t)P[i] + d] + 0.0;
    if (!(xb < xa)) return;
    int64_t h = i - 1;
    while (h > 0 && K[h - 1] == K[i] && i - h < PRESORT_RUN_CAP) --h;
    if (h > 0 && K[h - 1] == K[i]) {
        atomicOr(flag, 1u);
        return;
    }
    need[(int64_t)d * N + h] = 1;
}

// Insertion sort of each marked run by (x, id), by the thread of its first position.
__global__ void k_presort_fix(int64_t N, const uint32_t *__restrict__ keys,
                              uint32_t *__restrict__ perm, const double *__restrict__ pos,
                              const uint8_t *__restrict__ need, uint32_t *__restrict__ flag) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * N || !need[r]) return;
    const int d = (int)(r / N);
    const int64_t h = r - (int64_t)d * N;
    const uint32_t *K = keys + (int64_t)d * N;
    uint32_t *P = perm + (int64_t)d * N;
    int64_t e = h + 1;
    while (e < N && K[e] == K[h] && e - h < PRESORT_RUN_CAP) ++e;
    if (e < N && K[e] == K[h]) {
        atomicOr(flag, 1u);
        return;
    }
    for (int64_t a = h + 1; a < e; ++a) {
        const uint32_t v = P[a];
        const double xv = pos[3 * (int64_t)v + d] + 0.0;
        int64_t b = a - 1;
        while (b >= h) {
            const uint32_t u = P[b];
            const double xu = pos[3 * (int64_t)u + d] + 0.0;
            if (xu < xv || (xu == xv && u < v)) break;
            P[b + 1] = u;
            --b;
        }
        P[b + 1] = v;
    }
}

// rank4[p] = (r_x, r_y, r_z, -) with perm_d[r_d] = p: a particle's three ranks in
// one 16-byte word, so that the lists below gather each particle once per list
__global__ void k_ranks(int64_t N, const uint32_t *__restrict__ perm, uint32_t *__restrict__ rank4) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    for (int d = 0; d < 3; ++d) rank4[4 * (int64_t)perm[d * N + r] + d] = (uint32_t)r;
}

// Level-0 lists: rk[(3 d + e) N + r] = r_e of the particle at rank r on axis d
// (its own rank r_d = r).
__global__ void k_lists_init(int64_t N, const uint32_t *__restrict__ perm,
                             const uint4 *__restrict__ rank4, uint32_t *__restrict__ rk) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    for (int d = 0; d < 3; ++d) {
        const uint4 t = rank4[perm[d * N + r]];
        rk[(3 * d + 0) * N + r] = t.x;
        rk[(3 * d + 1) * N + r] = t.y;
        rk[(3 * d + 2) * N + r] = t.z;
    }
}

// Bounding box (clusters.jl:51-70), split axis (clusters.jl:39) and pivot rank
// of the existing nodes of one level; ext records the particle attaining each
// bound (argmin/argmax, lowest id on ties).
__global__ void k_level_nodes(int nnodes, const int32_t *__restrict__ nodes,
                              const int32_t *__restrict__ lo, const int32_t *__restrict__ cnt,
                              const uint8_t *__restrict__ flags, const uint32_t *__restrict__ rk,
                              const uint32_t *__restrict__ perm, int64_t N,
                              const double *__restrict__ pos, const double *__restrict__ pavg4,
                              double *__restrict__ bbox, int8_t *__restrict__ axis,
                              uint32_t *__restrict__ pivot, int32_t *__restrict__ ext) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnodes) return;
    int32_t k = nodes[t];
    int64_t a = lo[k], b = (int64_t)lo[k] + cnt[k] - 1;
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        const uint32_t pa = perm[d * N + rk[(4 * d) * N + a]];  // list d, own rank
        const uint32_t pb = perm[d * N + rk[(4 * d) * N + b]];
        mn[d] = pos[3 * (int64_t)pa + d] + 0.0;
        mx[d] = pos[3 * (int64_t)pb + d] + 0.0;
        bbox[6 * (int64_t)k + d] = mn[d];
        bbox[6 * (int64_t)k + 3 + d] = mx[d];
        ext[6 * (int64_t)k + d] = (int32_t)pa;
        ext[6 * (int64_t)k + 3 + d] = (int32_t)pb;
    }
    if ((flags[k] & 3) == 1) {
        double s[3] = {1.0, 1.0, pavg4[3]};
        int dir = 0;
        double best = s[0] * (mx[0] - mn[0]);
        for (int d = 1; d < 3; ++d) {
            double v = s[d] * (mx[d] - mn[d]);
            if (v > best) {
                best = v;
                dir = d;
            }
        }
        axis[k] = (int8_t)dir;
        pivot[k] = rk[(4 * dir) * N + a + ((cnt[k] + 1) >> 1)];  // first right entry of list dir
    }
}

// Per-block counts of left entries of each list (blockIdx.y = list).
__global__ __launch_bounds__(PT_THREADS) void k_part_count(
    int64_t N, int l, const int32_t *__restrict__ seg, const uint8_t *__restrict__ flags,
    const int8_t *__restrict__ axis, const uint32_t *__restrict__ pivot,
    const uint32_t *__restrict__ rk, uint32_t *__restrict__ bsum, int nblk) {
    __shared__ uint32_t s[PT_WAVES];
    const int d = blockIdx.y;
    const int64_t base = (int64_t)blockIdx.x * PT_TILE;
    // every entry's segment first, then its cluster's split data, then only the one
    // rank the side needs (r_a of the cluster's axis): 8 bytes per entry instead of 16
    int32_t kk[PT_ITEMS], ax[PT_ITEMS];
    uint32_t pv[PT_ITEMS], ra[PT_ITEMS];
    const uint32_t *r0 = rk + (int64_t)(3 * d) * N;
#pragma unroll
    for (int j = 0; j < PT_ITEMS; ++j) {
        const int64_t i = base + j * PT_THREADS + threadIdx.x;
        kk[j] = i < N ? seg[i] : -1;
    }
#pragma unroll
    for (int j = 0; j < PT_ITEMS; ++j) {
        const int32_t k = kk[j];
        ax[j] = -1;
        pv[j] = 0u;
        if (k >= 0 && splitting(k, l, flags)) {
            ax[j] = axis[k];
            pv[j] = pivot[k];
        }
    }
#pragma unroll
    for (int j = 0; j < PT_ITEMS; ++j) {
        const int64_t i = base + j * PT_THREADS + threadIdx.x;
        ra[j] = ax[j] >= 0 ? r0[ax[j] * N + i] : 0u;
    }
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < PT_ITEMS; ++j) {
        const uint32_t f = (ax[j] >= 0 && ra[j] < pv[j]) ? 1u : 0u;  // r_a below the pivot rank
        c += (uint32_t)__popcll(__ballot(f));
    }
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < PT_WAVES; ++w) t += s[w];
        bsum[(int64_t)d * nblk + blockIdx.x] = t;
    }
}

// Stable segmented partition of the three lists (blockIdx.y = list): left
// entries of a splitting cluster keep their order at lo + (left entries before
// them), right entries at lo + ceil(c/2) + (right entries before them); others
// stay.  bsum holds the exclusive prefix of the per-block counts, leftbase[k]
// the number of left entries before cluster k.  List 0's blocks also write the
// next level's segment ids.
__global__ __launch_bounds__(PT_THREADS) void k_part_scatter(
    int64_t N, int l, const int32_t *__restrict__ seg, const uint8_t *__restrict__ flags,
    const int8_t *__restrict__ axis, const uint32_t *__restrict__ pivot,
    const int32_t *__restrict__ lo, const int32_t *__restrict__ cnt,
    const int32_t *__restrict__ leftbase, const uint32_t *__restrict__ rk,
    const uint32_t *__restrict__ bsum, int nblk, uint32_t *__restrict__ rk_out,
    int32_t *__restrict__ seg_out) {
    __shared__ uint32_t s[PT_ITEMS][PT_WAVES];
    const int d = blockIdx.y;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * PT_TILE;
    const uint32_t *r0 = rk + (int64_t)(3 * d) * N;
    // every entry's segment and rank triple is loaded once, up front, and kept in
    // registers through the side decision and the move
    int32_t kk[PT_ITEMS];
    uint32_t v[PT_ITEMS][3];
#pragma unroll
    for (int j = 0; j < PT_ITEMS; ++j) {
        const int64_t i = base + j * PT_THREADS + threadIdx.x;
        kk[j] = -1;
        v[j][0] = v[j][1] = v[j][2] = 0u;
        if (i < N) {
            kk[j] = seg[i];
#pragma unroll
            for (int e = 0; e < 3; ++e) v[j][e] = r0[e * N + i];
        }
    }
    uint64_t bal[PT_ITEMS];
#pragma unroll
    for (int j = 0; j < PT_ITEMS; ++j) {
        const int32_t k = kk[j];
        uint32_t f = 0;
        if (k >= 0 && splitting(k, l, flags)) {  // r_a below the cluster's pivot rank
            const int a = axis[k];
            const uint32_t ra = a == 0 ? v[j][0] : (a == 1 ? v[j][1] : v[j][2]);
            f = ra < pivot[k] ? 1u : 0u;
        }
        bal[j] = __ballot(f);
        if (lane == 0) s[j][w] = (uint32_t)__popcll(bal[j]);
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    int64_t run = bsum[(int64_t)d * nblk + blockIdx.x];
#pragma unroll
    for (int j = 0; j < PT_ITEMS; ++j) {
        const int64_t i = base + j * PT_THREADS + threadIdx.x;
        uint32_t before = 0, total = 0;
        for (int u = 0; u < PT_WAVES; ++u) {
            before += u < w ? s[j][u] : 0u;
            total += s[j][u];
        }
        if (i < N) {
            const int32_t k = kk[j];
            int64_t np = i;
            if (splitting(k, l, flags)) {
                const int64_t lft = run + before + (uint32_t)__popcll(bal[j] & below) - leftbase[k];
                const int64_t nl = (cnt[k] + 1) >> 1;
                np = ((bal[j] >> lane) & 1ull) ? lo[k] + lft : lo[k] + nl + (i - lo[k]) - lft;
                if (d == 0) seg_out[i] = (i - lo[k]) < nl ? 2 * k + 1 : 2 * k + 2;
            } else if (d == 0) {
                seg_out[i] = k;
            }
#pragma unroll
            for (int e = 0; e < 3; ++e) rk_out[(3 * d + e) * N + np] = v[j][e];
        }
        run += total;
    }
}

// ------------------------------------------------ middle levels, per cluster
// Once every cluster of a level holds at most PL_MAXC particles, one workgroup per
// (cluster, list) partitions the cluster's segment of that list by itself: it
// walks the segment in chunks of PL_CHUNK entries in order, carrying the count of
// left entries, so neither the per-block count pass and its global scan nor the
// segment ids are needed; the workgroup also forms its cluster's bounding box, axis
// and pivot (as k_level_nodes; the three workgroups of a cluster compute the same
// values from the same entries, and the list-0 one stores them).  Entries of a
// leaf are copied unchanged.
constexpr int PL_I = 8, PL_MAXC = 1 << 16;

// PL_T threads (1024, or 256 for clusters of at most 4096 particles, whose one or two
// chunks would leave most of a larger block idle).  The first chunk's entries are
// loaded before the cluster's split is known, so those loads overlap the dependent
// gathers of the bounding box (six lanes, one per axis and end).
template <int PL_T>
__global__ __launch_bounds__(PL_T) void k_part_local(
    const int32_t *__restrict__ nodes, const int32_t *__restrict__ lo,
    const int32_t *__restrict__ cnt, const uint8_t *__restrict__ flags,
    const uint32_t *__restrict__ perm, int64_t N, const double *__restrict__ pos,
    const double *__restrict__ pavg4, const uint32_t *__restrict__ rk,
    uint32_t *__restrict__ rk_out, double *__restrict__ bbox, int8_t *__restrict__ axis,
    int32_t *__restrict__ ext) {
    constexpr int NW = PL_T / 64, PL_CHUNK = PL_T * PL_I;
    __shared__ uint32_t s[2][PL_I][NW];
    __shared__ double sx[6];
    __shared__ int32_t sp[6];
    const int d = blockIdx.y, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int32_t k = nodes[blockIdx.x];
    const int64_t b = lo[k];
    const int c = cnt[k];
    const bool split = (flags[k] & 3) == 1;
    const uint32_t *in = rk + (int64_t)(3 * d) * N;
    uint32_t *out = rk_out + (int64_t)(3 * d) * N;
    uint32_t v[PL_I][3];  // the current chunk's entries
    auto load = [&](int base) {
#pragma unroll
        for (int j = 0; j < PL_I; ++j) {
            const int i = base + j * PL_T + tid;
#pragma unroll
            for (int e = 0; e < 3; ++e) v[j][e] = i < c ? in[e * N + b + i] : 0u;
        }
    };
    load(0);
    if (tid < 6) {  // clusters.jl:51-70: lane 2 e + end, the particle at list e's end
        const int e = tid >> 1;
        const int64_t at = (tid & 1) ? b + c - 1 : b;
        const int32_t p = (int32_t)perm[e * N + rk[(4 * e) * N + at]];  // list e, own rank
        sp[tid] = p;
        sx[tid] = pos[3 * (int64_t)p + e] + 0.0;
    }
    __syncthreads();
    int dir = 0;  // clusters.jl:39, the same in every lane
    {
        const double sc[3] = {1.0, 1.0, pavg4[3]};
        double best = sc[0] * (sx[1] - sx[0]);
        for (int e = 1; e < 3; ++e) {
            const double q = sc[e] * (sx[2 * e + 1] - sx[2 * e]);
            if (q > best) {
                best = q;
                dir = e;
            }
        }
    }
    if (d == 0 && tid < 6) {
        const int e = tid >> 1, hi = tid & 1;
        bbox[6 * (int64_t)k + 3 * hi + e] = sx[tid];
        ext[6 * (int64_t)k + 3 * hi + e] = sp[tid];
        if (tid == 0 && split) axis[k] = (int8_t)dir;
    }
    if (!split) {
        for (int base = 0; base < c; base += PL_CHUNK) {
            if (base > 0) load(base);
#pragma unroll
            for (int j = 0; j < PL_I; ++j) {
                const int i = base + j * PL_T + tid;
                if (i < c)
#pragma unroll
                    for (int e = 0; e < 3; ++e) out[e * N + b + i] = v[j][e];
            }
        }
        return;
    }
    const uint32_t pv = rk[(4 * dir) * N + b + ((c + 1) >> 1)];  // first right entry of list dir
    const int64_t nl = (c + 1) >> 1;
    const uint64_t below = (1ull << lane) - 1ull;
    int64_t carry = 0;  // left entries before the chunk
    for (int base = 0, ph = 0; base < c; base += PL_CHUNK, ph ^= 1) {
        if (base > 0) load(base);
        uint64_t bal[PL_I];
#pragma unroll
        for (int j = 0; j < PL_I; ++j) {
            const int i = base + j * PL_T + tid;
            const uint32_t ra = dir == 0 ? v[j][0] : (dir == 1 ? v[j][1] : v[j][2]);
            bal[j] = __ballot(i < c && ra < pv);
            if (lane == 0) s[ph][j][w] = (uint32_t)__popcll(bal[j]);
        }
        __syncthreads();  // s[ph] complete (s[ph ^ 1] was last read before the previous barrier)
        int64_t run = carry;
#pragma unroll
        for (int j = 0; j < PL_I; ++j) {
            uint32_t before = 0, total = 0;
#pragma unroll
            for (int u = 0; u < NW; ++u) {
                const uint32_t t = s[ph][j][u];
                before += u < w ? t : 0u;
                total += t;
            }
            const int i = base + j * PL_T + tid;
            if (i < c) {
                const int64_t lft = run + before + (uint32_t)__popcll(bal[j] & below);
                const int64_t np = ((bal[j] >> lane) & 1ull) ? lft : nl + i - lft;
#pragma unroll
                for (int e = 0; e < 3; ++e) out[e * N + b + np] = v[j][e];
            }
            run += total;
        }
        carry = run;
    }
}

// ------------------------------------------------------- deep levels in LDS
// Once every cluster of a level L0 holds at most TF_CAP particles, one workgroup
// per cluster finishes its subtree in LDS: for each level l = L0 .. ml the
// bounding boxes, axes and argmin/argmax particles of the subtree's nodes
// (as k_level_nodes), and for l < ml the same stable partition of the three rank
// lists as k_part_scatter; then list x's own ranks (all the leaf sort reads) go
// back to rk.  At most TF_LEVELS levels below L0 (TF_SUB nodes per level).  One
// thread per position: the kernel waits on its dependent gathers (node geometry) and
// barriers, and more waves per cluster hide them (256 -> 1024 threads: the whole tree
// build 4.15 -> 4.00 ms at N = 2^22).
constexpr int TF_CAP = 1024, TF_T = 1024, TF_IT = TF_CAP / TF_T;
constexpr int TF_LEVELS = 6, TF_SUB = 1 << TF_LEVELS;

__global__ __launch_bounds__(TF_T) void k_tree_finish(
    const int32_t *__restrict__ nodes, int L0, int ml, const int32_t *__restrict__ lo,
    const int32_t *__restrict__ cnt, const uint8_t *__restrict__ flags, uint32_t *__restrict__ rk,
    const uint32_t *__restrict__ perm, int64_t N, const double *__restrict__ pos,
    const double *__restrict__ pavg4, double *__restrict__ bbox, int8_t *__restrict__ axis,
    int32_t *__restrict__ ext) {
    __shared__ uint32_t Ls[9][TF_CAP];   // list d, axis e at Ls[3 d + e]
    __shared__ uint32_t tmp[3][TF_CAP];  // one list's next order
    __shared__ uint16_t sseg[TF_CAP];    // node (index within the level) of each position
    __shared__ uint16_t pf[TF_CAP];      // exclusive prefix of the side flags
    __shared__ uint32_t tsum[TF_T / 64];
    __shared__ int32_t nlo[TF_SUB], ncnt[TF_SUB];
    __shared__ uint32_t npiv[TF_SUB];
    __shared__ int8_t nax[TF_SUB], nspl[TF_SUB];
    constexpr uint16_t NONE = 0xFFFF;
    const int tid = threadIdx.x;
    const int32_t k0 = nodes[blockIdx.x];
    const int32_t a0 = lo[k0], c0 = cnt[k0];
    for (int idx = tid; idx < 9 * c0; idx += TF_T) {
        const int e9 = idx / c0, j = idx - e9 * c0;
        Ls[e9][j] = rk[(int64_t)e9 * N + a0 + j];
    }
    for (int j = tid; j < c0; j += TF_T) sseg[j] = 0;
    __syncthreads();
    const double sz = pavg4[3];
    for (int l = L0; l <= ml; ++l) {
        const int m = l - L0, nsub = 1 << m;
        const int64_t first = ((int64_t)k0 + 1) * nsub - 1;
        for (int u = tid; u < nsub; u += TF_T) {  // clusters.jl:39,51-70 per node
            const int64_t k = first + u;
            nspl[u] = 0;
            ncnt[u] = 0;
            if (!(flags[k] & 1)) continue;
            const int32_t b = lo[k] - a0, c = cnt[k];
            nlo[u] = b;
            ncnt[u] = c;
            double mn[3], mx[3];
            for (int d = 0; d < 3; ++d) {
                const uint32_t pa = perm[d * N + Ls[4 * d][b]];
                const uint32_t pb = perm[d * N + Ls[4 * d][b + c - 1]];
                mn[d] = pos[3 * (int64_t)pa + d] + 0.0;
                mx[d] = pos[3 * (int64_t)pb + d] + 0.0;
                bbox[6 * k + d] = mn[d];
                bbox[6 * k + 3 + d] = mx[d];
                ext[6 * k + d] = (int32_t)pa;
                ext[6 * k + 3 + d] = (int32_t)pb;
            }
            if ((flags[k] & 3) == 1) {
                const double sd[3] = {1.0, 1.0, sz};
                int dir = 0;
                double best = sd[0] * (mx[0] - mn[0]);
                for (int d = 1; d < 3; ++d) {
                    const double v = sd[d] * (mx[d] - mn[d]);
                    if (v > best) {
                        best = v;
                        dir = d;
                    }
                }
                axis[k] = (int8_t)dir;
                nax[u] = (int8_t)dir;
                nspl[u] = 1;
                npiv[u] = Ls[4 * dir][b + ((c + 1) >> 1)];
            }
        }
        __syncthreads();
        if (l == ml) break;
        for (int d = 0; d < 3; ++d) {
            // side flags of this thread's TF_IT consecutive positions, block prefix sum
            uint32_t f[TF_IT], run = 0;
#pragma unroll
            for (int r = 0; r < TF_IT; ++r) {
                const int j = tid * TF_IT + r;
                f[r] = 0;
                if (j < c0) {
                    const uint16_t u = sseg[j];
                    if (u != NONE && nspl[u]) f[r] = Ls[3 * d + nax[u]][j] < npiv[u] ? 1u : 0u;
                }
                run += f[r];
            }
            // block prefix sum: wave scans by shuffles, then the waves' totals
            uint32_t x = run;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o);
                if ((tid & 63) >= o) x += y;
            }
            if ((tid & 63) == 63) tsum[tid >> 6] = x;
            __syncthreads();
            uint32_t pre = x - run;
            for (int w = 0; w < (tid >> 6); ++w) pre += tsum[w];
#pragma unroll
            for (int r = 0; r < TF_IT; ++r) {
                const int j = tid * TF_IT + r;
                if (j < c0) pf[j] = (uint16_t)pre;
                pre += f[r];
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < TF_IT; ++r) {
                const int j = tid * TF_IT + r;
                if (j >= c0) continue;
                const uint16_t u = sseg[j];
                int np = j;
                if (u != NONE && nspl[u]) {
                    const int b = nlo[u], lft = (int)pf[j] - (int)pf[b];
                    const int nl = (ncnt[u] + 1) >> 1;
                    np = f[r] ? b + lft : b + nl + (j - b) - lft;
                }
#pragma unroll
                for (int e = 0; e < 3; ++e) tmp[e][np] = Ls[3 * d + e][j];
            }
            __syncthreads();
            for (int j = tid; j < c0; j += TF_T)
#pragma unroll
                for (int e = 0; e < 3; ++e) Ls[3 * d + e][j] = tmp[e][j];
            __syncthreads();
        }
        // positions of the next level's nodes (children 2u, 2u + 1); leaves end here
        for (int j = tid; j < c0; j += TF_T) {
            const uint16_t u = sseg[j];
            if (u == NONE) continue;
            if (nspl[u]) sseg[j] = (uint16_t)(2 * u + ((j - nlo[u]) < ((ncnt[u] + 1) >> 1) ? 0 : 1));
            else sseg[j] = NONE;
        }
        __syncthreads();
    }
    for (int j = tid; j < c0; j += TF_T) rk[a0 + j] = Ls[0][j];
}

// Canonical within-leaf order: ascending particle id (rank sort in LDS); the same
// workgroup then gathers the leaf's particles in that order (k_gather's copy).
__device__ inline void gather_particle(int64_t i, int64_t p, const double *__restrict__ pos,
                                       const double *__restrict__ mom, double *__restrict__ sp8,
                                       float *__restrict__ sp4f) {
    const double x0 = pos[3 * p], x1 = pos[3 * p + 1], x2 = pos[3 * p + 2];
    const double m0 = mom[3 * p], m1 = mom[3 * p + 1], m2 = mom[3 * p + 2];
    const double g = sqrt(1.0 + ((m0 * m0 + m1 * m1) + m2 * m2));  // interaction.jl:67
    double4 *o = reinterpret_cast<double4 *>(sp8 + 8 * i);
    o[0] = make_double4(x0, x1, x2, g);
    o[1] = make_double4(m0, m1, m2, 0.0);
    if (sp4f)
        reinterpret_cast<float4 *>(sp4f)[i] = make_float4((float)g, (float)m0, (float)m1, (float)m2);
}

__global__ __launch_bounds__(256) void k_leaf_sort(const int32_t *__restrict__ leaves,
                                                   const int32_t *__restrict__ lo,
                                                   const int32_t *__restrict__ cnt,
                                                   const uint32_t *__restrict__ rkx,
                                                   const uint32_t *__restrict__ permx,
                                                   int32_t *__restrict__ parind,
                                                   int32_t *__restrict__ leafof,
                                                   const double *__restrict__ pos,
                                                   const double *__restrict__ mom,
                                                   double *__restrict__ sp8, float *__restrict__ sp4f) {
    __shared__ uint32_t s[MAX_N0];
    const int32_t k = leaves[blockIdx.x];
    const int32_t a = lo[k], c = cnt[k];
    for (int j = threadIdx.x; j < c; j += blockDim.x) s[j] = permx[rkx[a + j]];
    __syncthreads();
    for (int j = threadIdx.x; j < c; j += blockDim.x) {
        uint32_t v = s[j];
        int r = 0;
        for (int m = 0; m < c; ++m) r += s[m] < v;
        parind[a + r] = (int32_t)v;
        leafof[a + j] = k;
        gather_particle((int64_t)a + r, v, pos, mom, sp8, sp4f);
    }
}

// Particles in tree order: sp8 = (x, y, z, gamma, px, py, pz, 0) and, for the
// mixed-precision path, sp4f = (gamma, px, py, pz) in FP32.
__global__ void k_gather(int64_t N, const int32_t *__restrict__ parind,
                         const double *__restrict__ pos, const double *__restrict__ mom,
                         double *__restrict__ sp8, float *__restrict__ sp4f) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    gather_particle(i, parind[i], pos, mom, sp8, sp4f);
}

// ------------------------------------------------------------------ refit
// Bounding boxes of an existing tree for new positions: the particle order,
// cluster ranges and split structure stay.  A leaf's box is the min/max of its
// particles' coordinates (clusters.jl:51-70; one wave per leaf over the
// tree-ordered copy), a split node's box the union of its two children's (they
// partition its particles), level by level upward.  Min and max are exact, so
// unchanged positions give the boxes of the full build.
__global__ __launch_bounds__(64) void k_refit_leaves(const int32_t *__restrict__ leaves,
                                                     const int32_t *__restrict__ lo,
                                                     const int32_t *__restrict__ cnt,
                                                     const double *__restrict__ sp8,
                                                     double *__restrict__ bbox) {
    const int32_t k = leaves[blockIdx.x];
    const int64_t a = lo[k];
    const int c = cnt[k];
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        mn[d] = __builtin_huge_val();
        mx[d] = -__builtin_huge_val();
    }
    for (int p = threadIdx.x; p < c; p += 64) {
        const double *x = sp8 + 8 * (a + p);
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], x[d]);
            mx[d] = fmax(mx[d], x[d]);
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], __shfl_xor(mn[d], off));
            mx[d] = fmax(mx[d], __shfl_xor(mx[d], off));
        }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            bbox[6 * (int64_t)k + d] = mn[d] + 0.0;  // -0.0 -> +0.0, as the build
            bbox[6 * (int64_t)k + 3 + d] = mx[d] + 0.0;
        }
}

__global__ void k_refit_parents(int nnodes, const int32_t *__restrict__ nodes,
                                double *__restrict__ bbox) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnodes) return;
    const int64_t k = nodes[t];
    const double *l = bbox + 6 * (2 * k + 1), *r = bbox + 6 * (2 * k + 2);
    for (int d = 0; d < 3; ++d) {
        bbox[6 * k + d] = fmin(l[d], r[d]);
        bbox[6 * k + 3 + d] = fmax(l[3 + d], r[3 + d]);
    }
}

// ------------------------------------------------------------- host side
static inline unsigned nb(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void build_tree_device(Ctx &c, const double *pos, const double *mom, int64_t N, int64_t N0,
                       bool exact) {
    (void)N0;
    const Shape &S = c.shape;
    hipStream_t st = c.stream;
    uint64_t *keys = c.d_keys.as<uint64_t>(), *keys_alt = c.d_keys_alt.as<uint64_t>();
    uint32_t *vals = c.d_vals.as<uint32_t>(), *vals_alt = c.d_vals_alt.as<uint32_t>();
    int32_t *seg = c.d_seg.as<int32_t>(), *seg_alt = c.d_seg_alt.as<int32_t>();
    uint32_t *rk = c.d_rk.as<uint32_t>(), *rk_alt = rk + 9 * (size_t)N;
    uint32_t *flag = c.d_tcnt.as<uint32_t>() + TCNT_FLAG;
    const uint32_t *perm;
    if (exact) {  // 64-bit keys of the exact coordinate order
        hipLaunchKernelGGL(k_make_keys, dim3(nb(N, 256)), dim3(256), 0, st, pos, N, keys, vals,
                           seg, flag);
        FMM_LAUNCH_CHECK();
        const bool alt = radix_sort(keys, keys_alt, vals, vals_alt, N, 3, 0, 64,
                                    c.d_rs_scratch.as<uint32_t>(), st);
        perm = alt ? vals_alt : vals;
    } else {
        uint32_t *k32 = reinterpret_cast<uint32_t *>(keys);
        uint32_t *k32a = reinterpret_cast<uint32_t *>(keys_alt);
        uint8_t *need = reinterpret_cast<uint8_t *>(seg_alt);  // 3N bytes; the partition rewrites it
        hipLaunchKernelGGL(k_make_keys32, dim3(nb(N, 256)), dim3(256), 0, st, pos, N,
                           c.d_pavg.as<double>(), k32, vals, seg, need, flag);
        FMM_LAUNCH_CHECK();
        const bool alt = radix_sort32(k32, k32a, vals, vals_alt, N, 3, 0, PRESORT_BITS,
                                      c.d_rs_scratch.as<uint32_t>(), st);
        const uint32_t *ks = alt ? k32a : k32;
        uint32_t *pm = alt ? vals_alt : vals;
        hipLaunchKernelGGL(k_presort_mark, dim3(nb(3 * N, 256)), dim3(256), 0, st, N, ks, pm, pos,
                           need, flag);
        FMM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_presort_fix, dim3(nb(3 * N, 256)), dim3(256), 0, st, N, ks, pm, pos,
                           need, flag);
        FMM_LAUNCH_CHECK();
        perm = pm;
    }
    // the sorted keys are not needed any more: their buffer (24 B per particle) holds
    // the 16-byte rank words
    uint32_t *rank4 = reinterpret_cast<uint32_t *>(keys);
    hipLaunchKernelGGL(k_ranks, dim3(nb(N, 256)), dim3(256), 0, st, N, perm, rank4);
    FMM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_lists_init, dim3(nb(N, 256)), dim3(256), 0, st, N, perm,
                       reinterpret_cast<const uint4 *>(rank4), rk);
    FMM_LAUNCH_CHECK();

    const int nblk = (int)nb(N, PT_TILE);
    const int32_t *lo = c.d_lo.as<int32_t>(), *cnt = c.d_cnt.as<int32_t>();
    const uint8_t *flags = c.d_flags.as<uint8_t>();
    int8_t *axis = c.d_axis.as<int8_t>();
    uint32_t *pivot = c.d_pivot.as<uint32_t>();
    // first level whose clusters all fit the LDS finish (at most TF_LEVELS above ml)
    int L0 = S.ml;
    for (int l = std::max(0, S.ml - TF_LEVELS); l <= S.ml; ++l) {
        int mx = 0;
        for (int64_t q = S.lvl_off[l]; q < S.lvl_off[l + 1]; ++q)
            mx = std::max(mx, S.cnt[S.lvl_nodes[q]]);
        if (mx <= TF_CAP) {
            L0 = l;
            break;
        }
    }
    // first level whose clusters all fit one workgroup's chunked pass (k_part_local)
    int L1 = L0;
    for (int l = 0; l < L0; ++l) {
        int mx = 0;
        for (int64_t q = S.lvl_off[l]; q < S.lvl_off[l + 1]; ++q)
            mx = std::max(mx, S.cnt[S.lvl_nodes[q]]);
        if (mx <= PL_MAXC) {
            L1 = l;
            break;
        }
    }
    for (int l = 0; l < L1; ++l) {
        int nnodes = (int)(S.lvl_off[l + 1] - S.lvl_off[l]);
        if (nnodes > 0) {
            hipLaunchKernelGGL(k_level_nodes, dim3(nb(nnodes, 128)), dim3(128), 0, st, nnodes,
                               c.d_lvl_nodes.as<int32_t>() + S.lvl_off[l], lo, cnt, flags, rk,
                               perm, N, pos, c.d_pavg.as<double>(), c.d_bbox.as<double>(), axis,
                               pivot, c.d_ext.as<int32_t>());
            FMM_LAUNCH_CHECK();
        }
        int nsplit = (int)(S.split_off[l + 1] - S.split_off[l]);
        if (nsplit == 0) continue;
        hipLaunchKernelGGL(k_part_count, dim3(nblk, 3), dim3(PT_THREADS), 0, st, N, l, seg, flags,
                           axis, pivot, rk, c.d_bsum.as<uint32_t>(), nblk);
        FMM_LAUNCH_CHECK();
        scan_rows(c.d_bsum.as<uint32_t>(), 3, nblk, c.d_bsum.as<uint32_t>() + 3 * (size_t)nblk, st);
        hipLaunchKernelGGL(k_part_scatter, dim3(nblk, 3), dim3(PT_THREADS), 0, st, N, l, seg,
                           flags, axis, pivot, lo, cnt, c.d_leftbase.as<int32_t>(), rk,
                           c.d_bsum.as<uint32_t>(), nblk, rk_alt, seg_alt);
        FMM_LAUNCH_CHECK();
        std::swap(rk, rk_alt);
        std::swap(seg, seg_alt);
    }
    for (int l = L1; l < L0; ++l) {
        const int nnodes = (int)(S.lvl_off[l + 1] - S.lvl_off[l]);
        if (nnodes == 0) continue;
        int mx = 0;
        for (int64_t q = S.lvl_off[l]; q < S.lvl_off[l + 1]; ++q)
            mx = std::max(mx, S.cnt[S.lvl_nodes[q]]);
        if (mx <= 2 * 256 * PL_I)
            hipLaunchKernelGGL(k_part_local<256>, dim3(nnodes, 3), dim3(256), 0, st,
                               c.d_lvl_nodes.as<int32_t>() + S.lvl_off[l], lo, cnt, flags, perm,
                               N, pos, c.d_pavg.as<double>(), rk, rk_alt, c.d_bbox.as<double>(),
                               axis, c.d_ext.as<int32_t>());
        else
            hipLaunchKernelGGL(k_part_local<1024>, dim3(nnodes, 3), dim3(1024), 0, st,
                               c.d_lvl_nodes.as<int32_t>() + S.lvl_off[l], lo, cnt, flags, perm,
                               N, pos, c.d_pavg.as<double>(), rk, rk_alt, c.d_bbox.as<double>(),
                               axis, c.d_ext.as<int32_t>());
        FMM_LAUNCH_CHECK();
        std::swap(rk, rk_alt);
    }
    {
        const int nnodes = (int)(S.lvl_off[L0 + 1] - S.lvl_off[L0]);
        hipLaunchKernelGGL(k_tree_finish, dim3(nnodes), dim3(TF_T), 0, st,
                           c.d_lvl_nodes.as<int32_t>() + S.lvl_off[L0], L0, S.ml, lo, cnt, flags,
                           rk, perm, N, pos, c.d_pavg.as<double>(), c.d_bbox.as<double>(), axis,
                           c.d_ext.as<int32_t>());
        FMM_LAUNCH_CHECK();
    }
    const int nleaf = (int)S.leaves.size();
    // one lane per particle of the largest leaf (a 64-particle leaf: one wave, where a
    // 256-thread block would leave three idle)
    const int bt_leaf = std::min(256, std::max(64, 64 * ((S.maxleaf + 63) / 64)));
    hipLaunchKernelGGL(k_leaf_sort, dim3(nleaf), dim3(bt_leaf), 0, st, c.d_leaves.as<int32_t>(), lo,
                       cnt, rk, perm, c.d_parind.as<int32_t>(), c.d_leafof.as<int32_t>(), pos, mom,
                       c.d_sp8.as<double>(),
                       c.precision == FMM4RB_MIXED ? c.d_sp8f.as<float>() : (float *)nullptr);
    FMM_LAUNCH_CHECK();
}

void refit_tree_device(Ctx &c, const double *pos, const double *mom) {
    const Shape &S = c.shape;
    hipStream_t st = c.stream;
    const int64_t N = c.N;
    hipLaunchKernelGGL(k_gather, dim3(nb(N, 256)), dim3(256), 0, st, N, c.d_parind.as<int32_t>(),
                       pos, mom, c.d_sp8.as<double>(),
                       c.precision == FMM4RB_MIXED ? c.d_sp8f.as<float>() : (float *)nullptr);
    FMM_LAUNCH_CHECK();
    const int nleaf = (int)S.leaves.size();
    hipLaunchKernelGGL(k_refit_leaves, dim3(nleaf), dim3(64), 0, st, c.d_leaves.as<int32_t>(),
                       c.d_lo.as<int32_t>(), c.d_cnt.as<int32_t>(), c.d_sp8.as<double>(),
                       c.d_bbox.as<double>());
    FMM_LAUNCH_CHECK();
    for (int l = S.ml - 1; l >= 0; --l) {  // split nodes of level l: split_off[l] .. split_off[l+1]
        const int64_t b = S.split_off[l], e = S.split_off[l + 1];
        if (e == b) continue;
        hipLaunchKernelGGL(k_refit_parents, dim3(nb(e - b, 128)), dim3(128), 0, st, (int)(e - b),
                           c.d_lvl_split.as<int32_t>() + b, c.d_bbox.as<double>());
        FMM_LAUNCH_CHECK();
    }
}

}  // namespace fmm4rb
