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
//      (coordinate_d, id)  -> lists L_d;
//   2. at each level the lists stay segment-wise sorted, so a cluster's bbox
//      is read off the first/last entry of each list, and the median split
//      along its axis a is "first ceil(c/2) entries of L_a";
//   3. a per-particle left/right flag then drives a stable segmented
//      partition (global prefix sum) of all three lists.
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

// ----------------------------------------------------------------- p_avg
// Deterministic fixed-order sum of the momenta and a finiteness check of all
// inputs.  update_particles_field.jl:77 (reduce(+, momenta)/N).
__global__ __launch_bounds__(PAVG_THREADS) void k_pavg_partial(const double *__restrict__ pos,
                                                               const double *__restrict__ mom,
                                                               int64_t N, double *__restrict__ part,
                                                               int *__restrict__ bad) {
    __shared__ double s[3][PAVG_THREADS];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    int b = 0;
    for (int64_t i = (int64_t)blockIdx.x * PAVG_THREADS + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * PAVG_THREADS) {
        double m0 = mom[3 * i], m1 = mom[3 * i + 1], m2 = mom[3 * i + 2];
        double x0 = pos[3 * i], x1 = pos[3 * i + 1], x2 = pos[3 * i + 2];
        a0 += m0;
        a1 += m1;
        a2 += m2;
        if (!(isfinite(m0) && isfinite(m1) && isfinite(m2) && isfinite(x0) && isfinite(x1) &&
              isfinite(x2)))
            b = 1;
    }
    s[0][threadIdx.x] = a0;
    s[1][threadIdx.x] = a1;
    s[2][threadIdx.x] = a2;
    for (int off = PAVG_THREADS / 2; off > 0; off >>= 1) {
        __syncthreads();
        if (threadIdx.x < (unsigned)off)
            for (int d = 0; d < 3; ++d) s[d][threadIdx.x] += s[d][threadIdx.x + off];
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) part[3 * blockIdx.x + d] = s[d][0];
    if (b) atomicOr(bad, 1);
}

__global__ __launch_bounds__(PAVG_THREADS) void k_pavg_final(const double *__restrict__ part,
                                                             int nblk, int64_t N,
                                                             double *__restrict__ pavg4) {
    __shared__ double s[3][PAVG_THREADS];
    double a[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nblk; i += PAVG_THREADS)
        for (int d = 0; d < 3; ++d) a[d] += part[3 * i + d];
    for (int d = 0; d < 3; ++d) s[d][threadIdx.x] = a[d];
    for (int off = PAVG_THREADS / 2; off > 0; off >>= 1) {
        __syncthreads();
        if (threadIdx.x < (unsigned)off)
            for (int d = 0; d < 3; ++d) s[d][threadIdx.x] += s[d][threadIdx.x + off];
    }
    if (threadIdx.x == 0) {
        double p0 = s[0][0] / (double)N, p1 = s[1][0] / (double)N, p2 = s[2][0] / (double)N;
        pavg4[0] = p0;
        pavg4[1] = p1;
        pavg4[2] = p2;
        pavg4[3] = sqrt(1.0 + ((p0 * p0 + p1 * p1) + p2 * p2));  // update_particles_field.jl:78
    }
}

void launch_pavg(const double *pos, const double *mom, int64_t N, double *part, double *pavg4,
                 int *bad, hipStream_t st) {
    int nblk = (int)std::min<int64_t>(PAVG_MAXBLK, (N + PAVG_THREADS - 1) / PAVG_THREADS);
    hipLaunchKernelGGL(k_pavg_partial, dim3(nblk), dim3(PAVG_THREADS), 0, st, pos, mom, N, part,
                       bad);
    FMM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pavg_final, dim3(1), dim3(PAVG_THREADS), 0, st, part, nblk, N, pavg4);
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
                            uint32_t *__restrict__ vals, int32_t *__restrict__ seg) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int d = 0; d < 3; ++d) {
        keys[d * N + i] = ordered_key(pos[3 * i + d]);
        vals[d * N + i] = (uint32_t)i;
    }
    seg[i] = 0;
}

// Bounding box (clusters.jl:51-70) and split axis (clusters.jl:39) of the
// existing nodes of one level.
__global__ void k_level_nodes(int nnodes, const int32_t *__restrict__ nodes,
                              const int32_t *__restrict__ lo, const int32_t *__restrict__ cnt,
                              const uint8_t *__restrict__ flags, const uint32_t *__restrict__ L,
                              int64_t N, const double *__restrict__ pos,
                              const double *__restrict__ pavg4, double *__restrict__ bbox,
                              int8_t *__restrict__ axis, int32_t *__restrict__ ext) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnodes) return;
    int32_t k = nodes[t];
    int64_t a = lo[k], b = (int64_t)lo[k] + cnt[k] - 1;
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        const uint32_t pa = L[d * N + a], pb = L[d * N + b];  // argmin / argmax particle
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
    }
}

__global__ void k_split_flags(int64_t N, int l, const int32_t *__restrict__ seg,
                              const uint8_t *__restrict__ flags, const int8_t *__restrict__ axis,
                              const int32_t *__restrict__ lo, const int32_t *__restrict__ cnt,
                              const uint32_t *__restrict__ L, uint8_t *__restrict__ flag,
                              int32_t *__restrict__ seg_out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int32_t k = seg[i];
    if (splitting(k, l, flags)) {
        int a = axis[k];
        int32_t nl = (cnt[k] + 1) >> 1;
        bool left = (i - lo[k]) < nl;
        flag[L[a * N + i]] = left ? 1 : 0;
        seg_out[i] = left ? 2 * k + 1 : 2 * k + 2;
    } else {
        seg_out[i] = k;
    }
}

__global__ __launch_bounds__(PT_THREADS) void k_part_count(
    int64_t N, int l, const int32_t *__restrict__ seg, const uint8_t *__restrict__ flags,
    const uint32_t *__restrict__ L, const uint8_t *__restrict__ flag, uint8_t *__restrict__ fbuf,
    uint32_t *__restrict__ bsum, int nblk) {
    __shared__ uint32_t s[PT_THREADS];
    const int d = blockIdx.y;
    const int64_t base = (int64_t)blockIdx.x * PT_TILE + PT_ITEMS * threadIdx.x;
    uint32_t c = 0;
    for (int j = 0; j < PT_ITEMS; ++j) {
        int64_t i = base + j;
        if (i < N) {
            uint8_t f = 0;
            if (splitting(seg[i], l, flags)) f = flag[L[d * N + i]];
            fbuf[d * N + i] = f;
            c += f;
        }
    }
    s[threadIdx.x] = c;
    for (int off = PT_THREADS / 2; off > 0; off >>= 1) {
        __syncthreads();
        if (threadIdx.x < (unsigned)off) s[threadIdx.x] += s[threadIdx.x + off];
    }
    if (threadIdx.x == 0) bsum[(int64_t)d * nblk + blockIdx.x] = s[0];
}

__global__ __launch_bounds__(PT_THREADS) void k_part_scatter(
    int64_t N, int l, const int32_t *__restrict__ seg, const uint8_t *__restrict__ flags,
    const int32_t *__restrict__ lo, const int32_t *__restrict__ cnt,
    const int32_t *__restrict__ leftbase, const uint32_t *__restrict__ L,
    const uint8_t *__restrict__ fbuf, const uint32_t *__restrict__ bsum, int nblk,
    uint32_t *__restrict__ Lout) {
    __shared__ uint32_t s[PT_THREADS];
    const int d = blockIdx.y;
    const int64_t base = (int64_t)blockIdx.x * PT_TILE + PT_ITEMS * threadIdx.x;
    uint8_t f[PT_ITEMS];
    uint32_t c = 0;
    for (int j = 0; j < PT_ITEMS; ++j) {
        int64_t i = base + j;
        f[j] = i < N ? fbuf[d * N + i] : 0;
        c += f[j];
    }
    // exclusive block scan of the per-thread counts
    s[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < PT_THREADS; off <<= 1) {
        uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    int64_t run = (int64_t)bsum[(int64_t)d * nblk + blockIdx.x] + s[threadIdx.x] - c;
    for (int j = 0; j < PT_ITEMS; ++j) {
        int64_t i = base + j;
        if (i >= N) break;
        int32_t k = seg[i];
        int64_t np = i;
        if (splitting(k, l, flags)) {
            int64_t lft = run - leftbase[k];
            int64_t nl = (cnt[k] + 1) >> 1;
            np = f[j] ? lo[k] + lft : lo[k] + nl + (i - lo[k]) - lft;
        }
        Lout[d * N + np] = L[d * N + i];
        run += f[j];
    }
}

// Canonical within-leaf order: ascending particle id (rank sort in LDS).
__global__ __launch_bounds__(256) void k_leaf_sort(const int32_t *__restrict__ leaves,
                                                   const int32_t *__restrict__ lo,
                                                   const int32_t *__restrict__ cnt,
                                                   const uint32_t *__restrict__ Lx,
                                                   int32_t *__restrict__ parind,
                                                   int32_t *__restrict__ leafof) {
    __shared__ uint32_t s[MAX_N0];
    const int32_t k = leaves[blockIdx.x];
    const int32_t a = lo[k], c = cnt[k];
    for (int j = threadIdx.x; j < c; j += blockDim.x) s[j] = Lx[a + j];
    __syncthreads();
    for (int j = threadIdx.x; j < c; j += blockDim.x) {
        uint32_t v = s[j];
        int r = 0;
        for (int m = 0; m < c; ++m) r += s[m] < v;
        parind[a + r] = (int32_t)v;
        leafof[a + j] = k;
    }
}

// Particles in tree order: sp8 = (x, y, z, gamma, px, py, pz, 0) and, for the
// mixed-precision path, sp4f = (gamma, px, py, pz) in FP32.
__global__ void k_gather(int64_t N, const int32_t *__restrict__ parind,
                         const double *__restrict__ pos, const double *__restrict__ mom,
                         double *__restrict__ sp8, float *__restrict__ sp4f) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int64_t p = parind[i];
    double x0 = pos[3 * p], x1 = pos[3 * p + 1], x2 = pos[3 * p + 2];
    double m0 = mom[3 * p], m1 = mom[3 * p + 1], m2 = mom[3 * p + 2];
    double g = sqrt(1.0 + ((m0 * m0 + m1 * m1) + m2 * m2));  // interaction.jl:67
    double4 *o = reinterpret_cast<double4 *>(sp8 + 8 * i);
    o[0] = make_double4(x0, x1, x2, g);
    o[1] = make_double4(m0, m1, m2, 0.0);
    if (sp4f)
        reinterpret_cast<float4 *>(sp4f)[i] = make_float4((float)g, (float)m0, (float)m1, (float)m2);
}

// ------------------------------------------------------------- host side
static inline unsigned nb(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void build_tree_device(Ctx &c, const double *pos, const double *mom, int64_t N, int64_t N0) {
    (void)mom;
    (void)N0;
    const Shape &S = c.shape;
    hipStream_t st = c.stream;
    uint64_t *keys = c.d_keys.as<uint64_t>(), *keys_alt = c.d_keys_alt.as<uint64_t>();
    uint32_t *vals = c.d_vals.as<uint32_t>(), *vals_alt = c.d_vals_alt.as<uint32_t>();
    int32_t *seg = c.d_seg.as<int32_t>(), *seg_alt = c.d_seg_alt.as<int32_t>();

    hipLaunchKernelGGL(k_make_keys, dim3(nb(N, 256)), dim3(256), 0, st, pos, N, keys, vals, seg);
    FMM_LAUNCH_CHECK();
    bool alt = radix_sort(keys, keys_alt, vals, vals_alt, N, 3, 0, 64,
                          c.d_rs_scratch.as<uint32_t>(), st);
    uint32_t *L = alt ? vals_alt : vals, *Lo = alt ? vals : vals_alt;

    const int nblk = (int)nb(N, PT_TILE);
    const int32_t *lo = c.d_lo.as<int32_t>(), *cnt = c.d_cnt.as<int32_t>();
    const uint8_t *flags = c.d_flags.as<uint8_t>();
    for (int l = 0; l <= S.ml; ++l) {
        int nnodes = (int)(S.lvl_off[l + 1] - S.lvl_off[l]);
        if (nnodes > 0) {
            hipLaunchKernelGGL(k_level_nodes, dim3(nb(nnodes, 128)), dim3(128), 0, st, nnodes,
                               c.d_lvl_nodes.as<int32_t>() + S.lvl_off[l], lo, cnt, flags, L, N,
                               pos, c.d_pavg.as<double>(), c.d_bbox.as<double>(),
                               c.d_axis.as<int8_t>(), c.d_ext.as<int32_t>());
            FMM_LAUNCH_CHECK();
        }
        int nsplit = (int)(S.split_off[l + 1] - S.split_off[l]);
        if (l == S.ml || nsplit == 0) continue;
        hipLaunchKernelGGL(k_split_flags, dim3(nb(N, 256)), dim3(256), 0, st, N, l, seg, flags,
                           c.d_axis.as<int8_t>(), lo, cnt, L, c.d_flag.as<uint8_t>(), seg_alt);
        FMM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_part_count, dim3(nblk, 3), dim3(PT_THREADS), 0, st, N, l, seg, flags,
                           L, c.d_flag.as<uint8_t>(), c.d_fbuf.as<uint8_t>(),
                           c.d_bsum.as<uint32_t>(), nblk);
        FMM_LAUNCH_CHECK();
        scan_rows(c.d_bsum.as<uint32_t>(), 3, nblk, c.d_bsum.as<uint32_t>() + 3 * (size_t)nblk, st);
        hipLaunchKernelGGL(k_part_scatter, dim3(nblk, 3), dim3(PT_THREADS), 0, st, N, l, seg,
                           flags, lo, cnt, c.d_leftbase.as<int32_t>(), L, c.d_fbuf.as<uint8_t>(),
                           c.d_bsum.as<uint32_t>(), nblk, Lo);
        FMM_LAUNCH_CHECK();
        std::swap(L, Lo);
        std::swap(seg, seg_alt);
    }
    const int nleaf = (int)S.leaves.size();
    hipLaunchKernelGGL(k_leaf_sort, dim3(nleaf), dim3(256), 0, st, c.d_leaves.as<int32_t>(), lo,
                       cnt, L, c.d_parind.as<int32_t>(), c.d_leafof.as<int32_t>());
    FMM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gather, dim3(nb(N, 256)), dim3(256), 0, st, N, c.d_parind.as<int32_t>(),
                       pos, mom, c.d_sp8.as<double>(),
                       c.precision == FMM4RB_MIXED ? c.d_sp8f.as<float>() : (float *)nullptr);
    FMM_LAUNCH_CHECK();
}

}  // namespace fmm4rb
