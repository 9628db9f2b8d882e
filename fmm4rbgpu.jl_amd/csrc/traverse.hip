// traverse.hip -- dual-tree traversal into flat P2P / M2L lists on gfx950.
//
// Reference: dualtraversefill! and admissible (src/fmm/interaction_lists.jl:10-49),
// diameter (src/utils/bbox.jl:9-11), and the target-grouped CSR of
// InteractionListsGPU (src/gpu_fmm/gpu_interaction_lists.jl:14-40).
//
// The reference recursion from (root, root) decides each pair (t, s) from
// t's and s's data alone, so the visited pairs form a tree whose levels can
// be expanded breadth-first: iteration i holds the pairs whose depths sum to
// i, and at most 2*maxlevel+1 iterations run.  Each pair becomes a P2P pair
// (both leaves), an M2L pair (admissible), or two child pairs (split the leaf
// -less side, else the side with the larger diameter, ties split the source).
// Outputs are appended with workgroup-aggregated atomics as keys (t << sb) | s,
// then radix-sorted, which gives the canonical order (target, then source)
// and the CSR offsets per target node.
#include "fmm4rb_internal.hpp"

#pragma clang fp contract(off)

namespace fmm4rb {

constexpr int TR_THREADS = 256;
constexpr int TR_BLOCKS = 2048;

__device__ inline double norm3(double a, double b, double c) { return sqrt((a * a + b * b) + c * c); }

// interaction_lists.jl:10-21 with precomputed centre and radius per node.
__device__ inline bool admissible(const double *gt, const double *gs, double g, double eta) {
    double v0 = 1.0 * (gt[0] - gs[0]);
    double v1 = 1.0 * (gt[1] - gs[1]);
    double v2 = g * (gt[2] - gs[2]);
    double R = norm3(v0, v1, v2);
    double r = gt[3] < gs[3] ? gs[3] : gt[3];
    return (r / R) < eta;
}

__global__ void k_trav_init(int2 *front, uint32_t *cnt, int ncnt) {
    int i = threadIdx.x;
    for (int j = i; j < ncnt; j += blockDim.x) cnt[j] = 0;
    if (i == 0) {
        front[0] = make_int2(0, 0);
    }
    __syncthreads();
    if (i == 0) cnt[2] = 1;
}

// TR_ITEMS pairs per thread per step (independent loads in flight, one atomic per
// category per TR_ITEMS x TR_THREADS pairs); the outputs of a step are placed in
// (item, wave, lane) order after the workgroup's reservation.
constexpr int TR_ITEMS = 4;

__global__ __launch_bounds__(TR_THREADS) void k_traverse(
    const int2 *__restrict__ fin, int2 *__restrict__ fout, uint32_t *__restrict__ cnt, int it,
    uint32_t cap_f, uint64_t *__restrict__ p2p, uint32_t cap_p2p, uint64_t *__restrict__ m2l,
    uint32_t cap_m2l, int sb, const double *__restrict__ geo, const uint8_t *__restrict__ flags,
    const double *__restrict__ pavg4, double eta) {
    constexpr int NW = TR_THREADS / 64;
    __shared__ uint32_t wc[3][TR_ITEMS][NW], bb[3];
    const uint32_t n = min(cnt[2 + it], cap_f);
    const double g = pavg4[3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    constexpr uint32_t STEP = TR_ITEMS * TR_THREADS;
    for (uint32_t base = blockIdx.x * STEP; base < n; base += gridDim.x * STEP) {
        int cat[TR_ITEMS], t[TR_ITEMS], s[TR_ITEMS];
        int2 pr[TR_ITEMS];
#pragma unroll
        for (int j = 0; j < TR_ITEMS; ++j) {
            const uint32_t i = base + j * TR_THREADS + threadIdx.x;
            pr[j] = i < n ? fin[i] : make_int2(-1, -1);
        }
#pragma unroll
        for (int j = 0; j < TR_ITEMS; ++j) {
            t[j] = pr[j].x;
            s[j] = pr[j].y;
            cat[j] = -1;
            if (t[j] >= 0) {
                const bool tl = flags[t[j]] & 2, sl = flags[s[j]] & 2;
                const double *gt = geo + 8 * (int64_t)t[j], *gs = geo + 8 * (int64_t)s[j];
                if (tl && sl) cat[j] = 0;
                else if (admissible(gt, gs, g, eta)) cat[j] = 1;
                else if (tl) cat[j] = 3;
                else if (sl) cat[j] = 2;
                else cat[j] = gt[4] > gs[4] ? 2 : 3;  // diameter(t) > diameter(s)
            }
        }
        uint64_t m[3][TR_ITEMS];
#pragma unroll
        for (int j = 0; j < TR_ITEMS; ++j) {
            m[0][j] = __ballot(cat[j] == 0);
            m[1][j] = __ballot(cat[j] == 1);
            m[2][j] = __ballot(cat[j] >= 2);
            if (lane == 0) {
                wc[0][j][wave] = (uint32_t)__popcll(m[0][j]);
                wc[1][j][wave] = (uint32_t)__popcll(m[1][j]);
                wc[2][j][wave] = 2u * (uint32_t)__popcll(m[2][j]);
            }
        }
        __syncthreads();
        if (threadIdx.x < 3) {  // one atomic per category per step
            const int k = threadIdx.x;
            uint32_t tot = 0;
            for (int j = 0; j < TR_ITEMS; ++j)
                for (int w = 0; w < NW; ++w) tot += wc[k][j][w];
            uint32_t *ctr = k == 2 ? &cnt[2 + it + 1] : &cnt[k];
            bb[k] = tot ? atomicAdd(ctr, tot) : 0u;
        }
        __syncthreads();
        uint32_t o[3] = {bb[0], bb[1], bb[2]};
#pragma unroll
        for (int j = 0; j < TR_ITEMS; ++j) {
            uint32_t pre[3] = {0u, 0u, 0u}, tot[3] = {0u, 0u, 0u};
            for (int w = 0; w < NW; ++w)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    pre[k] += w < wave ? wc[k][j][w] : 0u;
                    tot[k] += wc[k][j][w];
                }
            if (cat[j] == 0) {
                const uint32_t q = o[0] + pre[0] + (uint32_t)__popcll(m[0][j] & lt);
                if (q < cap_p2p) p2p[q] = ((uint64_t)t[j] << sb) | (uint64_t)s[j];
            } else if (cat[j] == 1) {
                const uint32_t q = o[1] + pre[1] + (uint32_t)__popcll(m[1][j] & lt);
                if (q < cap_m2l) m2l[q] = ((uint64_t)t[j] << sb) | (uint64_t)s[j];
            } else if (cat[j] >= 2) {
                const uint32_t q = o[2] + pre[2] + 2u * (uint32_t)__popcll(m[2][j] & lt);
                if (q + 1 < cap_f) {
                    if (cat[j] == 2) {
                        fout[q] = make_int2(2 * t[j] + 1, s[j]);
                        fout[q + 1] = make_int2(2 * t[j] + 2, s[j]);
                    } else {
                        fout[q] = make_int2(t[j], 2 * s[j] + 1);
                        fout[q + 1] = make_int2(t[j], 2 * s[j] + 2);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) o[k] += tot[k];
        }
        __syncthreads();  // wc and bb are rewritten in the next step
    }
}

// off[t] = first index whose target is >= t (binary search), t in [0, nc].
__global__ void k_csr(const uint64_t *__restrict__ keys, int64_t n, int sb, int64_t nc,
                      int32_t *__restrict__ off) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > nc) return;
    const uint64_t key = (uint64_t)t << sb;
    int64_t a = 0, b = n;
    while (a < b) {
        int64_t m = (a + b) >> 1;
        if (keys[m] < key) a = m + 1;
        else b = m;
    }
    off[t] = (int32_t)a;
}

// The two lists side by side, each padded to L keys with pad (above every pair key).
__global__ void k_pack_pairs(const uint64_t *__restrict__ p2p, int64_t np2p,
                             const uint64_t *__restrict__ m2l, int64_t nm2l, int64_t L,
                             uint64_t pad, uint64_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    out[i] = i < np2p ? p2p[i] : pad;
    out[L + i] = i < nm2l ? m2l[i] : pad;
}

void sort_pairs_csr(Ctx &c, uint64_t *keys, uint64_t *keys_alt, int64_t n, int32_t *off,
                    uint64_t **sorted, uint32_t *vals, uint32_t *vals_alt, uint32_t **vsorted) {
    hipStream_t st = c.stream;
    bool alt = false;
    if (n > 0)
        alt = radix_sort(keys, keys_alt, vals, vals_alt, n, 1, 0, 2 * c.sb,
                         c.d_rs_scratch.as<uint32_t>(), st);
    uint64_t *r = alt ? keys_alt : keys;
    *sorted = r;
    if (vsorted) *vsorted = alt ? vals_alt : vals;
    const int64_t nc = c.shape.nc;
    hipLaunchKernelGGL(k_csr, dim3((unsigned)((nc + 1 + 255) / 256)), dim3(256), 0, st, r, n, c.sb,
                       nc, off);
    FMM_LAUNCH_CHECK();
}

// pos[fwd[e]] = e: the transposed position of every forward pair
__global__ void k_scatter_pos(const uint32_t *__restrict__ fwd, int64_t n,
                              uint32_t *__restrict__ pos) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) pos[fwd[e]] = (uint32_t)e;
}

// Both lists transposed ((t << sb) | s -> (s << sb) | t) side by side, each padded to
// L keys with pad, and each pair's position within its list as the value.
__global__ void k_transpose_pack(const uint64_t *__restrict__ p2p, int64_t np2p,
                                 const uint64_t *__restrict__ m2l, int64_t nm2l, int64_t L,
                                 int sb, uint64_t pad, uint64_t *__restrict__ out,
                                 uint32_t *__restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    const uint64_t mask = (1ull << sb) - 1ull;
    const uint64_t a = i < np2p ? p2p[i] : 0ull, b = i < nm2l ? m2l[i] : 0ull;
    out[i] = i < np2p ? ((a & mask) << sb) | (a >> sb) : pad;
    out[L + i] = i < nm2l ? ((b & mask) << sb) | (b >> sb) : pad;
    idx[i] = (uint32_t)i;
    idx[L + i] = (uint32_t)i;
}

// Pairs regrouped by source, both lists in one two-segment sort (as the forward
// lists): keys (s << sb) | t, CSR offsets per source; with pos, pos[i] = the position
// in the transposed M2L list of forward M2L pair i.
void transpose_lists_joint(Ctx &c, DevBuf *pos) {
    hipStream_t st = c.stream;
    const int64_t L = std::max<int64_t>(std::max(c.np2p, c.nm2l), 1), nc = c.shape.nc;
    c.d_p2pT.ensure(2 * 8 * (size_t)L);
    c.d_p2pT_alt.ensure(2 * 8 * (size_t)L);
    c.d_vals.ensure(2 * 4 * (size_t)L);
    c.d_vals_alt.ensure(2 * 4 * (size_t)L);
    c.d_p2pT_off.ensure(4 * (size_t)(nc + 1));
    c.d_m2lT_off.ensure(4 * (size_t)(nc + 1));
    c.d_rs_scratch.ensure(4 * radix_scratch_words(L, 2));
    uint64_t *k = c.d_p2pT.as<uint64_t>(), *ka = c.d_p2pT_alt.as<uint64_t>();
    uint32_t *v = c.d_vals.as<uint32_t>(), *va = c.d_vals_alt.as<uint32_t>();
    hipLaunchKernelGGL(k_transpose_pack, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, st,
                       c.p2p_sorted, c.np2p, c.m2l_sorted, c.nm2l, L, c.sb,
                       (1ull << (2 * c.sb)) - 1ull, k, v);
    FMM_LAUNCH_CHECK();
    const bool alt = radix_sort(k, ka, v, va, L, 2, 0, 2 * c.sb, c.d_rs_scratch.as<uint32_t>(), st);
    uint64_t *r = alt ? ka : k;
    const uint32_t *vs = alt ? va : v;
    c.p2pT_sorted = r;
    c.m2lT_sorted = r + L;
    hipLaunchKernelGGL(k_csr, dim3((unsigned)((nc + 1 + 255) / 256)), dim3(256), 0, st, r, L, c.sb,
                       nc, c.d_p2pT_off.as<int32_t>());
    FMM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_csr, dim3((unsigned)((nc + 1 + 255) / 256)), dim3(256), 0, st, r + L, L,
                       c.sb, nc, c.d_m2lT_off.as<int32_t>());
    FMM_LAUNCH_CHECK();
    if (pos && c.nm2l > 0) {
        pos->ensure(4 * (size_t)c.nm2l);
        hipLaunchKernelGGL(k_scatter_pos, dim3((unsigned)((c.nm2l + 255) / 256)), dim3(256), 0, st,
                           vs + L, c.nm2l, pos->as<uint32_t>());
        FMM_LAUNCH_CHECK();
    }
}

bool build_lists_device(Ctx &c, double eta) {
    const Shape &S = c.shape;
    hipStream_t st = c.stream;
    const int iters = 2 * S.ml + 1;
    const int ncnt = iters + 3;
    for (int attempt = 0;; ++attempt) {
        int2 *f0 = c.d_front0.as<int2>(), *f1 = c.d_front1.as<int2>();
        uint32_t *cnt = c.d_tcnt.as<uint32_t>();
        hipLaunchKernelGGL(k_trav_init, dim3(1), dim3(64), 0, st, f0, cnt, ncnt);
        FMM_LAUNCH_CHECK();
        for (int it = 0; it < iters; ++it) {
            const int2 *fin = (it & 1) ? f1 : f0;
            int2 *fout = (it & 1) ? f0 : f1;
            hipLaunchKernelGGL(k_traverse, dim3(TR_BLOCKS), dim3(TR_THREADS), 0, st, fin, fout, cnt,
                               it, (uint32_t)c.cap_front, c.d_p2p.as<uint64_t>(),
                               (uint32_t)c.cap_p2p, c.d_m2l.as<uint64_t>(), (uint32_t)c.cap_m2l,
                               c.sb, c.d_geo.as<double>(), c.d_flags.as<uint8_t>(),
                               c.d_pavg.as<double>(), eta);
            FMM_LAUNCH_CHECK();
        }
        uint32_t h[TCNT_WORDS];  // ncnt <= 2 * maxlevel + 4 <= 62 counters, then the tree flag
        FMM_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
        FMM_HIP(hipStreamSynchronize(st));
        if (h[TCNT_FLAG]) return false;  // the tree is not the reference's: build it exactly
        uint64_t maxf = 0;
        for (int it = 0; it <= iters; ++it) maxf = std::max<uint64_t>(maxf, h[2 + it]);
        bool over = h[0] > (uint64_t)c.cap_p2p || h[1] > (uint64_t)c.cap_m2l ||
                    maxf > (uint64_t)c.cap_front;
        if (!over) {
            c.np2p = h[0];
            c.nm2l = h[1];
            c.trav_iters = iters;
            break;
        }
        if (attempt >= 8) fail(FMM4RB_ENOMEM, "interaction list capacity did not converge");
        // grow what overflowed (2x the count seen) and traverse again
        if (h[0] >= (uint64_t)c.cap_p2p) {
            c.cap_p2p = std::max<int64_t>(2 * (int64_t)h[0], 2 * c.cap_p2p);
            c.d_p2p.ensure(8 * (size_t)c.cap_p2p);
        }
        if (h[1] >= (uint64_t)c.cap_m2l) {
            c.cap_m2l = std::max<int64_t>(2 * (int64_t)h[1], 2 * c.cap_m2l);
            c.d_m2l.ensure(8 * (size_t)c.cap_m2l);
        }
        if (maxf >= (uint64_t)c.cap_front) {
            c.cap_front = std::max<int64_t>(2 * (int64_t)maxf, 2 * c.cap_front);
            c.d_front0.ensure(8 * (size_t)c.cap_front);
            c.d_front1.ensure(8 * (size_t)c.cap_front);
        }
        if (c.cap_front > 0x7fffffffLL || c.cap_p2p > 0x7fffffffLL || c.cap_m2l > 0x7fffffffLL)
            fail(FMM4RB_ENOMEM, "interaction lists exceed 2^31 entries");
    }
    // Both lists sorted in one radix sort of two segments of L keys each (the shorter
    // padded with keys above every pair), half the launches of two sorts; the
    // buffers and scratch grow here only after a capacity retry (the stream is idle).
    const int64_t L = std::max<int64_t>(std::max(c.np2p, c.nm2l), 1);
    c.pair_stride = L;
    c.d_pairs.ensure(2 * 8 * (size_t)L);
    c.d_pairs_alt.ensure(2 * 8 * (size_t)L);
    c.d_rs_scratch.ensure(4 * radix_scratch_words(L, 2));
    uint64_t *pr = c.d_pairs.as<uint64_t>(), *pra = c.d_pairs_alt.as<uint64_t>();
    hipLaunchKernelGGL(k_pack_pairs, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, st,
                       c.d_p2p.as<uint64_t>(), c.np2p, c.d_m2l.as<uint64_t>(), c.nm2l, L,
                       (1ull << (2 * c.sb)) - 1ull, pr);
    FMM_LAUNCH_CHECK();
    const bool alt = radix_sort(pr, pra, nullptr, nullptr, L, 2, 0, 2 * c.sb,
                                c.d_rs_scratch.as<uint32_t>(), st);
    uint64_t *r = alt ? pra : pr;
    c.p2p_sorted = r;
    c.m2l_sorted = r + L;
    const int64_t nc = S.nc;
    hipLaunchKernelGGL(k_csr, dim3((unsigned)((nc + 1 + 255) / 256)), dim3(256), 0, st, r, L, c.sb,
                       nc, c.d_p2p_off.as<int32_t>());
    FMM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_csr, dim3((unsigned)((nc + 1 + 255) / 256)), dim3(256), 0, st, r + L, L,
                       c.sb, nc, c.d_m2l_off.as<int32_t>());
    FMM_LAUNCH_CHECK();
    return true;
}

}  // namespace fmm4rb
