// p2p.hip -- near-field particle-particle sums on gfx950 (FP64 and mixed precision).
//
// Reference: P2P! (src/fmm/interaction.jl:49-75) and gpu_P2P!
// (src/gpu_fmm/gpu_interaction.jl:43-86): for every leaf pair (t, s)
//   E_i += sum_j gamma_j K(x_i - x_j; p_j),  B_i += sum_j p_j x K(x_i - x_j; p_j)
// for i in t and j in s (the self pair included: R = 0 gives K = 0).
//
// One workgroup owns one target leaf, walks the leaf's CSR source list and
// stages the source leaves in LDS (x, y, z in FP64; gamma, px, py, pz), several
// consecutive leaves per tile (up to 4 times the block size, 10 KB per wave), so that
// the per-tile work (barriers, the FP32 tile sums' compensated update) is shared
// by more pairs.  A wave
// covers 16 TPL targets as 16 target lanes x 4 source groups: lane (ti, g) holds
// the TPL targets ti + 16 q and takes the staged sources j = g (mod 4), so one
// LDS read serves TPL pairs (one target per lane moves 40 bytes of LDS data per
// pair and is bound by LDS bandwidth).  Each staged tile's partial sums are added
// into the lane's FP64 sums; the 4 groups' sums meet in two xor-shuffles at the
// end, in a fixed order.  TPL = 4 (64 targets per wave), or 8 for FP64 with
// leaves above 64 particles (reducing per tile instead, to keep fewer FP64 sums
// per lane, measured 1.3x slower).  A wave with fewer targets left than 16 TPL (the
// last of a leaf) runs ceil(left / 16) targets per lane: a 98-particle leaf takes
// 64 + 48 target slots instead of 128 (N = 25.6M, mixed P2P 81.0 -> 73.9 ms, FP64
// 133.4 -> 114.8 ms).
// Results go to a tree-order near-field array that L2P adds to the far field.
// Mixed precision: R = x_i - x_j is formed in FP64 and rounded once to FP32
// (the closest pairs dominate the field of their particles, so R must keep
// its relative accuracy however small it is); the kernel then runs in FP32
// with v_rsq_f32, and each staged tile's FP32 partial sums are added into
// Kahan-compensated FP32 accumulators, combined in FP64 at the end.
#include "fmm4rb_internal.hpp"

namespace fmm4rb {

constexpr int P2P_MAXT = 512;  // threads per block: up to 8 waves
constexpr int P2P_TILE_MUL = 4;        // staged sources per tile: up to 4 x the block size
constexpr size_t P2P_TILE_LDS_WAVE = 10 << 10;  // and at most 10 KB of LDS per wave (so
                                                 // that LDS allows 4 waves per SIMD)

template <typename T>
__device__ inline void p2p_pair(const double *xi, double sx, double sy, double sz, T g, T px,
                                T py, T pz, T *a) {
    const T R0 = (T)(xi[0] - sx), R1 = (T)(xi[1] - sy), R2 = (T)(xi[2] - sz);
    const T rr = fma(R2, R2, fma(R1, R1, fma(R0, R0, (T)EPS64)));
    const T pr = fma(pz, R2, fma(py, R1, px * R0));
    const T d = fma(pr, pr, rr);
    const T y3 = rsq3<T>(d);
    const T u0 = y3 * R0, u1 = y3 * R1, u2 = y3 * R2;  // K = d^-3/2 R
    a[0] = fma(g, u0, a[0]);                            // E += g K
    a[1] = fma(g, u1, a[1]);
    a[2] = fma(g, u2, a[2]);
    a[3] = fma(-pz, u1, fma(py, u2, a[3]));             // B += p x K
    a[4] = fma(-px, u2, fma(pz, u0, a[4]));
    a[5] = fma(-py, u0, fma(px, u1, a[5]));
}

template <bool MIXED, int TPL>
__global__ __launch_bounds__(P2P_MAXT) void k_p2p(
    const int32_t *__restrict__ leaves, const int32_t *__restrict__ off,
    const uint64_t *__restrict__ keys, int sb, const int32_t *__restrict__ lo,
    const int32_t *__restrict__ cnt, const double *__restrict__ sp8, const float *__restrict__ sp4f,
    double *__restrict__ nearf, int TC) {
    using T = typename std::conditional<MIXED, float, double>::type;
    using Src = P2PSrc<MIXED>;
    // LDS sized by the launch (TC records, a multiple of blockDim.x), not by
    // P2P_MAXT: a one-wave block holds 10 KB (mixed, 256 sources), so LDS allows 4
    // waves per SIMD, as the registers do
    extern __shared__ double2 p2p_lds[];
    const int BT = blockDim.x;  // TC: tile capacity
    Src *ts = reinterpret_cast<Src *>(p2p_lds);
    const int tid = threadIdx.x;
    const int lane = tid & 63, ti = lane & 15, g = lane >> 4;
    const int tw = (tid >> 6) * 16 * TPL;  // first target of this wave within a pass
    const int per_pass = (BT >> 6) * 16 * TPL;
    const int32_t t = leaves[xcd_item<XCD_RUN_NEAR>(blockIdx.x, gridDim.x)];
    const int32_t a = lo[t], ct = cnt[t];
    const int32_t beg = off[t], end = off[t + 1];
    const uint64_t mask = (1ull << sb) - 1ull;
    const double4 *p8 = reinterpret_cast<const double4 *>(sp8);
    const float4 *q4 = reinterpret_cast<const float4 *>(sp4f);
    for (int tb = 0; tb < ct; tb += per_pass) {
        double xi[TPL][3];
        // FP64: the sums.  MIXED: Kahan-compensated FP32 sums of the tiles' partial
        // sums (hi, cmp), half the registers of FP64 sums
        double acc[MIXED ? 1 : TPL][6];
        float hi[MIXED ? TPL : 1][6], cmp[MIXED ? TPL : 1][6];
#pragma unroll
        for (int q = 0; q < TPL; ++q) {
            const int i = min(tb + tw + ti + 16 * q, ct - 1);
            const double4 x4 = p8[2 * (int64_t)(a + i)];
            xi[q][0] = x4.x;
            xi[q][1] = x4.y;
            xi[q][2] = x4.z;
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                if constexpr (MIXED) hi[q][m] = cmp[q][m] = 0.0f;
                else acc[q][m] = 0.0;
            }
        }
        // Targets of this wave in this pass: 16 nq slots cover them (the last wave of a
        // leaf whose size is not a multiple of 16 TPL runs fewer targets per lane, e.g.
        // 64 + 48 slots for a 98-particle leaf instead of 128); wave-uniform.
        const int nq = max(0, min(TPL, (ct - tb - tw + 15) / 16));
        // The tile's m staged sources against the lane's first Q targets; the tile's
        // FP32 sums go into the Kahan pairs (MIXED) or straight into the sums.
        auto tile_q = [&](auto QC, int m) __attribute__((always_inline)) {
            constexpr int Q = decltype(QC)::value;
            if constexpr (MIXED) {
                T part[Q][6];
#pragma unroll
                for (int q = 0; q < Q; ++q)
#pragma unroll
                    for (int k = 0; k < 6; ++k) part[q][k] = 0;
                // m is a multiple of 4 (the last tile is padded): the same trip count in
                // every lane, one address increment per step
                for (int it = 0; it < (m >> 2); ++it) {
                    const Src sr = ts[4 * it + g];
#pragma unroll
                    for (int q = 0; q < Q; ++q)
                        p2p_pair<T>(xi[q], sr.x, sr.y, sr.z, sr.g, sr.px, sr.py, sr.pz, part[q]);
                }
#pragma unroll
                for (int q = 0; q < Q; ++q)
#pragma unroll
                    for (int k = 0; k < 6; ++k) {  // Kahan: hi - cmp is the running sum
                        const float y = part[q][k] - cmp[q][k];
                        const float t = hi[q][k] + y;
                        cmp[q][k] = (t - hi[q][k]) - y;
                        hi[q][k] = t;
                    }
            } else {
                for (int it = 0; it < (m >> 2); ++it) {
                    const Src sr = ts[4 * it + g];
#pragma unroll
                    for (int q = 0; q < Q; ++q)
                        p2p_pair<T>(xi[q], sr.x, sr.y, sr.z, sr.g, sr.px, sr.py, sr.pz, acc[q]);
                }
            }
        };
        auto run_tile = [&](int m) __attribute__((always_inline)) {
            using std::integral_constant;
            if (nq == TPL) tile_q(integral_constant<int, TPL>{}, m);
            else if (nq == 1) tile_q(integral_constant<int, 1>{}, m);
            else if (nq == 2) tile_q(integral_constant<int, 2>{}, m);
            else if (nq == 3) tile_q(integral_constant<int, 3>{}, m);
            if constexpr (TPL > 4) {
                if (nq == 4) tile_q(integral_constant<int, 4>{}, m);
                else if (nq == 5) tile_q(integral_constant<int, (TPL > 5 ? 5 : 1)>{}, m);
                else if (nq == 6) tile_q(integral_constant<int, (TPL > 6 ? 6 : 1)>{}, m);
                else if (nq == 7) tile_q(integral_constant<int, (TPL > 7 ? 7 : 1)>{}, m);
            }
        };
        // Fill the tile with consecutive source leaves (a leaf larger than the room
        // left is split across tiles); fill is the same for every thread.
        int fill = 0;
        __syncthreads();  // the previous pass's last tile has been read
        for (int32_t e = beg; e < end; ++e) {
            const int32_t s = (int32_t)(keys[e] & mask);
            const int32_t sa = lo[s], cs = cnt[s];
            for (int s0 = 0; s0 < cs;) {
                const int take = min(cs - s0, TC - fill);
#pragma unroll 1
                for (int j = tid; j < take; j += BT) {
                    const int64_t src = sa + s0 + j;
                    const double4 x4 = p8[2 * src];
                    Src sr;
                    sr.x = x4.x;
                    sr.y = x4.y;
                    sr.z = x4.z;
                    if constexpr (MIXED) {
                        const float4 q = q4[src];  // gamma, px, py, pz
                        sr.g = q.x;
                        sr.px = q.y;
                        sr.py = q.z;
                        sr.pz = q.w;
                    } else {
                        const double4 p4 = p8[2 * src + 1];  // px, py, pz, 0
                        sr.g = x4.w;
                        sr.px = p4.x;
                        sr.py = p4.y;
                        sr.pz = p4.z;
                    }
                    ts[fill + j] = sr;
                }
                fill += take;
                s0 += take;
                if (fill == TC) {
                    __syncthreads();
                    run_tile(TC);
                    __syncthreads();
                    fill = 0;
                }
            }
        }
        if (fill > 0) {
            // pad to a multiple of 4 with sources that add exactly zero
            const int fill4 = (fill + 3) & ~3;  // <= TC, a multiple of 64
            if (tid < fill4 - fill) ts[fill + tid] = Src::pad();
            __syncthreads();
            run_tile(fill4);
        }
        // Sum the 4 source groups (lanes ti + 16 g), fixed order, in FP64.
        double tot[TPL][6];
#pragma unroll
        for (int q = 0; q < TPL; ++q)
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                double v;
                if constexpr (MIXED) v = (double)hi[q][k] - (double)cmp[q][k];
                else v = acc[q][k];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                tot[q][k] = v;
            }
        // lane (ti, g) writes targets ti + 16 (g + 4 r)
#pragma unroll
        for (int r = 0; r < TPL / 4; ++r) {
            double out[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                out[k] = tot[4 * r][k];
#pragma unroll
                for (int u = 1; u < 4; ++u)
                    if (g == u) out[k] = tot[4 * r + u][k];
            }
            const int i = tb + tw + ti + 16 * (g + 4 * r);
            if (i < ct) {
                double *o = nearf + 6 * (int64_t)(a + i);
#pragma unroll
                for (int k = 0; k < 6; ++k) o[k] = out[k];
            }
        }
    }
}

void launch_p2p(Ctx &c, const int32_t *off, const uint64_t *keys, int sb, hipStream_t st) {
    const unsigned nleaf = (unsigned)c.shape.leaves.size();
    const bool mixed = c.precision == FMM4RB_MIXED;
    const float *q4 = mixed ? c.d_sp8f.as<float>() : (const float *)nullptr;
    // targets per lane: 4 (a wave per 64 targets); FP64 with leaves above 64
    // particles 8 (a wave per 128 targets).  Measured at 25.6M particles (~98 per
    // leaf): FP64 P2P 134 ms with 8 vs 144 ms with 4; mixed 99 ms with 8 (256 VGPRs)
    // vs 92 ms with 4.
    const int tpl = (!mixed && c.shape.maxleaf > 64) ? 8 : 4;
    int bt = 64 * ((c.shape.maxleaf + 16 * tpl - 1) / (16 * tpl));
    bt = std::max(64, std::min(bt, P2P_MAXT));
    // per source one record: x, y, z (24 B), gamma and p (16 B mixed, 32 B FP64): 40 B
    // read per source and lane in mixed precision instead of 48 with a padded double4
    const size_t per = mixed ? sizeof(P2PSrc<true>) : sizeof(P2PSrc<false>);
    int tc = bt;  // a multiple of the block size, >= 1 block
    while (tc + bt <= P2P_TILE_MUL * bt &&
           (size_t)(tc + bt) * per + 16 <= P2P_TILE_LDS_WAVE * (size_t)(bt / 64))
        tc += bt;
    const size_t lds = (size_t)tc * per;
#define P2P_L(MX, TP)                                                                            \
    hipLaunchKernelGGL((k_p2p<MX, TP>), dim3(nleaf), dim3(bt), lds, st,                          \
                       c.d_leaves.as<int32_t>(), off, keys, sb, c.d_lo.as<int32_t>(),             \
                       c.d_cnt.as<int32_t>(), c.d_sp8.as<double>(), q4, c.d_near.as<double>(), tc)
    if (mixed) {
        if (tpl == 4) P2P_L(true, 4);
        else P2P_L(true, 8);
    } else {
        if (tpl == 4) P2P_L(false, 4);
        else P2P_L(false, 8);
    }
#undef P2P_L
    FMM_LAUNCH_CHECK();
}

}  // namespace fmm4rb
