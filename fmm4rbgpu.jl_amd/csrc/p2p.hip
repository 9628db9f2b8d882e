// p2p.hip -- near-field particle-particle sums on gfx950 (FP64 and mixed precision).
//
// Reference: P2P! (src/fmm/interaction.jl:49-75) and gpu_P2P!
// (src/gpu_fmm/gpu_interaction.jl:43-86): for every leaf pair (t, s)
//   E_i += sum_j gamma_j K(x_i - x_j; p_j),  B_i += sum_j p_j x K(x_i - x_j; p_j)
// for i in t and j in s (the self pair included: R = 0 gives K = 0).
//
// One workgroup owns one target leaf (one thread per target particle, a
// wave64 multiple of threads), walks the leaf's CSR source list and stages
// each source leaf in LDS (x, y, z, gamma, px, py, pz) as broadcast reads.
// Results go to a tree-order near-field array that L2P adds to the far field.
// Mixed precision: R = x_i - x_j is formed in FP64 and rounded once to FP32
// (the closest pairs dominate the field of their particles, so R must keep
// its relative accuracy however small it is); the kernel then runs in FP32
// with v_rsq_f32, and each staged source tile's FP32 partial sums are added
// into FP64 accumulators.
#include "fmm4rb_internal.hpp"

namespace fmm4rb {

constexpr int P2P_MAXT = 256;

template <bool MIXED>
__global__ __launch_bounds__(P2P_MAXT) void k_p2p(
    const int32_t *__restrict__ leaves, const int32_t *__restrict__ off,
    const uint64_t *__restrict__ keys, int sb, const int32_t *__restrict__ lo,
    const int32_t *__restrict__ cnt, const double *__restrict__ sp8, const float *__restrict__ sp4f,
    double *__restrict__ nearf) {
    using T = typename std::conditional<MIXED, float, double>::type;
    using V4 = typename std::conditional<MIXED, float4, double4>::type;
    __shared__ double4 tx[P2P_MAXT];  // source x, y, z (FP64)
    __shared__ V4 tq[P2P_MAXT];       // source gamma, px, py, pz
    const int BT = blockDim.x;
    const int tid = threadIdx.x;
    const int32_t t = leaves[blockIdx.x];
    const int32_t a = lo[t], ct = cnt[t];
    const int32_t beg = off[t], end = off[t + 1];
    const uint64_t mask = (1ull << sb) - 1ull;
    const double4 *p8 = reinterpret_cast<const double4 *>(sp8);
    const float4 *q4 = reinterpret_cast<const float4 *>(sp4f);
    for (int tb = 0; tb < ct; tb += BT) {
        const int i = tb + tid;
        const bool act = i < ct;
        const double4 xi = p8[2 * (int64_t)(a + (act ? i : 0))];
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int32_t e = beg; e < end; ++e) {
            const int32_t s = (int32_t)(keys[e] & mask);
            const int32_t sa = lo[s], cs = cnt[s];
            for (int s0 = 0; s0 < cs; s0 += BT) {
                __syncthreads();
                const int j = s0 + tid;
                if (j < cs) {
                    const double4 x4 = p8[2 * (int64_t)(sa + j)];
                    tx[tid] = x4;
                    if constexpr (MIXED) {
                        tq[tid] = q4[sa + j];
                    } else {
                        const double4 p4 = p8[2 * (int64_t)(sa + j) + 1];  // px, py, pz, 0
                        tq[tid] = make_double4(x4.w, p4.x, p4.y, p4.z);
                    }
                }
                __syncthreads();
                const int m = min(BT, cs - s0);
                if (act) {
                    T e0 = 0, e1 = 0, e2 = 0, b0 = 0, b1 = 0, b2 = 0;
                    for (int jj = 0; jj < m; ++jj) {
                        const double4 xs = tx[jj];
                        const V4 q = tq[jj];  // gamma, px, py, pz
                        const T R0 = (T)(xi.x - xs.x), R1 = (T)(xi.y - xs.y), R2 = (T)(xi.z - xs.z);
                        const T rr = fma(R2, R2, fma(R1, R1, R0 * R0));
                        const T pr = fma(q.w, R2, fma(q.z, R1, q.y * R0));
                        const T d = fma(pr, pr, rr) + (T)EPS64;
                        T y;
                        if (MIXED) y = __builtin_amdgcn_rsqf((float)d);
                        else y = (T)rsqrt_nr((double)d);
                        const T y3 = (y * y) * y;
                        const T gw = q.x * y3;
                        e0 = fma(gw, R0, e0);
                        e1 = fma(gw, R1, e1);
                        e2 = fma(gw, R2, e2);
                        const T c0 = fma(q.z, R2, -(q.w * R1));
                        const T c1 = fma(q.w, R0, -(q.y * R2));
                        const T c2 = fma(q.y, R1, -(q.z * R0));
                        b0 = fma(y3, c0, b0);
                        b1 = fma(y3, c1, b1);
                        b2 = fma(y3, c2, b2);
                    }
                    acc[0] += (double)e0;
                    acc[1] += (double)e1;
                    acc[2] += (double)e2;
                    acc[3] += (double)b0;
                    acc[4] += (double)b1;
                    acc[5] += (double)b2;
                }
            }
        }
        if (act) {
            double *o = nearf + 6 * (int64_t)(a + i);
#pragma unroll
            for (int m = 0; m < 6; ++m) o[m] = acc[m];
        }
    }
}

void launch_p2p(Ctx &c, const int32_t *off, const uint64_t *keys, int sb) {
    const unsigned nleaf = (unsigned)c.shape.leaves.size();
    int bt = ((c.shape.maxleaf + 63) / 64) * 64;
    bt = std::max(64, std::min(bt, P2P_MAXT));
    if (c.precision == FMM4RB_MIXED)
        hipLaunchKernelGGL(k_p2p<true>, dim3(nleaf), dim3(bt), 0, c.stream, c.d_leaves.as<int32_t>(),
                           off, keys, sb, c.d_lo.as<int32_t>(), c.d_cnt.as<int32_t>(),
                           c.d_sp8.as<double>(), c.d_sp8f.as<float>(), c.d_near.as<double>());
    else
        hipLaunchKernelGGL(k_p2p<false>, dim3(nleaf), dim3(bt), 0, c.stream,
                           c.d_leaves.as<int32_t>(), off, keys, sb, c.d_lo.as<int32_t>(),
                           c.d_cnt.as<int32_t>(), c.d_sp8.as<double>(), (const float *)nullptr,
                           c.d_near.as<double>());
    FMM_LAUNCH_CHECK();
}

}  // namespace fmm4rb
