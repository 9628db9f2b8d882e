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
// Mixed precision: positions are stored relative to their own leaf centre
// in FP32; per leaf pair the centre offset is formed in FP64 and rounded once,
// so R = (u_i + (c_t - c_s)) - v_j keeps its relative accuracy for close
// pairs.  The kernel runs in FP32 with v_rsq_f32; each staged source tile's
// FP32 partial sums are added into FP64 accumulators.
#include "fmm4rb_internal.hpp"

namespace fmm4rb {

constexpr int P2P_MAXT = 256;

template <bool MIXED>
__global__ __launch_bounds__(P2P_MAXT) void k_p2p(
    const int32_t *__restrict__ leaves, const int32_t *__restrict__ off,
    const uint64_t *__restrict__ keys, int sb, const int32_t *__restrict__ lo,
    const int32_t *__restrict__ cnt, const double *__restrict__ sp8, const float *__restrict__ sp8f,
    const double *__restrict__ geo, double *__restrict__ nearf) {
    using T = typename std::conditional<MIXED, float, double>::type;
    using V4 = typename std::conditional<MIXED, float4, double4>::type;
    __shared__ V4 tile[P2P_MAXT][2];
    const int BT = blockDim.x;
    const int tid = threadIdx.x;
    const int32_t t = leaves[blockIdx.x];
    const int32_t a = lo[t], ct = cnt[t];
    const int32_t beg = off[t], end = off[t + 1];
    const uint64_t mask = (1ull << sb) - 1ull;
    const V4 *src = MIXED ? reinterpret_cast<const V4 *>(sp8f) : reinterpret_cast<const V4 *>(sp8);
    for (int tb = 0; tb < ct; tb += BT) {
        const int i = tb + tid;
        const bool act = i < ct;
        V4 xi = src[2 * (int64_t)(a + (act ? i : 0))];
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int32_t e = beg; e < end; ++e) {
            const int32_t s = (int32_t)(keys[e] & mask);
            const int32_t sa = lo[s], cs = cnt[s];
            T w0 = xi.x, w1 = xi.y, w2 = xi.z;
            if (MIXED) {  // u_i + (c_t - c_s), offset formed in FP64
                w0 += (T)(geo[8 * (int64_t)t + 0] - geo[8 * (int64_t)s + 0]);
                w1 += (T)(geo[8 * (int64_t)t + 1] - geo[8 * (int64_t)s + 1]);
                w2 += (T)(geo[8 * (int64_t)t + 2] - geo[8 * (int64_t)s + 2]);
            }
            for (int s0 = 0; s0 < cs; s0 += BT) {
                __syncthreads();
                const int j = s0 + tid;
                if (j < cs) {
                    tile[tid][0] = src[2 * (int64_t)(sa + j)];
                    tile[tid][1] = src[2 * (int64_t)(sa + j) + 1];
                }
                __syncthreads();
                const int m = min(BT, cs - s0);
                if (act) {
                    T e0 = 0, e1 = 0, e2 = 0, b0 = 0, b1 = 0, b2 = 0;
                    for (int jj = 0; jj < m; ++jj) {
                        const V4 xs = tile[jj][0], ps = tile[jj][1];
                        const T R0 = w0 - xs.x, R1 = w1 - xs.y, R2 = w2 - xs.z;
                        const T rr = fma(R2, R2, fma(R1, R1, R0 * R0));
                        const T pr = fma(ps.z, R2, fma(ps.y, R1, ps.x * R0));
                        const T d = fma(pr, pr, rr) + (T)EPS64;
                        T y;
                        if (MIXED) y = __builtin_amdgcn_rsqf((float)d);
                        else y = (T)rsqrt((double)d);
                        const T y3 = (y * y) * y;
                        const T gw = xs.w * y3;
                        e0 = fma(gw, R0, e0);
                        e1 = fma(gw, R1, e1);
                        e2 = fma(gw, R2, e2);
                        const T c0 = fma(ps.y, R2, -(ps.z * R1));
                        const T c1 = fma(ps.z, R0, -(ps.x * R2));
                        const T c2 = fma(ps.x, R1, -(ps.y * R0));
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
                           c.d_sp8.as<double>(), c.d_sp8f.as<float>(), c.d_geo.as<double>(),
                           c.d_near.as<double>());
    else
        hipLaunchKernelGGL(k_p2p<false>, dim3(nleaf), dim3(bt), 0, c.stream,
                           c.d_leaves.as<int32_t>(), off, keys, sb, c.d_lo.as<int32_t>(),
                           c.d_cnt.as<int32_t>(), c.d_sp8.as<double>(), (const float *)nullptr,
                           c.d_geo.as<double>(), c.d_near.as<double>());
    FMM_LAUNCH_CHECK();
}

}  // namespace fmm4rb
