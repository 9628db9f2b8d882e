// p2p.hip -- near-field particle-particle sums on gfx950 (FP64 and mixed precision).
//
// Reference: P2P! (src/fmm/interaction.jl:49-75) and gpu_P2P!
// (src/gpu_fmm/gpu_interaction.jl:43-86): for every leaf pair (t, s)
//   E_i += sum_j gamma_j K(x_i - x_j; p_j),  B_i += sum_j p_j x K(x_i - x_j; p_j)
// for i in t and j in s (the self pair included: R = 0 gives K = 0).
//
// One workgroup owns one target leaf, walks the leaf's CSR source list and
// stages each source leaf in LDS (x, y, z in FP64; gamma, px, py, pz).  A wave
// covers 64 targets as (64 / NG) target lanes x NG source groups: lane (ti, g)
// holds the NG targets ti + (64 / NG) q and takes the staged sources
// j = g (mod NG), so one LDS read serves NG pairs (with NG = 1, a broadcast per
// target moves 40 bytes of LDS data per pair and the kernel is bound by LDS
// bandwidth).  The groups' partial sums are combined with xor-shuffles at the
// end, in a fixed order.
// Results go to a tree-order near-field array that L2P adds to the far field.
// Mixed precision: R = x_i - x_j is formed in FP64 and rounded once to FP32
// (the closest pairs dominate the field of their particles, so R must keep
// its relative accuracy however small it is); the kernel then runs in FP32
// with v_rsq_f32, and each staged tile's FP32 partial sums are added into FP64
// accumulators.
#include "fmm4rb_internal.hpp"

namespace fmm4rb {

constexpr int P2P_MAXT = 256;  // threads per block: up to 4 waves, 256 targets per pass
constexpr int P2P_NG = 4;      // source groups per wave (1 and 2 measured slower at N = 2^22)

template <typename T>
__device__ inline void p2p_pair(const double *xi, double sx, double sy, double sz, T g, T px,
                                T py, T pz, T *a) {
    const T R0 = (T)(xi[0] - sx), R1 = (T)(xi[1] - sy), R2 = (T)(xi[2] - sz);
    const T rr = fma(R2, R2, fma(R1, R1, R0 * R0));
    const T pr = fma(pz, R2, fma(py, R1, px * R0));
    const T d = fma(pr, pr, rr) + (T)EPS64;
    T y;
    if constexpr (sizeof(T) == 4) y = __builtin_amdgcn_rsqf(d);
    else y = rsqrt_nr(d);
    const T y3 = (y * y) * y;
    const T gw = g * y3;
    a[0] = fma(gw, R0, a[0]);
    a[1] = fma(gw, R1, a[1]);
    a[2] = fma(gw, R2, a[2]);
    const T c0 = fma(py, R2, -(pz * R1));
    const T c1 = fma(pz, R0, -(px * R2));
    const T c2 = fma(px, R1, -(py * R0));
    a[3] = fma(y3, c0, a[3]);
    a[4] = fma(y3, c1, a[4]);
    a[5] = fma(y3, c2, a[5]);
}

template <bool MIXED, int NG>
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
    constexpr int TL = 64 / NG;  // target lanes per wave
    const int lane = tid & 63, ti = lane % TL, g = lane / TL;
    const int wbase = tid - lane;  // first target of this wave within a pass
    const int32_t t = leaves[xcd_item(blockIdx.x, gridDim.x)];
    const int32_t a = lo[t], ct = cnt[t];
    const int32_t beg = off[t], end = off[t + 1];
    const uint64_t mask = (1ull << sb) - 1ull;
    const double4 *p8 = reinterpret_cast<const double4 *>(sp8);
    const float4 *q4 = reinterpret_cast<const float4 *>(sp4f);
    for (int tb = 0; tb < ct; tb += BT) {
        double xi[NG][3];
        double acc[NG][6];
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const int i = min(tb + wbase + ti + TL * q, ct - 1);
            const double4 x4 = p8[2 * (int64_t)(a + i)];
            xi[q][0] = x4.x;
            xi[q][1] = x4.y;
            xi[q][2] = x4.z;
#pragma unroll
            for (int m = 0; m < 6; ++m) acc[q][m] = 0.0;
        }
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
                T part[NG][6];
#pragma unroll
                for (int q = 0; q < NG; ++q)
#pragma unroll
                    for (int k = 0; k < 6; ++k) part[q][k] = 0;
                for (int jj = g; jj < m; jj += NG) {
                    const double4 xs = tx[jj];
                    const V4 sq = tq[jj];  // gamma, px, py, pz
#pragma unroll
                    for (int q = 0; q < NG; ++q)
                        p2p_pair<T>(xi[q], xs.x, xs.y, xs.z, sq.x, sq.y, sq.z, sq.w, part[q]);
                }
#pragma unroll
                for (int q = 0; q < NG; ++q)
#pragma unroll
                    for (int k = 0; k < 6; ++k) acc[q][k] += (double)part[q][k];
            }
        }
        // Sum the source groups (lanes ti + TL g), fixed order.
#pragma unroll
        for (int q = 0; q < NG; ++q)
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                if (NG >= 4) acc[q][k] += __shfl_xor(acc[q][k], 16);
                if (NG >= 2) acc[q][k] += __shfl_xor(acc[q][k], 32);
            }
        // Lane (ti, g) writes its target ti + TL g.
        double out[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            out[k] = acc[0][k];
#pragma unroll
            for (int q = 1; q < NG; ++q)
                if (g == q) out[k] = acc[q][k];
        }
        const int i = tb + wbase + ti + TL * g;
        if (i < ct) {
            double *o = nearf + 6 * (int64_t)(a + i);
#pragma unroll
            for (int k = 0; k < 6; ++k) o[k] = out[k];
        }
    }
}

void launch_p2p(Ctx &c, const int32_t *off, const uint64_t *keys, int sb) {
    const unsigned nleaf = (unsigned)c.shape.leaves.size();
    // one wave per 64 targets of the largest leaf, at most P2P_MAXT threads
    int bt = ((c.shape.maxleaf + 63) / 64) * 64;
    bt = std::max(64, std::min(bt, P2P_MAXT));
    const bool mixed = c.precision == FMM4RB_MIXED;
    const float *q4 = mixed ? c.d_sp8f.as<float>() : (const float *)nullptr;
#define P2P_L(MX)                                                                                \
    hipLaunchKernelGGL((k_p2p<MX, P2P_NG>), dim3(nleaf), dim3(bt), 0, c.stream,                  \
                       c.d_leaves.as<int32_t>(), off, keys, sb, c.d_lo.as<int32_t>(),             \
                       c.d_cnt.as<int32_t>(), c.d_sp8.as<double>(), q4, c.d_near.as<double>())
    if (mixed) P2P_L(true);
    else P2P_L(false);
#undef P2P_L
    FMM_LAUNCH_CHECK();
}

}  // namespace fmm4rb
