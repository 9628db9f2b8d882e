// m2l.hip -- multipole-to-local transfer on gfx950 (FP64 and mixed precision).
//
// Reference: M2L! (src/fmm/interaction.jl:3-46) and gpu_M2L!
// (src/gpu_fmm/gpu_interaction.jl:3-41): for every admissible pair (t, s)
//   E_t(x) += sum_y gamma_s(y) K(x - y; p_avg),  B_t(x) += sum_y m_s(y) x K(x - y; p_avg)
// over the (n+1)^3 interpolation nodes x of t and y of s, with
//   K(R; p) = R / sqrt(R.R + (p.R)^2 + eps)^3.
//
// One workgroup owns one target cluster and walks its CSR source list; one
// thread owns one target node.  Each source's multipoles and node coordinates
// are the same for all lanes: the mixed-precision kernel stages them in LDS and
// reads broadcasts; the FP64 kernel reads them with wave-uniform addresses,
// which become scalar loads (the scalar cache instead of LDS bandwidth).  The node grid is a tensor product, so for a fixed
// source row (j, k) the y/z parts of R.R and p.R are hoisted and the i loop
// needs three operations of geometry per evaluation (two in the quadratic form); E and B are accumulated
// through the row sums G = sum y3 g, H = sum y3 dx g, A = sum y3 m,
// U = sum y3 dx m (y3 = d^-3/2), which cost 9 multiply-adds per evaluation
// instead of 12.  In the mixed-precision form the row geometry is formed in
// FP64 and rounded once to FP32, the kernel and row sums run in FP32
// (v_rsq_f32), and each source pair's FP32 sum is added into FP64
// accumulators.  (The reference stages the source in shared memory,
// gpu_interaction.jl:15-38.)  When |px| <= 0.1 (a beam moving mostly across x,
// the inner loop's axis) the mixed kernel forms d = R.R + (p.R)^2 + eps as a
// quadratic in dx, d = ((1 + px^2) dx^2 + c) + (2 px dx) q with the row's q = p.R_yz
// and c = R_yz.R_yz + q^2 + eps, and the two dx terms tabulated per lane and source
// cluster: an add and a multiply-add per node pair, both in encodings that issue two
// at a time, instead of three multiply-adds.  The cross term is at most 0.1 of the
// other two (2 |px q dx| <= |px| (dx^2 + q^2)), which are non-negative, so d stays
// within 2 FP32 ulps.

#include "fmm4rb_internal.hpp"

namespace fmm4rb {

template <int NP, bool MIXED>
struct M2LShape {
    static constexpr int NN = NP * NP * NP;
    // NP = 7, mixed precision: 343 target nodes fill 6 waves to 89%; two lanes per
    // node, each taking half of the source planes (4 and 3), fill 11 waves to 97%
    // (156.6 -> 149.3 ms at N = 2^22).  FP64 (scalar-load source, 130 VGPRs) ran 25%
    // slower that way (281 -> 353 ms) and keeps one lane per node; so did mixed NP = 6
    // (216 nodes, 4 waves at 84% -> 7 at 96%: 52.3 -> 54.2 ms, fewer resident waves).
    static constexpr int S = (NP == 7 && MIXED) ? 2 : 1;
    static constexpr int BLK = S * NN >= 1024 ? 1024 : ((S * NN + 63) / 64) * 64;
    static constexpr int TPT = (NN + BLK - 1) / BLK;
};


// One source cluster's contribution to one target node.  Per node pair: the
// geometry (3 operations on row-hoisted terms; QUAD: 2), d^-1/2, y3 and 8
// products/sums.
// Running sums over the whole source cluster: H (E_x), U_y, U_z; per plane (dz
// fixed): G, A_x, A_y; per row (dy fixed): G, A_x, A_z.  So E = (H, sum_rows dy G,
// sum_planes dz G), B = (sum dz A_y - sum dy A_z, U_z - sum dz A_x, sum dy A_x - U_y).
// The first node of every row is peeled so that no row sum needs a reset.
template <int NP, typename T, typename V4, bool QUAD = false>
__device__ inline void m2l_pair(double tx, double ty, double tz, const double *sn, const V4 *smp,
                                T px, T py, T pz, double *acc, int sk0 = 0, int sk1 = NP) {
    static_assert(!QUAD || sizeof(T) == 4, "the quadratic form is for the FP32 kernel");
    const T qa = fma(px, px, (T)1), qb = px + px;  // QUAD: (1 + px^2), 2 px
    // FP64 with n > 5: squares formed in the multiply-adds and rows not unrolled
    // (122 instead of 164 VGPRs: FP64 M2L at n = 6 took 307 ms instead of 421 ms at
    // N = 2^22); otherwise precomputed squares and unrolled rows (faster there).
    constexpr bool LEAN = sizeof(T) == 8 && NP > 6;
    constexpr int ROWU = LEAN ? 1 : NP;
    T dxr[NP], dyr[NP], dx2[LEAN ? 1 : NP], dy2[LEAN ? 1 : NP];
    // QUAD: per source cluster and lane, the node's terms of d in dx: qa dx^2 and qb dx
    T qal[QUAD ? NP : 1], qbl[QUAD ? NP : 1];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        dxr[i] = (T)(tx - sn[i]);
        dyr[i] = (T)(ty - sn[NP + i]);
        if constexpr (!LEAN) {
            dx2[i] = dxr[i] * dxr[i];
            dy2[i] = fma(dyr[i], dyr[i], (T)EPS64);
        }
        if constexpr (QUAD) {
            qal[i] = (qa * dxr[i]) * dxr[i];
            qbl[i] = qb * dxr[i];
        }
    }
    T H = 0, Uy = 0, Uz = 0, e1 = 0, e2 = 0, b0 = 0, b1 = 0, b2 = 0;
    for (int sk = sk0; sk < sk1; ++sk) {
        const T dz = (T)(tz - sn[2 * NP + sk]);
        const T dz2 = LEAN ? fma(dz, dz, (T)EPS64) : dz * dz, pzdz = pz * dz;
        T Gp = 0, Axp = 0, Ayp = 0;
#pragma unroll ROWU
        for (int sj = 0; sj < NP; ++sj) {
            const T dy = dyr[sj];
            T rrjk;
            if constexpr (LEAN) rrjk = fma(dy, dy, dz2);
            else rrjk = dy2[sj] + dz2;
            const T prjk = fma(py, dy, pzdz);
            T qc = 0;  // QUAD: the row's constant term
            if constexpr (QUAD) qc = fma(prjk, prjk, rrjk);
            const V4 *row = smp + NP * (sj + NP * sk);
            T G, Ax, Az;
#pragma unroll
            for (int si = 0; si < NP; ++si) {
                const T dx = dxr[si];
                T d;
                if constexpr (QUAD) {
                    // (qa dx^2 + qc) + (qb dx) q: an add and an accumulate, both in the
                    // two-operand encodings that gfx950 issues two at a time
                    d = fma(qbl[si], prjk, qal[si] + qc);
                } else {
                    T rr;
                    if constexpr (LEAN) rr = fma(dx, dx, rrjk);
                    else rr = dx2[si] + rrjk;
                    const T pr = fma(px, dx, prjk);
                    d = fma(pr, pr, rr);
                }
                const T y3 = rsq3<T>(d);
                const V4 m = row[si];
                const T u = y3 * dx;
                if (si == 0) {
                    G = y3 * m.x;
                    Ax = y3 * m.y;
                    Az = y3 * m.w;
                } else {
                    G = fma(y3, m.x, G);
                    Ax = fma(y3, m.y, Ax);
                    Az = fma(y3, m.w, Az);
                }
                Ayp = fma(y3, m.z, Ayp);
                H = fma(u, m.x, H);
                Uy = fma(u, m.z, Uy);
                Uz = fma(u, m.w, Uz);
            }
            e1 = fma(dy, G, e1);
            b0 = fma(-dy, Az, b0);
            b2 = fma(dy, Ax, b2);
            Gp += G;
            Axp += Ax;
        }
        e2 = fma(dz, Gp, e2);
        b0 = fma(dz, Ayp, b0);
        b1 = fma(-dz, Axp, b1);
    }
    acc[0] += (double)H;
    acc[1] += (double)e1;
    acc[2] += (double)e2;
    acc[3] += (double)b0;
    acc[4] += (double)(b1 + Uz);
    acc[5] += (double)(b2 - Uy);
}

// MIXED: the compiler is held to 80 VGPRs, 6 waves per SIMD instead of 4 (120 VGPRs):
// gfx950 issues two independent FP32 VALU instructions per cycle (two-operand
// encodings, VGPR operands, from different waves), so more resident waves pair more
// of them: 16.29 -> 15.54 ms alone at N = 2^22 with the table form of d (8 waves
// spilled and took 17.6 ms).  NP = 7 likewise (128 -> 80 VGPRs, 92 bytes spilled outside
// the inner loop): its 11-wave blocks fit twice per CU at 6 waves per SIMD and once at 4
// or 5, 136 -> 124 ms alone at N = 2^22, n = 6.  NP = 6 ran 3% slower held to 6 waves, and
// the FP64 kernel at NP = 7 held to 5 waves 12% slower.
template <int NP, bool MIXED>
constexpr int m2l_waves() { return MIXED && (NP == 5 || NP == 7) ? 6 : 1; }  // (n = 3: 39 VGPRs spilled)

template <int NP, bool MIXED, bool QUAD>
__global__ __launch_bounds__((M2LShape<NP, MIXED>::BLK))
__attribute__((amdgpu_waves_per_eu(m2l_waves<NP, MIXED>()))) void k_m2l(
    const int32_t *__restrict__ off, const uint64_t *__restrict__ keys, int sb,
    const double *__restrict__ ncoord, const void *__restrict__ mpv,
    const double *__restrict__ pavg4, double *__restrict__ loc) {
    using T = typename std::conditional<MIXED, float, double>::type;
    using V4 = typename std::conditional<MIXED, float4, double4>::type;
    using Shp = M2LShape<NP, MIXED>;
    constexpr int NN = Shp::NN, BLK = Shp::BLK, TPT = Shp::TPT, S = Shp::S;
    static_assert(S == 1 || TPT == 1, "the plane split assumes one node per lane");
    // MIXED: the source staged in LDS (16 B per node read per pair); FP64: read
    // with wave-uniform addresses through the scalar cache (32 B per node per pair
    // would take most of the LDS bandwidth)
    __shared__ V4 lmp[MIXED ? NN : 1];
    __shared__ double lsn[MIXED ? 3 * NP : 1];
    __shared__ double comb[S > 1 ? NN : 1][6];  // the second half's sums, per node
    const V4 *mp = reinterpret_cast<const V4 *>(mpv);
    const int tid = threadIdx.x;
    // lane tid: node tid % NN, planes [sk0, sk1) of part tid / NN (S = 1: all planes)
    const int part = S > 1 ? tid / NN : 0;
    const int sk0 = S > 1 ? (part == 0 ? 0 : (NP + 1) / 2) : 0;
    const int sk1 = S > 1 ? (part == 0 ? (NP + 1) / 2 : NP) : NP;
    const int lnode = S > 1 ? tid - part * NN : tid;  // this lane's node (S > 1)
    const int32_t t = xcd_item<XCD_RUN_FAR>(blockIdx.x, gridDim.x);
    const int32_t beg = off[t], end = off[t + 1];
    const uint64_t mask = (1ull << sb) - 1ull;
    const T px = (T)pavg4[0], py = (T)pavg4[1], pz = (T)pavg4[2];
    double tx[TPT], ty[TPT], tz[TPT], acc[TPT][6];
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        const int node = S > 1 ? min(lnode, NN - 1) : min(tid + q * BLK, NN - 1);
        const int i = node % NP, j = (node / NP) % NP, k = node / (NP * NP);
        tx[q] = ncoord[((int64_t)t * 3 + 0) * NP + i];
        ty[q] = ncoord[((int64_t)t * 3 + 1) * NP + j];
        tz[q] = ncoord[((int64_t)t * 3 + 2) * NP + k];
#pragma unroll
        for (int m = 0; m < 6; ++m) acc[q][m] = 0.0;
    }
    for (int32_t e = beg; e < end; ++e) {
        const int64_t s = (int64_t)(keys[e] & mask);
        const double *sn = ncoord + s * 3 * NP;
        const V4 *smp = mp + s * NN;
        if constexpr (MIXED) {
            __syncthreads();
            for (int j = tid; j < NN; j += BLK) lmp[j] = smp[j];
            if (tid < 3 * NP) lsn[tid] = sn[tid];
            __syncthreads();
            sn = lsn;
            smp = lmp;
        }
        if constexpr (S > 1) {
            if (part < S)
                m2l_pair<NP, T, V4, QUAD>(tx[0], ty[0], tz[0], sn, smp, px, py, pz, acc[0], sk0, sk1);
        } else {
#pragma unroll
            for (int q = 0; q < TPT; ++q)
                if (tid + q * BLK < NN)
                    m2l_pair<NP, T, V4, QUAD>(tx[q], ty[q], tz[q], sn, smp, px, py, pz, acc[q]);
        }
    }
    if constexpr (S > 1) {  // second half into the first, fixed order
        if (part == 1)
            for (int m = 0; m < 6; ++m) comb[lnode][m] = acc[0][m];
        __syncthreads();
        if (part == 0) {
            double *o = loc + ((int64_t)t * NN + lnode) * 6;
            for (int m = 0; m < 6; ++m) o[m] = acc[0][m] + comb[lnode][m];
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        const int node = tid + q * BLK;
        if (node < NN) {
            double *o = loc + ((int64_t)t * NN + node) * 6;
#pragma unroll
            for (int m = 0; m < 6; ++m) o[m] = acc[q][m];
        }
    }
}

void launch_m2l(Ctx &c, const int32_t *off, const uint64_t *keys, int sb, int64_t ngroups_grid) {
    hipStream_t st = c.stream;
    const bool mixed = c.precision == FMM4RB_MIXED;
    // the quadratic form in dx needs |px| <= 0.1 (p_avg is on the host since the
    // input check of the call)
    const bool quad = mixed && std::fabs(c.pavg_host[0]) <= 0.1;
    const void *mp = mixed ? (const void *)c.d_mp4f.p : (const void *)c.d_mp4.p;
    const unsigned grid = (unsigned)ngroups_grid;
#define M2L_K(NP, MX, QD)                                                                        \
    hipLaunchKernelGGL((k_m2l<NP, MX, QD>), dim3(grid), dim3(M2LShape<NP, MX>::BLK), 0, st, off, \
                       keys, sb, c.d_ncoord.as<double>(), mp, c.d_pavg.as<double>(),            \
                       c.d_loc.as<double>())
#define M2L_L(NP)                    \
    if (quad) {                      \
        M2L_K(NP, true, true);       \
    } else if (mixed) {              \
        M2L_K(NP, true, false);      \
    } else {                         \
        M2L_K(NP, false, false);     \
    }
    switch (c.np) {
        case 2: M2L_L(2); break;
        case 3: M2L_L(3); break;
        case 4: M2L_L(4); break;
        case 5: M2L_L(5); break;
        case 6: M2L_L(6); break;
        case 7: M2L_L(7); break;
        case 8: M2L_L(8); break;
        case 9: M2L_L(9); break;
        case 10: M2L_L(10); break;
        case 11: M2L_L(11); break;
        default: fail(FMM4RB_EINVAL, "n out of range");
    }
    FMM_LAUNCH_CHECK();
}

}  // namespace fmm4rb
