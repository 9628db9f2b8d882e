// interp.hip -- interpolation geometry and the P2M / M2M / L2L / L2P operators.
//
// Reference: cheb2 (src/utils/cheb2.jl:3-5), lgweight (src/fmm/upwardpass.jl:4-15)
// with the singular rule of the GPU kernels (src/gpu_fmm/gpu_upwardpass.jl:52-54:
// the last node within eps of x gets weight 1), P2M (upwardpass.jl:17-52),
// M2M (upwardpass.jl:56-99), L2L (downwardpass.jl:3-45), L2P
// (downwardpass.jl:47-80) and the final amplitude scaling
// (update_particles_field.jl:142-147).
//
// The interpolation weights factor per axis, so M2M and L2L apply three 1-D
// (n+1)x(n+1) transfers in turn (O((n+1)^4) per component) instead of the
// reference's O((n+1)^6) loops; node coordinates are computed once per
// cluster (FP64) and reused by every operator and by M2L.
#include <algorithm>

#include "fmm4rb_internal.hpp"

#pragma clang fp contract(off)

namespace fmm4rb {

struct CosTab {
    double v[MAX_NP];
};

__device__ inline double norm3g(double a, double b, double c) { return sqrt((a * a + b * b) + c * c); }

template <int NP>
__device__ inline double bary(int i) {
    double w = (i & 1) ? -1.0 : 1.0;
    if (i == 0) w = 0.5;
    if (i == NP - 1) w = 0.5 * w;
    return w;
}

// Barycentric Lagrange weights of the NP nodes at x (upwardpass.jl:4-15).
template <int NP>
__device__ inline void lgw(double x, const double *nodes, double *w) {
    int flag = -1;
#pragma unroll
    for (int i = 0; i < NP; ++i)
        if (fabs(x - nodes[i]) < EPS64) flag = i;
    if (flag >= 0) {
#pragma unroll
        for (int i = 0; i < NP; ++i) w[i] = (i == flag) ? 1.0 : 0.0;
        return;
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) w[i] = bary<NP>(i) / (x - nodes[i]);
    double sum = w[0];
#pragma unroll
    for (int i = 1; i < NP; ++i) sum = sum + w[i];
#pragma unroll
    for (int i = 0; i < NP; ++i) w[i] = w[i] / sum;
}

// 1/v in FP64: the hardware estimate v_rcp_f64 refined by two Newton steps.
__device__ inline double rcp_nr(double v) {
    double r = __builtin_amdgcn_rcp(v);
    r = fma(r, fma(-v, r, 1.0), r);
    return fma(r, fma(-v, r, 1.0), r);
}

// lgw with reciprocals instead of divisions (within a few ulps of it): the particle
// weights of P2M and L2P, which are latency-bound on the division sequences.
template <int NP>
__device__ inline void lgw_rcp(double x, const double *nodes, double *w) {
    int flag = -1;
#pragma unroll
    for (int i = 0; i < NP; ++i)
        if (fabs(x - nodes[i]) < EPS64) flag = i;
    if (flag >= 0) {
#pragma unroll
        for (int i = 0; i < NP; ++i) w[i] = (i == flag) ? 1.0 : 0.0;
        return;
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) w[i] = bary<NP>(i) * rcp_nr(x - nodes[i]);
    double sum = w[0];
#pragma unroll
    for (int i = 1; i < NP; ++i) sum = sum + w[i];
    const double inv = rcp_nr(sum);
#pragma unroll
    for (int i = 0; i < NP; ++i) w[i] = w[i] * inv;
}

// Per-node centre, admissibility radius and diameter (interaction_lists.jl:15-19,
// bbox.jl:9-11) and the Chebyshev nodes of each axis (cheb2.jl:3-5).
__global__ void k_node_geo(int64_t nc, const uint8_t *__restrict__ flags,
                           const double *__restrict__ bbox, const double *__restrict__ pavg4, int np,
                           CosTab ct, double *__restrict__ geo, double *__restrict__ ncoord) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc || !(flags[k] & 1)) return;
    const double s[3] = {1.0, 1.0, pavg4[3]};
    double mn[3], mx[3], c[3], v[3], dv[3];
    for (int d = 0; d < 3; ++d) {
        mn[d] = bbox[6 * k + d];
        mx[d] = bbox[6 * k + 3 + d];
        c[d] = (mn[d] + mx[d]) / 2.0;
        v[d] = s[d] * (mx[d] - c[d]);
        dv[d] = (s[d] * (mx[d] - mn[d])) / 2.0;
    }
    double *g = geo + 8 * k;
    g[0] = c[0];
    g[1] = c[1];
    g[2] = c[2];
    g[3] = norm3g(v[0], v[1], v[2]);
    g[4] = norm3g(dv[0], dv[1], dv[2]);
    g[5] = g[6] = g[7] = 0.0;
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < np; ++i)
            ncoord[(k * 3 + d) * np + i] = (ct.v[i] * (mx[d] - mn[d])) / 2.0 + (mn[d] + mx[d]) / 2.0;
}

// ------------------------------------------------------------------- P2M
constexpr int OP_T = 128;

// The staged particles' weights per axis, W[axis][node][slot], rows padded by one
// double so that lanes reading different nodes of one slot hit different banks.
constexpr int OP_TP = OP_T + 1;

template <int NP>
__global__ __launch_bounds__(OP_T) void k_p2m(const int32_t *__restrict__ leaves,
                                              const int32_t *__restrict__ lo,
                                              const int32_t *__restrict__ cnt,
                                              const double *__restrict__ sp8,
                                              const double *__restrict__ ncoord,
                                              double *__restrict__ mp4) {
    constexpr int NN = NP * NP * NP, TPT = (NN + OP_T - 1) / OP_T;
    __shared__ double W[OP_T][3 * NP];
    __shared__ double Q[OP_T][4];
    __shared__ double nd[3 * NP];
    const int tid = threadIdx.x;
    const int32_t k = leaves[blockIdx.x];
    const int32_t a = lo[k], c = cnt[k];
    if (tid < 3 * NP) nd[tid] = ncoord[(int64_t)k * 3 * NP + tid];
    double acc[TPT][4];
#pragma unroll
    for (int q = 0; q < TPT; ++q)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[q][m] = 0.0;
    for (int p0 = 0; p0 < c; p0 += OP_T) {
        __syncthreads();
        const int p = p0 + tid;
        if (p < c) {
            const double *x = sp8 + 8 * (int64_t)(a + p);
            lgw<NP>(x[0], nd, &W[tid][0]);
            lgw<NP>(x[1], nd + NP, &W[tid][NP]);
            lgw<NP>(x[2], nd + 2 * NP, &W[tid][2 * NP]);
            Q[tid][0] = x[3];
            Q[tid][1] = x[4];
            Q[tid][2] = x[5];
            Q[tid][3] = x[6];
        }
        __syncthreads();
        const int m = min(OP_T, c - p0);
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const int node = tid + q * OP_T;
            if (node < NN) {
                const int i = node % NP, j = (node / NP) % NP, kk = node / (NP * NP);
                for (int pp = 0; pp < m; ++pp) {
                    const double w = (W[pp][i] * W[pp][NP + j]) * W[pp][2 * NP + kk];
                    acc[q][0] += Q[pp][0] * w;
                    acc[q][1] += Q[pp][1] * w;
                    acc[q][2] += Q[pp][2] * w;
                    acc[q][3] += Q[pp][3] * w;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        const int node = tid + q * OP_T;
        if (node < NN) {
            double *o = mp4 + ((int64_t)k * NN + node) * 4;
            o[0] = acc[q][0];
            o[1] = acc[q][1];
            o[2] = acc[q][2];
            o[3] = acc[q][3];
        }
    }
}

// ------------------------------------------------------- M2M (gather)
// Components transferred together per barrier phase: the largest divisor of NC
// whose three LDS buffers of NN doubles fit in 48 KB.
template <int NP, int NC>
struct TransferGroup {
    static constexpr int NN = NP * NP * NP;
    static constexpr int FIT = (48 * 1024) / (3 * NN * 8);
    static constexpr int CG = FIT >= NC ? NC
                              : (NC % 3 == 0 && FIT >= 3) ? 3
                              : (NC % 2 == 0 && FIT >= 2) ? 2
                                                          : 1;
};

// parent = transfer(child 1) + transfer(child 2), one parent per workgroup; when all
// four components go through one transfer pass, both children's multipoles are
// loaded into registers at the start (element e = tid + OP_T q), as in k_l2l.
template <int NP>
__global__ __launch_bounds__(OP_T) void k_m2m(const int32_t *__restrict__ parents,
                                              const double *__restrict__ ncoord,
                                              double *__restrict__ mp4) {
    constexpr int NN = NP * NP * NP, TPT = (NN + OP_T - 1) / OP_T, NE = 4 * NN;
    constexpr int CG = TransferGroup<NP, 4>::CG;
    constexpr bool PRE = CG == 4 && NE <= 8 * OP_T;
    constexpr int EPT = PRE ? (NE + OP_T - 1) / OP_T : 1;
    __shared__ double Wm[3][NP][NP];  // [axis][child node][parent node]
    __shared__ double pn[3 * NP], cn[3 * NP];
    __shared__ double bin[CG][NN], ba[CG][NN], bb[CG][NN];
    const int tid = threadIdx.x;
    const int32_t P = parents[blockIdx.x];
    double vc[2][EPT];
    if constexpr (PRE) {
#pragma unroll
        for (int ch = 0; ch < 2; ++ch)
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int e = tid + OP_T * q;
                vc[ch][q] = e < NE ? mp4[(int64_t)(2 * P + 1 + ch) * NE + e] : 0.0;
            }
    }
    if (tid < 3 * NP) pn[tid] = ncoord[(int64_t)P * 3 * NP + tid];
    double acc[TPT][4];
#pragma unroll
    for (int q = 0; q < TPT; ++q)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[q][m] = 0.0;
    for (int ch = 0; ch < 2; ++ch) {
        const int32_t C = 2 * P + 1 + ch;
        __syncthreads();
        if (tid < 3 * NP) cn[tid] = ncoord[(int64_t)C * 3 * NP + tid];
        __syncthreads();
        if (tid < 3 * NP) {
            const int d = tid / NP, ci = tid % NP;
            lgw<NP>(cn[tid], pn + d * NP, &Wm[d][ci][0]);
        }
        for (int c0 = 0; c0 < 4; c0 += CG) {
            __syncthreads();
            if constexpr (PRE) {
#pragma unroll
                for (int q = 0; q < EPT; ++q) {
                    const int e = tid + OP_T * q;
                    if (e < NE) bin[e % 4][e / 4] = vc[ch][q];
                }
            } else {
                for (int e = tid; e < CG * NN; e += OP_T) {  // contiguous when CG = 4
                    const int n = e / CG, cc = e % CG;
                    bin[cc][n] = mp4[((int64_t)C * NN + n) * 4 + c0 + cc];
                }
            }
            __syncthreads();
            for (int e = tid; e < CG * NN; e += OP_T) {  // contract x: ba[Pi, cj, ck]
                const int cc = e / NN, n = e % NN, Pi = n % NP, r = n / NP;
                double s = 0.0;
                for (int ci = 0; ci < NP; ++ci) s += Wm[0][ci][Pi] * bin[cc][ci + NP * r];
                ba[cc][n] = s;
            }
            __syncthreads();
            for (int e = tid; e < CG * NN; e += OP_T) {  // contract y: bb[Pi, Pj, ck]
                const int cc = e / NN, n = e % NN;
                const int Pi = n % NP, Pj = (n / NP) % NP, ck = n / (NP * NP);
                double s = 0.0;
                for (int cj = 0; cj < NP; ++cj) s += Wm[1][cj][Pj] * ba[cc][Pi + NP * (cj + NP * ck)];
                bb[cc][n] = s;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < TPT; ++q) {  // contract z
                const int e = tid + q * OP_T;
                if (e < NN) {
                    const int ij = e % (NP * NP), Pk = e / (NP * NP);
#pragma unroll
                    for (int cc = 0; cc < CG; ++cc) {
                        double s = 0.0;
                        for (int ck = 0; ck < NP; ++ck) s += Wm[2][ck][Pk] * bb[cc][ij + NP * NP * ck];
                        acc[q][c0 + cc] += s;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        const int e = tid + q * OP_T;
        if (e < NN) {
            double *o = mp4 + ((int64_t)P * NN + e) * 4;
            o[0] = acc[q][0];
            o[1] = acc[q][1];
            o[2] = acc[q][2];
            o[3] = acc[q][3];
        }
    }
}

// ------------------------------------------------------------------ L2L
// child += transfer(parent), one child per workgroup.  The kernel waits on memory:
// when all six components go through one transfer pass, both expansions are loaded
// into registers at the start (element e = tid + OP_T q), before the node weights.
template <int NP>
__global__ __launch_bounds__(OP_T) void k_l2l(const int32_t *__restrict__ nodes,
                                              const double *__restrict__ ncoord,
                                              double *__restrict__ loc) {
    constexpr int NN = NP * NP * NP, NE = 6 * NN;
    constexpr int CG = TransferGroup<NP, 6>::CG;
    constexpr bool PRE = CG == 6 && NE <= 8 * OP_T;
    constexpr int EPT = PRE ? (NE + OP_T - 1) / OP_T : 1;
    __shared__ double Wm[3][NP][NP];  // [axis][child node][parent node]
    __shared__ double pn[3 * NP], cn[3 * NP];
    __shared__ double bin[CG][NN], ba[CG][NN], bb[CG][NN];
    const int tid = threadIdx.x;
    const int32_t C = nodes[blockIdx.x];
    const int32_t P = (C - 1) / 2;
    double *lc = loc + (int64_t)C * NE;
    const double *lp = loc + (int64_t)P * NE;
    double vl[EPT], vp[EPT];
    if constexpr (PRE) {
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + OP_T * q;
            vl[q] = e < NE ? lc[e] : 0.0;
            vp[q] = e < NE ? lp[e] : 0.0;
        }
    }
    if (tid < 3 * NP) {
        pn[tid] = ncoord[(int64_t)P * 3 * NP + tid];
        cn[tid] = ncoord[(int64_t)C * 3 * NP + tid];
    }
    __syncthreads();
    if (tid < 3 * NP) {
        const int d = tid / NP, ci = tid % NP;
        lgw<NP>(cn[tid], pn + d * NP, &Wm[d][ci][0]);
    }
    for (int c0 = 0; c0 < 6; c0 += CG) {
        if constexpr (PRE) {
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int e = tid + OP_T * q;
                if (e < NE) bin[e % 6][e / 6] = vp[q];
            }
        } else {
            __syncthreads();
            for (int e = tid; e < CG * NN; e += OP_T) {  // contiguous when CG = 6
                const int n = e / CG, cc = e % CG;
                bin[cc][n] = lp[(int64_t)n * 6 + c0 + cc];
            }
        }
        __syncthreads();
        for (int e = tid; e < CG * NN; e += OP_T) {  // ba[ci, Pj, Pk]
            const int cc = e / NN, n = e % NN, ci = n % NP, r = n / NP;
            double s = 0.0;
            for (int Pi = 0; Pi < NP; ++Pi) s += Wm[0][ci][Pi] * bin[cc][Pi + NP * r];
            ba[cc][n] = s;
        }
        __syncthreads();
        for (int e = tid; e < CG * NN; e += OP_T) {  // bb[ci, cj, Pk]
            const int cc = e / NN, n = e % NN;
            const int ci = n % NP, cj = (n / NP) % NP, Pk = n / (NP * NP);
            double s = 0.0;
            for (int Pj = 0; Pj < NP; ++Pj) s += Wm[1][cj][Pj] * ba[cc][ci + NP * (Pj + NP * Pk)];
            bb[cc][n] = s;
        }
        __syncthreads();
        if constexpr (PRE) {  // child += contract z, element e = n * 6 + cc
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int e = tid + OP_T * q;
                if (e < NE) {
                    const int n = e / 6, cc = e % 6;
                    const int ij = n % (NP * NP), ck = n / (NP * NP);
                    double s = 0.0;
                    for (int Pk = 0; Pk < NP; ++Pk) s += Wm[2][ck][Pk] * bb[cc][ij + NP * NP * Pk];
                    lc[e] = vl[q] + s;
                }
            }
        } else {
            for (int e = tid; e < CG * NN; e += OP_T) {  // child += contract z (contiguous)
                const int n = e / CG, cc = e % CG;
                const int ij = n % (NP * NP), ck = n / (NP * NP);
                double s = 0.0;
                for (int Pk = 0; Pk < NP; ++Pk) s += Wm[2][ck][Pk] * bb[cc][ij + NP * NP * Pk];
                lc[(int64_t)n * 6 + c0 + cc] += s;
            }
        }
    }
}

// ---------------------------------------------------- L2P and final sum
// One leaf per workgroup: the leaf's final local expansion is its M2L sum plus its
// parent's final expansion transferred down (the same transfer as k_l2l), kept in
// LDS for the evaluation at the leaf's particles (downwardpass.jl:47-80) and also
// stored (the reverse mode's pullback reads it); then E and B, near plus far field
// scaled by amp (update_particles_field.jl:142-147), go to the particle's original
// index.  The kernel waits on memory, not on arithmetic, so every global load of
// the first particle pass (both expansions, coordinates, near field, indices) is
// issued at the start, before the transfer.  Per particle the weights factor per
// axis: for every (j, k) row the NP values along x are contracted with wx first,
// then weighted by wy wz.  When the leaf fits half the block, two lanes share a
// particle, each taking half of the rows, and their sums meet through LDS.
template <int NP>
__global__ __launch_bounds__(OP_T) void k_leaf_l2p(const int32_t *__restrict__ leaves,
                                                   const int32_t *__restrict__ lo,
                                                   const int32_t *__restrict__ cnt,
                                                   const int32_t *__restrict__ parind,
                                                   const double *__restrict__ sp8,
                                                   const double *__restrict__ ncoord,
                                                   double *__restrict__ loc,
                                                   const double *__restrict__ nearf, double amp,
                                                   double *__restrict__ E, double *__restrict__ B) {
    constexpr int NN = NP * NP * NP, NR = NP * NP, NE = 6 * NN;
    constexpr int CG = TransferGroup<NP, 6>::CG;
    // the expansions' elements e = tid + OP_T q in registers (all six components in
    // one transfer pass)
    constexpr bool PRE = CG == 6 && NE <= 8 * OP_T;
    constexpr int EPT = PRE ? (NE + OP_T - 1) / OP_T : 1;
    constexpr int NBUF = 3 * CG * NN, NWB = 3 * NP * OP_TP + (OP_T / 2) * 6;
    __shared__ double Ls[NE];
    __shared__ double sm[NBUF > NWB ? NBUF : NWB];  // transfer buffers, then weights
    __shared__ double Wm[3][NP][NP];                 // [axis][child node][parent node]
    __shared__ double pn[3 * NP], nd[3 * NP];
    double (*bin)[NN] = reinterpret_cast<double (*)[NN]>(sm);
    double (*ba)[NN] = reinterpret_cast<double (*)[NN]>(sm + CG * NN);
    double (*bb)[NN] = reinterpret_cast<double (*)[NN]>(sm + 2 * CG * NN);
    double (*W)[NP][OP_TP] = reinterpret_cast<double (*)[NP][OP_TP]>(sm);
    double (*red)[6] = reinterpret_cast<double (*)[6]>(sm + 3 * NP * OP_TP);
    const int tid = threadIdx.x;
    const int32_t C = leaves[blockIdx.x];
    const int32_t a = lo[C], c = cnt[C];
    const int32_t P = C > 0 ? (C - 1) / 2 : 0;
    double *lc = loc + (int64_t)C * NE;
    const double *lp = loc + (int64_t)P * NE;
    const int PT = 2 * c <= OP_T ? OP_T / 2 : OP_T;  // particles per pass
    const int slot = tid % PT, part = tid / PT;
    // ---- loads of the first pass
    double vl[EPT], vp[EPT];
    if constexpr (PRE) {
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + OP_T * q;
            vl[q] = e < NE ? lc[e] : 0.0;
            vp[q] = (C > 0 && e < NE) ? lp[e] : 0.0;
        }
    }
    if (tid < 3 * NP) {
        nd[tid] = ncoord[(int64_t)C * 3 * NP + tid];
        if (C > 0) pn[tid] = ncoord[(int64_t)P * 3 * NP + tid];
    }
    // weight tasks (particle, axis) of the first pass: e = tid + OP_T u < 3 m
    const int m0 = min(PT, c);
    double xt[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int e = tid + OP_T * u, d = e / max(m0, 1);
        xt[u] = e < 3 * m0 ? sp8[8 * ((int64_t)a + e - d * m0) + d] : 0.0;
    }
    double nf[6];
    int64_t qo = 0;
    if (part == 0 && slot < m0) {
        const int64_t i = (int64_t)a + slot;
        qo = parind[i];
#pragma unroll
        for (int d = 0; d < 6; ++d) nf[d] = nearf[6 * i + d];
    }
    // ---- final local expansion of the leaf
    __syncthreads();
    if (C == 0) {  // the root is the only leaf: no parent
        for (int e = tid; e < NE; e += OP_T) Ls[e] = lc[e];
    } else {
        if (tid < 3 * NP) {
            const int d = tid / NP, ci = tid % NP;
            lgw<NP>(nd[tid], pn + d * NP, &Wm[d][ci][0]);
        }
        for (int c0 = 0; c0 < 6; c0 += CG) {
            if constexpr (PRE) {
#pragma unroll
                for (int q = 0; q < EPT; ++q) {
                    const int e = tid + OP_T * q;
                    if (e < NE) bin[e % 6][e / 6] = vp[q];
                }
            } else {
                __syncthreads();
                for (int e = tid; e < CG * NN; e += OP_T) {
                    const int n = e / CG, cc = e % CG;
                    bin[cc][n] = lp[(int64_t)n * 6 + c0 + cc];
                }
            }
            __syncthreads();
            for (int e = tid; e < CG * NN; e += OP_T) {  // ba[ci, Pj, Pk]
                const int cc = e / NN, n = e % NN, ci = n % NP, r = n / NP;
                double s = 0.0;
                for (int Pi = 0; Pi < NP; ++Pi) s += Wm[0][ci][Pi] * bin[cc][Pi + NP * r];
                ba[cc][n] = s;
            }
            __syncthreads();
            for (int e = tid; e < CG * NN; e += OP_T) {  // bb[ci, cj, Pk]
                const int cc = e / NN, n = e % NN;
                const int ci = n % NP, cj = (n / NP) % NP, Pk = n / (NP * NP);
                double s = 0.0;
                for (int Pj = 0; Pj < NP; ++Pj) s += Wm[1][cj][Pj] * ba[cc][ci + NP * (Pj + NP * Pk)];
                bb[cc][n] = s;
            }
            __syncthreads();
            if constexpr (PRE) {  // M2L sum + contract z, element e = n * 6 + cc
#pragma unroll
                for (int q = 0; q < EPT; ++q) {
                    const int e = tid + OP_T * q;
                    if (e < NE) {
                        const int n = e / 6, cc = e % 6;
                        const int ij = n % (NP * NP), ck = n / (NP * NP);
                        double s = 0.0;
                        for (int Pk = 0; Pk < NP; ++Pk) s += Wm[2][ck][Pk] * bb[cc][ij + NP * NP * Pk];
                        Ls[e] = vl[q] + s;
                        lc[e] = vl[q] + s;
                    }
                }
            } else {
                for (int e = tid; e < CG * NN; e += OP_T) {
                    const int n = e / CG, cc = e % CG;
                    const int ij = n % (NP * NP), ck = n / (NP * NP);
                    double s = 0.0;
                    for (int Pk = 0; Pk < NP; ++Pk) s += Wm[2][ck][Pk] * bb[cc][ij + NP * NP * Pk];
                    const int64_t o = (int64_t)n * 6 + c0 + cc;
                    Ls[o] = lc[o] + s;
                    lc[o] = Ls[o];
                }
                __syncthreads();  // bin is refilled by the next pass
            }
        }
    }
    // ---- L2P
    const int r0 = PT == OP_T ? 0 : (part == 0 ? 0 : (NR + 1) / 2);
    const int r1 = PT == OP_T ? NR : (part == 0 ? (NR + 1) / 2 : NR);
    for (int p0 = 0; p0 < c; p0 += PT) {
        const int m = min(PT, c - p0);
        __syncthreads();  // Ls complete; the previous pass's weights and sums read
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int e = tid + OP_T * u;
            if (e < 3 * m) {
                const int d = e / m, sl = e - d * m;
                const double x = p0 == 0 ? xt[u] : sp8[8 * ((int64_t)a + p0 + sl) + d];
                double w[NP];
                lgw_rcp<NP>(x, nd + d * NP, w);
#pragma unroll
                for (int i = 0; i < NP; ++i) W[d][i][sl] = w[i];
            }
        }
        __syncthreads();
        double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (slot < m) {
            double wx[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i) wx[i] = W[0][i][slot];
#pragma unroll 1
            for (int r = r0; r < r1; ++r) {
                const double *l = Ls + 6 * NP * r;
                double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int i = 0; i < NP; ++i)
#pragma unroll
                    for (int d = 0; d < 6; ++d) t[d] = fma(l[6 * i + d], wx[i], t[d]);
                const double wyz = W[1][r % NP][slot] * W[2][r / NP][slot];
#pragma unroll
                for (int d = 0; d < 6; ++d) f[d] = fma(wyz, t[d], f[d]);
            }
        }
        if (PT < OP_T) {  // second half of the rows into the first, fixed order
            if (part == 1 && slot < m)
#pragma unroll
                for (int d = 0; d < 6; ++d) red[slot][d] = f[d];
            __syncthreads();
            if (part == 0 && slot < m)
#pragma unroll
                for (int d = 0; d < 6; ++d) f[d] += red[slot][d];
        }
        if (part == 0 && slot < m) {
            const int64_t i = (int64_t)a + p0 + slot;
            if (p0 > 0) {
                qo = parind[i];
#pragma unroll
                for (int d = 0; d < 6; ++d) nf[d] = nearf[6 * i + d];
            }
            for (int d = 0; d < 3; ++d) {
                E[3 * qo + d] = (nf[d] + f[d]) * amp;
                B[3 * qo + d] = (nf[3 + d] + f[3 + d]) * amp;
            }
        }
    }
}

__global__ void k_mp_to_f32(int64_t n, const double *__restrict__ mp4, float *__restrict__ mp4f) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double4 v = reinterpret_cast<const double4 *>(mp4)[e];
    reinterpret_cast<float4 *>(mp4f)[e] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
}

// ------------------------------------------------------------ dispatch
#define FMM_NP_SWITCH(np, F)                 \
    switch (np) {                            \
        case 2: F(2); break;                 \
        case 3: F(3); break;                 \
        case 4: F(4); break;                 \
        case 5: F(5); break;                 \
        case 6: F(6); break;                 \
        case 7: F(7); break;                 \
        case 8: F(8); break;                 \
        case 9: F(9); break;                 \
        case 10: F(10); break;               \
        case 11: F(11); break;               \
        default: fail(FMM4RB_EINVAL, "n out of range"); \
    }

void launch_node_geo(Ctx &c) {
    CosTab ct{};
    for (int i = 0; i < c.np; ++i) ct.v[i] = cos(M_PI * (double)i / (double)c.n);  // cheb2.jl:4
    const int64_t nc = c.shape.nc;
    hipLaunchKernelGGL(k_node_geo, dim3((unsigned)((nc + 127) / 128)), dim3(128), 0, c.stream, nc,
                       c.d_flags.as<uint8_t>(), c.d_bbox.as<double>(), c.d_pavg.as<double>(), c.np,
                       ct, c.d_geo.as<double>(), c.d_ncoord.as<double>());
    FMM_LAUNCH_CHECK();
}

void launch_upward(Ctx &c, hipStream_t st) {
    const Shape &S = c.shape;
    const unsigned nleaf = (unsigned)S.leaves.size();
#define P2M_L(NP)                                                                            \
    hipLaunchKernelGGL(k_p2m<NP>, dim3(nleaf), dim3(OP_T), 0, st, c.d_leaves.as<int32_t>(), \
                       c.d_lo.as<int32_t>(), c.d_cnt.as<int32_t>(), c.d_sp8.as<double>(),    \
                       c.d_ncoord.as<double>(), c.d_mp4.as<double>())
    FMM_NP_SWITCH(c.np, P2M_L);
    FMM_LAUNCH_CHECK();
    // M2M: parents on level l-1 for l = ml .. 1
    for (int l = S.ml; l >= 1; --l) {
        const int64_t b = S.split_off[l - 1], e = S.split_off[l];
        if (e == b) continue;
        const int32_t *par = c.d_lvl_split.as<int32_t>() + b;
#define M2M_L(NP)                                                                          \
    hipLaunchKernelGGL(k_m2m<NP>, dim3((unsigned)(e - b)), dim3(OP_T), 0, st, par,        \
                       c.d_ncoord.as<double>(), c.d_mp4.as<double>())
        FMM_NP_SWITCH(c.np, M2M_L);
        FMM_LAUNCH_CHECK();
    }
}

void launch_mp_to_f32(Ctx &c, hipStream_t st) {
    const int64_t n = c.shape.nc * (int64_t)c.nn;
    hipLaunchKernelGGL(k_mp_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                       c.d_mp4.as<double>(), c.d_mp4f.as<float>());
    FMM_LAUNCH_CHECK();
}

void launch_l2l(Ctx &c, bool leaves) {
    const Shape &S = c.shape;
    hipStream_t st = c.stream;
    for (int l = 1; l <= S.ml; ++l) {
        // every node of level l, or only its split nodes (the leaves' transfer is
        // then part of k_leaf_l2p)
        const int64_t b = leaves ? S.lvl_off[l] : S.split_off[l];
        const int64_t e = leaves ? S.lvl_off[l + 1] : S.split_off[l + 1];
        if (e == b) continue;
        const int32_t *nodes = (leaves ? c.d_lvl_nodes : c.d_lvl_split).as<int32_t>() + b;
#define L2L_L(NP)                                                                          \
    hipLaunchKernelGGL(k_l2l<NP>, dim3((unsigned)(e - b)), dim3(OP_T), 0, st, nodes,      \
                       c.d_ncoord.as<double>(), c.d_loc.as<double>())
        FMM_NP_SWITCH(c.np, L2L_L);
        FMM_LAUNCH_CHECK();
    }
}

void launch_leaf_l2p(Ctx &c, double amp, double *E, double *B) {
    const unsigned nleaf = (unsigned)c.shape.leaves.size();
    hipStream_t st = c.stream;
#define L2P_L(NP)                                                                             \
    hipLaunchKernelGGL(k_leaf_l2p<NP>, dim3(nleaf), dim3(OP_T), 0, st,                       \
                       c.d_leaves.as<int32_t>(), c.d_lo.as<int32_t>(), c.d_cnt.as<int32_t>(), \
                       c.d_parind.as<int32_t>(), c.d_sp8.as<double>(), c.d_ncoord.as<double>(), \
                       c.d_loc.as<double>(), c.d_near.as<double>(), amp, E, B)
    FMM_NP_SWITCH(c.np, L2P_L);
    FMM_LAUNCH_CHECK();
}

void launch_downward(Ctx &c, double amp, double *E, double *B) {
    launch_l2l(c, false);
    launch_leaf_l2p(c, amp, E, B);
}

}  // namespace fmm4rb
