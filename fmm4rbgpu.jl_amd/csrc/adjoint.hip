// adjoint.hip -- hand-written reverse mode (vector-Jacobian product) of the FMM
// field map (x, p) -> (E, B) at fixed tree topology and interaction lists.
// There is no counterpart in the reference; it differentiates the forward
// operators restated from src/fmm/*.jl and src/gpu_fmm/*.jl.
//
// Forward quantities (unscaled, amp = 2.8179403699772166e-15 q / lambda applied last):
//   K(R; p) = R d^-3/2,  d = R.R + (p.R)^2 + eps
//   near:  E_i += g_j K(x_i - x_j; p_j),  B_i += p_j x K
//   far:   multipoles M (P2M, M2M), locals L (M2L with p_avg, L2L), L2P
// Cotangents o_i = amp (Ebar_i, Bbar_i).  For one kernel term with source
// "charge" (g, m) and cotangent (e, b):  u = g e + b x m, and
//   dl/dR = J^T u = d^-3/2 u - 3 d^-5/2 (R.u) (R + (p.R) p)
//   dl/dp (through d) = -3 d^-5/2 (p.R) (R.u) R
//   dl/dg = e.K,  dl/dm = K x b.
// Interpolation weights w_a(x; X) (barycentric, last-match singular rule)
// depend on the point x and on the cluster's nodes X_m = (c_m (b-a))/2 + (a+b)/2,
// hence on the cluster's bounding box, hence on the particles that attain its
// minimum and maximum.  For l = sum_a c_a w_a:
//   dl/dx   = sum_a c_a w'_a,   w'_a = w_a (sum_k w_k/(x-X_k) - 1/(x-X_a))
//   dl/dX_m = w_m/(x-X_m) (c_m - sum_a c_a w_a)
// and at a node X_f (singular case): w'_a = (bw_a/bw_f)/(X_f-X_a) (a != f),
// w'_f = -sum_{a!=f} w'_a, dl/dX_f = -dl/dx, dl/dX_m = 0 otherwise.
// Order (reverse of the forward): L2P^T, L2L^T (bottom-up), M2L^T (target side
// and source side over transposed lists), M2M^T (top-down), P2M^T, P2P^T (target
// side and source side), then node-coordinate gradients -> bounding boxes ->
// the extreme particles, and the mean-momentum gradient to every particle.
// Every accumulation has one owner per output (no atomics): results are
// bit-reproducible.  The self pair (i = j) contributes nothing (K(x_i - x_i) = 0
// for every x_i); coincident distinct particles (R = 0, i != j) contribute the
// derivative of the regularised kernel there, d^-3/2(eps) u.
#include "fmm4rb_internal.hpp"

namespace fmm4rb {

template <int NP>
__device__ inline double abary(int i) {
    double w = (i & 1) ? -1.0 : 1.0;
    if (i == 0) w = 0.5;
    if (i == NP - 1) w = 0.5 * w;
    return w;
}

// Weights w and derivatives dw = dw/dx of the NP nodes X at x, and ri = 1/(x - X)
// (0 in the singular case); returns the singular node (last within eps) or -1.
template <int NP>
__device__ inline int lgw_d(double x, const double *X, double *w, double *dw, double *ri) {
    int flag = -1;
#pragma unroll
    for (int i = 0; i < NP; ++i)
        if (fabs(x - X[i]) < EPS64) flag = i;
    if (flag >= 0) {
        double s = 0.0;
        const double bf = abary<NP>(flag);
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            w[i] = (i == flag) ? 1.0 : 0.0;
            dw[i] = 0.0;
            ri[i] = 0.0;
            if (i != flag) {
                dw[i] = (abary<NP>(i) / bf) / (X[flag] - X[i]);
                s += dw[i];
            }
        }
        dw[flag] = -s;
        return flag;
    }
    double S = 0.0;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        ri[i] = 1.0 / (x - X[i]);
        w[i] = abary<NP>(i) * ri[i];  // abary = +-1 or +-1/2: exact
        S += w[i];
    }
    const double iS = 1.0 / S;
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        w[i] *= iS;
        t = fma(w[i], ri[i], t);
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) dw[i] = w[i] * (t - ri[i]);
    return -1;
}

// Given l = sum_a c_a w_a(x; X): returns dl/dx and adds dl/dX_m to Xb[m].
template <int NP>
__device__ inline double axis_adj(const double *w, const double *dw, const double *ri, int flag,
                                  const double *c, double *Xb) {
    double xb = 0.0, cw = 0.0;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        xb = fma(c[i], dw[i], xb);
        cw = fma(c[i], w[i], cw);
    }
    if (flag >= 0) {
#pragma unroll
        for (int m = 0; m < NP; ++m)
            if (m == flag) Xb[m] -= xb;
    } else {
#pragma unroll
        for (int m = 0; m < NP; ++m) Xb[m] = fma(w[m] * ri[m], c[m] - cw, Xb[m]);
    }
    return xb;
}

constexpr int AT = 128;  // threads of the leaf / transfer adjoint kernels

// Sum per-thread partials red[T][W] of a T-thread block into out[W] (+=), fixed order.
template <int W, int T = AT>
__device__ inline void block_sum_into(double (*red)[W], double *out) {
    __syncthreads();
    for (int j = threadIdx.x; j < W; j += T) {
        double s = 0.0;
        for (int t = 0; t < T; ++t) s += red[t][j];
        out[j] += s;
    }
    __syncthreads();
}

// ------------------------------------------------------------- leaf adjoints
// MODE 0 (L2P^T): A = forward locals (6 comps), o_i = cotangents ob[i];
//   Lbar[leaf][a] = sum_i o_i W_i[a];  gx_i += d/dx sum_a (o_i.A[a]) W_i[a]; Xbar[leaf] += d/dX.
// MODE 1 (P2M^T): A = Mbar (4 comps), o_i = (g_i, p_i);
//   qbar_i = sum_a A[a] W_i[a] -> gp_i += qbar_p + qbar_g p_i/g_i;  gx_i, Xbar as above.
// PART bit 1: the gradient (geometry) terms; bit 2: the Lbar values (MODE 0).
// LT threads: the geometry takes one thread per particle, so leaves of at most 64
// particles are launched with one wave (LT = 64) and larger ones with two.
template <int NP, int MODE, int PART, int LT = AT>
__global__ __launch_bounds__(LT) void k_adj_leaf(const int32_t *__restrict__ leaves,
                                                 const int32_t *__restrict__ lo,
                                                 const int32_t *__restrict__ cnt,
                                                 const double *__restrict__ sp8,
                                                 const double *__restrict__ ncoord,
                                                 const double *__restrict__ A,
                                                 const double *__restrict__ ob,
                                                 double *__restrict__ Abar, double *__restrict__ gx,
                                                 double *__restrict__ gp,
                                                 double *__restrict__ Xbar) {
    constexpr int NN = NP * NP * NP, NC = MODE == 0 ? 6 : 4, TPT = (NN + LT - 1) / LT;
    // the leaf's NN x NC coefficients are read by every particle: staged in LDS (n <= 5)
    constexpr bool STAGE = NN * NC * 8 <= 8192;
    __shared__ double nd[3 * NP];
    __shared__ double red[LT][3 * NP];  // Xbar partials, then the weight tiles
    __shared__ double Ot[LT][6];
    __shared__ double sA[STAGE ? NN * NC : 1];
    double(*Wt)[3 * NP] = red;
    const int tid = threadIdx.x;
    const int32_t k = leaves[blockIdx.x];
    const int32_t a = lo[k], c = cnt[k];
    if (tid < 3 * NP) nd[tid] = ncoord[(int64_t)k * 3 * NP + tid];
    double xbl[3 * NP];
#pragma unroll
    for (int j = 0; j < 3 * NP; ++j) xbl[j] = 0.0;
    const double *Ak = A + (int64_t)k * NN * NC;
    if (STAGE && (PART & 1)) {
        for (int j = tid; j < NN * NC; j += LT) sA[j] = Ak[j];
        Ak = sA;
    }
    __syncthreads();
    // per-particle gradient terms
    for (int p = tid; (PART & 1) && p < c; p += LT) {
        const int64_t i = a + p;
        const double *x = sp8 + 8 * i;
        // the weights only: their derivatives and 1/(x - X) are not kept across the
        // node loop (registers, and so resident waves) but formed again below, with
        // the same operations
        double w[3][NP];
        int fl[3];
        for (int d = 0; d < 3; ++d) {
            double dw[NP], ri[NP];
            fl[d] = lgw_d<NP>(x[d], nd + d * NP, w[d], dw, ri);
        }
        double o[NC];
        if (MODE == 0) {
#pragma unroll
            for (int m = 0; m < NC; ++m) o[m] = ob[6 * i + m];
        } else {
            o[0] = x[3];
            o[1] = x[4];
            o[2] = x[5];
            o[3] = x[6];
        }
        double T[3][NP];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int m = 0; m < NP; ++m) T[d][m] = 0.0;
        double qb[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) qb[m] = 0.0;
        // T[d][m] = sum over the nodes with index m on axis d of s W / w_d[m], s = o.A[node]:
        // T0 per node, T1 and T2 from the row sums R1 = sum_ix s w0[ix]; qb per node
        auto row = [&](int kz, int jy) {
            const double wyz = w[1][jy] * w[2][kz];
            double R1 = 0.0;
#pragma unroll
            for (int ix = 0; ix < NP; ++ix) {
                const double *Av = Ak + (ix + NP * (jy + NP * kz)) * NC;
                double s = o[0] * Av[0];
#pragma unroll
                for (int m = 1; m < NC; ++m) s = fma(o[m], Av[m], s);
                T[0][ix] = fma(s, wyz, T[0][ix]);
                R1 = fma(s, w[0][ix], R1);
                if (MODE == 1) {
                    const double W3 = w[0][ix] * wyz;
#pragma unroll
                    for (int m = 0; m < NC; ++m) qb[m] = fma(Av[m], W3, qb[m]);
                }
            }
            T[1][jy] = fma(R1, w[2][kz], T[1][jy]);
            T[2][kz] = fma(R1, w[1][jy], T[2][kz]);
        };
        if constexpr (MODE == 0) {  // rolled: 6 components unrolled over the rows spill
#pragma unroll 1
            for (int kz = 0; kz < NP; ++kz)
#pragma unroll 1
                for (int jy = 0; jy < NP; ++jy) row(kz, jy);
        } else {  // planes rolled, rows unrolled
#pragma unroll 1
            for (int kz = 0; kz < NP; ++kz)
#pragma unroll
                for (int jy = 0; jy < NP; ++jy) row(kz, jy);
        }
        double g[3];
        for (int d = 0; d < 3; ++d) {
            double wd[NP], dw[NP], ri[NP];
            lgw_d<NP>(x[d], nd + d * NP, wd, dw, ri);
            g[d] = axis_adj<NP>(w[d], dw, ri, fl[d], T[d], xbl + d * NP);
        }
        for (int d = 0; d < 3; ++d) gx[3 * i + d] += g[d];
        if (MODE == 1) {
            const double gi = qb[0] / x[3];
            for (int d = 0; d < 3; ++d) gp[3 * i + d] += qb[1 + d] + gi * x[4 + d];
        }
    }
    if (PART & 1) {
#pragma unroll
        for (int j = 0; j < 3 * NP; ++j) red[tid][j] = xbl[j];
        block_sum_into<3 * NP, LT>(red, Xbar + (int64_t)k * 3 * NP);
    }
    if (MODE == 0 && (PART & 2)) {  // Lbar[leaf][a] = sum_i o_i W_i[a]
        double acc[TPT][6];
#pragma unroll
        for (int q = 0; q < TPT; ++q)
#pragma unroll
            for (int m = 0; m < 6; ++m) acc[q][m] = 0.0;
        for (int p0 = 0; p0 < c; p0 += LT) {
            __syncthreads();
            const int p = p0 + tid;
            if (p < c) {
                const int64_t i = a + p;
                const double *x = sp8 + 8 * i;
                double dw[NP], ri[NP];
                for (int d = 0; d < 3; ++d) lgw_d<NP>(x[d], nd + d * NP, &Wt[tid][d * NP], dw, ri);
                for (int m = 0; m < 6; ++m) Ot[tid][m] = ob[6 * i + m];
            }
            __syncthreads();
            const int mm = min(LT, c - p0);
#pragma unroll
            for (int q = 0; q < TPT; ++q) {
                const int node = tid + q * LT;
                if (node < NN) {
                    const int ix = node % NP, jy = (node / NP) % NP, kz = node / (NP * NP);
                    for (int pp = 0; pp < mm; ++pp) {
                        const double W3 = (Wt[pp][ix] * Wt[pp][NP + jy]) * Wt[pp][2 * NP + kz];
#pragma unroll
                        for (int m = 0; m < 6; ++m) acc[q][m] = fma(Ot[pp][m], W3, acc[q][m]);
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const int node = tid + q * LT;
            if (node < NN)
#pragma unroll
                for (int m = 0; m < 6; ++m) Abar[((int64_t)k * NN + node) * 6 + m] = acc[q][m];
        }
    }
}

// One contraction pass of a transfer kernel: dst[cc][n] = term(cc, n) for the E
// elements (cc, n) of the block, AT threads.  Every element of the thread is
// evaluated before any is stored: the source and destination tiles share one LDS
// array, so interleaved stores would hold back the next element's loads.
template <int NN, int E, class F>
__device__ inline void lds_pass(double (*dst)[NN], F term) {
    constexpr int EPT = (E + AT - 1) / AT;
    double sv[EPT];
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
        const int e = (int)threadIdx.x + r * AT;
        sv[r] = e < E ? term(e / NN, e % NN) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
        const int e = (int)threadIdx.x + r * AT;
        if (e < E) dst[e / NN][e % NN] = sv[r];
    }
}

// ------------------------------------------------------- transfer adjoints
// Per parent P with children c = 2P+1, 2P+2; Om_c(a,b) = prod_d w_{P,d,b_d}(X_{c,d,a_d}).
// MODE 0 (L2L^T): Lbar_P[b] = sum_c sum_a Om_c(a,b) Lbar_c[a]; geometry with
//                 U_c = Lbar_c (child side), V = L_P (forward parent locals), 6 comps.
// MODE 1 (M2M^T): Mbar_c[a] += sum_b Om_c(a,b) Mbar_P[b]; geometry with
//                 U_c = M_c (forward child multipoles), V = Mbar_P, 4 comps.
// Geometry: l = sum_c sum_{a,b} (U_c[a].V[b]) Om_c(a,b); dl/dOm_x[ax][bx] =
// sum_{ay,az} U[a] . (sum_{by,bz} V[b] Om_y[ay][by] Om_z[az][bz]), likewise y, z.
// PART bit 1: geometry (Xbar); bit 2: the value propagation.  parity >= 0: the
// launch covers the parents of every level and this one handles those on levels
// of that parity only (the geometry needs no level order; within one parity no
// cluster is both a parent and a child, so the Xbar updates do not overlap).
template <int NP, int MODE, int PART>
__global__ __launch_bounds__(AT) void k_adj_transfer(const int32_t *__restrict__ parents,
                                                     const double *__restrict__ ncoord,
                                                     const double *__restrict__ Uarr,
                                                     const double *__restrict__ Varr,
                                                     double *__restrict__ out,
                                                     double *__restrict__ Xbar, int parity) {
    constexpr int NN = NP * NP * NP, NC = MODE == 0 ? 6 : 4, TPT = (NN + AT - 1) / AT;
    // components handled together per pass (LDS: 4 arrays of CG x NN doubles <= 48 KB)
    constexpr int CG0 = (48 * 1024) / (4 * NN * 8);
    constexpr int CG = CG0 >= NC ? NC : (CG0 >= 2 ? 2 : 1);
    static_assert(NC % CG == 0, "component groups must tile the components");
    __shared__ double Om[3][NP][NP], dOm[3][NP][NP];  // [axis][child node a][parent node b]
    __shared__ int fl[3][NP];
    __shared__ double pn[3 * NP], cn[3 * NP];
    // staging and contraction buffers; the per-thread partials (red) reuse them
    // once a child's contractions are done (a barrier separates the two uses)
    constexpr int BUFW = 4 * CG * NN, REDW = AT * 3 * NP;
    __shared__ double ubuf[BUFW > REDW ? BUFW : REDW];
    double(*bU)[NN] = reinterpret_cast<double(*)[NN]>(ubuf);
    double(*bV)[NN] = reinterpret_cast<double(*)[NN]>(ubuf + CG * NN);
    double(*t1)[NN] = reinterpret_cast<double(*)[NN]>(ubuf + 2 * CG * NN);
    double(*t2)[NN] = reinterpret_cast<double(*)[NN]>(ubuf + 3 * CG * NN);
    double(*red)[3 * NP] = reinterpret_cast<double(*)[3 * NP]>(ubuf);
    const int tid = threadIdx.x;
    const int32_t P = parents[blockIdx.x];
    if (parity >= 0 && ((31 - __clz(P + 1)) & 1) != parity) return;  // heap level of P
    if (tid < 3 * NP) pn[tid] = ncoord[(int64_t)P * 3 * NP + tid];
    // dl/dX_P terms per child: xPs[ch][d NP + a][m] from child node a on axis d
    __shared__ double xPs[2][3 * NP][NP];
    double accP[TPT][NC];  // MODE 0: Lbar_P
#pragma unroll
    for (int q = 0; q < TPT; ++q)
#pragma unroll
        for (int m = 0; m < NC; ++m) accP[q][m] = 0.0;
    for (int ch = 0; ch < 2; ++ch) {
        const int32_t C = 2 * P + 1 + ch;
        __syncthreads();
        if (tid < 3 * NP) cn[tid] = ncoord[(int64_t)C * 3 * NP + tid];
        __syncthreads();
        if (tid < 3 * NP) {
            const int d = tid / NP, ai = tid % NP;
            double ri[NP];
            fl[d][ai] = lgw_d<NP>(cn[tid], pn + d * NP, &Om[d][ai][0], &dOm[d][ai][0], ri);
        }
        // dl/dOm of the (a_d, b_d) pair go, summed by GO lanes (gl = 0 .. GO-1) that
        // take every GO-th term and meet in xor shuffles; lane gl = 0 keeps it
        constexpr int GO = NP * NP * 4 <= AT ? 4 : (NP * NP * 2 <= AT ? 2 : 1);
        double gO[3] = {0.0, 0.0, 0.0};
        const int go = tid / GO, gl = tid % GO;
        const int pa = go / NP, pb = go % NP;
        const bool own = go < NP * NP;
        double accC[TPT][NC];  // MODE 1: Mbar_c increment
#pragma unroll
        for (int q = 0; q < TPT; ++q)
#pragma unroll
            for (int m = 0; m < NC; ++m) accC[q][m] = 0.0;
        for (int c0 = 0; c0 < NC; c0 += CG) {
            __syncthreads();
            for (int e = tid; e < CG * NN; e += AT) {
                const int cc = e / NN, n = e % NN;
                bU[cc][n] = Uarr[((int64_t)C * NN + n) * NC + c0 + cc];
                bV[cc][n] = Varr[((int64_t)P * NN + n) * NC + c0 + cc];
            }
            __syncthreads();
            // ---- value propagation
            if (!(PART & 2)) {
            } else if (MODE == 0) {  // Lbar_P[b] += sum_a Om(a,b) Lbar_c[a]: contract a_x, a_y, a_z
                lds_pass<NN, CG * NN>(t1, [&](int cc, int n) {  // t1[bx, ay, az]
                    const int bx = n % NP, r = n / NP;
                    double s = 0.0;
#pragma unroll
                    for (int ax = 0; ax < NP; ++ax) s += Om[0][ax][bx] * bU[cc][ax + NP * r];
                    return s;
                });
                __syncthreads();
                lds_pass<NN, CG * NN>(t2, [&](int cc, int n) {  // t2[bx, by, az]
                    const int bx = n % NP, by = (n / NP) % NP, az = n / (NP * NP);
                    double s = 0.0;
#pragma unroll
                    for (int ay = 0; ay < NP; ++ay)
                        s += Om[1][ay][by] * t1[cc][bx + NP * (ay + NP * az)];
                    return s;
                });
                __syncthreads();
#pragma unroll
                for (int q = 0; q < TPT; ++q) {
                    const int n = tid + q * AT;
                    if (n < NN) {
                        const int ij = n % (NP * NP), bz = n / (NP * NP);
#pragma unroll
                        for (int cc = 0; cc < CG; ++cc) {
                            double s = 0.0;
                            for (int az = 0; az < NP; ++az)
                                s += Om[2][az][bz] * t2[cc][ij + NP * NP * az];
                            accP[q][c0 + cc] += s;
                        }
                    }
                }
            } else {  // Mbar_c[a] += sum_b Om(a,b) Mbar_P[b]: contract b_x, b_y, b_z
                lds_pass<NN, CG * NN>(t1, [&](int cc, int n) {  // t1[ax, by, bz]
                    const int ax = n % NP, r = n / NP;
                    double s = 0.0;
#pragma unroll
                    for (int bx = 0; bx < NP; ++bx) s += Om[0][ax][bx] * bV[cc][bx + NP * r];
                    return s;
                });
                __syncthreads();
                lds_pass<NN, CG * NN>(t2, [&](int cc, int n) {  // t2[ax, ay, bz]
                    const int ax = n % NP, ay = (n / NP) % NP, bz = n / (NP * NP);
                    double s = 0.0;
#pragma unroll
                    for (int by = 0; by < NP; ++by)
                        s += Om[1][ay][by] * t1[cc][ax + NP * (by + NP * bz)];
                    return s;
                });
                __syncthreads();
#pragma unroll
                for (int q = 0; q < TPT; ++q) {
                    const int n = tid + q * AT;
                    if (n < NN) {
                        const int ij = n % (NP * NP), az = n / (NP * NP);
#pragma unroll
                        for (int cc = 0; cc < CG; ++cc) {
                            double s = 0.0;
                            for (int bz = 0; bz < NP; ++bz)
                                s += Om[2][az][bz] * t2[cc][ij + NP * NP * bz];
                            accC[q][c0 + cc] += s;
                        }
                    }
                }
            }
            // ---- geometry: three axes
            // (unrolled: the axes are constants, and node indices move by fixed strides)
#pragma unroll
            for (int ax3 = 0; ax3 < 3; ++ax3) {
                if (!(PART & 1)) break;
                // V'[..] = V contracted over the two other axes with Om (b -> a)
                const int d1 = (ax3 + 1) % 3, d2 = (ax3 + 2) % 3;
                const int st1 = d1 == 0 ? 1 : (d1 == 1 ? NP : NP * NP);
                const int st2 = d2 == 0 ? 1 : (d2 == 1 ? NP : NP * NP);
                __syncthreads();
                lds_pass<NN, CG * NN>(t1, [&](int cc, int n) {  // contract axis d1: b_d1 -> a_d1
                    const int a1 = (n / st1) % NP, n0 = n - a1 * st1;
                    double s = 0.0;
#pragma unroll
                    for (int b1 = 0; b1 < NP; ++b1) s += Om[d1][a1][b1] * bV[cc][n0 + b1 * st1];
                    return s;
                });
                __syncthreads();
                lds_pass<NN, CG * NN>(t2, [&](int cc, int n) {  // contract axis d2
                    const int a2 = (n / st2) % NP, n0 = n - a2 * st2;
                    double s = 0.0;
#pragma unroll
                    for (int b2 = 0; b2 < NP; ++b2) s += Om[d2][a2][b2] * t1[cc][n0 + b2 * st2];
                    return s;
                });
                __syncthreads();
                // gO[ax3] += sum over the other a indices of U[a] t2[(a with a_ax3 -> b)]
                double s = 0.0;
#pragma unroll
                for (int r = 0; r < (NP * NP + GO - 1) / GO; ++r) {
                    const int u = gl + r * GO;
                    if (own && u < NP * NP) {
                        const int u1 = u / NP, u2 = u - u1 * NP;
                        int ia[3], ib[3];
                        ia[ax3] = pa;
                        ib[ax3] = pb;
                        ia[d1] = ib[d1] = u1;
                        ia[d2] = ib[d2] = u2;
                        const int na = ia[0] + NP * (ia[1] + NP * ia[2]);
                        const int nb = ib[0] + NP * (ib[1] + NP * ib[2]);
#pragma unroll
                        for (int cc = 0; cc < CG; ++cc) s += bU[cc][na] * t2[cc][nb];
                    }
                }
#pragma unroll
                for (int o = 1; o < GO; o <<= 1) s += __shfl_xor(s, o);
                gO[ax3] += s;
            }
        }
        // child's value output (MODE 1)
        if (MODE == 1 && (PART & 2)) {
#pragma unroll
            for (int q = 0; q < TPT; ++q) {
                const int e = tid + q * AT;
                if (e < NN)
#pragma unroll
                    for (int m = 0; m < NC; ++m) out[((int64_t)C * NN + e) * NC + m] += accC[q][m];
            }
        }
        // dl/dOm -> dl/dX_c (evaluation points) and dl/dX_P (nodes)
        __syncthreads();
        if ((PART & 1) && own && gl == 0)
            for (int d = 0; d < 3; ++d) red[go][d] = gO[d];  // red[(a,b)][axis]
        __syncthreads();
        if ((PART & 1) && tid < 3 * NP) {
            const int d = tid / NP, ai = tid % NP;
            double cvec[NP];
            for (int b = 0; b < NP; ++b) cvec[b] = red[ai * NP + b][d];
            double xP[NP];
            for (int m = 0; m < NP; ++m) xP[m] = 0.0;
            double ri[NP];  // 1/(x - X) for the node-coordinate terms
            if (fl[d][ai] < 0)
                for (int m = 0; m < NP; ++m) ri[m] = 1.0 / (cn[tid] - pn[d * NP + m]);
            const double gxc = axis_adj<NP>(&Om[d][ai][0], &dOm[d][ai][0], ri, fl[d][ai], cvec, xP);
            Xbar[(int64_t)C * 3 * NP + tid] += gxc;
            for (int m = 0; m < NP; ++m) xPs[ch][tid][m] = xP[m];
        }
    }
    if (MODE == 0 && (PART & 2)) {
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const int e = tid + q * AT;
            if (e < NN)
#pragma unroll
                for (int m = 0; m < NC; ++m) out[((int64_t)P * NN + e) * NC + m] = accP[q][m];
        }
    }
    if (PART & 1) {  // dl/dX_P[d][m]: over the child nodes a, both children per node
        __syncthreads();
        if (tid < 3 * NP) {
            const int d = tid / NP, m = tid % NP;
            double sum = 0.0;
            for (int a = 0; a < NP; ++a) sum += xPs[0][d * NP + a][m] + xPs[1][d * NP + a][m];
            Xbar[(int64_t)P * 3 * NP + tid] += sum;
        }
    }
}

// ------------------------------------------------------------- M2L adjoints
template <int NP>
struct AShape {
    static constexpr int NN = NP * NP * NP;
    static constexpr int BLK = NN >= 1024 ? 1024 : ((NN + 63) / 64) * 64;
    static constexpr int TPT = (NN + BLK - 1) / BLK;
};

// M2L^T.  For one node pair (target node x with local cotangent (e, b), source
// node y with multipole (g, m)), R = X - Y, c = d^-5/2 (R.u), u = g e + b x m:
//   dl/dR = d^-3/2 u - 3 c (R + (p.R) p),   R.u = g (e.R) + m.(R x b) = g (e.R) + b.(m x R),
//   dl/dg = d^-3/2 (e.R),  dl/dm = d^-3/2 (R x b),  dl/dp_avg = -3 c (p.R) R.
// Summed over one side with the other fixed, d^-3/2 u splits into products of
// fixed factors and kernel sums (e sum d^-3/2 g + b x sum d^-3/2 m on the target
// side, g sum d^-3/2 e + (sum d^-3/2 b) x m on the source side), and on the
// tensor grid of nodes every factor linear in R is hoisted per row (dy, dz fixed):
// sum c R = (sum c dx, dy sum c, dz sum c).  ~30 vector instructions per pair.

// Target side: per target node a, over its sources: dl/dY_t[a] and dl/dp_avg;
// with FWD the same pass also forms the forward M2L locals (row sums as in
// k_m2l) and stores them in loc.  With MIXED the node differences are formed in
// FP64 and rounded once, the terms run in FP32 (v_rsq_f32) and every source
// cluster's FP32 sums are added into FP64 (mp holds float4 multipoles then).
//
// Source-node coordinates (NP <= 7): the source cluster's dl/dY_s[b] is minus the
// same per-pair term J^T u summed over the target nodes, and only its sums over
// the nodes with index m on axis d are needed.  Per source cluster s, each lane
// groups its terms by the source node's x index (column sums), y index (row
// sums) and z index (plane sums); the block adds the lanes' 3 NP values through
// LDS and writes them to ysp at the pair's position in the transposed list
// (posT), so that the source side only reads them.
template <int NP>
struct YSplit {
    static constexpr bool on = NP <= 7;  // 3 NP x BLK doubles of LDS: 36 KB (NP 6), 63 KB (NP 7)
};

// Occupancy: the FP64 form with the source-node sums (4 <= NP <= 6) is held to 3 waves
// per SIMD (168 VGPRs, no spills): 167.5 -> 162.6 ms at N = 2^22, n = 4.  At NP = 7 the
// same bound made it 15% slower (rematerialisation), so larger NP keep the default (NP = 7
// is held to 3 waves per SIMD by its 78 KB of LDS anyway).  The mixed form at NP = 5
// without the forward sums (the pullback) is held to 4 waves per SIMD (128 instead of 152
// VGPRs, 16 bytes spilled outside the node loop; its 19 KB of LDS allow 4): 31.2 -> 29.5
// ms alone at N = 2^22, mixed pullback 82.9 -> 81.0 ms; with the forward sums the same
// bound spilled 76 bytes and gained nothing.
template <int NP, bool MIXED, bool FWD>
struct AdjTWaves {
    static constexpr int v =
        (!MIXED && NP >= 4 && NP <= 6) ? 3 : ((MIXED && !FWD && NP == 5) ? 4 : 1);
};

template <int NP, bool FWD, bool MIXED, bool QUAD = false>
__global__ __launch_bounds__(AShape<NP>::BLK)
__attribute__((amdgpu_waves_per_eu(AdjTWaves<NP, MIXED, FWD>::v))) void k_adj_m2l_t(
    const int32_t *__restrict__ off, const uint64_t *__restrict__ keys, int sb,
    const double *__restrict__ ncoord, const void *__restrict__ mpv,
    const double *__restrict__ lbar, const double *__restrict__ pavg4, double *__restrict__ Xbar,
    double *__restrict__ pbar_part, double *__restrict__ loc, const uint32_t *__restrict__ posT,
    double *__restrict__ ysp) {
    using T = typename std::conditional<MIXED, float, double>::type;
    using V4 = typename std::conditional<MIXED, float4, double4>::type;
    constexpr int NN = AShape<NP>::NN, BLK = AShape<NP>::BLK, TPT = AShape<NP>::TPT;
    // the final reductions (yb, pr3; BLK >= NN) and, with MIXED, the staged source
    constexpr int SH = (MIXED && 2 * NN > 3 * BLK) ? 2 * NN : 3 * BLK;
    constexpr int NA = 17;  // accumulators: F[6], G, A[3], sum c R [3], sum c pr, sum c pr R [3]
    constexpr bool YS = YSplit<NP>::on;
    static_assert(!YS || TPT == 1, "source-node sums assume one target node per thread");
    constexpr int NY = YS ? 3 * NP : 1, KS = YS ? BLK / 16 : 1;
    __shared__ double shb[SH];
    __shared__ double lsn[MIXED ? 3 * NP : 1];
    __shared__ double ybuf[NY][YS ? BLK : 1];  // per lane: column, row, plane sums
    __shared__ double ybr[NY][KS];             // sums over 16 lanes
    double(*yb)[3] = reinterpret_cast<double(*)[3]>(shb);
    double(*pr3)[3] = reinterpret_cast<double(*)[3]>(shb);
    const V4 *mp = reinterpret_cast<const V4 *>(mpv);
    const int tid = threadIdx.x;
    const int32_t t = xcd_item<XCD_RUN_FAR>(blockIdx.x, gridDim.x);
    const int32_t beg = off[t], end = off[t + 1];
    const uint64_t mask = (1ull << sb) - 1ull;
    const T px = (T)pavg4[0], py = (T)pavg4[1], pz = (T)pavg4[2];
    const T qa = fma(px, px, (T)1);  // QUAD: 1 + px^2
    double tx[TPT], ty[TPT], tz[TPT];
    T e[TPT][3], b[TPT][3], ac[TPT][NA];
    T u3c[TPT][3];       // e0/3, b1/3, b2/3: (J^T u)_x / 3 per pair for the column sums
    double ad[TPT][NA];  // FP64 sums (MIXED)
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        const int node = min(tid + q * BLK, NN - 1);
        tx[q] = ncoord[((int64_t)t * 3 + 0) * NP + node % NP];
        ty[q] = ncoord[((int64_t)t * 3 + 1) * NP + (node / NP) % NP];
        tz[q] = ncoord[((int64_t)t * 3 + 2) * NP + node / (NP * NP)];
        const double *lb = lbar + ((int64_t)t * NN + node) * 6;
        for (int d = 0; d < 3; ++d) {
            e[q][d] = (T)lb[d];
            b[q][d] = (T)lb[3 + d];
        }
        u3c[q][0] = (T)(lb[0] / 3.0);
        u3c[q][1] = (T)(lb[4] / 3.0);
        u3c[q][2] = (T)(lb[5] / 3.0);
        for (int k = 0; k < NA; ++k) ac[q][k] = 0, ad[q][k] = 0.0;
    }
    if (YS && tid >= NN)
        for (int j = 0; j < NY; ++j) ybuf[j][tid] = 0.0;
    for (int32_t ei = beg; ei < end; ++ei) {
        // the source's multipoles and node coordinates: MIXED stages them in LDS;
        // FP64 reads them with wave-uniform addresses (scalar loads)
        const int64_t s = (int64_t)(keys[ei] & mask);
        const V4 *smp = mp + s * NN;
        const double *sn = ncoord + s * 3 * NP;
        if constexpr (MIXED) {
            V4 *lmp = reinterpret_cast<V4 *>(shb);
            __syncthreads();
            for (int j = tid; j < NN; j += BLK) lmp[j] = smp[j];
            if (tid < 3 * NP) lsn[tid] = sn[tid];
            __syncthreads();
            smp = lmp;
            sn = lsn;
        }
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            if (tid + q * BLK >= NN) continue;
            T *A = ac[q];
            const T e0 = e[q][0], e1 = e[q][1], e2 = e[q][2];
            const T b0 = b[q][0], b1 = b[q][1], b2 = b[q][2];
            T dxv[NP], dx2v[NP];
            T Xs[NP], dx3v[NP];  // column sums / 3 (source x index), 3 dx; row sums in ybuf
#pragma unroll
            for (int si = 0; si < NP; ++si) {
                dxv[si] = (T)(tx[q] - sn[si]);
                dx2v[si] = dxv[si] * dxv[si];
                dx3v[si] = (T)3 * dxv[si];
                Xs[si] = 0;
                if (YS) ybuf[NP + si][tid] = 0.0;
            }
            // forward sums as in m2l_pair: H, U_y, U_z run over the source cluster;
            // per plane G, A_x, A_y, A_z (also the totals of the adjoint's e G + b x A);
            // per row G, A_x, A_z and the c sums (first node peeled: no resets)
            T H = 0, Uy = 0, Uz = 0;
#pragma unroll 1
            for (int sk = 0; sk < NP; ++sk) {
                const T dz = (T)(tz[q] - sn[2 * NP + sk]);
                T Gp = 0, Axp = 0, Ayp = 0, Azp = 0, C0p = 0, Cpp = 0;
                for (int sj = 0; sj < NP; ++sj) {
                    const T dy = (T)(ty[q] - sn[NP + sj]);
                    const T rrjk = fma(dy, dy, dz * dz) + (T)EPS64;
                    const T prjk = fma(py, dy, pz * dz);
                    T hj = 0, qc = 0;  // QUAD: px p.R_yz and R_yz.R_yz + (p.R_yz)^2 + eps
                    if constexpr (QUAD) hj = px * prjk, qc = fma(prjk, prjk, rrjk);
                    const T er = fma(dy, e1, dz * e2);        // e.R - dx e0
                    const T v0 = fma(dy, b2, -(dz * b1));     // (R x b)_0
                    const T v1r = dz * b0, v2r = dy * b0;     // (R x b)_1 + dx b2, (R x b)_2 - dx b1
                    T G, Ax, Az, C0, C1, C2;  // row sums: c, c dx, c dx^2
#pragma unroll
                    for (int si = 0; si < NP; ++si) {
                        const T dx = dxv[si];
                        T d, fx;  // fx = dx + px p.R (the x factor of the coordinate terms)
                        if constexpr (QUAD) {  // d as a quadratic in dx (as k_m2l with |px| <= 0.1)
                            fx = fma(qa, dx, hj);
                            d = fma(dx, fx + hj, qc);
                        } else {
                            const T pr = fma(px, dx, prjk);
                            d = fma(pr, pr, fma(dx, dx, rrjk));
                            fx = fma(pr, px, dx);
                        }
                        T y;
                        if constexpr (MIXED) y = __builtin_amdgcn_rsqf(d);
                        else y = rsqrt_nr(d);
                        const T y2 = y * y, y3 = y2 * y;
                        const V4 m = smp[si + NP * (sj + NP * sk)];
                        T Ru, u3;  // R.u; u_x / 3 (u = g e + b x m)
                        if constexpr (YS) {  // R.u = dx u_x + (dy u_y + dz u_z), the latter from row terms
                            u3 = fma(m.x, u3c[q][0], fma(u3c[q][1], m.w, -(u3c[q][2] * m.z)));
                            Ru = fma(dx3v[si], u3, fma(m.x, er, fma(m.y, v0, fma(m.z, v1r, -(m.w * v2r)))));
                        } else {
                            const T v1 = fma(-dx, b2, v1r), v2 = fma(dx, b1, -v2r);
                            Ru = fma(m.x, fma(dx, e0, er), fma(m.y, v0, fma(m.z, v1, m.w * v2)));
                            u3 = 0;
                        }
                        const T c = y3 * (y2 * Ru);
                        if (si == 0) {
                            G = y3 * m.x;
                            Ax = y3 * m.y;
                            Az = y3 * m.w;
                            C0 = c;
                            C1 = c * dx;
                            C2 = c * dx2v[si];
                        } else {
                            G = fma(y3, m.x, G);
                            Ax = fma(y3, m.y, Ax);
                            Az = fma(y3, m.w, Az);
                            C0 += c;
                            C1 = fma(c, dx, C1);
                            C2 = fma(c, dx2v[si], C2);
                        }
                        Ayp = fma(y3, m.z, Ayp);
                        if (FWD) {
                            const T uu = y3 * dx;
                            H = fma(uu, m.x, H);
                            Uy = fma(uu, m.z, Uy);
                            Uz = fma(uu, m.w, Uz);
                        }
                        if constexpr (YS)  // (J^T u)_x / 3 = y3 u_x / 3 - c (dx + pr px)
                            Xs[si] = fma(y3, u3, fma(-c, fx, Xs[si]));
                    }
                    // pr = px dx + prjk along the row: sum c pr and sum c pr dx from C0..C2
                    const T Cp = fma(px, C1, prjk * C0), Cpx = fma(px, C2, prjk * C1);
                    if (FWD) {
                        A[1] = fma(dy, G, A[1]);   // E_y
                        A[3] = fma(-dy, Az, A[3]);  // B_x (dy part)
                        A[5] = fma(dy, Ax, A[5]);   // B_z (dy part)
                    }
                    if constexpr (YS) {  // row: (J^T u)_y = e1 G + (b x A)_y - 3 (dy C0 + py Cp)
                        ybuf[NP + sj][tid] += (double)(fma(e1, G, fma(b2, Ax, -(b0 * Az))) -
                                                       (T)3 * fma(dy, C0, py * Cp));
                        C0p += C0;
                        Cpp += Cp;
                    }
                    Gp += G;
                    Axp += Ax;
                    Azp += Az;
                    A[10] += C1;
                    A[11] = fma(dy, C0, A[11]);
                    A[12] = fma(dz, C0, A[12]);
                    A[13] += Cp;
                    A[14] += Cpx;
                    A[15] = fma(dy, Cp, A[15]);
                    A[16] = fma(dz, Cp, A[16]);
                }
                if (FWD) {
                    A[2] = fma(dz, Gp, A[2]);   // E_z
                    A[3] = fma(dz, Ayp, A[3]);  // B_x (dz part)
                    A[4] = fma(-dz, Axp, A[4]);  // B_y (dz part)
                }
                if constexpr (YS)  // plane: (J^T u)_z = e2 G + (b x A)_z - 3 (dz C0 + pz Cp)
                    ybuf[2 * NP + sk][tid] = (double)(fma(e2, Gp, fma(b0, Ayp, -(b1 * Axp))) -
                                                      (T)3 * fma(dz, C0p, pz * Cpp));
                A[6] += Gp;
                A[7] += Axp;
                A[8] += Ayp;
                A[9] += Azp;
            }
            if (FWD) {
                A[0] += H;
                A[4] += Uz;
                A[5] -= Uy;
            }
            if constexpr (MIXED)
                for (int k = 0; k < NA; ++k) ad[q][k] += (double)A[k], A[k] = 0;
            if constexpr (YS) {
#pragma unroll
                for (int si = 0; si < NP; ++si) ybuf[si][tid] = 3.0 * (double)Xs[si];
            }
        }
        if constexpr (YS) {  // lanes -> sums over 16 lanes -> the pair's sums
            __syncthreads();
            for (int w = tid; w < NY * KS; w += BLK) {
                const int j = w / KS, k = w % KS;
                double v = 0.0;
#pragma unroll
                for (int l = 0; l < 16; ++l) v += ybuf[j][16 * k + l];
                ybr[j][k] = v;
            }
            __syncthreads();
            if (tid < NY) {
                double v = 0.0;
                for (int k = 0; k < KS; ++k) v += ybr[tid][k];
                ysp[(int64_t)posT[ei] * NY + tid] = v;
            }
        }
    }
    double pd[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        if constexpr (!MIXED)
            for (int k = 0; k < NA; ++k) ad[q][k] = ac[q][k];
        // dl/dX = e G + b x A - 3 (sum c R + p sum c pr);  dl/dp_avg = -3 sum c pr R
        const double *S = ad[q];
        const double E0 = e[q][0], E1 = e[q][1], E2 = e[q][2];
        const double B0 = b[q][0], B1 = b[q][1], B2 = b[q][2];
        const double p0 = pavg4[0], p1 = pavg4[1], p2 = pavg4[2];
        const double Yx = E0 * S[6] + (B1 * S[9] - B2 * S[8]) - 3.0 * (S[10] + p0 * S[13]);
        const double Yy = E1 * S[6] + (B2 * S[7] - B0 * S[9]) - 3.0 * (S[11] + p1 * S[13]);
        const double Yz = E2 * S[6] + (B0 * S[8] - B1 * S[7]) - 3.0 * (S[12] + p2 * S[13]);
        ad[q][0 + 6] = Yx;  // reuse slots for the node-coordinate gradient
        ad[q][1 + 6] = Yy;
        ad[q][2 + 6] = Yz;
        if (tid + q * BLK < NN) {
            pd[0] -= 3.0 * S[14];
            pd[1] -= 3.0 * S[15];
            pd[2] -= 3.0 * S[16];
        }
    }
    if (FWD) {
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const int node = tid + q * BLK;
            if (node < NN)
                for (int d = 0; d < 6; ++d) loc[((int64_t)t * NN + node) * 6 + d] = ad[q][d];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        const int node = tid + q * BLK;
        if (node < NN)
            for (int d = 0; d < 3; ++d) yb[node][d] = ad[q][6 + d];
    }
    __syncthreads();
    if (tid < 3 * NP) {  // dl/dX_{t,d,m} = sum over nodes with index m on axis d
        const int d = tid / NP, mi = tid % NP;
        double sum = 0.0;
        for (int u1 = 0; u1 < NP; ++u1)
            for (int u2 = 0; u2 < NP; ++u2) {
                int id[3];
                id[d] = mi;
                id[(d + 1) % 3] = u1;
                id[(d + 2) % 3] = u2;
                sum += yb[id[0] + NP * (id[1] + NP * id[2])][d];
            }
        Xbar[(int64_t)t * 3 * NP + tid] += sum;
    }
    __syncthreads();
    for (int d = 0; d < 3; ++d) pr3[tid][d] = pd[d];
    __syncthreads();
    if (tid < 3) {
        double sum = 0.0;
        for (int j = 0; j < BLK; ++j) sum += pr3[j][tid];
        pbar_part[3 * (int64_t)t + tid] = sum;
    }
}

// Source side: per source node b, over the targets that list it: Mbar_s[b]
// (= dl/dgamma, dl/dm) and dl/dY_s[b].  MIXED as in k_adj_m2l_t (the target
// clusters' local cotangents are staged as FP32).  For NP <= 7 the target side
// has already summed dl/dY_s per pair (ysp, in transposed-list order): only
// the multipole cotangents are evaluated here, with 8 sums per node pair.
template <int NP, bool MIXED, bool QUAD = false>
__global__ __launch_bounds__(AShape<NP>::BLK) void k_adj_m2l_s(
    const int32_t *__restrict__ offT, const uint64_t *__restrict__ keysT, int sb,
    const double *__restrict__ ncoord, const void *__restrict__ mpv,
    const void *__restrict__ lbarv, const double *__restrict__ pavg4, double *__restrict__ mbar,
    double *__restrict__ Xbar, const double *__restrict__ ysp) {
    using T = typename std::conditional<MIXED, float, double>::type;
    using V4 = typename std::conditional<MIXED, float4, double4>::type;
    constexpr int NN = AShape<NP>::NN, BLK = AShape<NP>::BLK, TPT = AShape<NP>::TPT;
    constexpr bool YS = YSplit<NP>::on;
    // gb[4]; without YS also sum y3 e [3], sum y3 b [3], sum c R [3], sum c pr
    constexpr int NA = YS ? 4 : 14;
    // MIXED: the target's cotangents staged in LDS as FP32; FP64: read with
    // wave-uniform addresses (scalar loads)
    constexpr int SL = (MIXED || !YS) ? 3 * NN : 1;  // staged FP32 cotangents / yb
    __shared__ double shl[SL];
    __shared__ double ltn[MIXED ? 3 * NP : 1];
    double(*yb)[3] = reinterpret_cast<double(*)[3]>(shl);
    const V4 *mp = reinterpret_cast<const V4 *>(mpv);
    const double *lbar = reinterpret_cast<const double *>(lbarv);
    const int tid = threadIdx.x;
    const int32_t s = xcd_item<XCD_RUN_FAR>(blockIdx.x, gridDim.x);
    const int32_t beg = offT[s], end = offT[s + 1];
    const uint64_t mask = (1ull << sb) - 1ull;
    const T px = (T)pavg4[0], py = (T)pavg4[1], pz = (T)pavg4[2];
    const T qa = fma(px, px, (T)1), qb = px + px;  // QUAD: 1 + px^2, 2 px
    double sx[TPT], sy[TPT], sz[TPT];
    T ac[TPT][NA];
    double ad[TPT][NA];
    V4 mq[TPT];
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        const int node = min(tid + q * BLK, NN - 1);
        sx[q] = ncoord[((int64_t)s * 3 + 0) * NP + node % NP];
        sy[q] = ncoord[((int64_t)s * 3 + 1) * NP + (node / NP) % NP];
        sz[q] = ncoord[((int64_t)s * 3 + 2) * NP + node / (NP * NP)];
        mq[q] = mp[(int64_t)s * NN + node];
        for (int k = 0; k < NA; ++k) ac[q][k] = 0, ad[q][k] = 0.0;
    }
    for (int32_t ei = beg; ei < end; ++ei) {
        // the target's node coordinates and local cotangents are the same for every
        // lane: wave-uniform addresses (scalar loads, no LDS staging)
        const int64_t t = (int64_t)(keysT[ei] & mask);
        const double *tn = ncoord + t * 3 * NP;
        const T *tlb;
        if constexpr (MIXED) {
            float *lf = reinterpret_cast<float *>(shl);
            __syncthreads();
            for (int j = tid; j < NN * 6; j += BLK) lf[j] = (float)lbar[t * NN * 6 + j];
            if (tid < 3 * NP) ltn[tid] = tn[tid];
            __syncthreads();
            tlb = lf;
            tn = ltn;
        } else {
            tlb = lbar + t * NN * 6;
        }
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            if (tid + q * BLK >= NN) continue;
            T *A = ac[q];
            const V4 m = mq[q];  // gamma, mx, my, mz
            T dxv[NP];
#pragma unroll
            for (int ti = 0; ti < NP; ++ti) dxv[ti] = (T)(tn[ti] - sx[q]);
            for (int tk = 0; tk < NP; ++tk) {
                const T dz = (T)(tn[2 * NP + tk] - sz[q]);
                for (int tj = 0; tj < NP; ++tj) {
                    const T dy = (T)(tn[NP + tj] - sy[q]);
                    const T rrjk = fma(dy, dy, fma(dz, dz, (T)EPS64));
                    const T prjk = fma(py, dy, pz * dz);
                    T qbj = 0, qc = 0;  // QUAD: 2 px p.R_yz and R_yz.R_yz + (p.R_yz)^2 + eps
                    if constexpr (QUAD) qbj = qb * prjk, qc = fma(prjk, prjk, rrjk);
                    const T w0 = fma(m.z, dz, -(m.w * dy));  // (m x R)_0
                    const T w1r = -(m.y * dz), w2r = m.y * dy;
                    T Sxe = 0, Ex = 0, Ey = 0, Ez = 0, Bx = 0, By = 0, Bz = 0, Sxby = 0, Sxbz = 0;
                    T C0 = 0, C1 = 0;
#pragma unroll
                    for (int ti = 0; ti < NP; ++ti) {
                        const T dx = dxv[ti];
                        T d;
                        if constexpr (QUAD) {
                            d = fma(dx, fma(qa, dx, qbj), qc);
                        } else {
                            const T pr = fma(px, dx, prjk);
                            d = fma(pr, pr, fma(dx, dx, rrjk));
                        }
                        T y;
                        if constexpr (MIXED) y = __builtin_amdgcn_rsqf(d);
                        else y = rsqrt_nr(d);
                        const T y2 = y * y, y3 = y2 * y, uu = y3 * dx;
                        const T *o = tlb + 6 * (ti + NP * (tj + NP * tk));
                        const T e0 = o[0], e1 = o[1], e2 = o[2], b0 = o[3], b1 = o[4], b2 = o[5];
                        Sxe = fma(uu, e0, Sxe);
                        Ey = fma(y3, e1, Ey);
                        Ez = fma(y3, e2, Ez);
                        Bx = fma(y3, b0, Bx);
                        By = fma(y3, b1, By);
                        Bz = fma(y3, b2, Bz);
                        Sxby = fma(uu, b1, Sxby);
                        Sxbz = fma(uu, b2, Sxbz);
                        if constexpr (!YS) {
                            Ex = fma(y3, e0, Ex);
                            const T w1 = fma(m.w, dx, w1r), w2 = fma(-m.z, dx, w2r);
                            const T Ru = fma(m.x, fma(dx, e0, fma(dy, e1, dz * e2)),
                                             fma(b0, w0, fma(b1, w1, b2 * w2)));
                            const T c = y3 * (y2 * Ru);
                            C0 += c;
                            C1 = fma(c, dx, C1);
                        }
                    }
                    A[0] += fma(dy, Ey, fma(dz, Ez, Sxe));  // dl/dgamma = sum d^-3/2 e.R
                    A[1] += fma(dy, Bz, -(dz * By));        // dl/dm = sum d^-3/2 R x b
                    A[2] += fma(dz, Bx, -Sxbz);
                    A[3] += fma(-dy, Bx, Sxby);
                    if constexpr (!YS) {
                        const T Cp = fma(px, C1, prjk * C0);  // sum c pr (pr = px dx + prjk)
                        A[4] += Ex;
                        A[5] += Ey;
                        A[6] += Ez;
                        A[7] += Bx;
                        A[8] += By;
                        A[9] += Bz;
                        A[10] += C1;
                        A[11] = fma(dy, C0, A[11]);
                        A[12] = fma(dz, C0, A[12]);
                        A[13] += Cp;
                    }
                }
            }
            if constexpr (MIXED)
                for (int k = 0; k < NA; ++k) ad[q][k] += (double)A[k], A[k] = 0;
        }
    }
    if constexpr (YS) {
        const int node = tid;
        if (node < NN) {  // (no LDS is read below: no barrier needed)
            const double *S = MIXED ? ad[0] : nullptr;
            double *o = mbar + ((int64_t)s * NN + node) * 4;
            for (int d = 0; d < 4; ++d) o[d] = MIXED ? S[d] : (double)ac[0][d];
        }
        if (tid < 3 * NP) {  // dl/dY_s = - sum over the pairs of the target side's sums
            double sum = 0.0;
            for (int32_t ei = beg; ei < end; ++ei) sum += ysp[(int64_t)ei * 3 * NP + tid];
            Xbar[(int64_t)s * 3 * NP + tid] -= sum;
        }
        return;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        if constexpr (!MIXED)
            for (int k = 0; k < NA; ++k) ad[q][k] = ac[q][k];
        const int node = tid + q * BLK;
        if (node < NN) {
            const double *S = ad[q];
            // dl/dY = -(g sum d^-3/2 e + (sum d^-3/2 b) x m) + 3 (sum c R + p sum c pr)
            const double g = mq[q].x, m0 = mq[q].y, m1 = mq[q].z, m2 = mq[q].w;
            const double p0 = pavg4[0], p1 = pavg4[1], p2 = pavg4[2];
            double *o = mbar + ((int64_t)s * NN + node) * 4;
            for (int d = 0; d < 4; ++d) o[d] = S[d];
            yb[node][0] = -(g * S[4] + (S[8] * m2 - S[9] * m1)) + 3.0 * (S[10] + p0 * S[13]);
            yb[node][1] = -(g * S[5] + (S[9] * m0 - S[7] * m2)) + 3.0 * (S[11] + p1 * S[13]);
            yb[node][2] = -(g * S[6] + (S[7] * m1 - S[8] * m0)) + 3.0 * (S[12] + p2 * S[13]);
        }
    }
    __syncthreads();
    if (tid < 3 * NP) {
        const int d = tid / NP, mi = tid % NP;
        double sum = 0.0;
        for (int u1 = 0; u1 < NP; ++u1)
            for (int u2 = 0; u2 < NP; ++u2) {
                int id[3];
                id[d] = mi;
                id[(d + 1) % 3] = u1;
                id[(d + 2) % 3] = u2;
                sum += yb[id[0] + NP * (id[1] + NP * id[2])][d];
            }
        Xbar[(int64_t)s * 3 * NP + tid] += sum;
    }
}

// ------------------------------------------------------------- P2P adjoints
// Both sides use the forward P2P's wave layout: 16 lanes x 4 groups, lane (li, g)
// owns the 4 particles li + 16 q of its 64-particle slice and takes the staged
// partners k = g (mod 4); the groups' sums meet in two xor-shuffles.  With MIXED
// the differences R are formed in FP64 and rounded once, the kernel terms run in
// FP32 (v_rsq_f32), y5 (R.u) is formed as y3 (y^2 (R.u)) so that no intermediate
// overflows, and each staged tile's FP32 sums are added into FP64.  The self
// pair (same particle) contributes nothing: y is set to 0.
constexpr int PT = 256;
constexpr int PNG = 4;

// Kernel value of one pair: y = d^-1/2 (0 for the self pair), pr = p.R, R rounded to T.
template <typename T>
__device__ inline T pair_geo(double R0d, double R1d, double R2d, T px, T py, T pz, bool self,
                             T *R, T &pr) {
    R[0] = (T)R0d;
    R[1] = (T)R1d;
    R[2] = (T)R2d;
    pr = fma(px, R[0], fma(py, R[1], pz * R[2]));
    const T d = fma(pr, pr, fma(R[0], R[0], fma(R[1], R[1], fma(R[2], R[2], (T)EPS64))));
    T y;
    if constexpr (sizeof(T) == 4) y = __builtin_amdgcn_rsqf(d);
    else y = rsqrt_nr(d);
    return self ? (T)0 : y;
}

// Target side: gx_i += sum_j J^T u_ij,  u_ij = g_j e_i + b_i x p_j.
// One staged tile for the lane's PNG targets; DIAG: the tile holds the targets'
// own leaf, and the self pair (source k == selfk + 16 q) is skipped.
// NT targets per lane, 64 / NT lanes per group; the first Q of them (the last wave of
// a leaf may have fewer targets than 16 NT).
// m: a multiple of NT (the tile is padded), the same trip count in every lane.
template <bool DIAG, int NT, int Q, typename T, typename Src>
__device__ __forceinline__ void adj_t_tile(const Src *ts, int g, int m, int selfk,
                                           const double (&xi)[NT][3], const T (&e)[NT][3],
                                           const T (&b)[NT][3], T (&part)[NT][3]) {
    for (int it = 0; it < m / NT; ++it) {
        const int k = NT * it + g;
        const Src sr = ts[k];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            T R[3], pr;
            const bool self = DIAG && k == selfk + (64 / NT) * q;
            const T y = pair_geo<T>(xi[q][0] - sr.x, xi[q][1] - sr.y, xi[q][2] - sr.z, sr.px,
                                    sr.py, sr.pz, self, R, pr);
            const T y2 = y * y, y3 = y2 * y;
            const T u0 = fma(sr.g, e[q][0], fma(b[q][1], sr.pz, -(b[q][2] * sr.py)));
            const T u1 = fma(sr.g, e[q][1], fma(b[q][2], sr.px, -(b[q][0] * sr.pz)));
            const T u2 = fma(sr.g, e[q][2], fma(b[q][0], sr.py, -(b[q][1] * sr.px)));
            const T Ru = fma(R[0], u0, fma(R[1], u1, R[2] * u2));
            const T c3 = (T)3 * (y3 * (y2 * Ru));
            part[q][0] = fma(y3, u0, fma(-c3, fma(pr, sr.px, R[0]), part[q][0]));
            part[q][1] = fma(y3, u1, fma(-c3, fma(pr, sr.py, R[1]), part[q][1]));
            part[q][2] = fma(y3, u2, fma(-c3, fma(pr, sr.pz, R[2]), part[q][2]));
        }
    }
}

// The mixed form is held to 128 VGPRs (4 waves per SIMD; it needs 130 unbounded).
template <bool MIXED>
__global__ __launch_bounds__(PT) __attribute__((amdgpu_waves_per_eu(MIXED ? 4 : 1)))
void k_adj_p2p_t(const int32_t *__restrict__ leaves,
                                                  const int32_t *__restrict__ off,
                                                  const uint64_t *__restrict__ keys, int sb,
                                                  const int32_t *__restrict__ lo,
                                                  const int32_t *__restrict__ cnt,
                                                  const double *__restrict__ sp8,
                                                  const double *__restrict__ ob,
                                                  double *__restrict__ gx, int TC) {
    using T = typename std::conditional<MIXED, float, double>::type;
    // targets per lane: 4 (16 lanes x 4 groups).  FP64 with 2 (32 lanes x 2 groups,
    // 118 instead of 220 VGPRs, 4 waves per SIMD instead of 2) ran slower, 22.8 ->
    // 24.0 ms at N = 2^22: twice the LDS reads per pair.
    constexpr int NT = PNG, TL = 64 / NT;
    using Src = P2PSrc<MIXED>;
    extern __shared__ double4 adj_lds[];  // sized by the launch: TC records
    const int BT = blockDim.x, tid = threadIdx.x;
    Src *ts = reinterpret_cast<Src *>(adj_lds);
    const int lane = tid & 63, li = lane % TL, g = lane / TL, wbase = tid - lane;
    const int32_t t = leaves[xcd_item<XCD_RUN_NEAR>(blockIdx.x, gridDim.x)];
    const int32_t a = lo[t], ct = cnt[t];
    const int32_t beg = off[t], end = off[t + 1];
    const uint64_t mask = (1ull << sb) - 1ull;
    const double4 *p8 = reinterpret_cast<const double4 *>(sp8);
    for (int tb = 0; tb < ct; tb += BT) {
        double xi[NT][3], acc[NT][3];
        T e[NT][3], b[NT][3];
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const int64_t ii = a + min(tb + wbase + li + TL * q, ct - 1);
            const double4 x4 = p8[2 * ii];
            xi[q][0] = x4.x;
            xi[q][1] = x4.y;
            xi[q][2] = x4.z;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                e[q][d] = (T)ob[6 * ii + d];
                b[q][d] = (T)ob[6 * ii + 3 + d];
                acc[q][d] = 0.0;
            }
        }
        // Sources staged in tiles of up to TC entries, consecutive leaves packed
        // together; the target leaf itself gets tiles of its own (the self pair).
        int fill = 0;
        auto stage = [&](int64_t src0, int slot0, int take) {
#pragma unroll 1
            for (int j = tid; j < take; j += BT) {
                const double4 x4 = p8[2 * (src0 + j)];
                const double4 p4 = p8[2 * (src0 + j) + 1];
                ts[slot0 + j] = Src{x4.x, x4.y, x4.z, (T)x4.w, (T)p4.x, (T)p4.y, (T)p4.z};
            }
        };
        // each tile: staged, then processed at this one call site
        int ei = beg, p0 = 0;  // next source leaf of the list, its first unstaged particle
        for (;;) {
            bool self = false;
            int selfk = 0;
            fill = 0;
            while (ei < end && fill < TC) {
                const int32_t s = (int32_t)(keys[ei] & mask);
                const int32_t sa = lo[s], cs = cnt[s];
                if (s == t && fill > 0) break;  // the target leaf starts a tile of its own
                const int take = s == t ? min(TC, cs - p0) : min(cs - p0, TC - fill);
                stage(sa + p0, fill, take);
                if (s == t) {
                    self = true;
                    selfk = tb + wbase + li - p0;
                }
                fill += take;
                p0 += take;
                if (p0 == cs) {
                    ++ei;
                    p0 = 0;
                }
                if (self) break;
            }
            if (fill == 0) break;
            // padded to a multiple of NT (TC is one) with sources that add zero
            const int fillp = (fill + NT - 1) / NT * NT;
            if (tid < fillp - fill) ts[fill + tid] = Src::pad();
            __syncthreads();
            T part[NT][3];
#pragma unroll
            for (int q = 0; q < NT; ++q) part[q][0] = part[q][1] = part[q][2] = 0;
            // targets of this wave in this pass: 16 nq slots (wave-uniform)
            const int nq = max(0, min(NT, (ct - tb - wbase + TL - 1) / TL));
            if (self) {
                if (nq == NT) adj_t_tile<true, NT, NT>(ts, g, fillp, selfk, xi, e, b, part);
                else if (nq == 3) adj_t_tile<true, NT, 3>(ts, g, fillp, selfk, xi, e, b, part);
                else if (nq == 2) adj_t_tile<true, NT, 2>(ts, g, fillp, selfk, xi, e, b, part);
                else if (nq == 1) adj_t_tile<true, NT, 1>(ts, g, fillp, selfk, xi, e, b, part);
            } else {
                if (nq == NT) adj_t_tile<false, NT, NT>(ts, g, fillp, 0, xi, e, b, part);
                else if (nq == 3) adj_t_tile<false, NT, 3>(ts, g, fillp, 0, xi, e, b, part);
                else if (nq == 2) adj_t_tile<false, NT, 2>(ts, g, fillp, 0, xi, e, b, part);
                else if (nq == 1) adj_t_tile<false, NT, 1>(ts, g, fillp, 0, xi, e, b, part);
            }
#pragma unroll
            for (int q = 0; q < NT; ++q)
#pragma unroll
                for (int d = 0; d < 3; ++d) acc[q][d] += (double)part[q][d];
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < NT; ++q)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (NT >= 4) acc[q][d] += __shfl_xor(acc[q][d], 16);
                acc[q][d] += __shfl_xor(acc[q][d], 32);
            }
        double o[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            o[d] = acc[0][d];
#pragma unroll
            for (int q = 1; q < NT; ++q)
                if (g == q) o[d] = acc[q][d];
        }
        const int i = tb + wbase + li + TL * g;
        if (i < ct) {
            const int64_t ii = a + i;
            for (int d = 0; d < 3; ++d) gx[3 * ii + d] += o[d];
        }
    }
}

// Source side: gx_j -= sum_i J^T u_ij;  gp_j += (e.K) p/g + K x b + dK/dp^T u.
// With g_j, p_j fixed per thread: sum_i d^-3/2 u = g sum d^-3/2 e + (sum d^-3/2 b) x p,
// R.u = g (e.R) + p.(R x b) (the R x b of dl/dm reused), c = d^-5/2 (R.u):
//   gx_j = -(g E3 + B3 x p) + 3 (sum c R + p sum c pr),
//   gp_j = (sum d^-3/2 e.R) p / g + sum d^-3/2 R x b - 3 sum c pr R.
template <bool MIXED>
__global__ __launch_bounds__(PT) void k_adj_p2p_s(const int32_t *__restrict__ leaves,
                                                  const int32_t *__restrict__ offT,
                                                  const uint64_t *__restrict__ keysT, int sb,
                                                  const int32_t *__restrict__ lo,
                                                  const int32_t *__restrict__ cnt,
                                                  const double *__restrict__ sp8,
                                                  const double *__restrict__ ob,
                                                  double *__restrict__ gx, double *__restrict__ gp) {
    using T = typename std::conditional<MIXED, float, double>::type;
    // sums: E3 [0-2], B3 [3-5], sum c R [6-8], sum c pr [9], sum y3 e.R [10],
    //       sum y3 R x b [11-13], sum c pr R [14-16]
    constexpr int NA = 17;
    // 2 sources per lane (32 lanes x 2 groups): FP64 keeps 17 FP64 sums per source;
    // MIXED 17 FP32 tile partials and 17 FP64 sums (4 sources per lane took 215
    // VGPRs, 2 waves per SIMD).  FP64 with one source per lane (94 VGPRs, 5 waves)
    // ran slower, 30.8 -> 33.4 ms at N = 2^22: twice the LDS reads per pair.
    constexpr int NG = 2, TL = 64 / NG;
    using Tgt = AdjTgt<MIXED>;
    extern __shared__ double4 adj_lds[];  // sized by the launch: blockDim.x records
    const int BT = blockDim.x, tid = threadIdx.x;
    Tgt *tt = reinterpret_cast<Tgt *>(adj_lds);
    const int lane = tid & 63, li = lane % TL, g = lane / TL, wbase = tid - lane;
    const int32_t s = leaves[xcd_item<XCD_RUN_NEAR>(blockIdx.x, gridDim.x)];
    const int32_t a = lo[s], cs = cnt[s];
    const int32_t beg = offT[s], end = offT[s + 1];
    const uint64_t mask = (1ull << sb) - 1ull;
    const double4 *p8 = reinterpret_cast<const double4 *>(sp8);
    for (int sb0 = 0; sb0 < cs; sb0 += BT) {
        // FP64: the sums of the lane's NG sources.  MIXED: each staged tile's FP32
        // partials meet across the groups (xor-shuffles) and the owning lane
        // (group g owns source li + TL g) adds them into its 17 FP64 sums.
        double xs[NG][3];
        double acc[MIXED ? 1 : NG][NA];
        T gs[NG], ps[NG][3];
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const int64_t jj = a + min(sb0 + wbase + li + TL * q, cs - 1);
            const double4 x4 = p8[2 * jj];
            const double4 p4 = p8[2 * jj + 1];
            xs[q][0] = x4.x;
            xs[q][1] = x4.y;
            xs[q][2] = x4.z;
            gs[q] = (T)x4.w;
            ps[q][0] = (T)p4.x;
            ps[q][1] = (T)p4.y;
            ps[q][2] = (T)p4.z;
        }
#pragma unroll
        for (int q = 0; q < (MIXED ? 1 : NG); ++q)
#pragma unroll
            for (int d = 0; d < NA; ++d) acc[q][d] = 0.0;
        for (int32_t ei = beg; ei < end; ++ei) {
            const int32_t t = (int32_t)(keysT[ei] & mask);
            const int32_t ta = lo[t], ct = cnt[t];
            for (int t0 = 0; t0 < ct; t0 += BT) {
                __syncthreads();
                const int i = t0 + tid;
                if (i < ct) {
                    const int64_t ii = ta + i;
                    const double4 x4 = p8[2 * ii];
                    tt[tid] = Tgt{x4.x, x4.y, x4.z, (T)ob[6 * ii], (T)ob[6 * ii + 1],
                                  (T)ob[6 * ii + 2], (T)ob[6 * ii + 3], (T)ob[6 * ii + 4],
                                  (T)ob[6 * ii + 5]};
                } else if (i == ct) {
                    tt[tid] = Tgt::pad();  // an odd tile padded to a multiple of NG
                }
                __syncthreads();
                const int m = min(BT, ct - t0);
                const int mp = (m + NG - 1) / NG * NG;  // <= BT (a multiple of 64)
                T part[NG][NA];
                if constexpr (MIXED)
#pragma unroll
                    for (int q = 0; q < NG; ++q)
#pragma unroll
                        for (int d = 0; d < NA; ++d) part[q][d] = 0;
                // the self pair can only occur when the staged leaf is the source leaf:
                // MIXED runs a check-free copy of the loop on the other leaves (FP64
                // keeps one copy, which needs fewer registers)
                const int selfk = s == t ? sb0 + wbase + li - t0 : -(1 << 30);
                auto tile = [&](auto diag) {
                for (int it = 0; it < mp / NG; ++it) {  // the same trip count in every lane
                    const int k = NG * it + g;
                    const Tgt tr = tt[k];
                    const T ex = tr.e0, ey = tr.e1, ez = tr.e2, bx = tr.b0, by = tr.b1, bz = tr.b2;
#pragma unroll
                    for (int q = 0; q < NG; ++q) {
                        const T px = ps[q][0], py = ps[q][1], pz = ps[q][2];
                        T R[3], pr;
                        const bool self = decltype(diag)::value && k == selfk + TL * q;
                        const T y = pair_geo<T>(tr.x - xs[q][0], tr.y - xs[q][1], tr.z - xs[q][2],
                                                px, py, pz, self, R, pr);
                        const T y2 = y * y, y3 = y2 * y;
                        const T s0 = fma(ex, R[0], fma(ey, R[1], ez * R[2]));
                        const T v0 = fma(R[1], bz, -(R[2] * by));
                        const T v1 = fma(R[2], bx, -(R[0] * bz));
                        const T v2 = fma(R[0], by, -(R[1] * bx));
                        const T Ru = fma(gs[q], s0, fma(px, v0, fma(py, v1, pz * v2)));
                        const T c = y3 * (y2 * Ru);
                        const T cp = c * pr;
                        T *P;
                        if constexpr (MIXED) P = part[q];
                        else P = acc[q];  // FP64: the sums themselves
                        P[0] = fma(y3, ex, P[0]);
                        P[1] = fma(y3, ey, P[1]);
                        P[2] = fma(y3, ez, P[2]);
                        P[3] = fma(y3, bx, P[3]);
                        P[4] = fma(y3, by, P[4]);
                        P[5] = fma(y3, bz, P[5]);
                        P[6] = fma(c, R[0], P[6]);
                        P[7] = fma(c, R[1], P[7]);
                        P[8] = fma(c, R[2], P[8]);
                        P[9] += cp;
                        P[10] = fma(y3, s0, P[10]);
                        P[11] = fma(y3, v0, P[11]);
                        P[12] = fma(y3, v1, P[12]);
                        P[13] = fma(y3, v2, P[13]);
                        P[14] = fma(cp, R[0], P[14]);
                        P[15] = fma(cp, R[1], P[15]);
                        P[16] = fma(cp, R[2], P[16]);
                    }
                }
                };
                if (!MIXED || s == t) tile(std::true_type{});
                else tile(std::false_type{});
                if constexpr (MIXED) {
#pragma unroll
                    for (int q = 0; q < NG; ++q)
#pragma unroll
                        for (int d = 0; d < NA; ++d) {
                            T v = part[q][d];
                            if (NG >= 4) v += __shfl_xor(v, 16);
                            v += __shfl_xor(v, 32);
                            if (g == q) acc[0][d] += (double)v;
                        }
                }
            }
        }
        double S[NA];
        if constexpr (MIXED) {
#pragma unroll
            for (int d = 0; d < NA; ++d) S[d] = acc[0][d];
        } else {
#pragma unroll
            for (int q = 0; q < NG; ++q)
#pragma unroll
                for (int d = 0; d < NA; ++d) {
                    if (NG >= 4) acc[q][d] += __shfl_xor(acc[q][d], 16);
                    if (NG >= 2) acc[q][d] += __shfl_xor(acc[q][d], 32);
                }
#pragma unroll
            for (int d = 0; d < NA; ++d) {
                S[d] = acc[0][d];
#pragma unroll
                for (int q = 1; q < NG; ++q)
                    if (g == q) S[d] = acc[q][d];
            }
        }
        const int j = sb0 + wbase + li + TL * g;
        if (j < cs) {
            const int64_t jj = a + j;
            const double4 x4 = p8[2 * jj];
            const double4 p4 = p8[2 * jj + 1];
            const double gj = x4.w, p0 = p4.x, p1 = p4.y, p2 = p4.z;
            const double cg = S[10] / gj;  // (sum d^-3/2 e.R) / gamma_j
            gx[3 * jj] += -(gj * S[0] + (S[4] * p2 - S[5] * p1)) + 3.0 * (S[6] + p0 * S[9]);
            gx[3 * jj + 1] += -(gj * S[1] + (S[5] * p0 - S[3] * p2)) + 3.0 * (S[7] + p1 * S[9]);
            gx[3 * jj + 2] += -(gj * S[2] + (S[3] * p1 - S[4] * p0)) + 3.0 * (S[8] + p2 * S[9]);
            gp[3 * jj] += fma(cg, p0, S[11]) - 3.0 * S[14];
            gp[3 * jj + 1] += fma(cg, p1, S[12]) - 3.0 * S[15];
            gp[3 * jj + 2] += fma(cg, p2, S[13]) - 3.0 * S[16];
        }
    }
}

__global__ void k_adj_prep(int64_t N, const int32_t *__restrict__ parind,
                           const double *__restrict__ Ebar, const double *__restrict__ Bbar,
                           double amp, double *__restrict__ ob, double *__restrict__ gx,
                           double *__restrict__ gp, double *__restrict__ gx2,
                           double *__restrict__ gp2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t q = parind[i];
    for (int d = 0; d < 3; ++d) {
        ob[6 * i + d] = amp * Ebar[3 * q + d];
        ob[6 * i + 3 + d] = amp * Bbar[3 * q + d];
        gx[3 * i + d] = 0.0;
        gp[3 * i + d] = 0.0;
        gx2[3 * i + d] = 0.0;
        gp2[3 * i + d] = 0.0;
    }
}

__global__ void k_fill0(double *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// dl/d(bbox) per node: a_d (min) and b_d (max) through X_m = (c_m (b-a))/2 + (a+b)/2.
__global__ void k_adj_bbox(int64_t nc, int np, const uint8_t *__restrict__ flags,
                           const double *__restrict__ Xbar, const double *__restrict__ Xbar2,
                           const double *__restrict__ Xbar3, const double *__restrict__ ctab,
                           double *__restrict__ abar) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    for (int d = 0; d < 3; ++d) {
        double sa = 0.0, sbm = 0.0;
        if (flags[k] & 1)
            for (int m = 0; m < np; ++m) {
                const int64_t j = (k * 3 + d) * np + m;
                const double xb = (Xbar[j] + Xbar2[j]) + Xbar3[j];
                sa += xb * ((1.0 - ctab[m]) / 2.0);
                sbm += xb * ((1.0 + ctab[m]) / 2.0);
            }
        abar[6 * k + d] = sa;
        abar[6 * k + 3 + d] = sbm;
    }
}

// Each particle collects the bbox gradients of the ancestors where it is the
// argmin / argmax (leaf to root), then the mean-momentum gradient; results go to the
// caller's original particle order.  One workgroup per leaf: the ancestors' extreme
// particles and their gradients (6 per level, the same for every particle of the
// leaf) are staged in LDS once.
constexpr int FIN_T = 64, FIN_MAXL = 32;  // leaf depth <= 29
__global__ __launch_bounds__(FIN_T) void k_adj_finish(
    const int32_t *__restrict__ leaves, const int32_t *__restrict__ lo,
    const int32_t *__restrict__ cnt, const int32_t *__restrict__ parind,
    const int32_t *__restrict__ ext, const double *__restrict__ abar,
    const double *__restrict__ pbar, int64_t N, const double *__restrict__ gx,
    const double *__restrict__ gp, const double *__restrict__ gx2, const double *__restrict__ gp2,
    double *__restrict__ gpos, double *__restrict__ gmom) {
    __shared__ int32_t sext[FIN_MAXL * 6];
    __shared__ double sab[FIN_MAXL * 6];
    const int32_t k = leaves[blockIdx.x];
    const int64_t a = lo[k];
    const int c = cnt[k];
    const int depth = 31 - __clz(k + 1);  // level of the leaf: ancestors 0 .. depth
    const int na = 6 * (depth + 1);
    for (int j = threadIdx.x; j < na; j += FIN_T) {  // entry j: ancestor j / 6 levels up
        int32_t anc = k;
        for (int u = 0; u < j / 6; ++u) anc = (anc - 1) / 2;
        sext[j] = ext[6 * (int64_t)anc + j % 6];
        sab[j] = abar[6 * (int64_t)anc + j % 6];
    }
    __syncthreads();
    const double p0 = pbar[0] / (double)N, p1 = pbar[1] / (double)N, p2 = pbar[2] / (double)N;
    for (int t = threadIdx.x; t < c; t += FIN_T) {
        const int64_t i = a + t;
        const int32_t p = parind[i];
        double g[3] = {gx[3 * i] + gx2[3 * i], gx[3 * i + 1] + gx2[3 * i + 1],
                       gx[3 * i + 2] + gx2[3 * i + 2]};
        for (int j = 0; j < na; ++j)
            if (sext[j] == p) g[j % 3] += sab[j];
        gpos[3 * (int64_t)p] = g[0];
        gpos[3 * (int64_t)p + 1] = g[1];
        gpos[3 * (int64_t)p + 2] = g[2];
        gmom[3 * (int64_t)p] = (gp[3 * i] + gp2[3 * i]) + p0;
        gmom[3 * (int64_t)p + 1] = (gp[3 * i + 1] + gp2[3 * i + 1]) + p1;
        gmom[3 * (int64_t)p + 2] = (gp[3 * i + 2] + gp2[3 * i + 2]) + p2;
    }
}

// Fixed-order sum of the per-target mean-momentum partials: 1024 lanes each sum a
// contiguous stretch with four running sums, then a tree over the lanes.
constexpr int PB_T = 1024;
__global__ __launch_bounds__(PB_T) void k_adj_pbar(const double *__restrict__ part, int64_t nc,
                                                   double *__restrict__ out) {
    __shared__ double s[3][PB_T];
    const int64_t per = (nc + PB_T - 1) / PB_T;
    const int64_t b = threadIdx.x * per, e = min(nc, b + per);
    double a[4][3] = {};
    int64_t k = b;
    for (; k + 4 <= e; k += 4)
#pragma unroll
        for (int u = 0; u < 4; ++u)
            for (int d = 0; d < 3; ++d) a[u][d] += part[3 * (k + u) + d];
    for (; k < e; ++k)
        for (int d = 0; d < 3; ++d) a[0][d] += part[3 * k + d];
    for (int d = 0; d < 3; ++d) s[d][threadIdx.x] = (a[0][d] + a[1][d]) + (a[2][d] + a[3][d]);
    for (int o = PB_T / 2; o > 0; o >>= 1) {
        __syncthreads();
        if (threadIdx.x < (unsigned)o)
            for (int d = 0; d < 3; ++d) s[d][threadIdx.x] += s[d][threadIdx.x + o];
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) out[d] = s[d][0];
}

// ----------------------------------------------------------------- driver
#define ADJ_NP_SWITCH(np, F)                            \
    switch (np) {                                       \
        case 2: F(2); break;                            \
        case 3: F(3); break;                            \
        case 4: F(4); break;                            \
        case 5: F(5); break;                            \
        case 6: F(6); break;                            \
        case 7: F(7); break;                            \
        case 8: F(8); break;                            \
        case 9: F(9); break;                            \
        case 10: F(10); break;                          \
        case 11: F(11); break;                          \
        default: fail(FMM4RB_EINVAL, "n out of range"); \
    }

static inline unsigned nbk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }


void run_adjoint(Ctx &c, const double *Ebar, const double *Bbar, double amp, double *gpos,
                 double *gmom, bool saved_locals) {
    const Shape &S = c.shape;
    hipStream_t st = c.stream;
    const int64_t N = c.N, nc = S.nc, NN = c.nn, np = c.np;
    const unsigned nleaf = (unsigned)S.leaves.size();
    c.d_ob.ensure(6 * (size_t)N * 8);
    c.d_gx.ensure(3 * (size_t)N * 8);
    c.d_gp.ensure(3 * (size_t)N * 8);
    c.d_gx2.ensure(3 * (size_t)N * 8);
    c.d_gp2.ensure(3 * (size_t)N * 8);
    c.d_lbar.ensure(6 * (size_t)NN * nc * 8);
    c.d_mbar.ensure(4 * (size_t)NN * nc * 8);
    c.d_xbar.ensure(3 * (size_t)np * nc * 8);
    c.d_xbar2.ensure(3 * (size_t)np * nc * 8);
    c.d_xbar3.ensure(3 * (size_t)np * nc * 8);
    c.d_pbar.ensure((3 * (size_t)nc + 3) * 8);
    c.d_abar.ensure(6 * (size_t)nc * 8);
    c.d_costab.ensure(MAX_NP * 8);
    for (int i = 0; i < MAX_NP; ++i)
        c.costab_host[i] = i < c.np ? cos(M_PI * (double)i / (double)c.n) : 0.0;  // cheb2.jl:4
    FMM_HIP(hipMemcpyAsync(c.d_costab.p, c.costab_host, sizeof c.costab_host, hipMemcpyHostToDevice,
                           st));
    // both lists regrouped by source in one sort; the target side's per-pair
    // source-node sums (n <= 6) go to the transposed position
    const bool ysplit = c.np <= 7;
    transpose_lists_joint(c, ysplit ? &c.d_m2lT_pos : nullptr);
    if (ysplit) c.d_ysp.ensure(3 * (size_t)c.np * (size_t)std::max<int64_t>(c.nm2l, 1) * 8);
    const uint32_t *posT = ysplit ? c.d_m2lT_pos.as<uint32_t>() : nullptr;
    double *ysp = ysplit ? c.d_ysp.as<double>() : nullptr;

    double *ob = c.d_ob.as<double>(), *gx = c.d_gx.as<double>(), *gp = c.d_gp.as<double>();
    double *gx2 = c.d_gx2.as<double>(), *gp2 = c.d_gp2.as<double>();
    double *lbar = c.d_lbar.as<double>(), *mbar = c.d_mbar.as<double>();
    double *xbar = c.d_xbar.as<double>();
    hipLaunchKernelGGL(k_adj_prep, dim3(nbk(N, 256)), dim3(256), 0, st, N,
                       c.d_parind.as<int32_t>(), Ebar, Bbar, amp, ob, gx, gp, gx2, gp2);
    FMM_LAUNCH_CHECK();
    const int32_t *leaves = c.d_leaves.as<int32_t>(), *lo = c.d_lo.as<int32_t>(),
                  *cnt = c.d_cnt.as<int32_t>();
    const double *sp8 = c.d_sp8.as<double>(), *ncoord = c.d_ncoord.as<double>();
    const bool mixed = c.precision == FMM4RB_MIXED;
    // P2P^T needs only the particles, the cotangents and the near-field lists: it
    // runs on the auxiliary stream into gx2/gp2 while this stream does the far field
    {
        hipStream_t sa = c.near_stream();
        FMM_HIP(hipEventRecord(c.evx[0], st));
        FMM_HIP(hipStreamWaitEvent(sa, c.evx[0], 0));
        FMM_HIP(hipEventRecord(c.evx[1], sa));
        int bt = ((S.maxleaf + 63) / 64) * 64;
        bt = std::max(64, std::min(bt, PT));
        const size_t per_t = mixed ? sizeof(P2PSrc<true>) : sizeof(P2PSrc<false>);
        const size_t per_s = mixed ? sizeof(AdjTgt<true>) : sizeof(AdjTgt<false>);
        // target side: tiles of up to 4 blocks of sources, within 12 KB of LDS per wave
        // (so that LDS does not cap the 4 waves per SIMD the registers allow)
        auto tcap = [&](size_t per) {
            int tc = bt;
            while (tc + bt <= 4 * bt && (size_t)(tc + bt) * per <= (size_t)(12 << 10) * (bt / 64))
                tc += bt;
            return tc;
        };
        // (the source side keeps one leaf per tile: packing leaves there raised its
        // registers from 136/156 to 166/240, mixed/FP64)
        const int tc_t = tcap(per_t);
        const size_t lds_t = (size_t)tc_t * per_t, lds_s = (size_t)bt * per_s;
        if (mixed) {
            hipLaunchKernelGGL(k_adj_p2p_t<true>, dim3(nleaf), dim3(bt), lds_t, sa, leaves,
                               c.d_p2p_off.as<int32_t>(), c.p2p_sorted, c.sb, lo, cnt, sp8, ob, gx2,
                               tc_t);
            FMM_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_adj_p2p_s<true>, dim3(nleaf), dim3(bt), lds_s, sa, leaves,
                               c.d_p2pT_off.as<int32_t>(), c.p2pT_sorted, c.sb, lo, cnt, sp8, ob,
                               gx2, gp2);
        } else {
            hipLaunchKernelGGL(k_adj_p2p_t<false>, dim3(nleaf), dim3(bt), lds_t, sa, leaves,
                               c.d_p2p_off.as<int32_t>(), c.p2p_sorted, c.sb, lo, cnt, sp8, ob, gx2,
                               tc_t);
            FMM_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_adj_p2p_s<false>, dim3(nleaf), dim3(bt), lds_s, sa, leaves,
                               c.d_p2pT_off.as<int32_t>(), c.p2pT_sorted, c.sb, lo, cnt, sp8, ob,
                               gx2, gp2);
        }
        FMM_LAUNCH_CHECK();
        FMM_HIP(hipEventRecord(c.evx[2], sa));
    }
    const int64_t nx = 3 * np * nc;
    hipLaunchKernelGGL(k_fill0, dim3(nbk(nx, 256)), dim3(256), 0, st, xbar, nx);
    FMM_LAUNCH_CHECK();
    double *xbar2 = c.d_xbar2.as<double>(), *xbar3 = c.d_xbar3.as<double>();
    hipLaunchKernelGGL(k_fill0, dim3(nbk(nx, 256)), dim3(256), 0, st, xbar2, nx);
    FMM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_fill0, dim3(nbk(nx, 256)), dim3(256), 0, st, xbar3, nx);
    FMM_LAUNCH_CHECK();
    double *loc = c.d_loc.as<double>(), *pbar = c.d_pbar.as<double>();
    // mixed precision: the M2L^T kernels read the float4 copy of the multipoles,
    // made by the upward pass
    const void *mpv = mixed ? (const void *)c.d_mp4f.p : (const void *)c.d_mp4.p;

    // 1. local cotangents (values only; they do not depend on the forward locals):
    //    L2P^T at the leaves, then L2L^T bottom-up
#define ADJ_L2PV(NP)                                                                                \
    hipLaunchKernelGGL((k_adj_leaf<NP, 0, 2>), dim3(nleaf), dim3(AT), 0, st, leaves, lo, cnt, sp8,  \
                       ncoord, loc, ob, lbar, gx, gp, xbar)
    ADJ_NP_SWITCH(c.np, ADJ_L2PV);
    FMM_LAUNCH_CHECK();
    for (int l = S.ml; l >= 1; --l) {
        const int64_t b = S.split_off[l - 1], e = S.split_off[l];
        if (e == b) continue;
        const int32_t *par = c.d_lvl_split.as<int32_t>() + b;
#define ADJ_L2LV(NP)                                                                                \
    hipLaunchKernelGGL((k_adj_transfer<NP, 0, 2>), dim3((unsigned)(e - b)), dim3(AT), 0, st, par, \
                       ncoord, lbar, loc, lbar, xbar, -1)
        ADJ_NP_SWITCH(c.np, ADJ_L2LV);
        FMM_LAUNCH_CHECK();
    }
    // 3. geometry of L2P^T and L2L^T (needs lbar and the final forward locals;
    //    latency-bound launches) on the third stream, into xbar3: with saved_locals
    //    the forward locals are in loc already, so it runs from here, beside P2P^T and
    //    the M2L^T target side; otherwise after the forward L2L (below), beside the
    //    M2L^T source side.  The L2L^T geometry of all levels in two launches (parents
    //    on even, then odd levels).
    hipStream_t sg = c.third_stream();
    const int64_t npar = S.split_off[S.ml];  // split nodes of levels 0 .. ml-1
    const int32_t *allpar = c.d_lvl_split.as<int32_t>();
    auto launch_geometry = [&]() {
        FMM_HIP(hipEventRecord(c.evy[0], st));
        FMM_HIP(hipStreamWaitEvent(sg, c.evy[0], 0));
#define ADJ_L2PG_T(NP, LT)                                                                         \
    hipLaunchKernelGGL((k_adj_leaf<NP, 0, 1, LT>), dim3(nleaf), dim3(LT), 0, sg, leaves, lo, cnt,  \
                       sp8, ncoord, loc, ob, lbar, gx, gp, xbar3)
#define ADJ_L2PG(NP)                    \
    if (S.maxleaf <= 64) {              \
        ADJ_L2PG_T(NP, 64);             \
    } else {                            \
        ADJ_L2PG_T(NP, AT);             \
    }
        ADJ_NP_SWITCH(c.np, ADJ_L2PG);
        FMM_LAUNCH_CHECK();
        for (int parity = 0; parity < 2 && npar > 0; ++parity) {
#define ADJ_L2LG(NP)                                                                               \
    hipLaunchKernelGGL((k_adj_transfer<NP, 0, 1>), dim3((unsigned)npar), dim3(AT), 0, sg, allpar, \
                       ncoord, lbar, loc, lbar, xbar3, parity)
            ADJ_NP_SWITCH(c.np, ADJ_L2LG);
            FMM_LAUNCH_CHECK();
        }
    };
    if (saved_locals) launch_geometry();
    // 2. forward M2L fused with the M2L^T target side, then the forward L2L; with
    //    saved_locals the final forward locals of the preceding field call are in
    //    loc already and only the M2L^T target side runs
    // the quadratic form of d in dx (|px| <= 0.1, as the forward M2L) for n <= 6
    const bool quad = std::fabs(c.pavg_host[0]) <= 0.1;
#define ADJ_M2LT_K(NP, FWD, MX, QD)                                                                \
    hipLaunchKernelGGL((k_adj_m2l_t<NP, FWD, MX, QD>), dim3((unsigned)nc), dim3(AShape<NP>::BLK), \
                       0, st, c.d_m2l_off.as<int32_t>(), c.m2l_sorted, c.sb, ncoord, mpv, lbar,    \
                       c.d_pavg.as<double>(), xbar, pbar, loc, posT, ysp)
#define ADJ_M2LT_F(NP, FWD)                          \
    if (mixed && quad)                               \
        ADJ_M2LT_K(NP, FWD, true, (NP <= 7));        \
    else if (mixed)                                  \
        ADJ_M2LT_K(NP, FWD, true, false);            \
    else if (quad)                                   \
        ADJ_M2LT_K(NP, FWD, false, (NP <= 7));       \
    else                                             \
        ADJ_M2LT_K(NP, FWD, false, false)
#define ADJ_M2LT(NP) ADJ_M2LT_F(NP, true)
#define ADJ_M2LT_SAVED(NP) ADJ_M2LT_F(NP, false)
    if (saved_locals) {
        ADJ_NP_SWITCH(c.np, ADJ_M2LT_SAVED);
        FMM_LAUNCH_CHECK();
    } else {
        ADJ_NP_SWITCH(c.np, ADJ_M2LT);
        FMM_LAUNCH_CHECK();
        launch_l2l(c);
        launch_geometry();
    }
    FMM_HIP(hipEventRecord(c.ev[8], st));
    FMM_HIP(hipStreamWaitEvent(sg, c.ev[8], 0));  // pbar: the target side's partials
    // the mean-momentum partials of the M2L^T target side, summed (one workgroup)
    hipLaunchKernelGGL(k_adj_pbar, dim3(1), dim3(PB_T), 0, sg, pbar, nc, pbar + 3 * nc);
    FMM_LAUNCH_CHECK();
    FMM_HIP(hipEventRecord(c.evy[1], sg));
    // 4. M2L^T source side
#define ADJ_M2LS_K(NP, MX, QD)                                                                    \
    hipLaunchKernelGGL((k_adj_m2l_s<NP, MX, QD>), dim3((unsigned)nc), dim3(AShape<NP>::BLK), 0, st, \
                       c.d_m2lT_off.as<int32_t>(), c.m2lT_sorted, c.sb, ncoord, mpv, lbar,          \
                       c.d_pavg.as<double>(), mbar, xbar2, ysp)
#define ADJ_M2LS(NP)                        \
    if (mixed && quad)                      \
        ADJ_M2LS_K(NP, true, (NP <= 7));    \
    else if (mixed)                         \
        ADJ_M2LS_K(NP, true, false);        \
    else if (quad)                          \
        ADJ_M2LS_K(NP, false, (NP <= 7));   \
    else                                    \
        ADJ_M2LS_K(NP, false, false)
    ADJ_NP_SWITCH(c.np, ADJ_M2LS);
    FMM_LAUNCH_CHECK();
    FMM_HIP(hipEventRecord(c.ev[9], st));
    // 5. M2M^T values top-down (parents on level l-1 for l = 1 .. ml), then the
    //    M2M^T geometry of all levels (needs the final Mbar) in two launches.  The
    //    geometry here and P2M^T (6.) add their node-coordinate terms into xbar2
    //    after the M2L^T source side and their particle terms into gx2/gp2 after
    //    P2P^T, so this stream does not wait for the third one until the
    //    bounding-box step; every buffer keeps one fixed order of additions.
    for (int l = 1; l <= S.ml; ++l) {
        const int64_t b = S.split_off[l - 1], e = S.split_off[l];
        if (e == b) continue;
        const int32_t *par = c.d_lvl_split.as<int32_t>() + b;
#define ADJ_M2M(NP)                                                                                 \
    hipLaunchKernelGGL((k_adj_transfer<NP, 1, 2>), dim3((unsigned)(e - b)), dim3(AT), 0, st, par, \
                       ncoord, c.d_mp4.as<double>(), mbar, mbar, xbar, -1)
        ADJ_NP_SWITCH(c.np, ADJ_M2M);
        FMM_LAUNCH_CHECK();
    }
    for (int parity = 0; parity < 2 && npar > 0; ++parity) {
#define ADJ_M2MG(NP)                                                                               \
    hipLaunchKernelGGL((k_adj_transfer<NP, 1, 1>), dim3((unsigned)npar), dim3(AT), 0, st, allpar, \
                       ncoord, c.d_mp4.as<double>(), mbar, mbar, xbar2, parity)
        ADJ_NP_SWITCH(c.np, ADJ_M2MG);
        FMM_LAUNCH_CHECK();
    }
    // 6. P2M^T
    FMM_HIP(hipStreamWaitEvent(st, c.evx[2], 0));  // join the P2P^T stream (gx2, gp2)
#define ADJ_P2M_T(NP, LT)                                                                          \
    hipLaunchKernelGGL((k_adj_leaf<NP, 1, 3, LT>), dim3(nleaf), dim3(LT), 0, st, leaves, lo, cnt,  \
                       sp8, ncoord, mbar, ob, (double *)nullptr, gx2, gp2, xbar2)
#define ADJ_P2M(NP)                     \
    if (S.maxleaf <= 64) {              \
        ADJ_P2M_T(NP, 64);              \
    } else {                            \
        ADJ_P2M_T(NP, AT);              \
    }
    ADJ_NP_SWITCH(c.np, ADJ_P2M);
    FMM_LAUNCH_CHECK();
    // bounding boxes -> extreme particles, mean momentum, output order
    FMM_HIP(hipStreamWaitEvent(st, c.evy[1], 0));  // join the third stream (xbar, gx, gp)
    hipLaunchKernelGGL(k_adj_bbox, dim3(nbk(nc, 128)), dim3(128), 0, st, nc, c.np,
                       c.d_flags.as<uint8_t>(), xbar, xbar2, xbar3, c.d_costab.as<double>(),
                       c.d_abar.as<double>());
    FMM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_adj_finish, dim3(nleaf), dim3(FIN_T), 0, st, leaves, lo, cnt,
                       c.d_parind.as<int32_t>(), c.d_ext.as<int32_t>(), c.d_abar.as<double>(),
                       pbar + 3 * nc, N, gx, gp, gx2, gp2, gpos, gmom);
    FMM_LAUNCH_CHECK();
}

}  // namespace fmm4rb
