// bruteforce.hip -- direct O(N * targets) field sum in FP64 on gfx950.
//
// Reference: update_particles_field!(particles, ::BruteForce; lambda)
// (src/update_particles_field.jl:21-38), the accuracy oracle of the FMM.
// Same expression order as the reference: K = R / sqrt(R.R + (p.R)^2 + eps)^3,
// E += (amp * gamma_j) * K, B += amp * (p_j x K), sources in index order.
// One thread per target; sources are staged through LDS in tiles.  With few
// targets (e.g. 4096 sampled targets of a large bunch) the source range is split
// into C chunks, one workgroup per (target block, chunk), so that the launch
// fills the GPU; the C partial sums of each target are added by a second kernel
// in chunk order (deterministic; the order of additions differs from the single
// running sum only by that final grouping).
#include <algorithm>

#include "fmm4rb_internal.hpp"

#pragma clang fp contract(off)

namespace fmm4rb {

constexpr int BF_T = 256;

// Block (x, c): targets x * BF_T .. + BF_T over sources [c * chunk, (c + 1) * chunk).
// One chunk (gridDim.y == 1): the sums go to E, B; else to part[c][a][6].
__global__ __launch_bounds__(BF_T) void k_bruteforce(const double *__restrict__ pos,
                                                     const double *__restrict__ mom, int64_t N,
                                                     const int64_t *__restrict__ targets, int64_t nt,
                                                     double amp, int64_t chunk,
                                                     double *__restrict__ E,
                                                     double *__restrict__ B,
                                                     double *__restrict__ part) {
    __shared__ double tile[BF_T][8];
    const int64_t a = (int64_t)blockIdx.x * BF_T + threadIdx.x;
    const bool act = a < nt;
    const int64_t i = act ? (targets ? targets[a] : a) : 0;
    const double xi0 = pos[3 * i], xi1 = pos[3 * i + 1], xi2 = pos[3 * i + 2];
    double e[3] = {0.0, 0.0, 0.0}, b[3] = {0.0, 0.0, 0.0};
    const int64_t jb = (int64_t)blockIdx.y * chunk, je = min(N, jb + chunk);
    for (int64_t j0 = jb; j0 < je; j0 += BF_T) {
        __syncthreads();
        const int64_t j = j0 + threadIdx.x;
        if (j < je) {
            const double m0 = mom[3 * j], m1 = mom[3 * j + 1], m2 = mom[3 * j + 2];
            tile[threadIdx.x][0] = pos[3 * j];
            tile[threadIdx.x][1] = pos[3 * j + 1];
            tile[threadIdx.x][2] = pos[3 * j + 2];
            tile[threadIdx.x][3] = sqrt(1.0 + ((m0 * m0 + m1 * m1) + m2 * m2));
            tile[threadIdx.x][4] = m0;
            tile[threadIdx.x][5] = m1;
            tile[threadIdx.x][6] = m2;
        }
        __syncthreads();
        const int m = (int)min<int64_t>(BF_T, je - j0);
        if (act) {
            for (int jj = 0; jj < m; ++jj) {
                const double *t = tile[jj];
                const double R0 = xi0 - t[0], R1 = xi1 - t[1], R2 = xi2 - t[2];
                const double rr = (R0 * R0 + R1 * R1) + R2 * R2;
                const double pr = (t[4] * R0 + t[5] * R1) + t[6] * R2;
                const double s = sqrt((rr + pr * pr) + EPS64);
                const double s3 = (s * s) * s;
                const double K0 = R0 / s3, K1 = R1 / s3, K2 = R2 / s3;
                const double ag = amp * t[3];
                e[0] += ag * K0;
                e[1] += ag * K1;
                e[2] += ag * K2;
                b[0] += amp * (t[5] * K2 - t[6] * K1);
                b[1] += amp * (t[6] * K0 - t[4] * K2);
                b[2] += amp * (t[4] * K1 - t[5] * K0);
            }
        }
    }
    if (!act) return;
    if (gridDim.y == 1) {
        for (int d = 0; d < 3; ++d) {
            E[3 * a + d] = e[d];
            B[3 * a + d] = b[d];
        }
    } else {
        double *o = part + ((int64_t)blockIdx.y * nt + a) * 6;
        for (int d = 0; d < 3; ++d) {
            o[d] = e[d];
            o[3 + d] = b[d];
        }
    }
}

// E, B of target a = the chunks' partial sums added in chunk order.
__global__ __launch_bounds__(BF_T) void k_bf_reduce(const double *__restrict__ part, int nch,
                                                    int64_t nt, double *__restrict__ E,
                                                    double *__restrict__ B) {
    const int64_t a = (int64_t)blockIdx.x * BF_T + threadIdx.x;
    if (a >= nt) return;
    double v[6];
    for (int d = 0; d < 6; ++d) v[d] = part[a * 6 + d];
    for (int c = 1; c < nch; ++c)
        for (int d = 0; d < 6; ++d) v[d] += part[((int64_t)c * nt + a) * 6 + d];
    for (int d = 0; d < 3; ++d) {
        E[3 * a + d] = v[d];
        B[3 * a + d] = v[3 + d];
    }
}

// Reverse mode of BruteForce: for every ordered pair (i, j), u = g_j e_i + b_i x p_j,
// G = d^-3/2 u - 3 d^-5/2 (R.u)(R + (p_j.R) p_j); gpos_i += G, gpos_j -= G,
// gmom_j += (e_i.K) p_j/g_j + K x b_i - 3 d^-5/2 (p_j.R)(R.u) R.  The self pair adds
// nothing; coincident distinct particles add the regularised kernel's derivative.
// ob holds amp * (Ebar, Bbar) per particle (original order).  Target side:
__global__ __launch_bounds__(BF_T) void k_bf_vjp_t(const double *__restrict__ pos,
                                                   const double *__restrict__ mom, int64_t N,
                                                   const double *__restrict__ ob,
                                                   double *__restrict__ gpos) {
    __shared__ double tile[BF_T][8];
    const int64_t i = (int64_t)blockIdx.x * BF_T + threadIdx.x;
    const bool act = i < N;
    const int64_t ii = act ? i : 0;
    const double x0 = pos[3 * ii], x1 = pos[3 * ii + 1], x2 = pos[3 * ii + 2];
    const double e0 = ob[6 * ii], e1 = ob[6 * ii + 1], e2 = ob[6 * ii + 2];
    const double b0 = ob[6 * ii + 3], b1 = ob[6 * ii + 4], b2 = ob[6 * ii + 5];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int64_t j0 = 0; j0 < N; j0 += BF_T) {
        __syncthreads();
        const int64_t j = j0 + threadIdx.x;
        if (j < N) {
            const double m0 = mom[3 * j], m1 = mom[3 * j + 1], m2 = mom[3 * j + 2];
            tile[threadIdx.x][0] = pos[3 * j];
            tile[threadIdx.x][1] = pos[3 * j + 1];
            tile[threadIdx.x][2] = pos[3 * j + 2];
            tile[threadIdx.x][3] = sqrt(1.0 + ((m0 * m0 + m1 * m1) + m2 * m2));
            tile[threadIdx.x][4] = m0;
            tile[threadIdx.x][5] = m1;
            tile[threadIdx.x][6] = m2;
        }
        __syncthreads();
        const int m = (int)min<int64_t>(BF_T, N - j0);
        if (act)
            for (int jj = 0; jj < m; ++jj) {
                const double *t = tile[jj];
                if (j0 + jj == i) continue;  // self pair
                const double R0 = x0 - t[0], R1 = x1 - t[1], R2 = x2 - t[2];
                const double px = t[4], py = t[5], pz = t[6], g = t[3];
                const double pr = (px * R0 + py * R1) + pz * R2;
                const double d = ((R0 * R0 + R1 * R1) + R2 * R2) + pr * pr + EPS64;
                const double y = 1.0 / sqrt(d);
                const double y3 = y * y * y, y5 = y3 * y * y;
                const double u0 = g * e0 + (b1 * pz - b2 * py);
                const double u1 = g * e1 + (b2 * px - b0 * pz);
                const double u2 = g * e2 + (b0 * py - b1 * px);
                const double Ru = (R0 * u0 + R1 * u1) + R2 * u2;
                const double c3 = 3.0 * y5 * Ru;
                g0 += y3 * u0 - c3 * (R0 + pr * px);
                g1 += y3 * u1 - c3 * (R1 + pr * py);
                g2 += y3 * u2 - c3 * (R2 + pr * pz);
            }
    }
    if (act) {
        gpos[3 * i] = g0;
        gpos[3 * i + 1] = g1;
        gpos[3 * i + 2] = g2;
    }
}

// Source side (adds to gpos written by the target side).
__global__ __launch_bounds__(BF_T) void k_bf_vjp_s(const double *__restrict__ pos,
                                                   const double *__restrict__ mom, int64_t N,
                                                   const double *__restrict__ ob,
                                                   double *__restrict__ gpos,
                                                   double *__restrict__ gmom) {
    __shared__ double tile[BF_T][9];
    const int64_t j = (int64_t)blockIdx.x * BF_T + threadIdx.x;
    const bool act = j < N;
    const int64_t jj = act ? j : 0;
    const double x0 = pos[3 * jj], x1 = pos[3 * jj + 1], x2 = pos[3 * jj + 2];
    const double px = mom[3 * jj], py = mom[3 * jj + 1], pz = mom[3 * jj + 2];
    const double g = sqrt(1.0 + ((px * px + py * py) + pz * pz));
    double gx0 = 0.0, gx1 = 0.0, gx2 = 0.0, gp0 = 0.0, gp1 = 0.0, gp2 = 0.0;
    for (int64_t i0 = 0; i0 < N; i0 += BF_T) {
        __syncthreads();
        const int64_t i = i0 + threadIdx.x;
        if (i < N) {
            for (int d = 0; d < 3; ++d) tile[threadIdx.x][d] = pos[3 * i + d];
            for (int d = 0; d < 6; ++d) tile[threadIdx.x][3 + d] = ob[6 * i + d];
        }
        __syncthreads();
        const int m = (int)min<int64_t>(BF_T, N - i0);
        if (act)
            for (int k = 0; k < m; ++k) {
                const double *t = tile[k];
                if (i0 + k == j) continue;  // self pair
                const double R0 = t[0] - x0, R1 = t[1] - x1, R2 = t[2] - x2;
                const double e0 = t[3], e1 = t[4], e2 = t[5], b0 = t[6], b1 = t[7], b2 = t[8];
                const double pr = (px * R0 + py * R1) + pz * R2;
                const double d = ((R0 * R0 + R1 * R1) + R2 * R2) + pr * pr + EPS64;
                const double y = 1.0 / sqrt(d);
                const double y3 = y * y * y, y5 = y3 * y * y;
                const double u0 = g * e0 + (b1 * pz - b2 * py);
                const double u1 = g * e1 + (b2 * px - b0 * pz);
                const double u2 = g * e2 + (b0 * py - b1 * px);
                const double Ru = (R0 * u0 + R1 * u1) + R2 * u2;
                const double c3 = 3.0 * y5 * Ru;
                gx0 -= y3 * u0 - c3 * (R0 + pr * px);
                gx1 -= y3 * u1 - c3 * (R1 + pr * py);
                gx2 -= y3 * u2 - c3 * (R2 + pr * pz);
                const double eK = y3 * ((e0 * R0 + e1 * R1) + e2 * R2);
                const double cg = eK / g, cp = -c3 * pr;
                gp0 += cg * px + cp * R0 + y3 * (R1 * b2 - R2 * b1);
                gp1 += cg * py + cp * R1 + y3 * (R2 * b0 - R0 * b2);
                gp2 += cg * pz + cp * R2 + y3 * (R0 * b1 - R1 * b0);
            }
    }
    if (act) {
        gpos[3 * j] += gx0;
        gpos[3 * j + 1] += gx1;
        gpos[3 * j + 2] += gx2;
        gmom[3 * j] = gp0;
        gmom[3 * j + 1] = gp1;
        gmom[3 * j + 2] = gp2;
    }
}

void launch_bruteforce_vjp(const double *pos, const double *mom, int64_t N, const double *ob,
                           double *gpos, double *gmom, hipStream_t st) {
    const unsigned g = (unsigned)((N + BF_T - 1) / BF_T);
    hipLaunchKernelGGL(k_bf_vjp_t, dim3(g), dim3(BF_T), 0, st, pos, mom, N, ob, gpos);
    FMM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bf_vjp_s, dim3(g), dim3(BF_T), 0, st, pos, mom, N, ob, gpos, gmom);
    FMM_LAUNCH_CHECK();
}

void launch_bruteforce(const double *pos, const double *mom, int64_t N, const int64_t *targets,
                       int64_t nt, double amp, double *E, double *B, DevBuf &part,
                       hipStream_t st) {
    const int64_t tb = (nt + BF_T - 1) / BF_T;
    // chunks of whole tiles: about 2048 workgroups in all (8 per CU), at most 1024
    // chunks, each at least one tile
    const int64_t tiles = (N + BF_T - 1) / BF_T;
    const int64_t want = std::min<int64_t>(1024, std::max<int64_t>(1, 2048 / tb));
    const int64_t tpc = (tiles + want - 1) / want;  // tiles per chunk
    const int64_t chunk = tpc * BF_T;
    const int nch = (int)((N + chunk - 1) / chunk);
    if (nch > 1) part.ensure((size_t)nch * (size_t)nt * 6 * sizeof(double));
    hipLaunchKernelGGL(k_bruteforce, dim3((unsigned)tb, (unsigned)nch), dim3(BF_T), 0, st, pos,
                       mom, N, targets, nt, amp, chunk, E, B,
                       nch > 1 ? part.as<double>() : (double *)nullptr);
    FMM_LAUNCH_CHECK();
    if (nch > 1) {
        hipLaunchKernelGGL(k_bf_reduce, dim3((unsigned)tb), dim3(BF_T), 0, st, part.as<double>(),
                           nch, nt, E, B);
        FMM_LAUNCH_CHECK();
    }
}

}  // namespace fmm4rb
