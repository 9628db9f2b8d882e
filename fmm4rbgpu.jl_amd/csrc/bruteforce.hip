// bruteforce.hip -- direct O(N * targets) field sum in FP64 on gfx950.
//
// Reference: update_particles_field!(particles, ::BruteForce; lambda)
// (src/update_particles_field.jl:21-38), the accuracy oracle of the FMM.
// Same expression order as the reference: K = R / sqrt(R.R + (p.R)^2 + eps)^3,
// E += (amp * gamma_j) * K, B += amp * (p_j x K), sources in index order.
// One thread per target; sources are staged through LDS in tiles.
#include "fmm4rb_internal.hpp"

#pragma clang fp contract(off)

namespace fmm4rb {

constexpr int BF_T = 256;

__global__ __launch_bounds__(BF_T) void k_bruteforce(const double *__restrict__ pos,
                                                     const double *__restrict__ mom, int64_t N,
                                                     const int64_t *__restrict__ targets, int64_t nt,
                                                     double amp, double *__restrict__ E,
                                                     double *__restrict__ B) {
    __shared__ double tile[BF_T][8];
    const int64_t a = (int64_t)blockIdx.x * BF_T + threadIdx.x;
    const bool act = a < nt;
    const int64_t i = act ? (targets ? targets[a] : a) : 0;
    const double xi0 = pos[3 * i], xi1 = pos[3 * i + 1], xi2 = pos[3 * i + 2];
    double e[3] = {0.0, 0.0, 0.0}, b[3] = {0.0, 0.0, 0.0};
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
    if (act)
        for (int d = 0; d < 3; ++d) {
            E[3 * a + d] = e[d];
            B[3 * a + d] = b[d];
        }
}

void launch_bruteforce(const double *pos, const double *mom, int64_t N, const int64_t *targets,
                       int64_t nt, double amp, double *E, double *B, hipStream_t st) {
    hipLaunchKernelGGL(k_bruteforce, dim3((unsigned)((nt + BF_T - 1) / BF_T)), dim3(BF_T), 0, st,
                       pos, mom, N, targets, nt, amp, E, B);
    FMM_LAUNCH_CHECK();
}

}  // namespace fmm4rb
