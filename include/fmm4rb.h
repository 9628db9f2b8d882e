/*
 * fmm4rb.h -- C ABI of the MI355X (gfx950) interpolation-based FMM for the
 * relativistic space-charge field of a charged-particle beam.
 *
 * It is the drop-in boundary for FMM4RBGPU.jl's
 *     update_particles_field!(beam, alg; lambda)
 * (src/update_particles_field.jl:1-148): a new algorithm type with the same
 * (eta, N0, n) parameters calls fmm4rb_field() through ccall.
 *
 * Conventions
 *   - Every entry point returns 0 (FMM4RB_OK) or a positive error code;
 *     fmm4rb_last_error() gives a thread-local message.  Nothing is written
 *     to outputs on error.
 *   - Particle arrays are AoS: 3 doubles per particle, exactly the memory of
 *     Julia's Vector{SVector{3,Float64}} (reinterpret(Float64, v)).
 *   - Tree nodes are 0-based heap ids: node k is the reference's node k+1,
 *     children of k are 2k+1 and 2k+2.  Particle ranges are 0-based inclusive.
 *   - Calls are synchronous.  One context is not thread-safe; distinct
 *     contexts are.  A context owns a HIP stream and its device scratch,
 *     grown on demand and kept between calls.
 */
#ifndef FMM4RB_H
#define FMM4RB_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define FMM4RB_OK 0
#define FMM4RB_EINVAL 1      /* invalid argument (sizes, parameters, NaN/Inf input) */
#define FMM4RB_ENOMEM 2      /* device allocation failed or exceeds the context cap */
#define FMM4RB_EDEVICE 3     /* HIP runtime or kernel launch error */

#define FMM4RB_FP64 0        /* all arithmetic in FP64 */
#define FMM4RB_MIXED 1       /* M2L and P2P sums in FP32, geometry and final sums in FP64 */

typedef struct fmm4rb_ctx fmm4rb_ctx;

/* Per-call counters and stage times (HIP events).  The near field (P2P, and
 * P2P^T in reverse mode) runs on a second stream of the context, concurrently
 * with the far-field stages, and the upward pass on a third one beside the list
 * build: ms_p2p and ms_upward are their own durations, the other stages include
 * the time they share the GPU, ms_downward includes waiting for P2P, and ms_total
 * is the wall time of the whole call. */
typedef struct {
    int64_t N, N0, ncluster, nleaf;
    int32_t n, precision, maxlevel, traversal_iterations;
    int64_t np2p, nm2l;                 /* leaf pairs, cluster pairs */
    int64_t nm2l_groups;                /* distinct M2L targets */
    int64_t p2p_particle_pairs;         /* sum over P2P pairs of n_t * n_s */
    double p_avg[3];                    /* mean momentum used (stretch, M2L kernel) */
    double g_avg;                       /* sqrt(1 + p_avg . p_avg) */
    float ms_total, ms_pavg, ms_tree, ms_lists, ms_upward, ms_m2l, ms_p2p, ms_downward;
} fmm4rb_stats;

const char *fmm4rb_version(void);
const char *fmm4rb_last_error(void);

/* No reference counterpart: the reference uses CUDA.jl's implicit context. */
/* A context owns three HIP streams (default priority): its own, one for the near
 * field, and one for the upward pass beside the list build (and the reverse mode's
 * L2P^T/L2L^T geometry), each forked from and joined back into the stream the call
 * runs on. */
int fmm4rb_create(fmm4rb_ctx **ctx, int device);
int fmm4rb_destroy(fmm4rb_ctx *ctx);
/* Run the context's work on a caller stream (hipStream_t; NULL restores the
 * context's own stream). */
int fmm4rb_set_stream(fmm4rb_ctx *ctx, void *stream);
/* No reference counterpart: on = 0 runs the helper streams' work in place on the
 * call's stream, one kernel at a time (results bit-identical to the default on = 1,
 * slower).  For timing and profiling single kernels, whose durations are otherwise
 * stretched by the kernels that share the GPU with them. */
int fmm4rb_set_overlap(fmm4rb_ctx *ctx, int on);

/*
 * Field evaluation, replaces update_particles_field!(beam, FMMGPU(n,N0,eta); lambda)
 * (src/update_particles_field.jl:66-148).  pos, mom: 3N doubles (host).
 * E, B: 3N doubles (host), overwritten with the fields scaled by
 * 2.8179403699772166e-15*charge/lambda.  Limits: 1 <= n <= 10,
 * 1 <= N0 <= 1024, eta >= 0, N >= 1.  stats may be NULL.
 */
int fmm4rb_field(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                 double charge, double lambda, int n, int64_t N0, double eta,
                 int precision, double *E, double *B, fmm4rb_stats *stats);

/* Same with device pointers (inputs already resident in HBM); the call is
 * ordered on the context stream and returns after the results are written. */
int fmm4rb_field_device(fmm4rb_ctx *ctx, int64_t N, const double *d_pos, const double *d_mom,
                        double charge, double lambda, int n, int64_t N0, double eta,
                        int precision, double *d_E, double *d_B, fmm4rb_stats *stats);

/*
 * Allocation-free variant of fmm4rb_field_device.  No reference counterpart: it
 * is for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc); like every entry point it synchronizes the
 * host (list capacities), so it cannot itself be graph-captured.  Same
 * arguments, semantics and results as fmm4rb_field_device (bit-identical), but
 * every device buffer of the call is carved from a static scratch array of
 * 256 MiB per device (each device the library runs on holds it until the
 * process ends); nothing calls hipMalloc, hipFree or another allocating
 * routine.  A call whose buffers exceed the scratch fails with FMM4RB_ENOMEM
 * before any kernel runs (about N <= 2^18 at n = 4, fewer at larger n); if the
 * interaction lists outgrow their capacity estimate (64 pairs per leaf) it can
 * fail the same way after the traversal.  Either way E and B are not written.
 * The scratch is shared by every context and thread on a device and keeps the
 * last call's contents: calls that overlap on one device race.  The context's
 * own buffers are untouched; its tree, lists and locals count as stale
 * afterwards (a following refit or pullback fails with FMM4RB_EINVAL).
 */
int fmm4rb_field_device_noalloc(fmm4rb_ctx *ctx, int64_t N, const double *d_pos,
                                const double *d_mom, double charge, double lambda, int n,
                                int64_t N0, double eta, int precision, double *d_E, double *d_B,
                                fmm4rb_stats *stats);

/*
 * Tracking steps (no reference counterpart: the reference rebuilds the tree on
 * every call, update_particles_field.jl:96-104).  The field for new positions and
 * momenta on the tree and interaction lists of this context's last completed
 * fmm4rb_field* or fmm4rb_field_vjp* call, which must have had the same N, N0 and
 * eta (else FMM4RB_EINVAL; n and precision may differ).  Kept: the particle
 * order, cluster ranges, splits and both lists.  Recomputed: p_avg, every
 * cluster's bounding box (min/max over its particles, clusters.jl:51-70), the
 * interpolation nodes and all expansions.  With the inputs of that call the
 * result is bit-identical to it; as particles drift the partition goes stale
 * and the error grows until the caller rebuilds with fmm4rb_field*.  Skips the
 * tree build and the traversal (about 3 of 26.6 ms at N = 2^22 mixed).
 */
int fmm4rb_field_refit(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                       double charge, double lambda, int n, int64_t N0, double eta,
                       int precision, double *E, double *B, fmm4rb_stats *stats);
int fmm4rb_field_device_refit(fmm4rb_ctx *ctx, int64_t N, const double *d_pos,
                              const double *d_mom, double charge, double lambda, int n,
                              int64_t N0, double eta, int precision, double *d_E, double *d_B,
                              fmm4rb_stats *stats);

/*
 * Reverse mode (no reference counterpart): the vector-Jacobian product of the
 * field map at fixed tree and lists.  Given cotangents Ebar, Bbar (3N each,
 * of the fields as fmm4rb_field returns them), writes gpos = sum Ebar.dE/dpos
 * + Bbar.dB/dpos and gmom likewise (3N each), including the dependence of the
 * interpolation nodes on the cluster bounding boxes and of the M2L kernel on
 * the mean momentum.  The tree and lists are rebuilt, not stored.  For a loss
 * L = sum_i w_i (|E_i|^2 + |B_i|^2), pass Ebar = 2 w E, Bbar = 2 w B.
 * FMM4RB_MIXED evaluates the M2L^T and P2P^T kernel terms in FP32 (differences
 * formed in FP64, sums added into FP64 per source block); the transfers, the
 * interpolation-weight and bounding-box terms stay FP64.  stats: ms_m2l =
 * transposed lists, local cotangents and the
 * forward M2L fused with the M2L^T target side plus the forward L2L; ms_p2p =
 * P2P^T (both sides, on the near-field stream); ms_downward = L2P^T/L2L^T
 * geometry (on the third stream) beside the M2L^T source side, then M2M^T, P2M^T
 * and the gradient assembly.
 */
int fmm4rb_field_vjp(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                     double charge, double lambda, int n, int64_t N0, double eta, int precision,
                     const double *Ebar, const double *Bbar, double *gpos, double *gmom,
                     fmm4rb_stats *stats);
int fmm4rb_field_vjp_device(fmm4rb_ctx *ctx, int64_t N, const double *d_pos, const double *d_mom,
                            double charge, double lambda, int n, int64_t N0, double eta,
                            int precision, const double *d_Ebar, const double *d_Bbar,
                            double *d_gpos, double *d_gmom, fmm4rb_stats *stats);
/*
 * Pullback after a forward call (no reference counterpart; the pullback of the
 * rrule): the same reverse mode on the tree, lists, multipoles and final local
 * expansions of this context's last call, which must have been a completed
 * fmm4rb_field or fmm4rb_field_device (not a refit) with the same inputs (checked by
 * a 64-bit digest of pos and mom), N, N0, n, eta and precision; else FMM4RB_EINVAL.
 * Skips the tree, list and upward-pass builds and the forward M2L and L2L that
 * fmm4rb_field_vjp recomputes.  Results equal
 * fmm4rb_field_vjp's up to the rounding of the forward locals (FP64 ~1e-15,
 * mixed ~1e-7 relative).  Any other call on the context between the two (except
 * another pullback) makes it fail.
 */
int fmm4rb_field_pullback(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                          double charge, double lambda, int n, int64_t N0, double eta,
                          int precision, const double *Ebar, const double *Bbar, double *gpos,
                          double *gmom, fmm4rb_stats *stats);
int fmm4rb_field_pullback_device(fmm4rb_ctx *ctx, int64_t N, const double *d_pos,
                                 const double *d_mom, double charge, double lambda, int n,
                                 int64_t N0, double eta, int precision, const double *d_Ebar,
                                 const double *d_Bbar, double *d_gpos, double *d_gmom,
                                 fmm4rb_stats *stats);
/* Reverse mode of BruteForce (all N targets), FP64, host pointers. */
int fmm4rb_bruteforce_vjp(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                          double charge, double lambda, const double *Ebar, const double *Bbar,
                          double *gpos, double *gmom);

/*
 * BruteForce on the GPU (src/update_particles_field.jl:21-38) for ntargets
 * particle indices (all N if targets == NULL, ntargets ignored).  Host
 * pointers; E, B hold 3 doubles per target.  FP64.
 */
int fmm4rb_bruteforce(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                      double charge, double lambda, int64_t ntargets, const int64_t *targets,
                      double *E, double *B);

/*
 * Cluster tree on the GPU: ClusterTree(particles; N0, stretch)
 * (src/cluster_tree/cluster_tree.jl:12-18, clusters.jl:21-70) with
 * stretch = (1, 1, sqrt(1+|p_avg|^2)) from the mean momentum
 * (update_particles_field.jl:77-79).  Host outputs, nc = 2^(maxlevel+1)-1:
 * parindices[N] (each leaf ascending), lo[nc], hi[nc] (-1 for nodes below
 * an early leaf), bboxes[6nc] {min xyz, max xyz}, children[2nc] (-1: leaf
 * or absent), p_avg[3].  Any output pointer may be NULL.
 */
int fmm4rb_tree(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom, int64_t N0,
                int64_t *parindices, int64_t *lo, int64_t *hi, double *bboxes,
                int64_t *children, double *p_avg);

/*
 * Interaction lists on the GPU: InteractionListsGPU(clusters; stretch, eta)
 * (src/gpu_fmm/gpu_interaction_lists.jl:27-40, traversal interaction_lists.jl:23-49)
 * for the tree fmm4rb_tree builds.  Pairs (target, source) sorted by
 * target then source.  Writes at most cap_* pairs; *np2p, *nm2l always get
 * the full counts.
 */
int fmm4rb_lists(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom, int64_t N0,
                 double eta, int64_t cap_p2p, int64_t *p2p, int64_t cap_m2l, int64_t *m2l,
                 int64_t *np2p, int64_t *nm2l);

/*
 * Operator entry points with the arguments of the reference kernels, on host
 * arrays in the reference's macro-particle layout ((n+1)^3 values per
 * cluster, x fastest; 3-vectors as 3 consecutive doubles).
 * M2L: gpu_M2L! (src/gpu_fmm/gpu_interaction.jl:3-41); le/lb (3*(n+1)^3*nc)
 *      are overwritten with the M2L sums of the listed pairs.
 * P2P: gpu_P2P! (src/gpu_fmm/gpu_interaction.jl:43-86); E/B (3N) are
 *      overwritten with the P2P sums (no amplitude factor).  The targets of
 *      the pairs must have disjoint particle ranges (else FMM4RB_EINVAL), as
 *      leaves do.  bboxes is accepted for symmetry with the tree hook and may
 *      be NULL.
 */
int fmm4rb_op_m2l(fmm4rb_ctx *ctx, int n, int precision, int64_t nc, const double *bboxes,
                  const double *gam, const double *mmom, int64_t npairs, const int64_t *pairs,
                  const double *p_avg, double *le, double *lb);
int fmm4rb_op_p2p(fmm4rb_ctx *ctx, int precision, int64_t N, const double *pos,
                  const double *mom, const int64_t *parindices, int64_t nc, const int64_t *lo,
                  const int64_t *hi, const double *bboxes, int64_t npairs, const int64_t *pairs,
                  double *E, double *B);

/*
 * Intermediate arrays of the last fmm4rb_field* call on this context, for
 * stage-by-stage parity checks (no reference counterpart).  Requires
 * fmm4rb_set_debug(ctx, 1) before that call.  what:
 *   1 multipoles: out = gam[(n+1)^3 nc] then mmom[3 (n+1)^3 nc]
 *   2 locals after M2L, before L2L: le then lb (3 (n+1)^3 nc each)
 *   3 locals after L2L: le then lb
 *   4 near field (P2P sums, particle order, no amplitude): E[3N] then B[3N]
 * count = number of doubles out can hold.
 */
int fmm4rb_set_debug(fmm4rb_ctx *ctx, int on);
int fmm4rb_debug_fetch(fmm4rb_ctx *ctx, int what, double *out, int64_t count);

#ifdef __cplusplus
}
#endif
#endif
