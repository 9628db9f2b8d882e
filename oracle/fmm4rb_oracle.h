/*
 * fmm4rb_oracle.h -- CPU restatement of FMM4RBGPU.jl's field-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  It is never linked into, or called by, the product path.
 *
 * Every function restates one function of the reference (src/... file:line
 * given per function).  Conventions that differ from the 1-based Julia code:
 *   - node k (0-based) is reference node k+1; children of k are 2k+1, 2k+2;
 *     parent of k is (k-1)/2 (root: -1);
 *   - particle indices and ranges are 0-based, ranges inclusive [lo, hi];
 *   - macro-particle arrays keep the reference's column-major layout
 *     (n+1, n+1, n+1, nc) with x fastest; 3-vectors are stored as 3
 *     consecutive doubles per entry (Array{SVector{3,Float64},4} in memory).
 * Arithmetic follows the reference's expression order; build with
 * -ffp-contract=off so no multiply-add is fused.
 */
#ifndef FMM4RB_ORACLE_H
#define FMM4RB_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* lgweight's rule when x lies within eps of interpolation nodes:
 * ORC_SINGULAR_ALL  = indicator of every matching node (src/fmm/upwardpass.jl:6-8)
 * ORC_SINGULAR_LAST = indicator of the last matching node (src/gpu_fmm/gpu_upwardpass.jl:52-54) */
enum { ORC_SINGULAR_ALL = 0, ORC_SINGULAR_LAST = 1 };

/* src/cluster_tree/treeinfo.jl:6-26 */
int64_t orc_maxlevel(int64_t N, int64_t N0);
int64_t orc_ncluster(int64_t N, int64_t N0);

/* src/utils/bbox.jl:9-15; bbox = {min x,y,z, max x,y,z} */
double orc_diameter(const double bbox[6], const double stretch[3]);
int orc_inside(const double child[6], const double parent[6]);

/* src/utils/partition.jl:10-22 (Lomuto) with an explicit pivot and a
 * (data, label) total order; data[label*stride] is the key of label. */
int64_t orc_partition(int64_t *labels, int64_t lo, int64_t hi, const double *data,
                      int64_t stride, int64_t pvtindex);
/* src/utils/partition.jl:29-43; returns k */
int64_t orc_kpermute(int64_t *labels, int64_t lo, int64_t hi, const double *data,
                     int64_t stride, int64_t k);

/* src/utils/cheb2.jl:3-5 */
void orc_cheb2(int n, double a, double b, double *out);
/* src/fmm/upwardpass.jl:4-15 (rule selects the singular case, see above) */
void orc_lgweight(double x, const double *nodes, int n, int rule, double *out);

/* p_avg = sum(momenta)/N with Julia's pairwise mapreduce order
 * (src/update_particles_field.jl:45). mom is AoS 3N. */
void orc_p_avg(int64_t N, const double *mom, double out[3]);

/* src/cluster_tree/cluster_tree.jl:12-18, clusters.jl:21-70.
 * pos AoS 3N.  Outputs sized nc = orc_ncluster(N,N0): lo/hi (inclusive,
 * -1 for nodes below an early leaf), levels, parents, children[2nc],
 * bboxes[6nc], exists[nc]; parindices[N] with each leaf's indices in
 * ascending order (the canonical order; the reference's is RNG-dependent). */
int orc_build_tree(int64_t N, const double *pos, int64_t N0, const double stretch[3],
                   int64_t *parindices, int64_t *lo, int64_t *hi, int64_t *levels,
                   int64_t *parents, int64_t *children, double *bboxes, uint8_t *exists);

/* src/fmm/interaction_lists.jl:10-21 */
int orc_admissible(const double b1[6], const double b2[6], const double stretch[3], double eta);
/* src/fmm/interaction_lists.jl:23-58: pairs (target, source) in DFS order.
 * Returns 0 and fills the lists if capacities suffice, else 1; *np2p and
 * *nm2l always receive the true counts. */
int orc_lists(int64_t nc, const int64_t *children, const double *bboxes,
              const double stretch[3], double eta, int64_t cap_p2p, int64_t *p2p,
              int64_t cap_m2l, int64_t *m2l, int64_t *np2p, int64_t *nm2l);

/* Operators (src/fmm/upwardpass.jl, interaction.jl, downwardpass.jl).
 * nodes: the node ids to process (P2M/L2P: leaves; M2M: children whose
 * parents are updated; L2L: children).  gam[(n+1)^3 nc], vec3 arrays 3x. */
void orc_p2m(int n, int rule, int64_t nnodes, const int64_t *nodes, const int64_t *parindices,
             const int64_t *lo, const int64_t *hi, const double *bboxes, const double *pos,
             const double *mom, double *gam, double *mmom);
void orc_m2m(int n, int rule, int64_t nnodes, const int64_t *nodes, const int64_t *parents,
             const double *bboxes, double *gam, double *mmom);
void orc_m2l(int n, int64_t npairs, const int64_t *pairs, const double *bboxes,
             const double p_avg[3], const double *gam, const double *mmom, double *le,
             double *lb);
void orc_p2p(int64_t npairs, const int64_t *pairs, const int64_t *parindices,
             const int64_t *lo, const int64_t *hi, const double *pos, const double *mom,
             double *E, double *B);
void orc_l2l(int n, int rule, int64_t nnodes, const int64_t *nodes, const int64_t *parents,
             const double *bboxes, double *le, double *lb);
void orc_l2p(int n, int rule, int64_t nnodes, const int64_t *nodes, const int64_t *parindices,
             const int64_t *lo, const int64_t *hi, const double *bboxes, const double *pos,
             const double *le, const double *lb, double *E, double *B);

/* Threads for the operators' outer loops (0 = the OpenMP default).  Results
 * are bit-identical for every thread count: groups own disjoint outputs and
 * keep the serial order of additions. */
void orc_set_threads(int n);

/* The serial FMM driver (src/update_particles_field.jl:40-64).  E/B are
 * overwritten (the reference accumulates into whatever they hold).  If
 * p_avg_in is non-NULL it replaces the computed mean momentum.  Returns 0. */
int orc_fmm(int64_t N, const double *pos, const double *mom, double charge, double lambda,
            int n, int64_t N0, double eta, const double *p_avg_in, int rule, double *E,
            double *B);

/* BruteForce (src/update_particles_field.jl:21-38) for the listed targets
 * (all N if targets == NULL).  E/B hold 3 doubles per listed target. */
void orc_bruteforce(int64_t N, const double *pos, const double *mom, double charge,
                    double lambda, int64_t ntargets, const int64_t *targets, int nthreads,
                    double *E, double *B);

#ifdef __cplusplus
}
#endif
#endif
