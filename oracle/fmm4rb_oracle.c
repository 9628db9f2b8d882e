/*
 * fmm4rb_oracle.c -- CPU restatement of FMM4RBGPU.jl's field-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product.  Each function cites the reference function it restates
 * (paths relative to the FMM4RBGPU.jl repository).  Compile with
 * -ffp-contract=off: the reference's arithmetic has no fused multiply-adds.
 *
 * Deviations from the reference, all deliberate and documented per function:
 *   - kpermute! orders keys by (value, particle index) and pivots on the
 *     middle element; the reference pivots at random (partition.jl:10), so
 *     its split membership is the same for distinct keys and random on ties;
 *   - each leaf's particle indices are finally sorted ascending (the
 *     reference's within-leaf order depends on its unseeded RNG);
 *   - bounding-box values get "+ 0.0" so -0.0 and +0.0 both read +0.0;
 *   - nodes below an early leaf (a cluster with <= N0 particles above the
 *     deepest level) are marked non-existent and skipped; the reference reads
 *     uninitialised entries there (upwardpass.jl:102, downwardpass.jl:87).
 */
#include "fmm4rb_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS64 DBL_EPSILON /* Julia eps() / eps(Float64) */
static const double AMP_RE = 2.8179403699772166e-15; /* update_particles_field.jl:28 */

/* ------------------------------------------------------------- threading */

/* The operators below run their outer loops on OpenMP threads over groups
 * that own disjoint outputs (a target cluster, a parent, a leaf), each group
 * visiting its inputs in the serial order: every output receives the same
 * additions in the same order as in the serial loops, so results are
 * bit-identical for any thread count.  When a call's outputs are not disjoint
 * between groups (overlapping particle ranges, a node listed with its
 * parent), the loop runs serially. */
static int g_threads = 0; /* 0: all the runtime's threads */

void orc_set_threads(int n) { g_threads = n < 0 ? 0 : n; }

static int nthreads_for(int64_t work) {
#ifdef _OPENMP
    int t = g_threads > 0 ? g_threads : omp_get_max_threads();
    if (work < 2) t = 1;
    return t;
#else
    (void)work;
    return 1;
#endif
}

typedef struct {
    int64_t key, pos;
} keypos;

static int cmp_keypos(const void *a, const void *b) {
    const keypos *x = (const keypos *)a, *y = (const keypos *)b;
    if (x->key != y->key) return (x->key > y->key) - (x->key < y->key);
    return (x->pos > y->pos) - (x->pos < y->pos);
}

/* Stable grouping of n items by key: order[] lists item positions grouped by
 * key, each group in the items' original order; goff[g..g+1] bounds group g.
 * Returns the number of groups (-1 if out of memory). */
static int64_t group_by_key(int64_t n, const int64_t *keys, int64_t stride, int64_t **order,
                            int64_t **goff) {
    keypos *kp = malloc(sizeof(keypos) * (size_t)(n > 0 ? n : 1));
    *order = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    *goff = malloc(sizeof(int64_t) * (size_t)(n + 1));
    if (!kp || !*order || !*goff) {
        free(kp); free(*order); free(*goff);
        *order = *goff = NULL;
        return -1;
    }
    for (int64_t i = 0; i < n; ++i) {
        kp[i].key = keys[i * stride];
        kp[i].pos = i;
    }
    qsort(kp, (size_t)n, sizeof(keypos), cmp_keypos);
    int64_t ng = 0;
    for (int64_t i = 0; i < n; ++i) {
        (*order)[i] = kp[i].pos;
        if (i == 0 || kp[i].key != kp[i - 1].key) (*goff)[ng++] = i;
    }
    (*goff)[ng] = n;
    free(kp);
    return ng;
}

/* 1 if the particle ranges [lo[k], hi[k]] of the listed nodes are pairwise
 * disjoint (a node listed twice counts as overlapping). */
static int ranges_disjoint(int64_t n, const int64_t *nodes, const int64_t *lo, const int64_t *hi) {
    int64_t *order = NULL, *goff = NULL;
    int64_t *l = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    if (!l) return 0;
    for (int64_t a = 0; a < n; ++a) l[a] = lo[nodes[a]];
    int64_t ng = group_by_key(n, l, 1, &order, &goff);
    int ok = ng == n;
    for (int64_t g = 1; ok && g < ng; ++g)
        ok = lo[nodes[order[goff[g]]]] > hi[nodes[order[goff[g - 1]]]];
    free(l); free(order); free(goff);
    return ok;
}

/* 1 if no listed node is the parent of another listed node. */
static int no_parent_listed(int64_t n, const int64_t *nodes, const int64_t *parents) {
    int64_t mx = 0;
    for (int64_t a = 0; a < n; ++a) mx = nodes[a] > mx ? nodes[a] : mx;
    uint8_t *in = calloc((size_t)mx + 1, 1);
    if (!in) return 0;
    for (int64_t a = 0; a < n; ++a) in[nodes[a]] = 1;
    int ok = 1;
    for (int64_t a = 0; ok && a < n; ++a) {
        int64_t p = parents[nodes[a]];
        ok = !(p >= 0 && p <= mx && in[p]);
    }
    free(in);
    return ok;
}

/* ---------------------------------------------------------------- treeinfo */

/* treeinfo.jl:6-8: ceil(log(N/N0)/log(2)) if N > N0 else 0 */
int64_t orc_maxlevel(int64_t N, int64_t N0) {
    if (N <= N0) return 0;
    double r = log((double)N / (double)N0) / log(2.0);
    return (int64_t)ceil(r);
}

/* treeinfo.jl:10-13 */
int64_t orc_ncluster(int64_t N, int64_t N0) {
    int64_t ml = orc_maxlevel(N, N0);
    return ((int64_t)1 << (ml + 1)) - 1;
}

/* -------------------------------------------------------------------- bbox */

/* StaticArrays norm of a 3-vector: sqrt((a1^2 + a2^2) + a3^2). */
static double norm3(double a, double b, double c) { return sqrt((a * a + b * b) + c * c); }

/* bbox.jl:9-11: norm(stretch .* (bmax - bmin) ./ 2.0) */
double orc_diameter(const double bb[6], const double s[3]) {
    double v[3];
    for (int d = 0; d < 3; ++d) v[d] = (s[d] * (bb[3 + d] - bb[d])) / 2.0;
    return norm3(v[0], v[1], v[2]);
}

/* bbox.jl:13-15 */
int orc_inside(const double c[6], const double p[6]) {
    for (int d = 0; d < 3; ++d)
        if (!(p[d] <= c[d])) return 0;
    for (int d = 0; d < 3; ++d)
        if (!(p[3 + d] >= c[3 + d])) return 0;
    return 1;
}

/* --------------------------------------------------------------- partition */

static int key_less(const double *data, int64_t stride, int64_t a, int64_t b) {
    double va = data[a * stride], vb = data[b * stride];
    return va < vb || (va == vb && a < b);
}

/* partition.jl:10-22 (Lomuto).  The comparison is the strict (value, label)
 * order, so keys that tie in value are ordered by particle index. */
int64_t orc_partition(int64_t *labels, int64_t lo, int64_t hi, const double *data,
                      int64_t stride, int64_t pvtindex) {
    int64_t pv = labels[pvtindex], t;
    labels[pvtindex] = labels[hi];
    labels[hi] = pv;
    int64_t j = lo;
    for (int64_t i = lo; i < hi; ++i) {
        if (key_less(data, stride, labels[i], pv)) {
            t = labels[j]; labels[j] = labels[i]; labels[i] = t;
            ++j;
        }
    }
    t = labels[hi]; labels[hi] = labels[j]; labels[j] = t;
    return j;
}

/* partition.jl:29-43 (quickselect); pivot = middle element. */
int64_t orc_kpermute(int64_t *labels, int64_t lo, int64_t hi, const double *data,
                     int64_t stride, int64_t k) {
    for (;;) {
        if (lo == hi) return lo;
        int64_t p = orc_partition(labels, lo, hi, data, stride, lo + (hi - lo) / 2);
        if (p == k) return k;
        if (p > k) hi = p - 1;
        else lo = p + 1;
    }
}

/* ------------------------------------------------------------ cheb2/weights */

/* cheb2.jl:3-5: [cos(pi*i/n) for i=0:n] .* (b-a) / 2.0 .+ (a+b)/2.0 */
void orc_cheb2(int n, double a, double b, double *out) {
    for (int i = 0; i <= n; ++i) {
        double c = cos(M_PI * (double)i / (double)n);
        out[i] = (c * (b - a)) / 2.0 + (a + b) / 2.0;
    }
}

/* upwardpass.jl:4-15 (barycentric Lagrange weights at x). */
void orc_lgweight(double x, const double *nodes, int n, int rule, double *out) {
    int any = 0, last = -1;
    for (int i = 0; i <= n; ++i) {
        if (fabs(x - nodes[i]) < EPS64) { any = 1; last = i; }
    }
    if (any) {
        for (int i = 0; i <= n; ++i) {
            int hit = fabs(x - nodes[i]) < EPS64;
            out[i] = (rule == ORC_SINGULAR_ALL) ? (hit ? 1.0 : 0.0) : (i == last ? 1.0 : 0.0);
        }
        return;
    }
    double sum = 0.0;
    for (int i = 0; i <= n; ++i) {
        double w = (i % 2 == 0) ? 1.0 : -1.0;
        if (i == 0) w = 0.5;
        if (i == n) w = 0.5 * w;
        out[i] = w / (x - nodes[i]);
    }
    sum = out[0];
    for (int i = 1; i <= n; ++i) sum = sum + out[i];
    for (int i = 0; i <= n; ++i) out[i] = out[i] / sum;
}

/* ------------------------------------------------------------------- p_avg */

/* Julia Base mapreduce_impl: pairwise halves down to blocks of <= 1024,
 * each summed left to right. */
static void pw_sum(const double *m, int64_t f, int64_t l, double v[3]) {
    if (f == l) {
        for (int d = 0; d < 3; ++d) v[d] = m[3 * f + d];
    } else if (l - f < 1024) {
        for (int d = 0; d < 3; ++d) v[d] = m[3 * f + d] + m[3 * (f + 1) + d];
        for (int64_t i = f + 2; i <= l; ++i)
            for (int d = 0; d < 3; ++d) v[d] = v[d] + m[3 * i + d];
    } else {
        int64_t mid = f + ((l - f) >> 1);
        double a[3], b[3];
        pw_sum(m, f, mid, a);
        pw_sum(m, mid + 1, l, b);
        for (int d = 0; d < 3; ++d) v[d] = a[d] + b[d];
    }
}

/* update_particles_field.jl:45: sum(momenta) / npar */
void orc_p_avg(int64_t N, const double *mom, double out[3]) {
    double s[3] = {0.0, 0.0, 0.0};
    if (N > 0) pw_sum(mom, 0, N - 1, s);
    for (int d = 0; d < 3; ++d) out[d] = s[d] / (double)N;
}

/* -------------------------------------------------------------------- tree */

typedef struct {
    const double *pos;
    int64_t N0;
    const double *stretch;
    int64_t *parindices, *lo, *hi, *levels, *parents, *children;
    double *bboxes;
    uint8_t *exists;
} tree_ctx;

/* clusters.jl:51-70 */
static void bbox_of(const double *pos, const int64_t *pi, int64_t lo, int64_t hi, double bb[6]) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = lo; i <= hi; ++i) {
        const double *x = pos + 3 * pi[i];
        for (int d = 0; d < 3; ++d) {
            mn[d] = x[d] < mn[d] ? x[d] : mn[d];
            mx[d] = x[d] > mx[d] ? x[d] : mx[d];
        }
    }
    for (int d = 0; d < 3; ++d) {
        bb[d] = mn[d] + 0.0;
        bb[3 + d] = mx[d] + 0.0;
    }
}

/* clusters.jl:21-49 */
static void fill(tree_ctx *c, int64_t node, int64_t lo, int64_t hi, int64_t level) {
    double *bb = c->bboxes + 6 * node;
    bbox_of(c->pos, c->parindices, lo, hi, bb);
    c->exists[node] = 1;
    c->levels[node] = level;
    c->lo[node] = lo;
    c->hi[node] = hi;
    c->parents[node] = node == 0 ? -1 : (node - 1) / 2;
    int64_t npar = hi - lo + 1;
    if (npar <= c->N0) {
        c->children[2 * node] = -1;
        c->children[2 * node + 1] = -1;
        return;
    }
    /* argmax(stretch .* (bmax - bmin)), first maximum wins */
    int dir = 0;
    double best = c->stretch[0] * (bb[3] - bb[0]);
    for (int d = 1; d < 3; ++d) {
        double v = c->stretch[d] * (bb[3 + d] - bb[d]);
        if (v > best) { best = v; dir = d; }
    }
    /* k = floor((lo+hi)/2) in 1-based terms: left gets ceil(npar/2) */
    int64_t k = lo + (npar + 1) / 2 - 1;
    orc_kpermute(c->parindices, lo, hi, c->pos + dir, 3, k);
    int64_t c1 = 2 * node + 1, c2 = 2 * node + 2;
    c->children[2 * node] = c1;
    c->children[2 * node + 1] = c2;
    fill(c, c1, lo, k, level + 1);
    fill(c, c2, k + 1, hi, level + 1);
}

static int cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* cluster_tree.jl:12-18 */
int orc_build_tree(int64_t N, const double *pos, int64_t N0, const double stretch[3],
                   int64_t *parindices, int64_t *lo, int64_t *hi, int64_t *levels,
                   int64_t *parents, int64_t *children, double *bboxes, uint8_t *exists) {
    if (N < 1 || N0 < 1) return 1;
    int64_t nc = orc_ncluster(N, N0);
    for (int64_t i = 0; i < N; ++i) parindices[i] = i;
    for (int64_t k = 0; k < nc; ++k) {
        lo[k] = hi[k] = levels[k] = parents[k] = -1;
        children[2 * k] = children[2 * k + 1] = -1;
        for (int d = 0; d < 6; ++d) bboxes[6 * k + d] = 0.0;
        exists[k] = 0;
    }
    tree_ctx c = {pos, N0, stretch, parindices, lo, hi, levels, parents, children, bboxes, exists};
    fill(&c, 0, 0, N - 1, 0);
    for (int64_t k = 0; k < nc; ++k)
        if (exists[k] && children[2 * k] < 0)
            qsort(parindices + lo[k], (size_t)(hi[k] - lo[k] + 1), sizeof(int64_t), cmp_i64);
    return 0;
}

/* ------------------------------------------------------- interaction lists */

/* interaction_lists.jl:10-21 */
int orc_admissible(const double b1[6], const double b2[6], const double s[3], double eta) {
    double c1[3], c2[3], v1[3], v2[3], v[3];
    for (int d = 0; d < 3; ++d) {
        c1[d] = (b1[d] + b1[3 + d]) / 2.0;
        c2[d] = (b2[d] + b2[3 + d]) / 2.0;
    }
    for (int d = 0; d < 3; ++d) {
        v1[d] = s[d] * (b1[3 + d] - c1[d]);
        v2[d] = s[d] * (b2[3 + d] - c2[d]);
        v[d] = s[d] * (c1[d] - c2[d]);
    }
    double r1 = norm3(v1[0], v1[1], v1[2]);
    double r2 = norm3(v2[0], v2[1], v2[2]);
    double R = norm3(v[0], v[1], v[2]);
    double r = r1 < r2 ? r2 : r1; /* Julia max for non-NaN values */
    return (r / R) < eta;
}

typedef struct {
    const int64_t *ch;
    const double *bb;
    const double *s;
    double eta;
    int64_t cap_p2p, cap_m2l, np2p, nm2l;
    int64_t *p2p, *m2l;
} lists_ctx;

/* interaction_lists.jl:23-49 */
static void dual(lists_ctx *c, int64_t t, int64_t s) {
    int64_t t1 = c->ch[2 * t], t2 = c->ch[2 * t + 1];
    int64_t s1 = c->ch[2 * s], s2 = c->ch[2 * s + 1];
    int tleaf = t1 < 0, sleaf = s1 < 0;
    if (tleaf && sleaf) {
        if (c->np2p < c->cap_p2p) { c->p2p[2 * c->np2p] = t; c->p2p[2 * c->np2p + 1] = s; }
        c->np2p++;
        return;
    }
    if (orc_admissible(c->bb + 6 * t, c->bb + 6 * s, c->s, c->eta)) {
        if (c->nm2l < c->cap_m2l) { c->m2l[2 * c->nm2l] = t; c->m2l[2 * c->nm2l + 1] = s; }
        c->nm2l++;
    } else if (tleaf) {
        dual(c, t, s1);
        dual(c, t, s2);
    } else if (sleaf) {
        dual(c, t1, s);
        dual(c, t2, s);
    } else if (orc_diameter(c->bb + 6 * t, c->s) > orc_diameter(c->bb + 6 * s, c->s)) {
        dual(c, t1, s);
        dual(c, t2, s);
    } else {
        dual(c, t, s1);
        dual(c, t, s2);
    }
}

/* interaction_lists.jl:51-58 */
int orc_lists(int64_t nc, const int64_t *children, const double *bboxes, const double stretch[3],
              double eta, int64_t cap_p2p, int64_t *p2p, int64_t cap_m2l, int64_t *m2l,
              int64_t *np2p, int64_t *nm2l) {
    (void)nc;
    lists_ctx c = {children, bboxes, stretch, eta, cap_p2p, cap_m2l, 0, 0, p2p, m2l};
    dual(&c, 0, 0);
    *np2p = c.np2p;
    *nm2l = c.nm2l;
    return (c.np2p > cap_p2p || c.nm2l > cap_m2l) ? 1 : 0;
}

/* --------------------------------------------------------------- operators */

#define IDX(i, j, k, node, np) ((i) + (np) * ((j) + (np) * ((k) + (np) * (int64_t)(node))))

/* upwardpass.jl:17-52 */
void orc_p2m(int n, int rule, int64_t nnodes, const int64_t *nodes, const int64_t *pind,
             const int64_t *lo, const int64_t *hi, const double *bboxes, const double *pos,
             const double *mom, double *gam, double *mmom) {
    int np = n + 1;
    /* each listed node owns its expansion: nodes in parallel when distinct */
    int par = ranges_disjoint(nnodes, nodes, lo, hi);
#pragma omp parallel for schedule(dynamic, 4) num_threads(par ? nthreads_for(nnodes) : 1)
    for (int64_t a = 0; a < nnodes; ++a) {
        double xc[11], yc[11], zc[11], wx[11], wy[11], wz[11];
        int64_t node = nodes[a];
        const double *bb = bboxes + 6 * node;
        orc_cheb2(n, bb[0], bb[3], xc);
        orc_cheb2(n, bb[1], bb[4], yc);
        orc_cheb2(n, bb[2], bb[5], zc);
        for (int64_t p = lo[node]; p <= hi[node]; ++p) {
            int64_t q = pind[p];
            const double *x = pos + 3 * q, *m = mom + 3 * q;
            double ga = sqrt(1.0 + ((m[0] * m[0] + m[1] * m[1]) + m[2] * m[2]));
            orc_lgweight(x[0], xc, n, rule, wx);
            orc_lgweight(x[1], yc, n, rule, wy);
            orc_lgweight(x[2], zc, n, rule, wz);
            for (int k = 0; k < np; ++k)
                for (int j = 0; j < np; ++j)
                    for (int i = 0; i < np; ++i) {
                        double w = (wx[i] * wy[j]) * wz[k];
                        int64_t id = IDX(i, j, k, node, np);
                        gam[id] += ga * w;
                        for (int d = 0; d < 3; ++d) mmom[3 * id + d] += m[d] * w;
                    }
        }
    }
}

/* upwardpass.jl:56-99 (scatter each child's nodes into its parent) */
void orc_m2m(int n, int rule, int64_t nnodes, const int64_t *nodes, const int64_t *parents,
             const double *bboxes, double *gam, double *mmom) {
    int np = n + 1;
    /* the nodes grouped by parent, each group in list order: one parent's
     * additions come from its group only, in the serial order */
    int64_t *order = NULL, *goff = NULL, *pk = malloc(sizeof(int64_t) * (size_t)(nnodes + 1));
    int64_t ng = -1;
    if (pk) {
        for (int64_t a = 0; a < nnodes; ++a) pk[a] = parents[nodes[a]];
        ng = group_by_key(nnodes, pk, 1, &order, &goff);
    }
    int par = ng >= 0 && no_parent_listed(nnodes, nodes, parents);
    if (!par) ng = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(par ? nthreads_for(ng) : 1)
    for (int64_t g = 0; g < ng; ++g)
    for (int64_t e = par ? goff[g] : 0; e < (par ? goff[g + 1] : nnodes); ++e) {
        double cx[11], cy[11], cz[11], px[11], py[11], pz[11], wx[11], wy[11], wz[11];
        int64_t a = par ? order[e] : e;
        int64_t node = nodes[a], pnode = parents[node];
        const double *cb = bboxes + 6 * node, *pb = bboxes + 6 * pnode;
        orc_cheb2(n, cb[0], cb[3], cx);
        orc_cheb2(n, cb[1], cb[4], cy);
        orc_cheb2(n, cb[2], cb[5], cz);
        orc_cheb2(n, pb[0], pb[3], px);
        orc_cheb2(n, pb[1], pb[4], py);
        orc_cheb2(n, pb[2], pb[5], pz);
        for (int ck = 0; ck < np; ++ck) {
            orc_lgweight(cz[ck], pz, n, rule, wz);
            for (int cj = 0; cj < np; ++cj) {
                orc_lgweight(cy[cj], py, n, rule, wy);
                for (int ci = 0; ci < np; ++ci) {
                    orc_lgweight(cx[ci], px, n, rule, wx);
                    int64_t cid = IDX(ci, cj, ck, node, np);
                    double cg = gam[cid], cm[3] = {mmom[3 * cid], mmom[3 * cid + 1], mmom[3 * cid + 2]};
                    for (int pk = 0; pk < np; ++pk)
                        for (int pj = 0; pj < np; ++pj)
                            for (int pi = 0; pi < np; ++pi) {
                                double w = (wx[pi] * wy[pj]) * wz[pk];
                                int64_t pid = IDX(pi, pj, pk, pnode, np);
                                gam[pid] += cg * w;
                                for (int d = 0; d < 3; ++d) mmom[3 * pid + d] += cm[d] * w;
                            }
                }
            }
        }
    }
    free(pk); free(order); free(goff);
}

/* K(R; p) = R / sqrt(R.R + (p.R)^2 + eps)^3 (interaction.jl:36, :69) */
static void kernel_K(const double R[3], const double p[3], double K[3]) {
    double rr = (R[0] * R[0] + R[1] * R[1]) + R[2] * R[2];
    double pr = (p[0] * R[0] + p[1] * R[1]) + p[2] * R[2];
    double s = sqrt((rr + pr * pr) + EPS64);
    double s3 = (s * s) * s;
    for (int d = 0; d < 3; ++d) K[d] = R[d] / s3;
}

static void add_cross(double *acc, const double a[3], const double b[3]) {
    acc[0] += a[1] * b[2] - a[2] * b[1];
    acc[1] += a[2] * b[0] - a[0] * b[2];
    acc[2] += a[0] * b[1] - a[1] * b[0];
}

/* interaction.jl:3-46 */
void orc_m2l(int n, int64_t npairs, const int64_t *pairs, const double *bboxes,
             const double p_avg[3], const double *gam, const double *mmom, double *le,
             double *lb) {
    int np = n + 1;
    /* pairs grouped by target, each group in list order (a target's
     * expansion receives the same additions in the same order as serially) */
    int64_t *order = NULL, *goff = NULL;
    int64_t ng = group_by_key(npairs, pairs, 2, &order, &goff);
    int par = ng >= 0;
    if (!par) ng = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(par ? nthreads_for(ng) : 1)
    for (int64_t g = 0; g < ng; ++g)
    for (int64_t e = par ? goff[g] : 0; e < (par ? goff[g + 1] : npairs); ++e) {
        double sx[11], sy[11], sz[11], tx[11], ty[11], tz[11];
        int64_t a = par ? order[e] : e;
        int64_t t = pairs[2 * a], s = pairs[2 * a + 1];
        const double *sb = bboxes + 6 * s, *tb = bboxes + 6 * t;
        orc_cheb2(n, sb[0], sb[3], sx);
        orc_cheb2(n, sb[1], sb[4], sy);
        orc_cheb2(n, sb[2], sb[5], sz);
        orc_cheb2(n, tb[0], tb[3], tx);
        orc_cheb2(n, tb[1], tb[4], ty);
        orc_cheb2(n, tb[2], tb[5], tz);
        for (int tk = 0; tk < np; ++tk)
            for (int tj = 0; tj < np; ++tj)
                for (int ti = 0; ti < np; ++ti) {
                    int64_t tid = IDX(ti, tj, tk, t, np);
                    for (int sk = 0; sk < np; ++sk)
                        for (int sj = 0; sj < np; ++sj)
                            for (int si = 0; si < np; ++si) {
                                int64_t sid = IDX(si, sj, sk, s, np);
                                double R[3] = {tx[ti] - sx[si], ty[tj] - sy[sj], tz[tk] - sz[sk]};
                                double K[3];
                                kernel_K(R, p_avg, K);
                                for (int d = 0; d < 3; ++d) le[3 * tid + d] += gam[sid] * K[d];
                                add_cross(lb + 3 * tid, mmom + 3 * sid, K);
                            }
                }
    }
    free(order); free(goff);
}

/* interaction.jl:49-75 */
void orc_p2p(int64_t npairs, const int64_t *pairs, const int64_t *pind, const int64_t *lo,
             const int64_t *hi, const double *pos, const double *mom, double *E, double *B) {
    /* pairs grouped by target cluster, each group in list order; in parallel
     * when the target clusters' particle ranges are disjoint */
    int64_t *order = NULL, *goff = NULL;
    int64_t ng = group_by_key(npairs, pairs, 2, &order, &goff);
    int par = ng >= 0;
    if (par) {
        int64_t *tg = malloc(sizeof(int64_t) * (size_t)(ng > 0 ? ng : 1));
        par = tg != NULL;
        if (par) {
            for (int64_t g = 0; g < ng; ++g) tg[g] = pairs[2 * order[goff[g]]];
            par = ranges_disjoint(ng, tg, lo, hi);
            free(tg);
        }
    }
    if (!par) ng = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(par ? nthreads_for(ng) : 1)
    for (int64_t g = 0; g < ng; ++g)
    for (int64_t e = par ? goff[g] : 0; e < (par ? goff[g + 1] : npairs); ++e) {
        int64_t a = par ? order[e] : e;
        int64_t t = pairs[2 * a], s = pairs[2 * a + 1];
        for (int64_t i = lo[t]; i <= hi[t]; ++i) {
            int64_t pi = pind[i];
            const double *xi = pos + 3 * pi;
            for (int64_t j = lo[s]; j <= hi[s]; ++j) {
                int64_t pj = pind[j];
                const double *xj = pos + 3 * pj, *m = mom + 3 * pj;
                double gj = sqrt(1.0 + ((m[0] * m[0] + m[1] * m[1]) + m[2] * m[2]));
                double R[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]}, K[3];
                kernel_K(R, m, K);
                for (int d = 0; d < 3; ++d) E[3 * pi + d] += gj * K[d];
                add_cross(B + 3 * pi, m, K);
            }
        }
    }
    free(order); free(goff);
}

/* downwardpass.jl:3-45 */
void orc_l2l(int n, int rule, int64_t nnodes, const int64_t *nodes, const int64_t *parents,
             const double *bboxes, double *le, double *lb) {
    int np = n + 1;
    /* each node owns its expansion and reads its parent's: nodes in parallel
     * when they are distinct and no parent is listed */
    int64_t *order = NULL, *goff = NULL;
    int64_t ng = group_by_key(nnodes, nodes, 1, &order, &goff);
    int par = ng == nnodes && no_parent_listed(nnodes, nodes, parents);
    free(order); free(goff);
#pragma omp parallel for schedule(dynamic, 4) num_threads(par ? nthreads_for(nnodes) : 1)
    for (int64_t a = 0; a < nnodes; ++a) {
        double cx[11], cy[11], cz[11], px[11], py[11], pz[11], wx[11], wy[11], wz[11];
        int64_t node = nodes[a], pnode = parents[node];
        const double *cb = bboxes + 6 * node, *pb = bboxes + 6 * pnode;
        orc_cheb2(n, cb[0], cb[3], cx);
        orc_cheb2(n, cb[1], cb[4], cy);
        orc_cheb2(n, cb[2], cb[5], cz);
        orc_cheb2(n, pb[0], pb[3], px);
        orc_cheb2(n, pb[1], pb[4], py);
        orc_cheb2(n, pb[2], pb[5], pz);
        for (int ck = 0; ck < np; ++ck) {
            orc_lgweight(cz[ck], pz, n, rule, wz);
            for (int cj = 0; cj < np; ++cj) {
                orc_lgweight(cy[cj], py, n, rule, wy);
                for (int ci = 0; ci < np; ++ci) {
                    orc_lgweight(cx[ci], px, n, rule, wx);
                    int64_t cid = IDX(ci, cj, ck, node, np);
                    for (int pk = 0; pk < np; ++pk)
                        for (int pj = 0; pj < np; ++pj)
                            for (int pi = 0; pi < np; ++pi) {
                                double w = (wx[pi] * wy[pj]) * wz[pk];
                                int64_t pid = IDX(pi, pj, pk, pnode, np);
                                for (int d = 0; d < 3; ++d) {
                                    le[3 * cid + d] += le[3 * pid + d] * w;
                                    lb[3 * cid + d] += lb[3 * pid + d] * w;
                                }
                            }
                }
            }
        }
    }
}

/* downwardpass.jl:47-80 */
void orc_l2p(int n, int rule, int64_t nnodes, const int64_t *nodes, const int64_t *pind,
             const int64_t *lo, const int64_t *hi, const double *bboxes, const double *pos,
             const double *le, const double *lb, double *E, double *B) {
    int np = n + 1;
    /* each node owns its particles' fields: in parallel when ranges are disjoint */
    int par = ranges_disjoint(nnodes, nodes, lo, hi);
#pragma omp parallel for schedule(dynamic, 4) num_threads(par ? nthreads_for(nnodes) : 1)
    for (int64_t a = 0; a < nnodes; ++a) {
        double xc[11], yc[11], zc[11], wx[11], wy[11], wz[11];
        int64_t node = nodes[a];
        const double *bb = bboxes + 6 * node;
        orc_cheb2(n, bb[0], bb[3], xc);
        orc_cheb2(n, bb[1], bb[4], yc);
        orc_cheb2(n, bb[2], bb[5], zc);
        for (int64_t p = lo[node]; p <= hi[node]; ++p) {
            int64_t q = pind[p];
            const double *x = pos + 3 * q;
            orc_lgweight(x[0], xc, n, rule, wx);
            orc_lgweight(x[1], yc, n, rule, wy);
            orc_lgweight(x[2], zc, n, rule, wz);
            for (int k = 0; k < np; ++k)
                for (int j = 0; j < np; ++j)
                    for (int i = 0; i < np; ++i) {
                        double w = (wx[i] * wy[j]) * wz[k];
                        int64_t id = IDX(i, j, k, node, np);
                        for (int d = 0; d < 3; ++d) {
                            E[3 * q + d] += le[3 * id + d] * w;
                            B[3 * q + d] += lb[3 * id + d] * w;
                        }
                    }
        }
    }
}

/* ------------------------------------------------------------------ driver */

/* update_particles_field.jl:40-64 (interact! = M2L! then P2P!, interaction.jl:77-80;
 * upwardpass! upwardpass.jl:101-108; downwardpass! downwardpass.jl:82-89). */
int orc_fmm(int64_t N, const double *pos, const double *mom, double charge, double lambda,
            int n, int64_t N0, double eta, const double *p_avg_in, int rule, double *E,
            double *B) {
    if (N < 1 || N0 < 1 || n < 1 || n > 10) return 1;
    double p_avg[3];
    if (p_avg_in) memcpy(p_avg, p_avg_in, sizeof p_avg);
    else orc_p_avg(N, mom, p_avg);
    double g_avg = sqrt(1.0 + ((p_avg[0] * p_avg[0] + p_avg[1] * p_avg[1]) + p_avg[2] * p_avg[2]));
    double stretch[3] = {1.0, 1.0, g_avg};
    int64_t ml = orc_maxlevel(N, N0), nc = orc_ncluster(N, N0);
    int np = n + 1;
    int64_t nn = (int64_t)np * np * np;
    int64_t *pind = malloc(sizeof(int64_t) * N), *lo = malloc(sizeof(int64_t) * nc),
            *hi = malloc(sizeof(int64_t) * nc), *lev = malloc(sizeof(int64_t) * nc),
            *par = malloc(sizeof(int64_t) * nc), *ch = malloc(sizeof(int64_t) * 2 * nc),
            *nodes = malloc(sizeof(int64_t) * nc);
    double *bb = malloc(sizeof(double) * 6 * nc);
    uint8_t *ex = malloc(nc);
    double *gam = calloc((size_t)(nn * nc), sizeof(double));
    double *mm = calloc((size_t)(3 * nn * nc), sizeof(double));
    double *le = calloc((size_t)(3 * nn * nc), sizeof(double));
    double *lb = calloc((size_t)(3 * nn * nc), sizeof(double));
    int rc = 0;
    if (!pind || !lo || !hi || !lev || !par || !ch || !nodes || !bb || !ex || !gam || !mm || !le || !lb) {
        rc = 2;
        goto done;
    }
    orc_build_tree(N, pos, N0, stretch, pind, lo, hi, lev, par, ch, bb, ex);
    /* upward pass: P2M on every leaf, then M2M level ml..1 */
    int64_t nl = 0;
    for (int64_t k = 0; k < nc; ++k)
        if (ex[k] && ch[2 * k] < 0) nodes[nl++] = k;
    orc_p2m(n, rule, nl, nodes, pind, lo, hi, bb, pos, mom, gam, mm);
    for (int64_t l = ml; l >= 1; --l) {
        int64_t m = 0;
        for (int64_t k = ((int64_t)1 << l) - 1; k < ((int64_t)1 << (l + 1)) - 1; ++k)
            if (ex[k]) nodes[m++] = k;
        orc_m2m(n, rule, m, nodes, par, bb, gam, mm);
    }
    /* interaction lists and interactions */
    int64_t np2p = 0, nm2l = 0;
    orc_lists(nc, ch, bb, stretch, eta, 0, NULL, 0, NULL, &np2p, &nm2l);
    int64_t *p2p = malloc(sizeof(int64_t) * 2 * (np2p + 1)), *m2l = malloc(sizeof(int64_t) * 2 * (nm2l + 1));
    if (!p2p || !m2l) { free(p2p); free(m2l); rc = 2; goto done; }
    orc_lists(nc, ch, bb, stretch, eta, np2p, p2p, nm2l, m2l, &np2p, &nm2l);
    for (int64_t i = 0; i < 3 * N; ++i) E[i] = B[i] = 0.0;
    orc_m2l(n, nm2l, m2l, bb, p_avg, gam, mm, le, lb);
    orc_p2p(np2p, p2p, pind, lo, hi, pos, mom, E, B);
    free(p2p);
    free(m2l);
    /* downward pass: L2L level 1..ml, then L2P on every leaf */
    for (int64_t l = 1; l <= ml; ++l) {
        int64_t m = 0;
        for (int64_t k = ((int64_t)1 << l) - 1; k < ((int64_t)1 << (l + 1)) - 1; ++k)
            if (ex[k]) nodes[m++] = k;
        orc_l2l(n, rule, m, nodes, par, bb, le, lb);
    }
    nl = 0;
    for (int64_t k = 0; k < nc; ++k)
        if (ex[k] && ch[2 * k] < 0) nodes[nl++] = k;
    orc_l2p(n, rule, nl, nodes, pind, lo, hi, bb, pos, le, lb, E, B);
    double amp = (AMP_RE * charge) / lambda;
    for (int64_t i = 0; i < 3 * N; ++i) {
        E[i] *= amp;
        B[i] *= amp;
    }
done:
    free(pind); free(lo); free(hi); free(lev); free(par); free(ch); free(nodes);
    free(bb); free(ex); free(gam); free(mm); free(le); free(lb);
    return rc;
}

/* -------------------------------------------------------------- brute force */

/* update_particles_field.jl:21-38 */
void orc_bruteforce(int64_t N, const double *pos, const double *mom, double charge,
                    double lambda, int64_t ntargets, const int64_t *targets, int nthreads,
                    double *E, double *B) {
    int64_t nt = targets ? ntargets : N;
    double amp = (AMP_RE * charge) / lambda;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (int64_t a = 0; a < nt; ++a) {
        int64_t i = targets ? targets[a] : a;
        const double *xi = pos + 3 * i;
        double e[3] = {0.0, 0.0, 0.0}, b[3] = {0.0, 0.0, 0.0};
        for (int64_t j = 0; j < N; ++j) {
            const double *xj = pos + 3 * j, *m = mom + 3 * j;
            double R[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]}, K[3];
            kernel_K(R, m, K);
            double g = sqrt(1.0 + ((m[0] * m[0] + m[1] * m[1]) + m[2] * m[2]));
            double ag = amp * g;
            for (int d = 0; d < 3; ++d) e[d] += ag * K[d];
            double c[3] = {m[1] * K[2] - m[2] * K[1], m[2] * K[0] - m[0] * K[2], m[0] * K[1] - m[1] * K[0]};
            for (int d = 0; d < 3; ++d) b[d] += amp * c[d];
        }
        for (int d = 0; d < 3; ++d) {
            E[3 * a + d] = e[d];
            B[3 * a + d] = b[d];
        }
    }
}
