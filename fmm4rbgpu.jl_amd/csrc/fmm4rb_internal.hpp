// fmm4rb_internal.hpp -- shared declarations of the gfx950 FMM library.
//
// Data layout in HBM (N particles, nc clusters, NP = n+1, NN = NP^3):
//   sp8   [N][8]  double  particles in tree order: x, y, z, gamma, px, py, pz, 0
//   sp8f  [N] float4      gamma, px, py, pz (sources of the mixed-precision P2P)
//   geo   [nc][8] double  centre xyz, admissibility radius, diameter, 0, 0, 0
//   ncoord[nc][3][NP]     Chebyshev (2nd kind) interpolation nodes per axis
//   mp4   [nc][NN] double4 multipoles (gamma, mx, my, mz)
//   mp4f  [nc][NN] float4  FP32 copy for the mixed-precision M2L
//   loc   [nc][NN][6]     local expansions (Ex, Ey, Ez, Bx, By, Bz)
//   near  [N][6]          P2P sums in tree order
// Nodes are 0-based heap ids (children 2k+1, 2k+2).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "fmm4rb.h"

namespace fmm4rb {

struct Error {
    int code;
    std::string msg;
};

#define FMM_HIP(x)                                                                                \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess)                                                                     \
            throw ::fmm4rb::Error{FMM4RB_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define FMM_LAUNCH_CHECK() FMM_HIP(hipGetLastError())

inline void fail(int code, const std::string &m) { throw Error{code, m}; }

// Grow-only device buffer.  Inside a call of the allocation-free entry point
// (an Arena is active on the calling thread) ensure() carves the buffer from
// that arena instead of calling hipMalloc, and the buffer is "borrowed": it is
// never passed to hipFree.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool borrowed = false;
    void ensure(size_t bytes);
    void release();
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// Bump allocator over a caller-provided device region (256-byte aligned carves).
struct Arena {
    unsigned char *base = nullptr;
    size_t size = 0, used = 0;
};
// The arena of the allocation-free call running on this thread (nullptr outside one).
extern thread_local Arena *t_arena;

constexpr double EPS64 = 2.220446049250313e-16;  // Julia eps()
constexpr float EPS32K = 2.220446049250313e-16f;  // the same softening in FP32
constexpr double AMP_RE = 2.8179403699772166e-15;
constexpr int MAX_NP = 11;      // n <= 10

// 1/sqrt(d) in FP64: the hardware estimate v_rsq_f64 (~5e-8 relative) refined by
// one Newton step, y (1 + (1 - d y^2)/2): ~4e-15 relative (measured on gfx950).
__device__ inline double rsqrt_nr(double d) {
    const double y = __builtin_amdgcn_rsq(d);
    const double r = fma(-d * y, y, 1.0);
    return fma(y * 0.5, r, y);
}
// d^-3/2, the kernel's scale, as y^3 with y = 1/sqrt(d): FP64 rsqrt_nr, FP32
// v_rsq_f32.  (y times v_rcp_f32(d) saves a multiply but ran slower: the second
// transcendental does not issue for free, M2L 15.7 -> 17.5 ms at N = 2^22.)
template <typename T>
__device__ inline T rsq3(T d) {
    T y;
    if constexpr (sizeof(T) == 4) y = __builtin_amdgcn_rsqf(d);
    else y = rsqrt_nr(d);
    return (y * y) * y;
}
// Particle records staged in LDS by the near-field kernels (forward P2P and both P2P^T
// sides): the position in FP64, then the per-particle terms (FP32 in the mixed
// kernels), one record per particle so that one address register walks a tile.  A
// tile is padded to the kernel's step with records at 1e30 whose terms are zero: they
// add exactly zero (R.R overflows to inf in FP32, so d^-3/2 = 0; in FP64 every
// product carries a zero factor).
template <bool MIXED>
struct P2PSrc {  // a source: gamma and p
    using V = typename std::conditional<MIXED, float, double>::type;
    double x, y, z;
    V g, px, py, pz;
    __device__ static P2PSrc pad() { return P2PSrc{1e30, 1e30, 1e30, 0, 0, 0, 0}; }
};
template <bool MIXED>
struct AdjTgt {  // a target of the P2P^T source side: the cotangents e = Ebar, b = Bbar
    using V = typename std::conditional<MIXED, float, double>::type;
    double x, y, z;
    V e0, e1, e2, b0, b1, b2;
    __device__ static AdjTgt pad() { return AdjTgt{1e30, 1e30, 1e30, 0, 0, 0, 0, 0, 0}; }
};
constexpr int MAX_N0 = 1024;
constexpr int TCNT_WORDS = 64, TCNT_FLAG = 63;  // 2 maxlevel + 4 <= 62 counters, then the flag

// Workgroup b of a launch runs on XCD b % 8 (round-robin dispatch over the
// MI355X's 8 XCDs, each with its own L2).  The kernels that walk interaction
// lists take work item xcd_item(b, G): within every chunk of 8 * XCD_RUN
// workgroups, each XCD owns a run of XCD_RUN consecutive targets, whose
// neighbours share most of their sources, so a source cluster or leaf is
// fetched into fewer L2s, while the chunks keep every XCD's share of each tree
// level (whole contiguous eighths of the grid would give some XCDs the light upper
// levels and others the heavy deep ones).  A bijection on [0, G).  Run lengths
// measured at N = 2^22 (kernels alone, same-box comparisons): M2L about 3% faster
// with runs of 512 targets than of 64 (1024 no better, 4096 3% slower than 64), P2P
// about 8% faster with 4096 than with 64 (2048 and 8192 in between); the mixed
// forward 28.2 -> 27.0 ms, the reverse modes 1-2% faster.
constexpr int NXCD = 8;
constexpr int XCD_RUN_FAR = 512;    // M2L and M2L^T (targets: clusters)
constexpr int XCD_RUN_NEAR = 4096;  // P2P and P2P^T (targets: leaves)
template <int XCD_RUN>
__device__ inline int32_t xcd_item(int32_t b, int32_t G) {
    constexpr int32_t CH = NXCD * XCD_RUN;
    const int32_t base = (b / CH) * CH, n = min(CH, G - base), l = b - base;
    const int32_t q = n / NXCD, r = n % NXCD, x = l % NXCD, k = l / NXCD;
    return base + (x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k);
}

// Tree shape: a closed-form function of (N, N0) (clusters.jl:21-49 splits a
// node of c particles into ceil(c/2) and floor(c/2) while c > N0).
struct Shape {
    int64_t N = -1, N0 = -1;
    int ml = 0;
    int64_t nc = 0;
    std::vector<int32_t> lo, cnt;       // per node (-1 / 0 if absent)
    std::vector<uint8_t> flags;         // bit0 exists, bit1 leaf
    std::vector<int32_t> leaves;        // existing leaves, ascending id
    std::vector<int32_t> lvl_nodes;     // existing nodes grouped by level
    std::vector<int64_t> lvl_off;       // ml+2 offsets into lvl_nodes
    std::vector<int32_t> lvl_split;     // existing internal nodes grouped by level
    std::vector<int64_t> split_off;     // ml+2 offsets into lvl_split
    std::vector<int32_t> leftbase;      // per splitting node: lefts before it on its level
    int maxleaf = 0;
};

struct Ctx;

// ---- radix sort (radix_sort.hip)
// Sort nseg independent arrays of n (key, value) pairs by key bits [bit0, bit1).
// Returns true if the result is in the *_alt buffers.
size_t radix_scratch_words(int64_t n, int nseg);
bool radix_sort(uint64_t *keys, uint64_t *keys_alt, uint32_t *vals, uint32_t *vals_alt,
                int64_t n, int nseg, int bit0, int bit1, uint32_t *scratch, hipStream_t st);
bool radix_sort32(uint32_t *keys, uint32_t *keys_alt, uint32_t *vals, uint32_t *vals_alt,
                  int64_t n, int nseg, int bit0, int bit1, uint32_t *scratch, hipStream_t st);
// Exclusive scan of nrows rows of len words each; row totals to tot[nrows].
void scan_rows(uint32_t *rows, int nrows, int len, uint32_t *tot, hipStream_t st);

// ---- tree (tree.hip)
void launch_pavg(const double *pos, const double *mom, int64_t N, double *part, double *pavg4,
                 int *bad, hipStream_t st);
// exact = false: the presort uses 32-bit quantised keys and raises the tree flag
// (word TCNT_FLAG of d_tcnt) if a run of equal keys was too long to order (then build
// again with exact = true, 64-bit keys of the exact coordinate order).
void build_tree_device(Ctx &c, const double *pos, const double *mom, int64_t N, int64_t N0,
                       bool exact);
// New positions on the existing tree: tree-ordered particles and every bounding box.
void refit_tree_device(Ctx &c, const double *pos, const double *mom);

// ---- traversal (traverse.hip)
// Returns false, with no lists built, if the tree's quantised presort raised its flag.
bool build_lists_device(Ctx &c, double eta);
// Sort (target << sb | source) keys, CSR offsets per target; vals (may be null)
// are permuted alongside, *vsorted = the buffer holding them sorted.
void sort_pairs_csr(Ctx &c, uint64_t *keys, uint64_t *keys_alt, int64_t n, int32_t *off,
                    uint64_t **sorted, uint32_t *vals = nullptr, uint32_t *vals_alt = nullptr,
                    uint32_t **vsorted = nullptr);

// ---- interpolation operators (interp.hip)
void launch_node_geo(Ctx &c);
void launch_upward(Ctx &c, hipStream_t st);
void launch_downward(Ctx &c, double amp, double *E, double *B);
// L2L level by level from the root; leaves = false skips the leaves, whose transfer
// k_leaf_l2p does (launch_leaf_l2p: the leaves' final locals, then E and B).
void launch_l2l(Ctx &c, bool leaves = true);
void launch_leaf_l2p(Ctx &c, double amp, double *E, double *B);
void launch_mp_to_f32(Ctx &c, hipStream_t st);

// ---- interactions (m2l.hip, p2p.hip)
void launch_m2l(Ctx &c, const int32_t *off, const uint64_t *keys, int sb, int64_t ngroups_grid);
void launch_p2p(Ctx &c, const int32_t *off, const uint64_t *keys, int sb, hipStream_t st);

// ---- adjoint (adjoint.hip, traverse.hip)
void run_adjoint(Ctx &c, const double *Ebar, const double *Bbar, double amp, double *gpos,
                 double *gmom, bool saved_locals);
void transpose_lists(Ctx &c, const uint64_t *sorted, int64_t n, DevBuf &kT, DevBuf &kT_alt,
                     DevBuf &offT, uint64_t **sortedT, DevBuf *pos = nullptr);
// Both of the context's lists regrouped by source in one two-segment sort:
// p2pT_sorted / m2lT_sorted and their CSR offsets per source; with pos, pos[i] = the
// transposed position of forward M2L pair i.
void transpose_lists_joint(Ctx &c, DevBuf *pos);
void launch_bruteforce_vjp(const double *pos, const double *mom, int64_t N, const double *ob,
                           double *gpos, double *gmom, hipStream_t st);

// ---- brute force (bruteforce.hip)
// part: scratch for the per-chunk partial sums (grown as needed)
void launch_bruteforce(const double *pos, const double *mom, int64_t N, const int64_t *targets,
                       int64_t nt, double amp, double *E, double *B, DevBuf &part,
                       hipStream_t st);

struct Ctx {
    int device = 0;
    hipStream_t own = nullptr, stream = nullptr;
    // the near field (P2P, P2P^T) runs on aux, concurrently with the far-field chain:
    // evx[0] fork, evx[1] P2P start, evx[2] P2P end (the join)
    hipStream_t aux = nullptr;
    hipEvent_t ev[10] = {}, evx[3] = {};
    // reverse mode: the L2P^T/L2L^T geometry runs on aux2 beside the M2L^T source
    // side (evy[0] fork, evy[1] join)
    hipStream_t aux2 = nullptr;
    hipEvent_t evy[2] = {};
    // the upward pass runs on aux2 beside the list build: evu[0] fork, [1] start, [2] end
    hipEvent_t evu[3] = {};
    // overlap = false runs the helper streams' work in place on the call's stream
    // (same order of results, bit-identical): every kernel then runs alone on the GPU
    bool overlap = true;
    hipStream_t near_stream() const { return overlap ? aux : stream; }
    hipStream_t third_stream() const { return overlap ? aux2 : stream; }
    bool debug = false;
    // the tree and lists on the device belong to a completed field or reverse-mode
    // call with these N, N0, eta (the refit entry points reuse them)
    bool tree_valid = false;
    int64_t tree_N = 0, tree_N0 = 0;
    double tree_eta = -1.0;
    // the final local expansions in d_loc belong to a completed field call with
    // these parameters and inputs (digest of pos and mom): the pullback reuses them
    bool loc_valid = false;
    uint64_t loc_digest = 0, in_digest = 0;  // in_digest: the current call's inputs
    int64_t loc_N = 0, loc_N0 = 0;
    int loc_n = 0, loc_prec = -1;
    double loc_eta = -1.0;

    // parameters of the current call
    int64_t N = 0, N0 = 0;
    int n = 0, np = 0, nn = 0, precision = 0;
    Shape shape;

    // shape arrays on the device
    DevBuf d_lo, d_cnt, d_flags, d_leaves, d_lvl_nodes, d_lvl_split, d_leftbase;

    // tree build scratch
    DevBuf d_part, d_pavg, d_bad;
    DevBuf d_keys, d_keys_alt, d_vals, d_vals_alt, d_rs_scratch;
    DevBuf d_seg, d_seg_alt, d_rk, d_pivot, d_bsum, d_axis;
    // tree outputs
    DevBuf d_parind, d_leafof, d_bbox, d_geo, d_ncoord, d_costab, d_ext;
    DevBuf d_sp8, d_sp8f;
    // lists
    // d_tcnt: the traversal's counters (words 0 .. 2 ml + 3) and the tree flag (word
    // TCNT_FLAG), read back in one copy
    DevBuf d_front0, d_front1, d_tcnt, d_p2p, d_p2p_alt, d_m2l, d_m2l_alt, d_p2p_off, d_m2l_off;
    // both lists side by side (stride pair_stride) for their joint sort
    DevBuf d_pairs, d_pairs_alt;
    int64_t pair_stride = 0;
    int64_t cap_front = 0, cap_p2p = 0, cap_m2l = 0;
    int64_t np2p = 0, nm2l = 0;
    int sb = 1;
    int trav_iters = 0;
    uint64_t *p2p_sorted = nullptr, *m2l_sorted = nullptr;
    // expansions and fields
    DevBuf d_mp4, d_mp4f, d_loc, d_near;
    // adjoint (field_vjp)
    DevBuf d_ob, d_gx, d_gp, d_lbar, d_mbar, d_xbar, d_pbar, d_abar;
    DevBuf d_gx2, d_gp2;  // P2P^T gradients (written on aux)
    DevBuf d_xbar2;       // node-coordinate gradients of the M2L^T source side
    DevBuf d_xbar3;       // node-coordinate gradients of the L2P^T / L2L^T geometry
    DevBuf d_p2pT, d_p2pT_alt, d_p2pT_off, d_m2lT_off;  // both transposed lists, side by side
    DevBuf d_m2lT_pos, d_ysp;  // transposed position of each M2L pair; per-pair source sums
    uint64_t *p2pT_sorted = nullptr, *m2lT_sorted = nullptr;
    // debug snapshots
    std::vector<double> snap_loc_m2l;
    // host-side p_avg of the last call
    double pavg_host[4] = {0, 0, 0, 1};
    double costab_host[MAX_NP] = {};
};

}  // namespace fmm4rb
