// api.hip -- the C ABI (include/fmm4rb.h): contexts, validation, the tree
// shape, device buffers and the field-evaluation pipeline.
//
// Pipeline (update_particles_field!(beam, FMMGPU(n, N0, eta); lambda),
// src/update_particles_field.jl:66-148):
//   p_avg + input check -> cluster tree -> node geometry -> dual traversal,
//   list sort/CSR -> { P2M, M2M -> M2L -> L2L  ||  P2P } -> L2P + scaling,
// where P2P runs on the context's second stream, concurrently with the
// far-field chain (whose per-level launches leave most of the GPU idle).
// Unlike the reference, the tree and the lists are built on the device and
// nothing but a handful of counters crosses PCIe between stages.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "fmm4rb_internal.hpp"

namespace fmm4rb {

static thread_local std::string g_err;

thread_local Arena *t_arena = nullptr;

void DevBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return;
    if (t_arena) {
        // allocation-free call: carve from the arena; the old region (if any) is
        // the arena's too and is simply left behind
        Arena &a = *t_arena;
        const size_t b = std::max<size_t>(bytes, 256);
        const size_t off = (a.used + 255) & ~(size_t)255;
        if (off > a.size || b > a.size - off)
            fail(FMM4RB_ENOMEM, "the call needs more than the " + std::to_string(a.size >> 20) +
                                    " MiB static scratch of the allocation-free entry point "
                                    "(at least " + std::to_string((off + b + ((1 << 20) - 1)) >> 20) +
                                    " MiB so far): use fmm4rb_field_device for this size");
        p = a.base + off;
        cap = b;
        borrowed = true;
        a.used = off + b;
        return;
    }
    release();
    size_t b = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        fail(FMM4RB_ENOMEM, "hipMalloc of " + std::to_string(b) + " bytes failed: " +
                                hipGetErrorString(e));
    }
    cap = b;
}

void DevBuf::release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    borrowed = false;
}

// ---- static scratch of the allocation-free entry point (fmm4rb_field_device_noalloc)
// Every device buffer of that entry point's pipeline is carved from this array.
// In host code its name is only a stand-in for the symbol: its device address is
// looked up once per device (scratch_init, from fmm4rb_create with the device
// current) and only that address is used.
constexpr size_t SCRATCH_BYTES = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

static void scratch_init(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return;  // noalloc then refuses
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        fail(FMM4RB_EDEVICE, std::string("hipGetSymbolAddress of the static scratch failed: ") +
                                 hipGetErrorString(e));
    }
    g_scratch_ptr[device] = ptr;
}

static void *scratch_ptr(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

// Every device buffer a context owns (destroy releases them; the allocation-free
// entry point swaps them for arena carves during its call).  Returns the count.
constexpr int CTX_NBUF_MAX = 80;
static int ctx_bufs(Ctx &c, DevBuf **out) {
    DevBuf *all[] = {&c.d_lo, &c.d_cnt, &c.d_flags, &c.d_leaves, &c.d_lvl_nodes, &c.d_lvl_split,
            &c.d_leftbase, &c.d_part, &c.d_pavg, &c.d_bad, &c.d_keys, &c.d_keys_alt,
            &c.d_vals, &c.d_vals_alt, &c.d_rs_scratch, &c.d_seg, &c.d_seg_alt,
            &c.d_rk, &c.d_pivot, &c.d_bsum, &c.d_axis, &c.d_parind, &c.d_leafof,
            &c.d_bbox, &c.d_geo, &c.d_ncoord, &c.d_costab, &c.d_ext, &c.d_sp8, &c.d_sp8f,
            &c.d_ob, &c.d_gx, &c.d_gp, &c.d_gx2, &c.d_gp2, &c.d_xbar2, &c.d_xbar3, &c.d_lbar, &c.d_mbar,
            &c.d_xbar, &c.d_pbar, &c.d_abar, &c.d_p2pT, &c.d_p2pT_alt,
            &c.d_p2pT_off, &c.d_m2lT_off, &c.d_front0, &c.d_front1, &c.d_tcnt, &c.d_pairs, &c.d_pairs_alt, &c.d_p2p,
            &c.d_p2p_alt, &c.d_m2l, &c.d_m2l_alt, &c.d_p2p_off, &c.d_m2l_off, &c.d_mp4,
            &c.d_mp4f, &c.d_loc, &c.d_near, &c.d_m2lT_pos, &c.d_ysp};
    constexpr int nb = (int)(sizeof(all) / sizeof(all[0]));
    static_assert(nb <= CTX_NBUF_MAX, "raise CTX_NBUF_MAX");
    for (int i = 0; i < nb; ++i) out[i] = all[i];
    return nb;
}

// Zero n 32-bit words (a kernel rather than hipMemsetAsync, so that no memset
// node is recorded if a caller captures the stream).
__global__ void k_zero_u32(uint32_t *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

static void zero_words(void *p, int64_t nwords, hipStream_t st) {
    if (nwords <= 0) return;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, st,
                       (uint32_t *)p, nwords);
    FMM_LAUNCH_CHECK();
}

// treeinfo.jl:6-8
static int64_t maxlevel_of(int64_t N, int64_t N0) {
    if (N <= N0) return 0;
    return (int64_t)std::ceil(std::log((double)N / (double)N0) / std::log(2.0));
}

static void compute_shape(Shape &S, int64_t N, int64_t N0) {
    S.N = N;
    S.N0 = N0;
    const int64_t ml = maxlevel_of(N, N0);
    if (ml > 29) fail(FMM4RB_EINVAL, "tree too deep (maxlevel > 29)");
    S.ml = (int)ml;
    S.nc = ((int64_t)1 << (ml + 1)) - 1;
    const int64_t nc = S.nc;
    S.lo.assign(nc, -1);
    S.cnt.assign(nc, 0);
    S.flags.assign(nc, 0);
    S.lo[0] = 0;
    S.cnt[0] = (int32_t)N;
    S.flags[0] = 1;
    S.maxleaf = 0;
    for (int64_t k = 0; k < nc; ++k) {
        if (!(S.flags[k] & 1)) continue;
        if (S.cnt[k] <= N0) {
            S.flags[k] |= 2;
            S.maxleaf = std::max(S.maxleaf, S.cnt[k]);
            continue;
        }
        const int64_t c1 = 2 * k + 1, c2 = 2 * k + 2;
        if (c2 >= nc)
            fail(FMM4RB_EINVAL, "cluster at the deepest level exceeds N0 (maxlevel rounding)");
        const int32_t nl = (S.cnt[k] + 1) / 2;
        S.lo[c1] = S.lo[k];
        S.cnt[c1] = nl;
        S.lo[c2] = S.lo[k] + nl;
        S.cnt[c2] = S.cnt[k] - nl;
        S.flags[c1] = S.flags[c2] = 1;
    }
    S.leaves.clear();
    S.lvl_nodes.clear();
    S.lvl_split.clear();
    S.lvl_off.assign(ml + 2, 0);
    S.split_off.assign(ml + 2, 0);
    S.leftbase.assign(nc, 0);
    for (int64_t l = 0; l <= ml; ++l) {
        S.lvl_off[l] = (int64_t)S.lvl_nodes.size();
        S.split_off[l] = (int64_t)S.lvl_split.size();
        int64_t run = 0;
        for (int64_t k = ((int64_t)1 << l) - 1; k < ((int64_t)1 << (l + 1)) - 1; ++k) {
            if (!(S.flags[k] & 1)) continue;
            S.lvl_nodes.push_back((int32_t)k);
            if (!(S.flags[k] & 2)) {
                S.lvl_split.push_back((int32_t)k);
                S.leftbase[k] = (int32_t)run;
                run += (S.cnt[k] + 1) / 2;
            }
        }
    }
    S.lvl_off[ml + 1] = (int64_t)S.lvl_nodes.size();
    S.split_off[ml + 1] = (int64_t)S.lvl_split.size();
    for (int64_t k = 0; k < nc; ++k)
        if ((S.flags[k] & 3) == 3) S.leaves.push_back((int32_t)k);
}

template <class T>
static void upload(DevBuf &b, const std::vector<T> &v, hipStream_t st) {
    b.ensure(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty())
        FMM_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

static void validate(int64_t N, int n, int64_t N0, double eta, int precision) {
    if (N < 1) fail(FMM4RB_EINVAL, "N must be >= 1");
    if (N > 0x7fffffffLL - 1) fail(FMM4RB_EINVAL, "N must be < 2^31 - 1");
    if (n < 1 || n > MAX_NP - 1) fail(FMM4RB_EINVAL, "n must be in [1, 10]");
    if (N0 < 1 || N0 > MAX_N0) fail(FMM4RB_EINVAL, "N0 must be in [1, 1024]");
    if (!(eta >= 0.0) || !std::isfinite(eta)) fail(FMM4RB_EINVAL, "eta must be finite and >= 0");
    if (precision != FMM4RB_FP64 && precision != FMM4RB_MIXED)
        fail(FMM4RB_EINVAL, "precision must be FMM4RB_FP64 or FMM4RB_MIXED");
}

// Set parameters, (re)compute the shape and size every buffer of the pipeline.
static void prepare(Ctx &c, int64_t N, int64_t N0, int n, int precision) {
    hipStream_t st = c.stream;
    // a call that failed after forking work to the helper streams may have left it
    // running: drain them before any buffer is reused (no cost when idle)
    FMM_HIP(hipStreamSynchronize(c.aux));
    FMM_HIP(hipStreamSynchronize(c.aux2));
    if (c.shape.N != N || c.shape.N0 != N0) {
        compute_shape(c.shape, N, N0);
        FMM_HIP(hipStreamSynchronize(st));  // host vectors may be in flight
        upload(c.d_lo, c.shape.lo, st);
        upload(c.d_cnt, c.shape.cnt, st);
        upload(c.d_flags, c.shape.flags, st);
        upload(c.d_leaves, c.shape.leaves, st);
        upload(c.d_lvl_nodes, c.shape.lvl_nodes, st);
        upload(c.d_lvl_split, c.shape.lvl_split, st);
        upload(c.d_leftbase, c.shape.leftbase, st);
        FMM_HIP(hipStreamSynchronize(st));
    }
    const Shape &S = c.shape;
    c.N = N;
    c.N0 = N0;
    c.n = n;
    c.np = n + 1;
    c.nn = c.np * c.np * c.np;
    c.precision = precision;
    c.sb = S.ml + 1;
    const size_t nc = (size_t)S.nc, NN = (size_t)c.nn, n3 = 3 * (size_t)N;
    const int64_t nleaf = (int64_t)S.leaves.size();
    // per block: 3 sums, a digest word and the position range
    c.d_part.ensure(10 * 1024 * sizeof(double));
    // p_avg, g_avg, input digest, the non-finite flag, 2 spare, position range
    c.d_pavg.ensure(16 * sizeof(double));
    c.d_bad.ensure(sizeof(int));
    c.d_keys.ensure(n3 * 8);
    c.d_keys_alt.ensure(n3 * 8);
    c.d_vals.ensure(n3 * 4);
    c.d_vals_alt.ensure(n3 * 4);
    c.d_seg.ensure((size_t)N * 4);
    c.d_seg_alt.ensure((size_t)N * 4);
    c.d_rk.ensure(2 * 3 * n3 * 4);
    c.d_pivot.ensure((size_t)nc * 4);
    c.d_bsum.ensure(4 * (3 * (size_t)((N + 2047) / 2048) + 3));
    c.d_axis.ensure(nc);
    c.d_ext.ensure(6 * nc * 4);
    c.d_parind.ensure((size_t)N * 4);
    c.d_leafof.ensure((size_t)N * 4);
    c.d_bbox.ensure(6 * nc * 8);
    c.d_geo.ensure(8 * nc * 8);
    c.d_ncoord.ensure(3 * (size_t)c.np * nc * 8);
    c.d_sp8.ensure(8 * (size_t)N * 8);
    if (precision == FMM4RB_MIXED) c.d_sp8f.ensure(4 * (size_t)N * 4);
    // list capacities from the leaf count on every call (a context first used at a
    // small N must not rely on the traversal's overflow retries at a large one)
    c.cap_front = std::max<int64_t>(c.cap_front, std::max<int64_t>(1 << 16, 48 * nleaf));
    c.cap_p2p = std::max<int64_t>(c.cap_p2p, std::max<int64_t>(1 << 16, 64 * nleaf));
    c.cap_m2l = std::max<int64_t>(c.cap_m2l, std::max<int64_t>(1 << 16, 64 * nleaf));
    // the sort scratch also covers the joint sort of the two lists up to their capacity
    c.d_rs_scratch.ensure(
        4 * std::max(radix_scratch_words(N, 3),
                     radix_scratch_words(std::max(c.cap_p2p, c.cap_m2l), 2)));
    c.d_front0.ensure(8 * (size_t)c.cap_front);
    c.d_front1.ensure(8 * (size_t)c.cap_front);
    c.d_tcnt.ensure(4 * (size_t)TCNT_WORDS);
    c.d_p2p.ensure(8 * (size_t)c.cap_p2p);
    c.d_m2l.ensure(8 * (size_t)c.cap_m2l);
    c.d_pairs.ensure(2 * 8 * (size_t)std::max(c.cap_p2p, c.cap_m2l));
    c.d_pairs_alt.ensure(2 * 8 * (size_t)std::max(c.cap_p2p, c.cap_m2l));
    c.d_p2p_off.ensure(4 * (nc + 1));
    c.d_m2l_off.ensure(4 * (nc + 1));
    c.d_mp4.ensure(4 * NN * nc * 8);
    if (precision == FMM4RB_MIXED) c.d_mp4f.ensure(4 * NN * nc * 4);
    c.d_loc.ensure(6 * NN * nc * 8);
    c.d_near.ensure(6 * (size_t)N * 8);
}

__global__ __launch_bounds__(256) void k_list_stats(const uint64_t *__restrict__ p2p, int64_t np2p,
                                                    const uint64_t *__restrict__ m2l, int64_t nm2l,
                                                    int sb, const int32_t *__restrict__ cnt,
                                                    unsigned long long *__restrict__ out) {
    __shared__ unsigned long long s[2][256];
    const uint64_t mask = (1ull << sb) - 1ull;
    unsigned long long pp = 0, g = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < np2p || i < nm2l;
         i += (int64_t)gridDim.x * 256) {
        if (i < np2p) {
            const uint64_t k = p2p[i];
            pp += (unsigned long long)cnt[k >> sb] * (unsigned long long)cnt[k & mask];
        }
        if (i < nm2l && (i == 0 || (m2l[i] >> sb) != (m2l[i - 1] >> sb))) g += 1;
    }
    s[0][threadIdx.x] = pp;
    s[1][threadIdx.x] = g;
    for (int off = 128; off > 0; off >>= 1) {
        __syncthreads();
        if (threadIdx.x < (unsigned)off) {
            s[0][threadIdx.x] += s[0][threadIdx.x + off];
            s[1][threadIdx.x] += s[1][threadIdx.x + off];
        }
    }
    if (threadIdx.x == 0) {
        atomicAdd(&out[0], s[0][0]);
        atomicAdd(&out[1], s[1][0]);
    }
}

// p_avg and the input check; returns after the check (one host sync).
// p_avg, g_avg, the input digest and the non-finite flag (words 0-5 of d_pavg) in
// one copy.
static void stage_pavg(Ctx &c, const double *pos, const double *mom) {
    hipStream_t st = c.stream;
    double *pv = c.d_pavg.as<double>();
    zero_words(pv + 5, 2, st);
    launch_pavg(pos, mom, c.N, c.d_part.as<double>(), pv, reinterpret_cast<int *>(pv + 5), st);
    double h[6];
    FMM_HIP(hipMemcpyAsync(h, pv, sizeof(h), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    for (int d = 0; d < 4; ++d) c.pavg_host[d] = h[d];
    std::memcpy(&c.in_digest, &h[4], sizeof(uint64_t));
    int bad = 0;
    std::memcpy(&bad, &h[5], sizeof(int));
    if (bad) fail(FMM4RB_EINVAL, "positions or momenta contain NaN or Inf");
}

// Tree, node geometry, the upward pass (forked) and the lists.  The tree's presort
// uses quantised keys; if it could not order a run of equal keys (reported at the
// list build's synchronisation), tree and lists are built again with exact keys.
static void tree_and_lists(Ctx &c, const double *pos, const double *mom, int64_t N, int64_t N0,
                           double eta, hipEvent_t ev_tree);

// The upward pass (P2M, M2M; for mixed precision also the FP32 copy of the
// multipoles) needs only the tree: it runs on the third stream, beside the list
// build on the call's stream, which joins it (evu[2]) before the far field.
static void fork_upward(Ctx &c) {
    hipStream_t s3 = c.third_stream();
    FMM_HIP(hipEventRecord(c.evu[0], c.stream));
    FMM_HIP(hipStreamWaitEvent(s3, c.evu[0], 0));
    FMM_HIP(hipEventRecord(c.evu[1], s3));
    launch_upward(c, s3);
    if (c.precision == FMM4RB_MIXED) launch_mp_to_f32(c, s3);
    FMM_HIP(hipEventRecord(c.evu[2], s3));
}

static void tree_and_lists(Ctx &c, const double *pos, const double *mom, int64_t N, int64_t N0,
                           double eta, hipEvent_t ev_tree) {
    for (int exact = 0; exact < 2; ++exact) {
        build_tree_device(c, pos, mom, N, N0, exact == 1);
        launch_node_geo(c);
        FMM_HIP(hipEventRecord(ev_tree, c.stream));
        fork_upward(c);
        if (build_lists_device(c, eta)) return;
    }
    fail(FMM4RB_EDEVICE, "tree presort: the exact keys left a run unordered");
}

// refit: keep the tree and lists of the last completed call with the same N, N0
// and eta (particle order, ranges, splits, lists); recompute p_avg, the bounding
// boxes and the interpolation nodes for the new inputs.
static void run_field(Ctx &c, int64_t N, const double *pos, const double *mom, double charge,
                      double lambda, int n, int64_t N0, double eta, int precision, double *E,
                      double *B, fmm4rb_stats *stats, bool refit = false) {
    validate(N, n, N0, eta, precision);
    if (!std::isfinite(charge) || !std::isfinite(lambda) || lambda == 0.0)
        fail(FMM4RB_EINVAL, "charge must be finite and lambda finite and non-zero");
    if (refit && !(c.tree_valid && c.tree_N == N && c.tree_N0 == N0 && c.tree_eta == eta))
        fail(FMM4RB_EINVAL,
             "no tree to refit: the context's last field or reverse-mode call must have "
             "completed with the same N, N0 and eta");
    if (!refit) c.tree_valid = false;
    c.loc_valid = false;
    prepare(c, N, N0, n, precision);
    hipStream_t st = c.stream;
    const double amp = (AMP_RE * charge) / lambda;  // update_particles_field.jl:142
    FMM_HIP(hipEventRecord(c.ev[0], st));
    stage_pavg(c, pos, mom);
    FMM_HIP(hipEventRecord(c.ev[1], st));
    if (refit) {
        refit_tree_device(c, pos, mom);
        launch_node_geo(c);
        FMM_HIP(hipEventRecord(c.ev[2], st));
        // no list build to run beside: the upward pass goes first on this stream, so
        // that M2L starts together with P2P and the two share the GPU
        FMM_HIP(hipEventRecord(c.ev[3], st));
        FMM_HIP(hipEventRecord(c.evu[1], st));
        launch_upward(c, st);
        if (c.precision == FMM4RB_MIXED) launch_mp_to_f32(c, st);
        FMM_HIP(hipEventRecord(c.evu[2], st));
    } else {
        tree_and_lists(c, pos, mom, N, N0, eta, c.ev[2]);
        FMM_HIP(hipEventRecord(c.ev[3], st));
    }
    // the near field needs only the tree order and the lists: it runs on the
    // auxiliary stream while this one forms the far field, and L2P waits for it
    FMM_HIP(hipEventRecord(c.evx[0], st));
    hipStream_t sn = c.near_stream();
    FMM_HIP(hipStreamWaitEvent(sn, c.evx[0], 0));
    FMM_HIP(hipEventRecord(c.evx[1], sn));
    launch_p2p(c, c.d_p2p_off.as<int32_t>(), c.p2p_sorted, c.sb, sn);
    FMM_HIP(hipEventRecord(c.evx[2], sn));
    // list statistics for the caller, also on the near-field stream (off the far
    // field's path)
    unsigned long long ls[2] = {0, 0};
    if (stats) {
        DevBuf &tmp = c.d_bad;  // reuse: needs 16 bytes
        tmp.ensure(16);
        zero_words(tmp.p, 4, sn);
        const int64_t mx = std::max(c.np2p, c.nm2l);
        if (mx > 0) {
            hipLaunchKernelGGL(k_list_stats, dim3((unsigned)std::min<int64_t>(1024, (mx + 255) / 256)),
                               dim3(256), 0, sn,
                               c.p2p_sorted, c.np2p, c.m2l_sorted, c.nm2l, c.sb,
                               c.d_cnt.as<int32_t>(), tmp.as<unsigned long long>());
            FMM_LAUNCH_CHECK();
        }
    }
    FMM_HIP(hipStreamWaitEvent(st, c.evu[2], 0));  // join the upward pass
    FMM_HIP(hipEventRecord(c.ev[4], st));
    launch_m2l(c, c.d_m2l_off.as<int32_t>(), c.m2l_sorted, c.sb, c.shape.nc);
    FMM_HIP(hipEventRecord(c.ev[5], st));
    if (c.debug) {
        c.snap_loc_m2l.resize(6 * (size_t)c.nn * c.shape.nc);
        FMM_HIP(hipMemcpyAsync(c.snap_loc_m2l.data(), c.d_loc.p, c.snap_loc_m2l.size() * 8,
                               hipMemcpyDeviceToHost, st));
        FMM_HIP(hipStreamSynchronize(st));
    }
    launch_l2l(c, false);
    FMM_HIP(hipStreamWaitEvent(st, c.evx[2], 0));
    FMM_HIP(hipEventRecord(c.ev[6], st));
    launch_leaf_l2p(c, amp, E, B);
    FMM_HIP(hipEventRecord(c.ev[7], st));
    // the copy to pageable host memory blocks the host until the near-field stream
    // reaches it, so it is issued only after every launch of the call
    if (stats) FMM_HIP(hipMemcpyAsync(ls, c.d_bad.p, 16, hipMemcpyDeviceToHost, sn));
    FMM_HIP(hipStreamSynchronize(st));
    if (stats) FMM_HIP(hipStreamSynchronize(sn));
    c.tree_valid = true;
    c.tree_N = N;
    c.tree_N0 = N0;
    c.tree_eta = eta;
    // a refit's locals belong to the stale tree, not to the one a pullback rebuilds
    c.loc_valid = !refit;
    c.loc_digest = c.in_digest;
    c.loc_N = N;
    c.loc_N0 = N0;
    c.loc_n = n;
    c.loc_prec = precision;
    c.loc_eta = eta;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->N = N;
        stats->N0 = N0;
        stats->ncluster = c.shape.nc;
        stats->nleaf = (int64_t)c.shape.leaves.size();
        stats->n = n;
        stats->precision = precision;
        stats->maxlevel = c.shape.ml;
        stats->traversal_iterations = c.trav_iters;
        stats->np2p = c.np2p;
        stats->nm2l = c.nm2l;
        stats->nm2l_groups = (int64_t)ls[1];
        stats->p2p_particle_pairs = (int64_t)ls[0];
        for (int d = 0; d < 3; ++d) stats->p_avg[d] = c.pavg_host[d];
        stats->g_avg = c.pavg_host[3];
        FMM_HIP(hipEventElapsedTime(&stats->ms_total, c.ev[0], c.ev[7]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_pavg, c.ev[0], c.ev[1]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_tree, c.ev[1], c.ev[2]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_lists, c.ev[2], c.ev[3]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_upward, c.evu[1], c.evu[2]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_m2l, c.ev[4], c.ev[5]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_p2p, c.evx[1], c.evx[2]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_downward, c.ev[5], c.ev[7]));
    }
}

// Reverse mode: tree, lists and upward pass rebuilt, then the adjoint.  The forward
// M2L and L2L are recomputed inside it, except for the pullback (saved_locals),
// which takes the final local expansions of the preceding field call on the same
// inputs from loc (verified by parameters and the input digest).  The pullback also
// keeps that call's tree, lists and multipoles when the context still holds them
// (tree_valid with the same N, N0 and eta): the build is deterministic, so
// rebuilding would reproduce them bit for bit, and the adjoint only reads them.
static void run_vjp(Ctx &c, int64_t N, const double *pos, const double *mom, double charge,
                    double lambda, int n, int64_t N0, double eta, int precision,
                    const double *Ebar, const double *Bbar, double *gpos, double *gmom,
                    fmm4rb_stats *stats, bool saved_locals = false) {
    validate(N, n, N0, eta, precision);
    if (!std::isfinite(charge) || !std::isfinite(lambda) || lambda == 0.0)
        fail(FMM4RB_EINVAL, "charge must be finite and lambda finite and non-zero");
    const char *nofwd =
        "no matching forward call: the pullback needs this context's last call to be a "
        "completed fmm4rb_field* with the same inputs, N, N0, n, eta and precision";
    if (saved_locals && !(c.loc_valid && c.loc_N == N && c.loc_N0 == N0 && c.loc_n == n &&
                          c.loc_prec == precision && c.loc_eta == eta))
        fail(FMM4RB_EINVAL, nofwd);
    if (!saved_locals) {
        c.loc_valid = false;
        c.tree_valid = false;
    }
    prepare(c, N, N0, n, precision);
    hipStream_t st = c.stream;
    const double amp = (AMP_RE * charge) / lambda;
    FMM_HIP(hipEventRecord(c.ev[0], st));
    stage_pavg(c, pos, mom);
    // a refused pullback leaves the context as it was (same parameters: prepare
    // changed nothing); past this point the tree is being rebuilt
    if (saved_locals && c.in_digest != c.loc_digest) fail(FMM4RB_EINVAL, nofwd);
    const bool keep_tree = saved_locals && c.tree_valid && c.tree_N == N && c.tree_N0 == N0 &&
                           c.tree_eta == eta;
    FMM_HIP(hipEventRecord(c.ev[1], st));
    if (keep_tree) {  // tree, lists and multipoles of the forward call (zero stage times)
        FMM_HIP(hipEventRecord(c.ev[2], st));
        FMM_HIP(hipEventRecord(c.ev[3], st));
        FMM_HIP(hipEventRecord(c.evu[1], st));
        FMM_HIP(hipEventRecord(c.evu[2], st));
    } else {
        c.tree_valid = false;
        tree_and_lists(c, pos, mom, N, N0, eta, c.ev[2]);
        FMM_HIP(hipEventRecord(c.ev[3], st));
        FMM_HIP(hipStreamWaitEvent(st, c.evu[2], 0));  // join the upward pass
    }
    FMM_HIP(hipEventRecord(c.ev[4], st));
    FMM_HIP(hipEventRecord(c.ev[5], st));
    run_adjoint(c, Ebar, Bbar, amp, gpos, gmom, saved_locals);
    FMM_HIP(hipEventRecord(c.ev[6], st));
    FMM_HIP(hipStreamSynchronize(st));
    c.tree_valid = true;
    c.tree_N = N;
    c.tree_N0 = N0;
    c.tree_eta = eta;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->N = N;
        stats->N0 = N0;
        stats->ncluster = c.shape.nc;
        stats->nleaf = (int64_t)c.shape.leaves.size();
        stats->n = n;
        stats->precision = precision;
        stats->maxlevel = c.shape.ml;
        stats->traversal_iterations = c.trav_iters;
        stats->np2p = c.np2p;
        stats->nm2l = c.nm2l;
        for (int d = 0; d < 3; ++d) stats->p_avg[d] = c.pavg_host[d];
        stats->g_avg = c.pavg_host[3];
        FMM_HIP(hipEventElapsedTime(&stats->ms_total, c.ev[0], c.ev[6]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_pavg, c.ev[0], c.ev[1]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_tree, c.ev[1], c.ev[2]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_upward, c.evu[1], c.evu[2]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_lists, c.ev[2], c.ev[3]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_m2l, c.ev[5], c.ev[8]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_p2p, c.evx[1], c.evx[2]));
        FMM_HIP(hipEventElapsedTime(&stats->ms_downward, c.ev[8], c.ev[6]));
    }
}

}  // namespace fmm4rb

using namespace fmm4rb;

// Host <-> device copies of the host-pointer entry points go through two pinned
// chunks: host threads copy one chunk between the caller's (pageable) memory and a
// pinned buffer while the DMA engine moves the other.  Measured on MI355X for 100 MB:
// pageable copies through the runtime 3.6 ms H2D and 6.1 ms D2H; pinned DMA 1.8 ms.
namespace {
constexpr size_t STAGE_CHUNK = (size_t)64 << 20;

void par_memcpy(void *dst, const void *src, size_t bytes) {
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t T = std::max<size_t>(1, std::min<size_t>(8, hw ? hw : 1));
    if (T == 1 || bytes < ((size_t)4 << 20)) {
        std::memcpy(dst, src, bytes);
        return;
    }
    // this thread copies [0, per); helpers take the following slices; whatever no
    // helper took (a thread that could not start) is copied here too
    const size_t per = (bytes + T - 1) / T;
    std::vector<std::thread> th;
    size_t next = per;
    try {
        while (next < bytes && th.size() + 1 < T) {
            const size_t o = next, m = std::min(per, bytes - o);
            th.emplace_back([=] { std::memcpy((char *)dst + o, (const char *)src + o, m); });
            next = o + m;
        }
    } catch (const std::system_error &) {
    }
    std::memcpy(dst, src, per);
    if (next < bytes) std::memcpy((char *)dst + next, (const char *)src + next, bytes - next);
    for (auto &x : th) x.join();
}
}  // namespace

struct HostStage {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {};
    bool used[2] = {false, false};
    void init() {
        if (buf[0]) return;
        for (int b = 0; b < 2; ++b) {
            FMM_HIP(hipHostMalloc(&buf[b], STAGE_CHUNK, hipHostMallocDefault));
            FMM_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
        }
    }
    void release() {
        for (int b = 0; b < 2; ++b) {
            if (ev[b]) (void)hipEventSynchronize(ev[b]);
            if (buf[b]) (void)hipHostFree(buf[b]);
            if (ev[b]) (void)hipEventDestroy(ev[b]);
            buf[b] = nullptr;
            ev[b] = nullptr;
            used[b] = false;
        }
    }
    // ordered on st; returns once the caller's memory may be reused
    void h2d(void *d, const void *h, size_t bytes, hipStream_t st) {
        init();
        for (size_t off = 0, k = 0; off < bytes; off += STAGE_CHUNK, ++k) {
            const size_t n = std::min(STAGE_CHUNK, bytes - off);
            const int b = (int)(k & 1);
            if (used[b]) FMM_HIP(hipEventSynchronize(ev[b]));  // its last DMA has read it
            par_memcpy(buf[b], (const char *)h + off, n);
            FMM_HIP(hipMemcpyAsync((char *)d + off, buf[b], n, hipMemcpyHostToDevice, st));
            FMM_HIP(hipEventRecord(ev[b], st));
            used[b] = true;
        }
    }
    // chunk k is moved by DMA while chunk k-1 is copied out to the caller
    void d2h(void *h, const void *d, size_t bytes, hipStream_t st) {
        init();
        size_t poff = 0, pn = 0;
        int pb = -1;
        for (size_t off = 0, k = 0; off < bytes; off += STAGE_CHUNK, ++k) {
            const size_t n = std::min(STAGE_CHUNK, bytes - off);
            const int b = (int)(k & 1);
            if (used[b]) FMM_HIP(hipEventSynchronize(ev[b]));  // e.g. an H2D on another stream
            FMM_HIP(hipMemcpyAsync(buf[b], (const char *)d + off, n, hipMemcpyDeviceToHost, st));
            FMM_HIP(hipEventRecord(ev[b], st));
            used[b] = true;
            if (pb >= 0) {
                FMM_HIP(hipEventSynchronize(ev[pb]));
                par_memcpy((char *)h + poff, buf[pb], pn);
            }
            pb = b;
            poff = off;
            pn = n;
        }
        if (pb >= 0) {
            FMM_HIP(hipEventSynchronize(ev[pb]));
            par_memcpy((char *)h + poff, buf[pb], pn);
        }
    }
};

struct fmm4rb_ctx {
    Ctx c;
    DevBuf h_pos, h_mom, h_E, h_B, h_tg, h_gpos, h_gmom, h_ob, h_bfpart;
    HostStage stage;
};

#define FMM_API_BEGIN try {
#define FMM_API_END                                   \
    }                                                 \
    catch (const fmm4rb::Error &e) {                  \
        g_err = e.msg;                                \
        return e.code;                                \
    }                                                 \
    catch (const std::exception &e) {                 \
        g_err = e.what();                             \
        return FMM4RB_EDEVICE;                        \
    }                                                 \
    g_err.clear();                                    \
    return FMM4RB_OK;

static void set_device(fmm4rb_ctx *ctx) { FMM_HIP(hipSetDevice(ctx->c.device)); }

#define FMM_NEED(cond, msg) \
    if (!(cond)) fail(FMM4RB_EINVAL, msg)

extern "C" {

const char *fmm4rb_version(void) { return "fmm4rb-hip 0.1.0 (gfx950)"; }

const char *fmm4rb_last_error(void) { return g_err.c_str(); }

int fmm4rb_create(fmm4rb_ctx **out, int device) {
    FMM_API_BEGIN
    FMM_NEED(out, "ctx pointer is NULL");
    int ndev = 0;
    FMM_HIP(hipGetDeviceCount(&ndev));
    FMM_NEED(device >= 0 && device < ndev, "device index out of range");
    FMM_HIP(hipSetDevice(device));
    scratch_init(device);  // once per device: the allocation-free entry point's memory
    fmm4rb_ctx *ctx = new fmm4rb_ctx();
    ctx->c.device = device;
    try {
        FMM_HIP(hipStreamCreateWithFlags(&ctx->c.own, hipStreamNonBlocking));
        FMM_HIP(hipStreamCreateWithFlags(&ctx->c.aux, hipStreamNonBlocking));
        FMM_HIP(hipStreamCreateWithFlags(&ctx->c.aux2, hipStreamNonBlocking));
        ctx->c.stream = ctx->c.own;
        for (auto &e : ctx->c.ev) FMM_HIP(hipEventCreate(&e));
        for (auto &e : ctx->c.evx) FMM_HIP(hipEventCreate(&e));
        for (auto &e : ctx->c.evy) FMM_HIP(hipEventCreate(&e));
        for (auto &e : ctx->c.evu) FMM_HIP(hipEventCreate(&e));
    } catch (...) {
        delete ctx;
        throw;
    }
    *out = ctx;
    FMM_API_END
}

int fmm4rb_destroy(fmm4rb_ctx *ctx) {
    FMM_API_BEGIN
    if (!ctx) return FMM4RB_OK;
    (void)hipSetDevice(ctx->c.device);
    if (ctx->c.own) (void)hipStreamSynchronize(ctx->c.own);
    if (ctx->c.aux) (void)hipStreamSynchronize(ctx->c.aux);
    if (ctx->c.aux2) (void)hipStreamSynchronize(ctx->c.aux2);
    Ctx &c = ctx->c;
    DevBuf *bufs[CTX_NBUF_MAX];
    const int nb = ctx_bufs(c, bufs);
    for (int i = 0; i < nb; ++i) bufs[i]->release();
    DevBuf *hbufs[] = {&ctx->h_pos, &ctx->h_mom, &ctx->h_E, &ctx->h_B,
                       &ctx->h_tg, &ctx->h_gpos, &ctx->h_gmom, &ctx->h_ob, &ctx->h_bfpart};
    for (DevBuf *b : hbufs) b->release();
    ctx->stage.release();
    for (auto &e : c.ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c.evx)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c.evy)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c.evu)
        if (e) (void)hipEventDestroy(e);
    if (c.own) (void)hipStreamDestroy(c.own);
    if (c.aux) (void)hipStreamDestroy(c.aux);
    if (c.aux2) (void)hipStreamDestroy(c.aux2);
    delete ctx;
    FMM_API_END
}

int fmm4rb_set_stream(fmm4rb_ctx *ctx, void *stream) {
    FMM_API_BEGIN
    FMM_NEED(ctx, "ctx is NULL");
    ctx->c.stream = stream ? (hipStream_t)stream : ctx->c.own;
    FMM_API_END
}

int fmm4rb_set_overlap(fmm4rb_ctx *ctx, int on) {
    FMM_API_BEGIN
    FMM_NEED(ctx, "ctx is NULL");
    ctx->c.overlap = on != 0;
    FMM_API_END
}

int fmm4rb_set_debug(fmm4rb_ctx *ctx, int on) {
    FMM_API_BEGIN
    FMM_NEED(ctx, "ctx is NULL");
    ctx->c.debug = on != 0;
    FMM_API_END
}

int fmm4rb_field_device(fmm4rb_ctx *ctx, int64_t N, const double *d_pos, const double *d_mom,
                        double charge, double lambda, int n, int64_t N0, double eta,
                        int precision, double *d_E, double *d_B, fmm4rb_stats *stats) {
    FMM_API_BEGIN
    FMM_NEED(ctx && d_pos && d_mom && d_E && d_B, "NULL argument");
    set_device(ctx);
    run_field(ctx->c, N, d_pos, d_mom, charge, lambda, n, N0, eta, precision, d_E, d_B, stats);
    FMM_API_END
}

// The allocation-free form of fmm4rb_field_device: the context's own buffers are
// set aside, every buffer of the call is carved from this device's static scratch
// (so nothing calls hipMalloc or hipFree), and afterwards the context gets its own
// buffers back, with its tree, lists and locals marked stale (the scratch is
// shared by every context on the device).
int fmm4rb_field_device_noalloc(fmm4rb_ctx *ctx, int64_t N, const double *d_pos,
                                const double *d_mom, double charge, double lambda, int n,
                                int64_t N0, double eta, int precision, double *d_E, double *d_B,
                                fmm4rb_stats *stats) {
    FMM_API_BEGIN
    FMM_NEED(ctx && d_pos && d_mom && d_E && d_B, "NULL argument");
    set_device(ctx);
    Ctx &c = ctx->c;
    void *base = scratch_ptr(c.device);
    if (!base) fail(FMM4RB_EDEVICE, "no static scratch address for this device");
    validate(N, n, N0, eta, precision);
    DevBuf *bufs[CTX_NBUF_MAX];
    DevBuf saved[CTX_NBUF_MAX];
    const int nb = ctx_bufs(c, bufs);
    const int64_t caps[3] = {c.cap_front, c.cap_p2p, c.cap_m2l};
    for (int i = 0; i < nb; ++i) {
        saved[i] = *bufs[i];
        *bufs[i] = DevBuf{};
    }
    c.cap_front = c.cap_p2p = c.cap_m2l = 0;  // sized from the leaf count by prepare()
    c.shape.N = -1;                           // the shape arrays are uploaded to the scratch
    Arena arena;
    arena.base = static_cast<unsigned char *>(base);
    arena.size = SCRATCH_BYTES;
    t_arena = &arena;
    auto restore = [&] {
        t_arena = nullptr;
        for (int i = 0; i < nb; ++i) *bufs[i] = saved[i];
        c.cap_front = caps[0];
        c.cap_p2p = caps[1];
        c.cap_m2l = caps[2];
        c.shape.N = -1;  // the context's own shape arrays are uploaded again on its next call
        c.tree_valid = false;
        c.loc_valid = false;
    };
    try {
        run_field(c, N, d_pos, d_mom, charge, lambda, n, N0, eta, precision, d_E, d_B, stats);
    } catch (...) {
        // no work of this call may still use the scratch when the next call starts
        (void)hipStreamSynchronize(c.stream);
        (void)hipStreamSynchronize(c.aux);
        (void)hipStreamSynchronize(c.aux2);
        restore();
        throw;
    }
    restore();
    FMM_API_END
}

// Host-pointer form: stage pos/mom, run, copy E/B back.
static void field_host(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                       double charge, double lambda, int n, int64_t N0, double eta, int precision,
                       double *E, double *B, fmm4rb_stats *stats, bool refit) {
    validate(N, n, N0, eta, precision);
    set_device(ctx);
    hipStream_t st = ctx->c.stream;
    const size_t bytes = 3 * (size_t)N * sizeof(double);
    ctx->h_pos.ensure(bytes);
    ctx->h_mom.ensure(bytes);
    ctx->h_E.ensure(bytes);
    ctx->h_B.ensure(bytes);
    ctx->stage.h2d(ctx->h_pos.p, pos, bytes, st);
    ctx->stage.h2d(ctx->h_mom.p, mom, bytes, st);
    run_field(ctx->c, N, ctx->h_pos.as<double>(), ctx->h_mom.as<double>(), charge, lambda, n, N0,
              eta, precision, ctx->h_E.as<double>(), ctx->h_B.as<double>(), stats, refit);
    ctx->stage.d2h(E, ctx->h_E.p, bytes, st);
    ctx->stage.d2h(B, ctx->h_B.p, bytes, st);
}

int fmm4rb_field(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom, double charge,
                 double lambda, int n, int64_t N0, double eta, int precision, double *E, double *B,
                 fmm4rb_stats *stats) {
    FMM_API_BEGIN
    FMM_NEED(ctx && pos && mom && E && B, "NULL argument");
    field_host(ctx, N, pos, mom, charge, lambda, n, N0, eta, precision, E, B, stats, false);
    FMM_API_END
}

int fmm4rb_field_refit(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                       double charge, double lambda, int n, int64_t N0, double eta, int precision,
                       double *E, double *B, fmm4rb_stats *stats) {
    FMM_API_BEGIN
    FMM_NEED(ctx && pos && mom && E && B, "NULL argument");
    field_host(ctx, N, pos, mom, charge, lambda, n, N0, eta, precision, E, B, stats, true);
    FMM_API_END
}

int fmm4rb_field_device_refit(fmm4rb_ctx *ctx, int64_t N, const double *d_pos,
                              const double *d_mom, double charge, double lambda, int n,
                              int64_t N0, double eta, int precision, double *d_E, double *d_B,
                              fmm4rb_stats *stats) {
    FMM_API_BEGIN
    FMM_NEED(ctx && d_pos && d_mom && d_E && d_B, "NULL argument");
    set_device(ctx);
    run_field(ctx->c, N, d_pos, d_mom, charge, lambda, n, N0, eta, precision, d_E, d_B, stats,
              true);
    FMM_API_END
}

int fmm4rb_field_vjp_device(fmm4rb_ctx *ctx, int64_t N, const double *d_pos, const double *d_mom,
                            double charge, double lambda, int n, int64_t N0, double eta,
                            int precision, const double *d_Ebar, const double *d_Bbar,
                            double *d_gpos, double *d_gmom, fmm4rb_stats *stats) {
    FMM_API_BEGIN
    FMM_NEED(ctx && d_pos && d_mom && d_Ebar && d_Bbar && d_gpos && d_gmom, "NULL argument");
    set_device(ctx);
    run_vjp(ctx->c, N, d_pos, d_mom, charge, lambda, n, N0, eta, precision, d_Ebar, d_Bbar,
            d_gpos, d_gmom, stats);
    FMM_API_END
}

int fmm4rb_field_pullback_device(fmm4rb_ctx *ctx, int64_t N, const double *d_pos,
                                 const double *d_mom, double charge, double lambda, int n,
                                 int64_t N0, double eta, int precision, const double *d_Ebar,
                                 const double *d_Bbar, double *d_gpos, double *d_gmom,
                                 fmm4rb_stats *stats) {
    FMM_API_BEGIN
    FMM_NEED(ctx && d_pos && d_mom && d_Ebar && d_Bbar && d_gpos && d_gmom, "NULL argument");
    set_device(ctx);
    run_vjp(ctx->c, N, d_pos, d_mom, charge, lambda, n, N0, eta, precision, d_Ebar, d_Bbar,
            d_gpos, d_gmom, stats, true);
    FMM_API_END
}

// Host-pointer reverse mode: stage inputs and cotangents, run, copy gradients back.
static void vjp_host(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                     double charge, double lambda, int n, int64_t N0, double eta, int precision,
                     const double *Ebar, const double *Bbar, double *gpos, double *gmom,
                     fmm4rb_stats *stats, bool saved_locals) {
    validate(N, n, N0, eta, precision);
    set_device(ctx);
    hipStream_t st = ctx->c.stream;
    const size_t bytes = 3 * (size_t)N * sizeof(double);
    DevBuf *b[] = {&ctx->h_pos, &ctx->h_mom, &ctx->h_E, &ctx->h_B, &ctx->h_gpos, &ctx->h_gmom};
    for (DevBuf *x : b) x->ensure(bytes);
    ctx->stage.h2d(ctx->h_pos.p, pos, bytes, st);
    ctx->stage.h2d(ctx->h_mom.p, mom, bytes, st);
    ctx->stage.h2d(ctx->h_E.p, Ebar, bytes, st);
    ctx->stage.h2d(ctx->h_B.p, Bbar, bytes, st);
    run_vjp(ctx->c, N, ctx->h_pos.as<double>(), ctx->h_mom.as<double>(), charge, lambda, n, N0,
            eta, precision, ctx->h_E.as<double>(), ctx->h_B.as<double>(),
            ctx->h_gpos.as<double>(), ctx->h_gmom.as<double>(), stats, saved_locals);
    ctx->stage.d2h(gpos, ctx->h_gpos.p, bytes, st);
    ctx->stage.d2h(gmom, ctx->h_gmom.p, bytes, st);
}

int fmm4rb_field_pullback(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                          double charge, double lambda, int n, int64_t N0, double eta,
                          int precision, const double *Ebar, const double *Bbar, double *gpos,
                          double *gmom, fmm4rb_stats *stats) {
    FMM_API_BEGIN
    FMM_NEED(ctx && pos && mom && Ebar && Bbar && gpos && gmom, "NULL argument");
    vjp_host(ctx, N, pos, mom, charge, lambda, n, N0, eta, precision, Ebar, Bbar, gpos, gmom,
             stats, true);
    FMM_API_END
}

int fmm4rb_field_vjp(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                     double charge, double lambda, int n, int64_t N0, double eta, int precision,
                     const double *Ebar, const double *Bbar, double *gpos, double *gmom,
                     fmm4rb_stats *stats) {
    FMM_API_BEGIN
    FMM_NEED(ctx && pos && mom && Ebar && Bbar && gpos && gmom, "NULL argument");
    vjp_host(ctx, N, pos, mom, charge, lambda, n, N0, eta, precision, Ebar, Bbar, gpos, gmom,
             stats, false);
    FMM_API_END
}

int fmm4rb_bruteforce_vjp(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                          double charge, double lambda, const double *Ebar, const double *Bbar,
                          double *gpos, double *gmom) {
    FMM_API_BEGIN
    FMM_NEED(ctx && pos && mom && Ebar && Bbar && gpos && gmom, "NULL argument");
    FMM_NEED(N >= 1 && N < 0x7fffffffLL, "N out of range");
    FMM_NEED(std::isfinite(charge) && std::isfinite(lambda) && lambda != 0.0,
             "charge must be finite and lambda finite and non-zero");
    set_device(ctx);
    hipStream_t st = ctx->c.stream;
    const double amp = (AMP_RE * charge) / lambda;
    const size_t bytes = 3 * (size_t)N * sizeof(double);
    std::vector<double> ob(6 * (size_t)N);
    for (int64_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) {
            ob[6 * i + d] = amp * Ebar[3 * i + d];
            ob[6 * i + 3 + d] = amp * Bbar[3 * i + d];
        }
    ctx->h_pos.ensure(bytes);
    ctx->h_mom.ensure(bytes);
    ctx->h_ob.ensure(2 * bytes);
    ctx->h_gpos.ensure(bytes);
    ctx->h_gmom.ensure(bytes);
    FMM_HIP(hipMemcpyAsync(ctx->h_pos.p, pos, bytes, hipMemcpyHostToDevice, st));
    FMM_HIP(hipMemcpyAsync(ctx->h_mom.p, mom, bytes, hipMemcpyHostToDevice, st));
    FMM_HIP(hipMemcpyAsync(ctx->h_ob.p, ob.data(), 2 * bytes, hipMemcpyHostToDevice, st));
    launch_bruteforce_vjp(ctx->h_pos.as<double>(), ctx->h_mom.as<double>(), N,
                          ctx->h_ob.as<double>(), ctx->h_gpos.as<double>(),
                          ctx->h_gmom.as<double>(), st);
    FMM_HIP(hipMemcpyAsync(gpos, ctx->h_gpos.p, bytes, hipMemcpyDeviceToHost, st));
    FMM_HIP(hipMemcpyAsync(gmom, ctx->h_gmom.p, bytes, hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    FMM_API_END
}

int fmm4rb_bruteforce(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                      double charge, double lambda, int64_t ntargets, const int64_t *targets,
                      double *E, double *B) {
    FMM_API_BEGIN
    FMM_NEED(ctx && pos && mom && E && B, "NULL argument");
    FMM_NEED(N >= 1, "N must be >= 1");
    FMM_NEED(std::isfinite(charge) && std::isfinite(lambda) && lambda != 0.0,
             "charge must be finite and lambda finite and non-zero");
    const int64_t nt = targets ? ntargets : N;
    FMM_NEED(nt >= 0, "ntargets must be >= 0");
    if (targets)
        for (int64_t a = 0; a < nt; ++a)
            FMM_NEED(targets[a] >= 0 && targets[a] < N, "target index out of range");
    set_device(ctx);
    hipStream_t st = ctx->c.stream;
    const size_t bytes = 3 * (size_t)N * sizeof(double), obytes = 3 * (size_t)nt * sizeof(double);
    ctx->h_pos.ensure(bytes);
    ctx->h_mom.ensure(bytes);
    ctx->h_E.ensure(std::max<size_t>(obytes, 8));
    ctx->h_B.ensure(std::max<size_t>(obytes, 8));
    FMM_HIP(hipMemcpyAsync(ctx->h_pos.p, pos, bytes, hipMemcpyHostToDevice, st));
    FMM_HIP(hipMemcpyAsync(ctx->h_mom.p, mom, bytes, hipMemcpyHostToDevice, st));
    const int64_t *dt = nullptr;
    if (targets && nt > 0) {
        ctx->h_tg.ensure(8 * (size_t)nt);
        FMM_HIP(hipMemcpyAsync(ctx->h_tg.p, targets, 8 * (size_t)nt, hipMemcpyHostToDevice, st));
        dt = ctx->h_tg.as<int64_t>();
    }
    if (nt > 0) {
        const double amp = (AMP_RE * charge) / lambda;
        launch_bruteforce(ctx->h_pos.as<double>(), ctx->h_mom.as<double>(), N, dt, nt, amp,
                          ctx->h_E.as<double>(), ctx->h_B.as<double>(), ctx->h_bfpart, st);
        FMM_HIP(hipMemcpyAsync(E, ctx->h_E.p, obytes, hipMemcpyDeviceToHost, st));
        FMM_HIP(hipMemcpyAsync(B, ctx->h_B.p, obytes, hipMemcpyDeviceToHost, st));
    }
    FMM_HIP(hipStreamSynchronize(st));
    FMM_API_END
}

// Shared front half of the tree / lists hooks: stage inputs, p_avg, tree.
static void hook_tree(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom,
                      int64_t N0, int n) {
    validate(N, n, N0, 0.0, FMM4RB_FP64);
    set_device(ctx);
    Ctx &c = ctx->c;
    c.tree_valid = false;
    c.loc_valid = false;
    hipStream_t st = c.stream;
    const size_t bytes = 3 * (size_t)N * sizeof(double);
    ctx->h_pos.ensure(bytes);
    ctx->h_mom.ensure(bytes);
    FMM_HIP(hipMemcpyAsync(ctx->h_pos.p, pos, bytes, hipMemcpyHostToDevice, st));
    FMM_HIP(hipMemcpyAsync(ctx->h_mom.p, mom, bytes, hipMemcpyHostToDevice, st));
    prepare(c, N, N0, n, FMM4RB_FP64);
    stage_pavg(c, ctx->h_pos.as<double>(), ctx->h_mom.as<double>());
    build_tree_device(c, ctx->h_pos.as<double>(), ctx->h_mom.as<double>(), N, N0, false);
    uint32_t tflag = 0;
    FMM_HIP(hipMemcpyAsync(&tflag, c.d_tcnt.as<uint32_t>() + TCNT_FLAG, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    if (tflag) build_tree_device(c, ctx->h_pos.as<double>(), ctx->h_mom.as<double>(), N, N0, true);
    launch_node_geo(c);
}

int fmm4rb_tree(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom, int64_t N0,
                int64_t *parindices, int64_t *lo, int64_t *hi, double *bboxes, int64_t *children,
                double *p_avg) {
    FMM_API_BEGIN
    FMM_NEED(ctx && pos && mom, "NULL argument");
    hook_tree(ctx, N, pos, mom, N0, 1);
    Ctx &c = ctx->c;
    const Shape &S = c.shape;
    const int64_t nc = S.nc;
    std::vector<int32_t> pi((size_t)N);
    std::vector<double> bb(6 * (size_t)nc);
    FMM_HIP(hipMemcpyAsync(pi.data(), c.d_parind.p, 4 * (size_t)N, hipMemcpyDeviceToHost, c.stream));
    FMM_HIP(hipMemcpyAsync(bb.data(), c.d_bbox.p, 8 * bb.size(), hipMemcpyDeviceToHost, c.stream));
    FMM_HIP(hipStreamSynchronize(c.stream));
    if (parindices)
        for (int64_t i = 0; i < N; ++i) parindices[i] = pi[i];
    for (int64_t k = 0; k < nc; ++k) {
        const bool ex = S.flags[k] & 1, leaf = S.flags[k] & 2;
        if (lo) lo[k] = ex ? S.lo[k] : -1;
        if (hi) hi[k] = ex ? (int64_t)S.lo[k] + S.cnt[k] - 1 : -1;
        if (children) {
            children[2 * k] = (ex && !leaf) ? 2 * k + 1 : -1;
            children[2 * k + 1] = (ex && !leaf) ? 2 * k + 2 : -1;
        }
        if (bboxes)
            for (int d = 0; d < 6; ++d) bboxes[6 * k + d] = ex ? bb[6 * k + d] : 0.0;
    }
    if (p_avg)
        for (int d = 0; d < 3; ++d) p_avg[d] = c.pavg_host[d];
    FMM_API_END
}

int fmm4rb_lists(fmm4rb_ctx *ctx, int64_t N, const double *pos, const double *mom, int64_t N0,
                 double eta, int64_t cap_p2p, int64_t *p2p, int64_t cap_m2l, int64_t *m2l,
                 int64_t *np2p, int64_t *nm2l) {
    FMM_API_BEGIN
    FMM_NEED(ctx && pos && mom && np2p && nm2l, "NULL argument");
    FMM_NEED(eta >= 0.0 && std::isfinite(eta), "eta must be finite and >= 0");
    hook_tree(ctx, N, pos, mom, N0, 1);
    Ctx &c = ctx->c;
    if (!build_lists_device(c, eta)) fail(FMM4RB_EDEVICE, "tree presort flag after the exact tree");
    const uint64_t mask = (1ull << c.sb) - 1ull;
    auto fetch = [&](const uint64_t *src, int64_t n, int64_t cap, int64_t *out) {
        const int64_t m = std::min(n, cap);
        if (m <= 0 || !out) return;
        std::vector<uint64_t> h((size_t)m);
        FMM_HIP(hipMemcpyAsync(h.data(), src, 8 * (size_t)m, hipMemcpyDeviceToHost, c.stream));
        FMM_HIP(hipStreamSynchronize(c.stream));
        for (int64_t i = 0; i < m; ++i) {
            out[2 * i] = (int64_t)(h[i] >> c.sb);
            out[2 * i + 1] = (int64_t)(h[i] & mask);
        }
    };
    fetch(c.p2p_sorted, c.np2p, cap_p2p, p2p);
    fetch(c.m2l_sorted, c.nm2l, cap_m2l, m2l);
    *np2p = c.np2p;
    *nm2l = c.nm2l;
    FMM_API_END
}

// Operator hooks build a throw-away shape from the caller's clusters; the
// cached shape of the field pipeline is invalidated.
static void hook_shape(Ctx &c, int64_t nc, const int64_t *lo, const int64_t *hi, int64_t N) {
    Shape &S = c.shape;
    S.N = -1;
    S.N0 = -1;
    S.nc = nc;
    int ml = 0;
    while ((((int64_t)1 << (ml + 1)) - 1) < nc) ++ml;
    S.ml = ml;
    S.lo.assign(nc, -1);
    S.cnt.assign(nc, 0);
    S.flags.assign(nc, 1);
    S.maxleaf = 0;
    if (lo && hi)
        for (int64_t k = 0; k < nc; ++k) {
            if (lo[k] < 0) continue;
            FMM_NEED(lo[k] <= hi[k] + 1 && hi[k] < N, "cluster range out of bounds");
            S.lo[k] = (int32_t)lo[k];
            S.cnt[k] = (int32_t)(hi[k] - lo[k] + 1);
            S.maxleaf = std::max(S.maxleaf, S.cnt[k]);
        }
    c.sb = 1;
    while (((int64_t)1 << c.sb) < nc) ++c.sb;
}

static void hook_pairs(Ctx &c, int64_t nc, int64_t npairs, const int64_t *pairs, DevBuf &keys,
                       DevBuf &keys_alt, DevBuf &off, uint64_t **sorted) {
    std::vector<uint64_t> h((size_t)std::max<int64_t>(npairs, 1));
    for (int64_t i = 0; i < npairs; ++i) {
        FMM_NEED(pairs[2 * i] >= 0 && pairs[2 * i] < nc && pairs[2 * i + 1] >= 0 &&
                     pairs[2 * i + 1] < nc,
                 "pair node out of range");
        h[i] = ((uint64_t)pairs[2 * i] << c.sb) | (uint64_t)pairs[2 * i + 1];
    }
    keys.ensure(8 * h.size());
    keys_alt.ensure(8 * h.size());
    off.ensure(4 * (size_t)(nc + 1));
    c.d_rs_scratch.ensure(4 * radix_scratch_words(std::max<int64_t>(npairs, 1), 1));
    FMM_HIP(hipMemcpyAsync(keys.p, h.data(), 8 * (size_t)npairs, hipMemcpyHostToDevice, c.stream));
    sort_pairs_csr(c, keys.as<uint64_t>(), keys_alt.as<uint64_t>(), npairs, off.as<int32_t>(), sorted);
    FMM_HIP(hipStreamSynchronize(c.stream));
}

int fmm4rb_op_m2l(fmm4rb_ctx *ctx, int n, int precision, int64_t nc, const double *bboxes,
                  const double *gam, const double *mmom, int64_t npairs, const int64_t *pairs,
                  const double *p_avg, double *le, double *lb) {
    FMM_API_BEGIN
    FMM_NEED(ctx && bboxes && gam && mmom && p_avg && le && lb, "NULL argument");
    FMM_NEED(npairs == 0 || pairs, "pairs is NULL");
    FMM_NEED(n >= 1 && n <= MAX_NP - 1, "n must be in [1, 10]");
    FMM_NEED(nc >= 1 && nc < (1 << 28), "nc out of range");
    FMM_NEED(precision == FMM4RB_FP64 || precision == FMM4RB_MIXED, "bad precision");
    set_device(ctx);
    Ctx &c = ctx->c;
    c.tree_valid = false;
    c.loc_valid = false;
    hipStream_t st = c.stream;
    hook_shape(c, nc, nullptr, nullptr, 0);
    c.n = n;
    c.np = n + 1;
    c.nn = c.np * c.np * c.np;
    c.precision = precision;
    const size_t NN = (size_t)c.nn;
    upload(c.d_flags, c.shape.flags, st);
    c.d_bbox.ensure(6 * (size_t)nc * 8);
    c.d_geo.ensure(8 * (size_t)nc * 8);
    c.d_ncoord.ensure(3 * (size_t)c.np * nc * 8);
    c.d_mp4.ensure(4 * NN * nc * 8);
    c.d_mp4f.ensure(4 * NN * nc * 4);
    c.d_loc.ensure(6 * NN * nc * 8);
    c.d_pavg.ensure(32);
    double pa[4] = {p_avg[0], p_avg[1], p_avg[2],
                    std::sqrt(1.0 + ((p_avg[0] * p_avg[0] + p_avg[1] * p_avg[1]) + p_avg[2] * p_avg[2]))};
    std::vector<double> mp(4 * NN * nc);
    for (size_t e = 0; e < NN * (size_t)nc; ++e) {
        mp[4 * e] = gam[e];
        for (int d = 0; d < 3; ++d) mp[4 * e + 1 + d] = mmom[3 * e + d];
    }
    FMM_HIP(hipMemcpyAsync(c.d_pavg.p, pa, 32, hipMemcpyHostToDevice, st));
    for (int d = 0; d < 4; ++d) c.pavg_host[d] = pa[d];  // launch_m2l reads it
    FMM_HIP(hipMemcpyAsync(c.d_bbox.p, bboxes, 48 * (size_t)nc, hipMemcpyHostToDevice, st));
    FMM_HIP(hipMemcpyAsync(c.d_mp4.p, mp.data(), 8 * mp.size(), hipMemcpyHostToDevice, st));
    launch_node_geo(c);
    if (precision == FMM4RB_MIXED) launch_mp_to_f32(c, c.stream);
    uint64_t *sorted = nullptr;
    hook_pairs(c, nc, npairs, pairs, c.d_m2l, c.d_m2l_alt, c.d_m2l_off, &sorted);
    c.cap_m2l = std::max<int64_t>(c.cap_m2l, (int64_t)(c.d_m2l.cap / 8));
    launch_m2l(c, c.d_m2l_off.as<int32_t>(), sorted, c.sb, nc);
    std::vector<double> loc(6 * NN * nc);
    FMM_HIP(hipMemcpyAsync(loc.data(), c.d_loc.p, 8 * loc.size(), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    for (size_t e = 0; e < NN * (size_t)nc; ++e)
        for (int d = 0; d < 3; ++d) {
            le[3 * e + d] = loc[6 * e + d];
            lb[3 * e + d] = loc[6 * e + 3 + d];
        }
    FMM_API_END
}

int fmm4rb_op_p2p(fmm4rb_ctx *ctx, int precision, int64_t N, const double *pos, const double *mom,
                  const int64_t *parindices, int64_t nc, const int64_t *lo, const int64_t *hi,
                  const double *bboxes, int64_t npairs, const int64_t *pairs, double *E,
                  double *B) {
    FMM_API_BEGIN
    FMM_NEED(ctx && pos && mom && parindices && lo && hi && E && B, "NULL argument");
    FMM_NEED(npairs == 0 || pairs, "pairs is NULL");
    FMM_NEED(N >= 1 && N < 0x7fffffffLL, "N out of range");
    FMM_NEED(nc >= 1 && nc < (1 << 28), "nc out of range");
    FMM_NEED(precision == FMM4RB_FP64 || precision == FMM4RB_MIXED, "bad precision");
    for (int64_t i = 0; i < N; ++i)
        FMM_NEED(parindices[i] >= 0 && parindices[i] < N, "parindices out of range");
    set_device(ctx);
    Ctx &c = ctx->c;
    c.tree_valid = false;
    c.loc_valid = false;
    hipStream_t st = c.stream;
    hook_shape(c, nc, lo, hi, N);
    c.precision = precision;
    c.N = N;
    Shape &S = c.shape;
    FMM_NEED(S.maxleaf <= MAX_N0, "cluster larger than 1024 particles");
    // the targets of the listed pairs act as the leaves
    std::vector<uint8_t> seen((size_t)nc, 0);
    S.leaves.clear();
    for (int64_t i = 0; i < npairs; ++i)
        for (int side = 0; side < 2; ++side) {
            const int64_t k = pairs[2 * i + side];
            FMM_NEED(k >= 0 && k < nc && S.lo[k] >= 0, "pair refers to a cluster without range");
            if (side == 0 && !seen[k]) {
                seen[k] = 1;
                S.leaves.push_back((int32_t)k);
            }
        }
    {
        // each target's workgroup overwrites its particles' rows: target ranges must
        // not overlap (e.g. a parent and its child both listed as targets)
        std::vector<std::pair<int32_t, int32_t>> r;
        for (int32_t k : S.leaves)
            if (S.cnt[k] > 0) r.emplace_back(S.lo[k], S.lo[k] + S.cnt[k] - 1);
        std::sort(r.begin(), r.end());
        for (size_t i = 1; i < r.size(); ++i)
            FMM_NEED(r[i].first > r[i - 1].second,
                     "pair targets have overlapping particle ranges");
    }
    (void)bboxes;  // cluster boxes are not needed: R is formed from absolute positions
    std::vector<double> sp(8 * (size_t)N);
    std::vector<float> spf(4 * (size_t)N);
    for (int64_t i = 0; i < N; ++i) {
        const int64_t q = parindices[i];
        const double *m = mom + 3 * q;
        const double g = std::sqrt(1.0 + ((m[0] * m[0] + m[1] * m[1]) + m[2] * m[2]));
        for (int d = 0; d < 3; ++d) {
            sp[8 * i + d] = pos[3 * q + d];
            sp[8 * i + 4 + d] = m[d];
            spf[4 * i + 1 + d] = (float)m[d];
        }
        sp[8 * i + 3] = g;
        spf[4 * i] = (float)g;
        sp[8 * i + 7] = 0.0;
    }
    upload(c.d_lo, S.lo, st);
    upload(c.d_cnt, S.cnt, st);
    upload(c.d_flags, S.flags, st);
    upload(c.d_leaves, S.leaves, st);
    c.d_sp8.ensure(8 * sp.size());
    c.d_sp8f.ensure(4 * spf.size());
    c.d_near.ensure(6 * (size_t)N * 8);
    FMM_HIP(hipMemcpyAsync(c.d_sp8.p, sp.data(), 8 * sp.size(), hipMemcpyHostToDevice, st));
    FMM_HIP(hipMemcpyAsync(c.d_sp8f.p, spf.data(), 4 * spf.size(), hipMemcpyHostToDevice, st));
    FMM_HIP(hipMemsetAsync(c.d_near.p, 0, 6 * (size_t)N * 8, st));
    uint64_t *sorted = nullptr;
    hook_pairs(c, nc, npairs, pairs, c.d_p2p, c.d_p2p_alt, c.d_p2p_off, &sorted);
    c.cap_p2p = std::max<int64_t>(c.cap_p2p, (int64_t)(c.d_p2p.cap / 8));
    if (!S.leaves.empty()) launch_p2p(c, c.d_p2p_off.as<int32_t>(), sorted, c.sb, c.stream);
    std::vector<double> nf(6 * (size_t)N);
    FMM_HIP(hipMemcpyAsync(nf.data(), c.d_near.p, 8 * nf.size(), hipMemcpyDeviceToHost, st));
    FMM_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < N; ++i) {
        const int64_t q = parindices[i];
        for (int d = 0; d < 3; ++d) {
            E[3 * q + d] = nf[6 * i + d];
            B[3 * q + d] = nf[6 * i + 3 + d];
        }
    }
    FMM_API_END
}

int fmm4rb_debug_fetch(fmm4rb_ctx *ctx, int what, double *out, int64_t count) {
    FMM_API_BEGIN
    FMM_NEED(ctx && out, "NULL argument");
    set_device(ctx);
    Ctx &c = ctx->c;
    FMM_NEED(c.shape.N > 0, "no field evaluation on this context yet");
    const size_t NN = (size_t)c.nn, nc = (size_t)c.shape.nc, N = (size_t)c.N;
    hipStream_t st = c.stream;
    if (what == 1) {
        FMM_NEED(count >= (int64_t)(4 * NN * nc), "out too small");
        std::vector<double> mp(4 * NN * nc);
        FMM_HIP(hipMemcpyAsync(mp.data(), c.d_mp4.p, 8 * mp.size(), hipMemcpyDeviceToHost, st));
        FMM_HIP(hipStreamSynchronize(st));
        for (size_t e = 0; e < NN * nc; ++e) {
            out[e] = mp[4 * e];
            for (int d = 0; d < 3; ++d) out[NN * nc + 3 * e + d] = mp[4 * e + 1 + d];
        }
    } else if (what == 2 || what == 3) {
        FMM_NEED(count >= (int64_t)(6 * NN * nc), "out too small");
        std::vector<double> loc;
        if (what == 2) {
            FMM_NEED(c.snap_loc_m2l.size() == 6 * NN * nc, "no snapshot: enable debug first");
            loc = c.snap_loc_m2l;
        } else {
            loc.resize(6 * NN * nc);
            FMM_HIP(hipMemcpyAsync(loc.data(), c.d_loc.p, 8 * loc.size(), hipMemcpyDeviceToHost, st));
            FMM_HIP(hipStreamSynchronize(st));
        }
        for (size_t e = 0; e < NN * nc; ++e)
            for (int d = 0; d < 3; ++d) {
                out[3 * e + d] = loc[6 * e + d];
                out[3 * NN * nc + 3 * e + d] = loc[6 * e + 3 + d];
            }
    } else if (what == 4) {
        FMM_NEED(count >= (int64_t)(6 * N), "out too small");
        std::vector<double> nf(6 * N);
        std::vector<int32_t> pi(N);
        FMM_HIP(hipMemcpyAsync(nf.data(), c.d_near.p, 8 * nf.size(), hipMemcpyDeviceToHost, st));
        FMM_HIP(hipMemcpyAsync(pi.data(), c.d_parind.p, 4 * N, hipMemcpyDeviceToHost, st));
        FMM_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < N; ++i)
            for (int d = 0; d < 3; ++d) {
                out[3 * (size_t)pi[i] + d] = nf[6 * i + d];
                out[3 * N + 3 * (size_t)pi[i] + d] = nf[6 * i + 3 + d];
            }
    } else {
        fail(FMM4RB_EINVAL, "unknown debug array");
    }
    FMM_API_END
}

}  // extern "C"
