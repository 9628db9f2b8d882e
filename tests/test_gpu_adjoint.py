"""Reverse mode (fmm4rb_field_vjp, fmm4rb_bruteforce_vjp) against central finite
differences of the forward maps, against each other where they must agree
exactly (eta = 0 leaves only P2P), and against the oracle's forward FMM.

Loss L = sum_i w_i (|E_i|^2 + |B_i|^2), w ~ U(0, 1): cotangents 2 w E, 2 w B.
Finite differences use h = 1e-6 of the coordinate scale and only coordinates
whose perturbation leaves the tree and the lists unchanged (the gradient is
that of the fixed-topology map)."""
import numpy as np
import pytest

from fmm4rb_hip import workloads as W

pytestmark = pytest.mark.gpu
LAM = 1e-3


def _pick(rng, N, k):
    return [(int(i), int(a), int(d)) for i, a, d in
            zip(rng.integers(0, N, k), rng.integers(0, 2, k), rng.integers(0, 3, k))]


def _fd(fields, w, pos, mom, i, arr, d, h):
    """Central difference of L = sum w (|E|^2 + |B|^2), formed as
    sum w (F+ - F-).(F+ + F-) so that the large L itself never cancels."""
    xs = [pos.copy(), mom.copy()]
    xs[arr][i, d] += h
    Ep, Bp = fields(*xs)
    xs[arr][i, d] -= 2 * h
    Em, Bm = fields(*xs)
    dl = (w[:, None] * ((Ep - Em) * (Ep + Em) + (Bp - Bm) * (Bp + Bm))).sum()
    return float(dl) / (2 * h)


def _steps(pos, mom):
    return (1e-6 * np.abs(pos).std(axis=0).max(), 1e-6 * max(np.abs(mom).std(axis=0).max(), 1e-2))


def test_bruteforce_vjp_matches_finite_differences(ctx):
    pos, mom = W.bunch(400, seed=50)
    w = np.random.default_rng(51).random(len(pos))
    E, B = ctx.bruteforce(pos, mom, -1.0, LAM)
    gx, gp = ctx.bruteforce_vjp(pos, mom, 2 * w[:, None] * E, 2 * w[:, None] * B, -1.0, LAM)
    fields = lambda x, p: ctx.bruteforce(x, p, -1.0, LAM)
    hs = _steps(pos, mom)
    errs = []
    for i, a, d in _pick(np.random.default_rng(52), len(pos), 24):
        g = (gx, gp)[a][i, d]
        fd = _fd(fields, w, pos, mom, i, a, d, hs[a])
        errs.append(abs(g - fd) / (abs(fd) + 1e-3 * np.sqrt(np.mean((gx, gp)[a] ** 2))))
    assert np.median(errs) <= 1e-7 and max(errs) <= 1e-4, errs


def test_coincident_particles_vjp(ctx):
    """Two distinct particles at the same position: the regularised kernel is smooth
    at R = 0 (K(R) ~ R eps^-3/2), so its derivative there enters the gradient; only the
    self pair is skipped.  BruteForce VJP vs FD on the duplicated coordinates, and the
    FMM's P2P-only VJP (eta = 0) equal to it."""
    pos, mom = W.bunch(300, seed=70)
    pos[1] = pos[0]
    w = np.random.default_rng(71).random(len(pos))
    E, B = ctx.bruteforce(pos, mom, -1.0, LAM)
    Eb, Bb = 2 * w[:, None] * E, 2 * w[:, None] * B
    gx, gp = ctx.bruteforce_vjp(pos, mom, Eb, Bb, -1.0, LAM)
    fields = lambda x, p: ctx.bruteforce(x, p, -1.0, LAM)
    # around R = 0 the kernel varies on the scale sqrt(eps) / |p| ~ 1.5e-10: position
    # steps of the duplicated pair must be far below it to see the derivative
    hs = (1e-14, _steps(pos, mom)[1])
    errs = []
    for i, a, d in [(0, 0, 0), (0, 0, 2), (1, 0, 1), (1, 0, 2), (0, 1, 0), (1, 1, 2)]:
        g = (gx, gp)[a][i, d]
        fd = _fd(fields, w, pos, mom, i, a, d, hs[a])
        errs.append(abs(g - fd) / (abs(fd) + 1e-3 * np.sqrt(np.mean((gx, gp)[a] ** 2))))
    assert max(errs) <= 1e-4, errs
    Ef, Bf, _ = ctx.field(pos, mom, -1.0, LAM, 4, 64, 0.0)
    gxf, gpf, _ = ctx.field_vjp(pos, mom, 2 * w[:, None] * Ef, 2 * w[:, None] * Bf, -1.0, LAM,
                                4, 64, 0.0)
    assert W.rel_l2(gxf, gx) <= 1e-12 and W.rel_l2(gpf, gp) <= 1e-12


def test_fmm_vjp_eta0_equals_bruteforce_vjp(ctx):
    pos, mom = W.bunch(3000, seed=53)
    w = np.random.default_rng(54).random(len(pos))
    E, B, _ = ctx.field(pos, mom, -1.0, LAM, 4, 64, 0.0)
    Eb, Bb = 2 * w[:, None] * E, 2 * w[:, None] * B
    gx, gp, _ = ctx.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, 4, 64, 0.0)
    gxb, gpb = ctx.bruteforce_vjp(pos, mom, Eb, Bb, -1.0, LAM)
    assert W.rel_l2(gx, gxb) <= 1e-12 and W.rel_l2(gp, gpb) <= 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fmm_vjp_eta0_large_leaves(ctx, prec):
    """N0 = 1024 (leaves of up to 1024 particles: several passes per block) with
    eta = 0: the FMM reverse mode is P2P^T only and equals the BruteForce VJP."""
    pos, mom = W.bunch(5000, seed=72)
    w = np.random.default_rng(73).random(len(pos))
    E, B = ctx.bruteforce(pos, mom, -1.0, LAM)
    Eb, Bb = 2 * w[:, None] * E, 2 * w[:, None] * B
    gxb, gpb = ctx.bruteforce_vjp(pos, mom, Eb, Bb, -1.0, LAM)
    gx, gp, _ = ctx.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, 4, 1024, 0.0, precision=prec)
    tol = 1e-12 if prec == "fp64" else 1e-5
    assert W.rel_l2(gx, gxb) <= tol and W.rel_l2(gp, gpb) <= tol


def _topology(ctx, pos, mom, N0, eta):
    t = ctx.tree(pos, mom, N0)
    p2p, m2l = ctx.lists(pos, mom, N0, eta)
    return t["parindices"], p2p, m2l


def _same_topology(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("n", [3, 4, 6, 7, 9])
def test_fmm_vjp_matches_finite_differences(ctx, n):
    _fmm_fd_check(ctx, n, 4096, 64, 55)


def test_fmm_vjp_matches_finite_differences_two_wave_leaves(ctx):
    """Leaves of 93-94 particles: the leaf geometry kernels (L2P^T, P2M^T) run two
    waves per leaf instead of one."""
    pos, _ = W.bunch(6000, seed=58)
    t = ctx.tree(pos, np.zeros_like(pos), 125)
    leaf = (t["lo"] >= 0) & (t["children"][:, 0] < 0)
    assert (t["hi"][leaf] - t["lo"][leaf] + 1).max() > 64
    _fmm_fd_check(ctx, 4, 6000, 125, 58)


def _fmm_fd_check(ctx, n, N, N0, seed):
    eta = 0.5
    pos, mom = W.bunch(N, seed=seed)
    w = np.random.default_rng(seed + 1).random(len(pos))
    E, B, _ = ctx.field(pos, mom, -1.0, LAM, n, N0, eta)
    gx, gp, st = ctx.field_vjp(pos, mom, 2 * w[:, None] * E, 2 * w[:, None] * B, -1.0, LAM, n,
                               N0, eta)
    fields = lambda x, p: ctx.field(x, p, -1.0, LAM, n, N0, eta)[:2]
    hs = _steps(pos, mom)
    base = _topology(ctx, pos, mom, N0, eta)
    rng = np.random.default_rng(seed + 2)
    picks = _pick(rng, len(pos), 40)
    # bounding-box extremes: particles whose coordinate is a leaf's min or max
    t = ctx.tree(pos, mom, N0)
    leaves = np.nonzero((t["lo"] >= 0) & (t["children"][:, 0] < 0))[0]
    for k in rng.choice(leaves, 12, replace=False):
        mem = t["parindices"][t["lo"][k]: t["hi"][k] + 1]
        d = int(rng.integers(0, 3))
        picks.append((int(mem[np.argmax(pos[mem, d])]), 0, d))
    errs, skipped = [], 0
    for i, a, d in picks:
        xs = [pos.copy(), mom.copy()]
        xs[a][i, d] += hs[a]
        up = _topology(ctx, *xs, N0, eta)
        xs[a][i, d] -= 2 * hs[a]
        dn = _topology(ctx, *xs, N0, eta)
        if not (_same_topology(base, up) and _same_topology(base, dn)):
            skipped += 1
            continue
        g = (gx, gp)[a][i, d]
        fd = _fd(fields, w, pos, mom, i, a, d, hs[a])
        scale = np.sqrt(np.mean((gx, gp)[a] ** 2))
        errs.append(abs(g - fd) / (abs(fd) + 1e-3 * scale))
    assert skipped <= len(picks) // 4
    errs = np.array(errs)
    print(f"n={n}: {len(errs)} coordinates, median rel err {np.median(errs):.2e}, "
          f"max {errs.max():.2e}, skipped {skipped}; vjp ms {st.ms_total:.1f}")
    assert np.median(errs) <= 1e-6 and errs.max() <= 1e-4


def test_fmm_vjp_matches_oracle_finite_differences(ctx, orc):
    n, N0, eta = 4, 40, 0.5
    pos, mom = W.bunch(1500, seed=58)
    w = np.random.default_rng(59).random(len(pos))
    E, B, _ = ctx.field(pos, mom, -1.0, LAM, n, N0, eta)
    gx, gp, _ = ctx.field_vjp(pos, mom, 2 * w[:, None] * E, 2 * w[:, None] * B, -1.0, LAM, n,
                              N0, eta)
    fields = lambda x, p: orc.fmm(x, p, -1.0, LAM, n, N0, eta)
    hs = _steps(pos, mom)
    base = _topology(ctx, pos, mom, N0, eta)
    errs = []
    for i, a, d in _pick(np.random.default_rng(60), len(pos), 12):
        xs = [pos.copy(), mom.copy()]
        xs[a][i, d] += hs[a]
        up = _topology(ctx, *xs, N0, eta)
        xs[a][i, d] -= 2 * hs[a]
        if not (_same_topology(base, up) and _same_topology(base, _topology(ctx, *xs, N0, eta))):
            continue
        g = (gx, gp)[a][i, d]
        fd = _fd(fields, w, pos, mom, i, a, d, hs[a])
        errs.append(abs(g - fd) / (abs(fd) + 1e-3 * np.sqrt(np.mean((gx, gp)[a] ** 2))))
    assert len(errs) >= 8 and np.median(errs) <= 1e-6 and max(errs) <= 1e-4


def test_fmm_vjp_close_to_bruteforce_vjp(ctx):
    pos, mom = W.bunch(4096, seed=61)
    w = np.random.default_rng(62).random(len(pos))
    E, B = ctx.bruteforce(pos, mom, -1.0, LAM)
    Eb, Bb = 2 * w[:, None] * E, 2 * w[:, None] * B
    gxb, gpb = ctx.bruteforce_vjp(pos, mom, Eb, Bb, -1.0, LAM)
    gx, gp, _ = ctx.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, 6, 64, 0.5)
    ex, ep = W.rel_l2(gx, gxb), W.rel_l2(gp, gpb)
    print(f"FMM (n=6) vs BruteForce adjoint: pos {ex:.2e} mom {ep:.2e}")
    assert ex <= 2e-2 and ep <= 2e-2


def test_vjp_deterministic(ctx):
    pos, mom = W.bunch(5000, seed=63)
    Eb, Bb = np.ones_like(pos), np.ones_like(pos)
    for prec in ("fp64", "mixed"):
        a = ctx.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, 4, 64, 0.5, precision=prec)
        b = ctx.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, 4, 64, 0.5, precision=prec)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("n", [2, 4, 6, 9])
def test_mixed_vjp_close_to_fp64(ctx, n):
    """Mixed-precision reverse mode within 1e-5 (relative L2) of the FP64 one, for
    the loss L = sum w (|E|^2 + |B|^2) on a 2^15 bunch."""
    pos, mom = W.bunch(1 << 15, seed=64)
    w = np.random.default_rng(65).random(len(pos))
    E, B, _ = ctx.field(pos, mom, -1.0, LAM, n, 125, 0.5)
    Eb, Bb = 2 * w[:, None] * E, 2 * w[:, None] * B
    gx, gp, _ = ctx.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, n, 125, 0.5)
    gxm, gpm, _ = ctx.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, n, 125, 0.5, precision="mixed")
    ex, ep = W.rel_l2(gxm, gx), W.rel_l2(gpm, gp)
    print(f"n={n}: mixed vs fp64 adjoint: pos {ex:.2e} mom {ep:.2e}")
    assert ex <= 1e-5 and ep <= 1e-5
