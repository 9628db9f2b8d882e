"""Parity of the HIP library (through the C ABI) with the CPU oracle.

Bars: bit-exact for the tree (permutation, ranges, bounding boxes) and the
interaction lists; relative L2 <= 1e-12 for every FP64 stage and the final
FP64 fields; mixed precision within 1e-5 of FP64.  The oracle receives the
mean momentum the GPU computed (its own pairwise sum may differ in the last
bit, which could only flip an exact tie of the split axis)."""
import numpy as np
import pytest

import fmm4rb_hip as F
from fmm4rb_hip import workloads as W
from jlcompat import isapprox

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-5


def _case(name):
    if name == "uniform8192":
        return W.uniform_cube(8192, seed=1), 125
    if name == "bunch32k":
        return W.bunch(1 << 15, seed=2), 125
    if name == "ties":
        rng = np.random.default_rng(3)
        pos = rng.integers(0, 6, size=(5000, 3)).astype(float)
        pos[::5, 0] = -0.0
        return (pos, np.zeros((5000, 3))), 30
    if name == "earlyleaf":
        return W.uniform_cube(251, seed=4), 125
    if name == "single":
        return (np.array([[0.1, 0.2, 0.3]]), np.array([[0.0, 0.0, 1.0]])), 125
    if name == "small_n0":
        return W.bunch(1000, seed=5), 7
    if name == "n0_1":  # one particle per leaf
        return W.bunch(300, seed=6), 1
    if name == "n0_1024":  # leaves of up to 1024 particles: several target passes per block
        return W.bunch(20000, seed=7), 1024
    if name == "near_ties":
        # copies of 512 particles moved by +-1e-13 per axis, below the presort's key
        # resolution (range / 2^32 = 2.3e-10): runs of equal keys with different
        # coordinates, half of them in descending id order, which the fix-up orders
        pos, mom = W.uniform_cube(4096, seed=8)
        sgn = np.random.default_rng(9).choice([-1.0, 1.0], size=(512, 3))
        pos[3584:] = pos[:512] + sgn * 1e-13
        return (pos, mom), 125
    if name == "outlier":
        # one particle 1e12 away on x: every other x shares one presort key (a run
        # longer than the fix-up takes), so the tree is built again from exact keys
        pos, mom = W.uniform_cube(4096, seed=10)
        pos[17, 0] = 1e12
        return (pos, mom), 125
    raise KeyError(name)


CASES = ["uniform8192", "bunch32k", "ties", "earlyleaf", "single", "small_n0", "n0_1",
         "n0_1024", "near_ties", "outlier"]


def _oracle_tree(orc, pos, N0, pavg):
    return orc.build_tree(pos, N0, orc.stretch_of(pavg))


@pytest.mark.parametrize("name", CASES)
def test_tree_bit_exact(ctx, orc, name):
    (pos, mom), N0 = _case(name)
    g = ctx.tree(pos, mom, N0)
    # summation order differs: compare at the scale of the summands
    scale = np.abs(mom).mean(axis=0) + 1e-300
    assert np.all(np.abs(g["p_avg"] - orc.p_avg(mom)) <= 1e-13 * scale)
    o = _oracle_tree(orc, pos, N0, g["p_avg"])
    assert np.array_equal(g["parindices"], o["parindices"])
    assert np.array_equal(g["lo"], o["lo"]) and np.array_equal(g["hi"], o["hi"])
    assert np.array_equal(g["children"], o["children"])
    assert np.array_equal(g["bboxes"].view(np.uint64), o["bboxes"].view(np.uint64))


@pytest.mark.parametrize("name,eta", [("uniform8192", 0.5), ("bunch32k", 0.5), ("ties", 0.7),
                                      ("earlyleaf", 0.5), ("single", 0.5), ("small_n0", 0.3),
                                      ("near_ties", 0.5), ("outlier", 0.5),
                                      ("uniform8192", 0.0), ("bunch32k", 999.0)])
def test_lists_exact(ctx, orc, name, eta):
    (pos, mom), N0 = _case(name)
    g = ctx.tree(pos, mom, N0)
    o = _oracle_tree(orc, pos, N0, g["p_avg"])
    op2p, om2l = orc.lists(o, orc.stretch_of(g["p_avg"]), eta)
    gp2p, gm2l = ctx.lists(pos, mom, N0, eta)
    key = lambda a: a[np.lexsort((a[:, 1], a[:, 0]))] if len(a) else a.reshape(0, 2)
    assert np.array_equal(gp2p, key(op2p))
    assert np.array_equal(gm2l, key(om2l))


def test_bruteforce_matches_oracle(ctx, orc):
    pos, mom = W.bunch(3000, seed=6)
    E, B = ctx.bruteforce(pos, mom, -1.0, 1e-3)
    Eo, Bo = orc.bruteforce(pos, mom, -1.0, 1e-3, nthreads=4)
    assert W.rel_l2(E, Eo) <= 1e-14 and W.rel_l2(B, Bo) <= 1e-14
    tg = np.array([0, 17, 2999, 1234])
    E2, B2 = ctx.bruteforce(pos, mom, -1.0, 1e-3, targets=tg)
    assert np.array_equal(E2, E[tg]) and np.array_equal(B2, B[tg])


def test_fp64_stages_match_oracle(ctx, orc):
    n, N0, eta = 4, 64, 0.5
    pos, mom = W.bunch(1 << 14, seed=7)
    ctx.set_debug(True)
    try:
        E, B, st = ctx.field(pos, mom, -1.0, 1e-3, n, N0, eta, "fp64")
        nc = st.ncluster
        gam_g, mm_g = ctx.debug_fetch(1, n, nc, len(pos))
        le_m2l_g, lb_m2l_g = ctx.debug_fetch(2, n, nc, len(pos))
        le_g, lb_g = ctx.debug_fetch(3, n, nc, len(pos))
        En_g, Bn_g = ctx.debug_fetch(4, n, nc, len(pos))
    finally:
        ctx.set_debug(False)
    pavg = np.array(st.p_avg)
    s = orc.stretch_of(pavg)
    t = orc.build_tree(pos, N0, s)
    gam, mm = orc.upward(n, t, pos, mom)
    ex = np.repeat(t["exists"].astype(bool), (n + 1) ** 3)
    assert W.rel_l2(gam_g[ex], gam[ex]) <= FP64_TOL and W.rel_l2(mm_g[ex], mm[ex]) <= FP64_TOL
    p2p, m2l = orc.lists(t, s, eta)
    le, lb = orc.m2l(n, t["bboxes"], m2l, pavg, gam, mm)
    assert W.rel_l2(le_m2l_g[ex], le[ex]) <= FP64_TOL
    assert W.rel_l2(lb_m2l_g[ex], lb[ex]) <= FP64_TOL
    for l in range(1, t["ml"] + 1):
        orc.l2l(n, t, orc.level_nodes(t, l), le, lb)
    assert W.rel_l2(le_g[ex], le[ex]) <= FP64_TOL and W.rel_l2(lb_g[ex], lb[ex]) <= FP64_TOL
    En, Bn = orc.p2p(p2p, t["parindices"], t["lo"], t["hi"], pos, mom)
    assert W.rel_l2(En_g, En) <= FP64_TOL and W.rel_l2(Bn_g, Bn) <= FP64_TOL
    Eo, Bo = orc.fmm(pos, mom, -1.0, 1e-3, n, N0, eta, pavg=pavg)
    assert W.rel_l2(E, Eo) <= FP64_TOL and W.rel_l2(B, Bo) <= FP64_TOL


@pytest.mark.parametrize("n", [2, 4, 6])
def test_fp64_fmm_uniform8192(ctx, orc, n):
    # the N = 8192 unit-cube case: n in {2,4,6}, N0=125, eta=0.5, lambda=1
    pos, mom = W.uniform_cube(8192, seed=1)
    E, B, st = ctx.field(pos, mom, -1.0, 1.0, n, 125, 0.5, "fp64")
    Eo, Bo = orc.fmm(pos, mom, -1.0, 1.0, n, 125, 0.5, pavg=np.array(st.p_avg))
    assert W.rel_l2(E, Eo) <= FP64_TOL
    assert np.array_equal(B, np.zeros_like(B)) and np.array_equal(Bo, np.zeros_like(Bo))


@pytest.mark.parametrize("name", ["bunch32k", "earlyleaf", "single", "small_n0", "ties", "n0_1",
                                  "n0_1024", "near_ties", "outlier"])
def test_fp64_fmm_cases(ctx, orc, name):
    (pos, mom), N0 = _case(name)
    E, B, st = ctx.field(pos, mom, -1.0, 1e-3, 4, N0, 0.5, "fp64")
    Eo, Bo = orc.fmm(pos, mom, -1.0, 1e-3, 4, N0, 0.5, pavg=np.array(st.p_avg))
    assert W.rel_l2(E, Eo) <= FP64_TOL and W.rel_l2(B, Bo) <= FP64_TOL
    Em, Bm, _ = ctx.field(pos, mom, -1.0, 1e-3, 4, N0, 0.5, "mixed")
    assert W.rel_l2(Em, E) <= MIXED_TOL and W.rel_l2(Bm, B) <= MIXED_TOL


@pytest.mark.parametrize("n", [2, 4, 6, 7])
def test_mixed_close_to_fp64(ctx, n):
    pos, mom = W.bunch(1 << 16, seed=8)
    E, B, _ = ctx.field(pos, mom, -1.0, 1e-3, n, 125, 0.5, "fp64")
    Em, Bm, _ = ctx.field(pos, mom, -1.0, 1e-3, n, 125, 0.5, "mixed")
    assert W.rel_l2(Em, E) <= MIXED_TOL and W.rel_l2(Bm, B) <= MIXED_TOL


@pytest.mark.parametrize("beam", ["along_z", "px_0.09", "along_x", "diagonal"])
def test_mixed_close_to_fp64_beam_directions(ctx, beam):
    """The mixed M2L forms R.R + (p.R)^2 as a quadratic in the x offset when
    |p_avg,x| <= 0.1 and in the general form otherwise: both within 1e-5 of FP64,
    including at the bound and for beams moving along x and along a diagonal."""
    pos, mom = W.bunch(1 << 15, seed=13)
    if beam == "px_0.09":
        mom[:, 0] += 0.09
    elif beam == "along_x":
        pos, mom = pos[:, ::-1].copy(), mom[:, ::-1].copy()
    elif beam == "diagonal":
        rot = np.linalg.qr(np.random.default_rng(14).standard_normal((3, 3)))[0]
        pos, mom = pos @ rot.T, mom @ rot.T
    E, B, st = ctx.field(pos, mom, -1.0, 1e-3, 4, 125, 0.5, "fp64")
    Em, Bm, _ = ctx.field(pos, mom, -1.0, 1e-3, 4, 125, 0.5, "mixed")
    assert (abs(st.p_avg[0]) <= 0.1) == (beam in ("along_z", "px_0.09"))
    assert W.rel_l2(Em, E) <= MIXED_TOL and W.rel_l2(Bm, B) <= MIXED_TOL


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_deterministic(ctx, prec):
    pos, mom = W.bunch(1 << 15, seed=9)
    a = ctx.field(pos, mom, -1.0, 1e-3, 4, 125, 0.5, prec)
    b = ctx.field(pos, mom, -1.0, 1e-3, 4, 125, 0.5, prec)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_reference_gpu_m2l_far_field(ctx, prec):
    # test/test_gpu_fmm/test_gpu_interaction.jl:2-48
    n, distance = 3, 10000.0
    bb = np.array([[-1.0, -1.0, -1.0, 1.0, 1.0, 1.0],
                   [-1.0, -1.0, -1.0 + distance, 1.0, 1.0, 1.0 + distance]])
    nn = (n + 1) ** 3
    gam, mm = np.ones(2 * nn), np.zeros((2 * nn, 3))
    le, lb = ctx.op_m2l(n, bb, gam, mm, [(0, 1), (1, 0)], [0.0, 0.0, 0.0], prec)
    far = (n + 1) ** 3 * np.array([0, 0, distance]) / distance ** 3
    assert isapprox(le[nn], far, rtol=1e-3) and isapprox(lb[nn], [0, 0, 0], rtol=1e-3)
    assert isapprox(le[0], -far, rtol=1e-3) and isapprox(lb[0], [0, 0, 0], rtol=1e-3)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_reference_gpu_p2p_far_field(ctx, prec):
    # test/test_gpu_fmm/test_gpu_interaction.jl:50-96
    npar, sep = 200, 10000.0
    pos = np.random.default_rng(10).random((npar, 3))
    pos[: npar // 2, 2] += sep
    mom = np.zeros((npar, 3))
    E, B = ctx.op_p2p(pos, mom, np.arange(npar), [0, 100], [99, 199], [(0, 1), (1, 0)],
                      precision=prec)
    far = npar / 2 * np.array([0, 0, sep]) / sep ** 3
    assert isapprox(E[0], far, rtol=1e-3) and isapprox(B[0], [0, 0, 0], rtol=1e-3)
    assert isapprox(E[npar - 1], -far, rtol=1e-3) and isapprox(B[npar - 1], [0, 0, 0], rtol=1e-3)


def test_op_p2p_rejects_overlapping_targets(ctx):
    # cluster 0 = particles 0..199 contains cluster 1 = 0..99: both as targets would
    # let two workgroups overwrite the same rows
    pos = np.random.default_rng(12).random((200, 3))
    mom = np.zeros((200, 3))
    with pytest.raises(F.FMMError) as ex:
        ctx.op_p2p(pos, mom, np.arange(200), [0, 0, 100], [199, 99, 199], [(0, 2), (1, 2)])
    assert ex.value.code == 1 and "overlapping" in str(ex.value)


def test_op_m2l_matches_oracle(ctx, orc):
    rng = np.random.default_rng(11)
    n, nc = 4, 6
    lo = rng.uniform(-1, 1, (nc, 3)) * 10
    bb = np.concatenate([lo, lo + rng.uniform(0.5, 2.0, (nc, 3))], axis=1)
    nn = (n + 1) ** 3
    gam, mm = rng.random(nn * nc), rng.standard_normal((nn * nc, 3))
    pairs = [(0, 3), (0, 4), (1, 5), (2, 0), (5, 1), (5, 2), (4, 3)]
    pavg = np.array([0.01, -0.02, 50.0])
    le, lb = ctx.op_m2l(n, bb, gam, mm, pairs, pavg, "fp64")
    leo, lbo = orc.m2l(n, bb, pairs, pavg, gam, mm)
    assert W.rel_l2(le, leo) <= FP64_TOL and W.rel_l2(lb, lbo) <= FP64_TOL
    lem, lbm = ctx.op_m2l(n, bb, gam, mm, pairs, pavg, "mixed")
    assert W.rel_l2(lem, leo) <= MIXED_TOL and W.rel_l2(lbm, lbo) <= MIXED_TOL


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 10])
def test_op_m2l_mixed_every_order(ctx, orc, n):
    """Mixed-precision M2L at every order (the general form of R.R + (p.R)^2: the
    mean momentum's x component is 0.3) on admissible pairs: 8 boxes on a grid of
    spacing 5, sides <= 1.5, all ordered pairs; FP64 locals and the oracle agree."""
    rng = np.random.default_rng(40 + n)
    nc = 8
    cen = np.array([[5.0 * (i % 2), 5.0 * ((i // 2) % 2), 5.0 * (i // 4)] for i in range(nc)])
    half = rng.uniform(0.2, 0.75, (nc, 3))
    bb = np.concatenate([cen - half, cen + half], axis=1) + 1e3
    nn = (n + 1) ** 3
    gam, mm = 1 + rng.random(nn * nc), rng.standard_normal((nn * nc, 3))
    pairs = [(t, s) for t in range(nc) for s in range(nc) if t != s]
    pavg = np.array([0.3, -0.2, 20.0])
    leo, lbo = orc.m2l(n, bb, pairs, pavg, gam, mm)
    le, lb = ctx.op_m2l(n, bb, gam, mm, pairs, pavg, "fp64")
    assert W.rel_l2(le, leo) <= FP64_TOL and W.rel_l2(lb, lbo) <= FP64_TOL
    lem, lbm = ctx.op_m2l(n, bb, gam, mm, pairs, pavg, "mixed")
    assert W.rel_l2(lem, leo) <= MIXED_TOL and W.rel_l2(lbm, lbo) <= MIXED_TOL


def test_fp64_fmm_ellipsoid_matches_oracle(ctx, orc):
    """The 25.6M-particle case's distribution (uniform ellipsoid, p_z ~ N(200, 2))
    at N = 2^15: FP64 FMM vs the oracle's serial FMM, and mixed vs FP64."""
    pos, mom = W.ellipsoid(1 << 15, seed=5)
    E, B, st = ctx.field(pos, mom, -1.0, 1e-3, 4, 125, 0.5, "fp64")
    Eo, Bo = orc.fmm(pos, mom, -1.0, 1e-3, 4, 125, 0.5, pavg=np.array(st.p_avg))
    assert W.rel_l2(E, Eo) <= FP64_TOL and W.rel_l2(B, Bo) <= FP64_TOL
    Em, Bm, _ = ctx.field(pos, mom, -1.0, 1e-3, 4, 125, 0.5, "mixed")
    assert W.rel_l2(Em, E) <= MIXED_TOL and W.rel_l2(Bm, B) <= MIXED_TOL


def test_host_api_overwrites(ctx):
    pos, mom = W.bunch(5000, seed=12)
    beam = F.Beam(pos=pos.T, mom=mom.T)
    alg = F.FMMHIP(eta=0.5, N0=125, n=4)
    F.update_particles_field(beam, alg, lambda_=1e-3, ctx=ctx)
    e1 = beam.efields.copy()
    F.update_particles_field(beam, alg, lambda_=1e-3, ctx=ctx)
    assert np.array_equal(beam.efields, e1)
    F.update_particles_field(beam, F.BruteForceHIP(), lambda_=1e-3, ctx=ctx)
    assert W.rel_l2(e1, beam.efields) < 1e-2


@pytest.mark.parametrize("bad", ["nan", "n0", "n", "n0big", "eta", "lambda"])
def test_errors(ctx, bad):
    pos, mom = W.bunch(500, seed=13)
    args = dict(charge=-1.0, lam=1e-3, n=4, N0=125, eta=0.5)
    if bad == "nan":
        pos = pos.copy()
        pos[7, 1] = np.nan
    elif bad == "n0":
        args["N0"] = 0
    elif bad == "n0big":
        args["N0"] = 2000
    elif bad == "n":
        args["n"] = 11
    elif bad == "eta":
        args["eta"] = -1.0
    elif bad == "lambda":
        args["lam"] = 0.0
    with pytest.raises(F.FMMError) as ei:
        ctx.field(pos, mom, **args)
    assert ei.value.code == 1
    E, B, _ = ctx.field(*W.bunch(500, seed=13), -1.0, 1e-3, 4, 125, 0.5)  # context still usable
    assert np.all(np.isfinite(E))


@pytest.mark.parametrize("name", ["A", "B"])
def test_lists_hand_derived_kat(ctx, name):
    # the GPU tree and traversal on the particles of a hand-derived traversal case
    # (tests/traversal_kats.py, from interaction_lists.jl:23-49): the tree is the
    # hand-built one and the lists are the hand-derived ones
    import traversal_kats as TK
    case = TK.CASES[name]
    g = ctx.tree(case["pos"], case["mom"], case["N0"])
    assert np.array_equal(g["children"], TK.CHILDREN)
    assert np.array_equal(g["bboxes"], case["bboxes"])
    gp2p, gm2l = ctx.lists(case["pos"], case["mom"], case["N0"], case["eta"])
    assert np.array_equal(gp2p, TK.sorted_pairs(case["p2p"]))
    assert np.array_equal(gm2l, TK.sorted_pairs(case["m2l"]))
