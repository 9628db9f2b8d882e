"""Pins of the oracle from what the mathematics fixes: brute-force checks of
the tree and the interaction lists, polynomial reproduction of the
interpolation operators, closed-form fields, and FMM-vs-BruteForce limits.
Each check is chosen so that a plausible slip (a transposed index, a dropped
term, a wrong sign, a wrong tie rule) fails it."""
import numpy as np
import pytest

from fmm4rb_hip import workloads as W


def _check_tree(orc, pos, N0, stretch):
    pos = np.ascontiguousarray(pos)
    N = len(pos)
    t = orc.build_tree(pos, N0, stretch)
    pi, lo, hi, ch, bb = t["parindices"], t["lo"], t["hi"], t["children"], t["bboxes"]
    assert sorted(pi.tolist()) == list(range(N))
    for k in range(t["nc"]):
        if not t["exists"][k]:
            continue
        mem = pi[lo[k]: hi[k] + 1]
        x = pos[mem]
        np.testing.assert_array_equal(bb[k, :3], x.min(axis=0) + 0.0)
        np.testing.assert_array_equal(bb[k, 3:], x.max(axis=0) + 0.0)
        c = len(mem)
        if c <= N0:
            assert ch[k, 0] == -1
            assert np.all(np.diff(mem) > 0), "leaf not in ascending index order"
            continue
        ax = int(np.argmax(np.asarray(stretch) * (x.max(axis=0) - x.min(axis=0))))
        order = np.lexsort((mem, x[:, ax]))            # by coordinate, ties by index
        left = set(mem[order[: (c + 1) // 2]].tolist())
        c1, c2 = ch[k]
        assert (c1, c2) == (2 * k + 1, 2 * k + 2)
        assert set(pi[lo[c1]: hi[c1] + 1].tolist()) == left
        assert hi[c1] - lo[c1] + 1 == (c + 1) // 2 and hi[c2] - lo[c2] + 1 == c // 2
        assert orc.inside(bb[c1], bb[k]) and orc.inside(bb[c2], bb[k])
    return t


def test_tree_uniform(orc):
    pos, _ = W.uniform_cube(3000, seed=21)
    _check_tree(orc, pos, 50, [1.0, 1.0, 1.0])


def test_tree_bunch_stretched(orc):
    pos, mom = W.bunch(2500, seed=22)
    _check_tree(orc, pos, 40, orc.stretch_of(orc.p_avg(mom)))


def test_tree_ties_broken_by_index(orc):
    rng = np.random.default_rng(23)
    pos = rng.integers(0, 4, size=(700, 3)).astype(float)   # many equal coordinates
    pos[::7, 1] = -0.0                                       # signed zeros tie with +0
    _check_tree(orc, pos, 20, [1.0, 1.0, 1.0])


def test_tree_early_leaf(orc):
    # N=251, N0=125: a level-1 cluster of 125 becomes a leaf above maxlevel 2
    pos, _ = W.uniform_cube(251, seed=24)
    t = _check_tree(orc, pos, 125, [1.0, 1.0, 1.0])
    assert t["ml"] == 2
    assert t["children"][2, 0] == -1 and t["exists"][5] == 0 and t["exists"][6] == 0


def _coverage(orc, pos, N0, eta, stretch):
    N = len(pos)
    t = orc.build_tree(pos, N0, stretch)
    p2p, m2l = orc.lists(t, stretch, eta)
    cov = np.zeros((N, N), np.int32)
    pi, lo, hi, ch, bb = t["parindices"], t["lo"], t["hi"], t["children"], t["bboxes"]
    for tt, ss in p2p:
        assert ch[tt, 0] == -1 and ch[ss, 0] == -1
        cov[np.ix_(pi[lo[tt]: hi[tt] + 1], pi[lo[ss]: hi[ss] + 1])] += 1
    for tt, ss in m2l:
        assert orc.admissible(bb[tt], bb[ss], stretch, eta)
        cov[np.ix_(pi[lo[tt]: hi[tt] + 1], pi[lo[ss]: hi[ss] + 1])] += 1
    return cov, p2p, m2l


@pytest.mark.parametrize("eta", [0.0, 0.3, 0.5, 0.9, 999.0])
def test_lists_cover_every_pair_once(orc, eta):
    pos, mom = W.bunch(900, seed=25)
    s = orc.stretch_of(orc.p_avg(mom))
    cov, p2p, m2l = _coverage(orc, pos, 30, eta, s)
    assert np.all(cov == 1)
    if eta == 0.0:
        assert len(m2l) == 0
        nleaf = len(orc.leaves(orc.build_tree(pos, 30, s)))
        assert len(p2p) == nleaf * nleaf


def test_lgweight_polynomial_reproduction(orc):
    rng = np.random.default_rng(26)
    for n in (1, 2, 4, 6, 10):
        a, b = sorted(rng.uniform(-2, 3, 2))
        nodes = orc.cheb2(n, a, b)
        coef = rng.standard_normal(n + 1)
        for x in rng.uniform(a, b, 20):
            w = orc.lgweight(x, nodes)
            assert abs(w.sum() - 1.0) < 1e-12
            assert abs(w @ np.polyval(coef, nodes) - np.polyval(coef, x)) < 1e-10
        w = orc.lgweight(nodes[2 % (n + 1)], nodes)            # on a node: indicator
        assert np.array_equal(w, np.eye(n + 1)[2 % (n + 1)])


def test_singular_rules_differ_only_on_degenerate_boxes(orc):
    nodes = orc.cheb2(4, 0.3, 0.3)           # zero-width box: all nodes coincide
    assert np.array_equal(orc.lgweight(0.3, nodes, orc.SINGULAR_ALL), np.ones(5))
    assert np.array_equal(orc.lgweight(0.3, nodes, orc.SINGULAR_LAST), np.eye(5)[4])
    nodes = orc.cheb2(4, 0.0, 1.0)
    for x in (0.0, 0.5, 1.0, 0.123):
        assert np.array_equal(orc.lgweight(x, nodes, orc.SINGULAR_ALL),
                              orc.lgweight(x, nodes, orc.SINGULAR_LAST))


def _node_coords(orc, n, bb):
    return [orc.cheb2(n, bb[d], bb[3 + d]) for d in range(3)]


def _moments(orc, n, t, gam, k, f):
    """sum over the nodes of cluster k of gam[node] * f(x, y, z)."""
    xs, ys, zs = _node_coords(orc, n, t["bboxes"][k])
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")   # x fastest like the layout
    nn = (n + 1) ** 3
    return float((gam[k * nn:(k + 1) * nn] * f(X.ravel(), Y.ravel(), Z.ravel())).sum())


def test_p2m_m2m_reproduce_polynomial_moments(orc):
    n, N0 = 3, 40
    pos, mom = W.bunch(1200, seed=27)
    pos = pos / 1e-3
    s = [1.0, 1.0, 1.0]
    t = orc.build_tree(pos, N0, s)
    gam, mm = orc.upward(n, t, pos, mom)
    g = np.sqrt(1.0 + (mom * mom).sum(axis=1))
    f = lambda x, y, z: x ** 3 * y ** 2 * z + 0.5 * x * z ** 3 - y  # degree <= 3 per axis
    # every cluster: its node moments equal its particles' moments
    for k in range(t["nc"]):
        mem = t["parindices"][t["lo"][k]: t["hi"][k] + 1]
        ref = float((g[mem] * f(*pos[mem].T)).sum())
        assert abs(_moments(orc, n, t, gam, k, f) - ref) <= 1e-10 * max(1.0, abs(ref)) * len(mem)


def test_l2l_l2p_reproduce_polynomials(orc):
    n, N0 = 4, 60
    pos, _ = W.uniform_cube(1000, seed=28)
    s = [1.0, 1.0, 1.0]
    t = orc.build_tree(pos, N0, s)
    nc, nn = t["nc"], (n + 1) ** 3
    f = lambda x, y, z: np.stack([x ** 4 - y * z, y ** 2 * z ** 3, x * y * z ** 4], axis=-1)
    le, lb = np.zeros((nn * nc, 3)), np.zeros((nn * nc, 3))
    xs, ys, zs = _node_coords(orc, n, t["bboxes"][0])
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    le[:nn] = f(X.ravel(), Y.ravel(), Z.ravel())
    lb[:nn] = 2.0 * le[:nn]
    for l in range(1, t["ml"] + 1):
        orc.l2l(n, t, orc.level_nodes(t, l), le, lb)
    E, B = orc.l2p(n, t, pos, le, lb)
    np.testing.assert_allclose(E, f(*pos.T), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(B, 2.0 * f(*pos.T), rtol=1e-9, atol=1e-12)


def test_bruteforce_closed_forms(orc):
    amp = 2.8179403699772166e-15 * (-1.0) / 0.5
    # one source at the origin, targets on the axes; p along z (gamma = sqrt(1+p^2))
    p = 3.0
    g = np.sqrt(1.0 + p * p)
    pos = np.array([[0, 0, 0], [2.0, 0, 0], [0, 0, 2.0]])
    mom = np.array([[0, 0, p], [0, 0, 0], [0, 0, 0]])
    E, B = orc.bruteforce(pos, mom, -1.0, 0.5, targets=[1, 2])
    # transverse target: E = amp g R/R^3, B = amp p x R / R^3 (p.R = 0)
    R = np.array([2.0, 0, 0])
    src1 = lambda q: q / 2.0 ** 3
    # target 1 also feels target 2 (momentum 0): Coulomb term from (0,0,2)
    R12 = np.array([2.0, 0, -2.0])
    np.testing.assert_allclose(E[0], amp * (g * src1(R) + R12 / np.linalg.norm(R12) ** 3), rtol=1e-14)
    np.testing.assert_allclose(B[0], amp * np.cross([0, 0, p], R) / 8.0, rtol=1e-14)
    # longitudinal target: d = R^2 (1 + p^2) -> E = amp R / (R^3 g^2) (Lorentz contraction)
    Rz = np.array([0, 0, 2.0])
    R21 = np.array([-2.0, 0, 2.0])
    np.testing.assert_allclose(E[1], amp * (Rz / (8.0 * g * g) + R21 / np.linalg.norm(R21) ** 3),
                               rtol=1e-14)
    np.testing.assert_allclose(B[1], 0.0, atol=1e-30)


def test_fmm_eta0_equals_bruteforce(orc):
    pos, mom = W.bunch(1500, seed=29)
    E, B = orc.fmm(pos, mom, -1.0, 1e-3, 4, 50, 0.0)
    Eb, Bb = orc.bruteforce(pos, mom, -1.0, 1e-3, nthreads=4)
    assert W.rel_l2(E, Eb) < 1e-13 and W.rel_l2(B, Bb) < 1e-13


def test_fmm_converges_to_bruteforce(orc):
    pos, mom = W.uniform_cube(3000, seed=30)
    Eb, _ = orc.bruteforce(pos, mom, -1.0, 1.0, nthreads=4)
    errs = [W.rel_l2(orc.fmm(pos, mom, -1.0, 1.0, n, 60, 0.3)[0], Eb) for n in (2, 4, 6, 8)]
    assert errs[0] < 1e-2
    assert all(b < a / 5 for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-8


def test_fmm_bunch_fields(orc):
    pos, mom = W.bunch(4000, seed=31)
    E, B = orc.fmm(pos, mom, -1.0, 1e-3, 4, 125, 0.5)
    Eb, Bb = orc.bruteforce(pos, mom, -1.0, 1e-3, nthreads=4)
    assert W.rel_l2(E, Eb) < 1e-2 and W.rel_l2(B, Bb) < 1e-2


def test_fmm_overwrites(orc):
    pos, mom = W.uniform_cube(800, seed=32)
    a = orc.fmm(pos, mom, -1.0, 1.0, 3, 40, 0.5)
    b = orc.fmm(pos, mom, -1.0, 1.0, 3, 40, 0.5)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_fmm_early_leaf(orc):
    pos, mom = W.uniform_cube(251, seed=33)
    E, _ = orc.fmm(pos, mom, -1.0, 1.0, 6, 125, 0.5)
    Eb, _ = orc.bruteforce(pos, mom, -1.0, 1.0)
    assert W.rel_l2(E, Eb) < 1e-2


def test_planar_beam_needs_last_match_rule(orc):
    # all particles share z: every cluster is flat in z and all z-nodes coincide
    pos, mom = W.uniform_cube(1200, seed=34)
    pos[:, 2] = 0.25
    Eb, _ = orc.bruteforce(pos, mom, -1.0, 1.0, nthreads=4)
    e_last = W.rel_l2(orc.fmm(pos, mom, -1.0, 1.0, 4, 50, 0.5, rule=orc.SINGULAR_LAST)[0], Eb)
    e_all = W.rel_l2(orc.fmm(pos, mom, -1.0, 1.0, 4, 50, 0.5, rule=orc.SINGULAR_ALL)[0], Eb)
    assert e_last < 1e-2 < e_all


def test_p_avg_matches_exact_mean(orc):
    _, mom = W.bunch(100000, seed=35)
    exact = np.array([np.sum(mom[:, d].astype(np.longdouble)) / len(mom) for d in range(3)],
                     dtype=float)
    np.testing.assert_allclose(orc.p_avg(mom), exact, rtol=1e-12, atol=1e-16)
