"""The oracle against the known-answer tests of the reference's own test suite.

Each test restates one reference testset (test/... file:line).  The reference
draws unseeded random inputs; here they are seeded, the asserted properties
and tolerances are the reference's.
"""
import numpy as np

from jlcompat import isapprox


def test_bbox_diameter_kat(orc):
    # test/test_utils/test_bbox.jl:2-9
    assert orc.diameter([0, 0, 0, 1, 1, 1]) == 0.8660254037844386


def test_bbox_inside(orc):
    # test/test_utils/test_bbox.jl:11-23
    b1 = [0, 0, 0, 1, 1, 1]
    b2 = [0, 0, 0, 0.5, 0.9, 1.0]
    b3 = [0, -0.2, 0, 0.5, 0.5, 0.5]
    assert orc.inside(b2, b1)
    assert not orc.inside(b3, b1)


def test_lomuto_partition(orc):
    # test/test_utils/test_partition.jl:2-9 (pivot = last element, value 3)
    data = np.array([1, 1, 5, 2, 4, 4, 3, 3], dtype=float)
    n = len(data)
    labels = np.arange(n, dtype=np.int64)
    orc.partition(labels, 0, n - 1, data, n - 1)
    assert np.all(data[labels[:3]] < 3)
    assert np.all(data[labels[3:]] >= 3)


def test_kpermute(orc):
    # test/test_utils/test_partition.jl:11-28
    data = np.array([2, 1, 3, 4, 4, 6], dtype=float)
    n = len(data)
    for k1, kval in ((3, 3), (4, 4)):
        labels = np.arange(n, dtype=np.int64)
        assert orc.kpermute(labels, 0, n - 1, data, k1 - 1) == k1 - 1
        assert np.all(data[labels[:k1]] <= kval)
        assert np.all(data[labels[k1:]] >= kval)


def test_cheb2_closed_form(orc):
    # test/test_utils/test_cheb2.jl:2-6
    n = 5
    c = np.array([np.cos(np.pi * i / n) for i in range(n + 1)])
    assert np.array_equal(orc.cheb2(n), c)
    assert np.array_equal(orc.cheb2(n, 0.0, 1.0), (c + 1.0) / 2.0)


def test_ncluster_kat(orc):
    # test/test_cluster_tree/test_treeinfo.jl:2-11
    assert orc.ncluster(4, 2) == 3
    assert orc.ncluster(11, 3) == 7
    assert orc.ncluster(18, 3) == 15
    assert orc.ncluster(2, 4) == 1


def test_clusters_leaves_at_maxlevel(orc):
    # test/test_cluster_tree/test_clusters.jl:10-28
    N, N0 = 10000, 32
    pos = np.random.default_rng(11).random((N, 3))
    t = orc.build_tree(pos, N0, [1.0, 1.0, 1.0])
    ml = orc.maxlevel(N, N0)
    lrg = np.arange((1 << ml) - 1, (1 << (ml + 1)) - 1)
    assert np.all(t["levels"][lrg] == ml)
    assert np.all(t["children"][lrg] == -1)
    assert len(t["levels"]) == orc.ncluster(N, N0)


def test_macroparticle_shape(orc):
    # test/test_fmm/test_macroparticles.jl:2-18
    n, nc = 3, orc.ncluster(10000, 32)
    gam, mom = orc.macro_zeros(n, nc)
    assert gam.reshape(nc, n + 1, n + 1, n + 1)[3].shape == (n + 1, n + 1, n + 1)
    assert mom.reshape(nc, n + 1, n + 1, n + 1, 3)[3].shape[:3] == (n + 1, n + 1, n + 1)


def test_p2m_momentum_conservation(orc):
    # test/test_fmm/test_upwardpass.jl:2-17: one leaf, sum of node momenta = N * p
    N, N0, n = 100, 200, 2
    pos = np.random.default_rng(12).random((N, 3))
    mom = np.zeros((N, 3))
    mom[:, 2] = 2.0
    t = orc.build_tree(pos, N0, [1.0, 1.0, 1.0])
    gam, mm = orc.p2m(n, t, pos, mom)
    assert isapprox(mm[: (n + 1) ** 3].sum(axis=0), N * mom[0], rtol=0.01)


def test_upward_pass_conservation(orc):
    # test/test_fmm/test_upwardpass.jl:19-37: root holds N*gamma and N*p
    N, n, N0 = 1000, 2, 80
    pos = np.random.default_rng(13).random((N, 3))
    mom = np.zeros((N, 3))
    gamma = 30.0
    mom[:, 2] = np.sqrt(gamma ** 2 - 1.0)
    t = orc.build_tree(pos, N0, [1.0, 1.0, 1.0])
    gam, mm = orc.upward(n, t, pos, mom)
    nn = (n + 1) ** 3
    assert isapprox(gam[:nn].sum(), N * gamma)      # Julia's default rtol = sqrt(eps)
    assert isapprox(mm[:nn].sum(axis=0), N * mom[0])


def test_admissible_kat(orc):
    # test/test_fmm/test_interaction_lists.jl:2-7
    b1 = [0, 0, 0, 0.5, 0.5, 0.5]
    b2 = [0.5, 0.5, 0.5, 1, 1, 1]
    assert orc.admissible(b1, b2, [1.0, 1.0, 1.0], 0.50) is False
    assert orc.admissible(b1, b2, [1.0, 1.0, 2.0], 0.55) is True


def test_lists_extreme_eta(orc):
    # test/test_fmm/test_interaction_lists.jl:9-25
    N, N0 = 1000, 400
    pos = np.random.default_rng(14).random((N, 3))
    s = [1.0, 1.0, 1.0]
    t = orc.build_tree(pos, N0, s)
    p2p, m2l = orc.lists(t, s, 999.0)
    assert len(m2l) == 2
    p2p, m2l = orc.lists(t, s, 0.0)
    assert len(p2p) == 16


def _two_clusters(n, distance):
    b1 = [-1.0, -1.0, -1.0, 1.0, 1.0, 1.0]
    b2 = [-1.0, -1.0, -1.0 + distance, 1.0, 1.0, 1.0 + distance]
    return np.array([b1, b2])


def test_m2l_far_field(orc):
    # test/test_fmm/test_interaction.jl:2-30
    n, distance = 3, 10000.0
    bb = _two_clusters(n, distance)
    gam, mm = orc.macro_zeros(n, 2)
    gam[:] = 1.0
    le, lb = orc.m2l(n, bb, [(0, 1), (1, 0)], [0.0, 0.0, 0.0], gam, mm)
    nn = (n + 1) ** 3
    far = (n + 1) ** 3 * np.array([0, 0, distance]) / distance ** 3
    assert isapprox(le[nn], far, rtol=1e-3)          # node (1,1,1) of cluster 2
    assert isapprox(lb[nn], [0, 0, 0], rtol=1e-3)
    assert isapprox(le[0], -far, rtol=1e-3)          # node (1,1,1) of cluster 1
    assert isapprox(lb[0], [0, 0, 0], rtol=1e-3)


def test_p2p_far_field(orc):
    # test/test_fmm/test_interaction.jl:32-60
    npar, sep = 200, 10000.0
    pos = np.random.default_rng(15).random((npar, 3))
    pos[100:, 2] += sep
    mom = np.zeros((npar, 3))
    E, B = orc.p2p([(0, 1), (1, 0)], np.arange(npar), [0, 100], [99, 199], pos, mom)
    far = npar / 2 * np.array([0, 0, sep]) / sep ** 3
    assert isapprox(E[0], -far, rtol=1e-3)
    assert isapprox(B[0], [0, 0, 0], rtol=1e-3)
    assert isapprox(E[100], far, rtol=1e-3)
    assert isapprox(B[100], [0, 0, 0], rtol=1e-3)


# ----------------------------------------------------------------------------
# Dual traversal: lists derived by hand from interaction_lists.jl:23-49 on two
# 7-node trees (tests/traversal_kats.py holds the trees and the derivation).
import traversal_kats as TK  # noqa: E402


def test_traversal_kat_hand_tree(orc):
    for name, case in TK.CASES.items():
        p2p, m2l = orc.lists(TK.hand_tree(case), case["stretch"], case["eta"])
        # the reference's push order (depth first, first child first)
        assert [tuple(p) for p in p2p] == case["p2p"], name
        assert [tuple(p) for p in m2l] == case["m2l"], name


def test_traversal_kat_tree_from_particles(orc):
    # the median-split tree of the case's particles is the hand-built tree
    for name, case in TK.CASES.items():
        s = orc.stretch_of(orc.p_avg(case["mom"]))
        assert np.array_equal(s, case["stretch"]), name
        t = orc.build_tree(case["pos"], case["N0"], s)
        assert np.array_equal(t["children"], TK.CHILDREN), name
        assert np.array_equal(t["bboxes"], case["bboxes"]), name
        p2p, m2l = orc.lists(t, s, case["eta"])
        assert [tuple(p) for p in p2p] == case["p2p"], name
        assert [tuple(p) for p in m2l] == case["m2l"], name


def test_traversal_kat_stretch_decides_split(orc):
    # case B without the stretch: node 2's diameter (.9) is below node 1's (1), so
    # the reference splits the source at (2, 1) and the lists change
    case = TK.CASE_B
    p2p, m2l = orc.lists(TK.hand_tree(case), [1.0, 1.0, 1.0], case["eta"])
    assert [tuple(p) for p in m2l] != case["m2l"] or [tuple(p) for p in p2p] != case["p2p"]
