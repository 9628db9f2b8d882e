"""Tracking-step refit (fmm4rb_field_refit / fmm4rb_field_device_refit): the field for
new inputs on the tree and lists of the context's last completed call.

Pinned by: bit-identity with the full call on the same inputs; bit-identity with a
full rebuild whenever a drift leaves the tree and both lists unchanged (the full
call is itself pinned to the oracle); eta = 0 (every leaf pair in P2P, so the FMM
is the direct sum on any partition) equal to BruteForce after a large drift; and
an accuracy bound against BruteForce on a stale partition."""
import numpy as np
import pytest

import fmm4rb_hip as F
from fmm4rb_hip import workloads as W

pytestmark = pytest.mark.gpu

LAM = 1e-3


def _topology(ctx, pos, mom, N0, eta):
    t = ctx.tree(pos, mom, N0)
    p2p, m2l = ctx.lists(pos, mom, N0, eta)
    return t["parindices"], p2p, m2l


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_refit_same_inputs_bit_identical(ctx, prec):
    pos, mom = W.bunch(1 << 15, seed=81)
    E, B, st = ctx.field(pos, mom, -1.0, LAM, 4, 125, 0.5, prec)
    E2, B2, st2 = ctx.field_refit(pos, mom, -1.0, LAM, 4, 125, 0.5, prec)
    assert np.array_equal(E, E2) and np.array_equal(B, B2)
    assert st2.ms_lists < st.ms_lists and st2.np2p == st.np2p and st2.nm2l == st.nm2l
    # n and precision may change on the same tree
    other = "mixed" if prec == "fp64" else "fp64"
    E3, B3, _ = ctx.field(pos, mom, -1.0, LAM, 3, 125, 0.5, other)
    E4, B4, _ = ctx.field_refit(pos, mom, -1.0, LAM, 3, 125, 0.5, other)
    assert np.array_equal(E3, E4) and np.array_equal(B3, B4)


def test_refit_after_reverse_mode(ctx):
    pos, mom = W.bunch(1 << 14, seed=82)
    E, B, _ = ctx.field(pos, mom, -1.0, LAM, 4, 64, 0.5)
    w = np.random.default_rng(83).random((len(pos), 1))
    ctx.field_vjp(pos, mom, 2 * w * E, 2 * w * B, -1.0, LAM, 4, 64, 0.5)
    E2, B2, _ = ctx.field_refit(pos, mom, -1.0, LAM, 4, 64, 0.5)
    assert np.array_equal(E, E2) and np.array_equal(B, B2)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_refit_equals_rebuild_when_topology_is_unchanged(ctx, prec):
    N0, eta = 125, 0.5
    pos, mom = W.bunch(1 << 15, seed=84)
    rng = np.random.default_rng(85)
    base = _topology(ctx, pos, mom, N0, eta)
    pos2 = pos + 1e-12 * pos.std(axis=0) * rng.standard_normal(pos.shape)
    mom2 = mom * (1.0 + 1e-12 * rng.standard_normal(mom.shape))
    assert _same(base, _topology(ctx, pos2, mom2, N0, eta))
    assert not np.array_equal(pos2, pos)
    ctx.field(pos, mom, -1.0, LAM, 4, N0, eta, prec)
    Er, Br, _ = ctx.field_refit(pos2, mom2, -1.0, LAM, 4, N0, eta, prec)
    Ef, Bf, _ = ctx.field(pos2, mom2, -1.0, LAM, 4, N0, eta, prec)
    assert np.array_equal(Er, Ef) and np.array_equal(Br, Bf)


def test_refit_eta0_is_the_direct_sum_after_drift(ctx):
    pos, mom = W.bunch(6000, seed=86)
    ctx.field(pos, mom, -1.0, LAM, 4, 64, 0.0)
    rng = np.random.default_rng(87)
    pos2 = pos + 0.3 * pos.std(axis=0) * rng.standard_normal(pos.shape)
    mom2 = mom + 0.1 * rng.standard_normal(mom.shape)
    E, B, _ = ctx.field_refit(pos2, mom2, -1.0, LAM, 4, 64, 0.0)
    Eb, Bb = ctx.bruteforce(pos2, mom2, -1.0, LAM)
    assert W.rel_l2(E, Eb) <= 1e-12 and W.rel_l2(B, Bb) <= 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_refit_accuracy_on_a_stale_tree(ctx, prec):
    """After a drift of 1% of the bunch size the stale partition still gives the FMM's
    accuracy: within 2x of a full rebuild's error against BruteForce."""
    pos, mom = W.bunch(1 << 15, seed=88)
    ctx.field(pos, mom, -1.0, LAM, 4, 125, 0.5, prec)
    rng = np.random.default_rng(89)
    pos2 = pos + 0.01 * pos.std(axis=0) * rng.standard_normal(pos.shape)
    Er, Br, _ = ctx.field_refit(pos2, mom, -1.0, LAM, 4, 125, 0.5, prec)
    Ef, Bf, _ = ctx.field(pos2, mom, -1.0, LAM, 4, 125, 0.5, prec)
    tg = rng.choice(len(pos), 2048, replace=False)
    Eb, Bb = ctx.bruteforce(pos2, mom, -1.0, LAM, targets=tg)
    er, ef = W.rel_l2(Er[tg], Eb), W.rel_l2(Ef[tg], Eb)
    br, bf = W.rel_l2(Br[tg], Bb), W.rel_l2(Bf[tg], Bb)
    print(f"{prec}: refit E {er:.2e} B {br:.2e}; rebuild E {ef:.2e} B {bf:.2e}")
    assert er <= 2.0 * ef and br <= 2.0 * bf
    assert not np.array_equal(Er, Ef)


def test_refit_device_matches_host(ctx):
    import torch

    pos, mom = W.bunch(1 << 14, seed=90)
    pos2 = pos + 0.01 * pos.std(axis=0) * np.random.default_rng(91).standard_normal(pos.shape)
    ctx.field(pos, mom, -1.0, LAM, 4, 125, 0.5, "mixed")
    Eh, Bh, _ = ctx.field_refit(pos2, mom, -1.0, LAM, 4, 125, 0.5, "mixed")
    dev = torch.device("cuda:0")
    dp, dm = torch.from_numpy(pos2).to(dev), torch.from_numpy(mom).to(dev)
    dE, dB = torch.empty_like(dp), torch.empty_like(dp)
    ctx.field_device_refit(len(pos), dp.data_ptr(), dm.data_ptr(), dE.data_ptr(), dB.data_ptr(),
                           -1.0, LAM, 4, 125, 0.5, "mixed")
    torch.cuda.synchronize()
    assert np.array_equal(dE.cpu().numpy(), Eh) and np.array_equal(dB.cpu().numpy(), Bh)


def test_refit_needs_a_matching_tree():
    pos, mom = W.bunch(4096, seed=92)
    c = F.Context(0)
    try:
        with pytest.raises(F.FMMError) as e:
            c.field_refit(pos, mom, -1.0, LAM, 4, 125, 0.5)
        assert e.value.code == 1 and "no tree to refit" in str(e.value)
        c.field(pos, mom, -1.0, LAM, 4, 125, 0.5)
        for N0, eta, N in ((64, 0.5, 4096), (125, 0.4, 4096), (125, 0.5, 4000)):
            with pytest.raises(F.FMMError):
                c.field_refit(pos[:N], mom[:N], -1.0, LAM, 4, N0, eta)
        c.field_refit(pos, mom, -1.0, LAM, 4, 125, 0.5)  # still valid after the refusals
        c.tree(pos, mom, 125)  # a parity hook rebuilds the tree without lists
        with pytest.raises(F.FMMError):
            c.field_refit(pos, mom, -1.0, LAM, 4, 125, 0.5)
        bad = pos.copy()
        bad[7, 1] = np.nan
        c.field(pos, mom, -1.0, LAM, 4, 125, 0.5)
        with pytest.raises(F.FMMError):
            c.field_refit(bad, mom, -1.0, LAM, 4, 125, 0.5)
    finally:
        c.close()


def _edge_case(name):
    if name == "ties":  # integer coordinates, many ties, signed zeros
        rng = np.random.default_rng(3)
        pos = rng.integers(0, 6, size=(5000, 3)).astype(float)
        pos[::5, 0] = -0.0
        return pos, np.zeros((5000, 3)), 30
    if name == "earlyleaf":  # leaves above the deepest level
        pos, mom = W.uniform_cube(251, seed=4)
        return pos, mom, 125
    if name == "single":
        return np.array([[0.1, 0.2, 0.3]]), np.array([[0.0, 0.0, 1.0]]), 125
    if name == "n0_1":
        pos, mom = W.bunch(300, seed=6)
        return pos, mom, 1
    if name == "n0_1024":
        pos, mom = W.bunch(20000, seed=7)
        return pos, mom, 1024
    raise KeyError(name)


@pytest.mark.parametrize("name", ["ties", "earlyleaf", "single", "n0_1", "n0_1024"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_refit_edge_trees_bit_identical(ctx, name, prec):
    """Every tree shape the build supports (leaves on several levels, one particle,
    one particle per leaf, 1024-particle leaves, signed zeros on box faces): a refit on
    the inputs of the full call reproduces its boxes and fields exactly."""
    pos, mom, N0 = _edge_case(name)
    E, B, _ = ctx.field(pos, mom, -1.0, 1.0, 4, N0, 0.5, prec)
    E2, B2, _ = ctx.field_refit(pos, mom, -1.0, 1.0, 4, N0, 0.5, prec)
    assert np.array_equal(E, E2) and np.array_equal(B, B2)
