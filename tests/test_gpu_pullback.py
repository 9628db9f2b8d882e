"""Pullback after a forward call (fmm4rb_field_pullback / _device): the reverse mode
on the tree, lists, multipoles and final local expansions of the preceding
fmm4rb_field call (a second pullback of the same call must give the same bits).  Pinned to fmm4rb_field_vjp (itself pinned to finite
differences, the oracle and the BruteForce adjoint): equal up to the rounding of the
forward locals; at eta = 0 (no far field) bit-identical; and it refuses to run on
anything but the inputs and parameters of the last completed forward call."""
import numpy as np
import pytest

import fmm4rb_hip as F
from fmm4rb_hip import workloads as W

pytestmark = pytest.mark.gpu

LAM = 1e-3


def _cot(E, B, seed):
    w = np.random.default_rng(seed).random((len(E), 1))
    return 2 * w * E, 2 * w * B


@pytest.mark.parametrize("n,prec,tol", [(4, "fp64", 1e-13), (6, "fp64", 1e-13), (2, "fp64", 1e-13),
                                        (4, "mixed", 1e-6), (6, "mixed", 1e-6)])
def test_pullback_equals_vjp(ctx, n, prec, tol):
    pos, mom = W.bunch(1 << 15, seed=101)
    E, B, _ = ctx.field(pos, mom, -1.0, LAM, n, 125, 0.5, prec)
    Eb, Bb = _cot(E, B, 102)
    gx, gp, _ = ctx.field_pullback(pos, mom, Eb, Bb, -1.0, LAM, n, 125, 0.5, prec)
    gx2, gp2, _ = ctx.field_pullback(pos, mom, Eb, Bb, -1.0, LAM, n, 125, 0.5, prec)  # again
    assert np.array_equal(gx, gx2) and np.array_equal(gp, gp2)
    vx, vp, _ = ctx.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, n, 125, 0.5, prec)
    ex, ep = W.rel_l2(gx, vx), W.rel_l2(gp, vp)
    print(f"n={n} {prec}: pullback vs vjp {ex:.2e} {ep:.2e}")
    assert ex <= tol and ep <= tol


def test_pullback_eta0_is_the_vjp(ctx):
    pos, mom = W.bunch(4096, seed=103)
    E, B, _ = ctx.field(pos, mom, -1.0, LAM, 4, 64, 0.0)
    Eb, Bb = _cot(E, B, 104)
    gx, gp, _ = ctx.field_pullback(pos, mom, Eb, Bb, -1.0, LAM, 4, 64, 0.0)
    vx, vp, _ = ctx.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, 4, 64, 0.0)
    assert np.array_equal(gx, vx) and np.array_equal(gp, vp)


def test_pullback_device_matches_host(ctx):
    import torch

    pos, mom = W.bunch(1 << 14, seed=105)
    E, B, _ = ctx.field(pos, mom, -1.0, LAM, 4, 125, 0.5, "mixed")
    Eb, Bb = _cot(E, B, 106)
    hx, hp, _ = ctx.field_pullback(pos, mom, Eb, Bb, -1.0, LAM, 4, 125, 0.5, "mixed")
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (pos, mom, Eb, Bb)]
    dE, dB = torch.empty_like(t[0]), torch.empty_like(t[0])
    gx, gp = torch.empty_like(t[0]), torch.empty_like(t[0])
    N = len(pos)
    ctx.field_device(N, t[0].data_ptr(), t[1].data_ptr(), dE.data_ptr(), dB.data_ptr(), -1.0, LAM,
                     4, 125, 0.5, "mixed")
    ctx.field_pullback_device(N, t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(),
                              t[3].data_ptr(), gx.data_ptr(), gp.data_ptr(), -1.0, LAM, 4, 125,
                              0.5, "mixed")
    torch.cuda.synchronize()
    assert np.array_equal(gx.cpu().numpy(), hx) and np.array_equal(gp.cpu().numpy(), hp)


def test_pullback_needs_the_forward_call():
    pos, mom = W.bunch(4096, seed=107)
    c = F.Context(0)
    try:
        E, B, _ = c.field(pos, mom, -1.0, LAM, 4, 125, 0.5)
        Eb, Bb = _cot(E, B, 108)
        c.field_pullback(pos, mom, Eb, Bb, -1.0, LAM, 4, 125, 0.5)  # valid
        moved = pos.copy()
        moved[17, 2] = np.nextafter(moved[17, 2], 1.0)  # one ulp of one coordinate
        bad = [dict(pos=moved), dict(n=5), dict(N0=64), dict(eta=0.4), dict(precision="mixed")]
        for b in bad:
            args = dict(pos=pos, mom=mom, n=4, N0=125, eta=0.5, precision="fp64")
            args.update(b)
            with pytest.raises(F.FMMError) as e:
                c.field_pullback(args["pos"], args["mom"], Eb, Bb, -1.0, LAM, args["n"],
                                 args["N0"], args["eta"], args["precision"])
            assert e.value.code == 1
        c.field_pullback(pos, mom, Eb, Bb, -1.0, LAM, 4, 125, 0.5)  # still valid after refusals
        with pytest.raises(F.FMMError):
            c.field_pullback(moved, mom, Eb, Bb, -1.0, LAM, 4, 125, 0.5)
        c.field_refit(pos, mom, -1.0, LAM, 4, 125, 0.5)  # the tree survives a refused pullback
        c.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, 4, 125, 0.5)  # recomputes the locals itself
        with pytest.raises(F.FMMError) as e:
            c.field_pullback(pos, mom, Eb, Bb, -1.0, LAM, 4, 125, 0.5)
        assert "no matching forward call" in str(e.value)
        c.field(pos, mom, -1.0, LAM, 4, 125, 0.5)
        c.field_refit(pos, mom, -1.0, LAM, 4, 125, 0.5)  # a refit's locals are not reused
        with pytest.raises(F.FMMError):
            c.field_pullback(pos, mom, Eb, Bb, -1.0, LAM, 4, 125, 0.5)
    finally:
        c.close()
