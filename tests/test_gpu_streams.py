"""The library's work is ordered on the caller's stream (fmm4rb_set_stream).  The
near field, the upward pass and the reverse mode's geometry terms fork from that
stream onto the context's helper streams and join back into it, so results on a
caller stream equal the default-stream results bit for bit, and inputs written on
that stream just before the call (behind slow work) are the ones the call reads."""
import numpy as np
import pytest

from fmm4rb_hip import workloads as W

pytestmark = pytest.mark.gpu

LAM = 1e-3


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_caller_stream_forward_and_reverse(ctx, prec):
    import torch

    pos, mom = W.bunch(1 << 16, seed=71)
    N = len(pos)
    dev = torch.device("cuda:0")
    dpos, dmom = torch.from_numpy(pos).to(dev), torch.from_numpy(mom).to(dev)
    args = (-1.0, LAM, 4, 125, 0.5, prec)
    ctx.set_stream(None)
    E0, B0 = torch.empty_like(dpos), torch.empty_like(dpos)
    ctx.field_device(N, dpos.data_ptr(), dmom.data_ptr(), E0.data_ptr(), B0.data_ptr(), *args)
    w = torch.rand(N, 1, dtype=torch.float64, device=dev,
                   generator=torch.Generator(device=dev).manual_seed(72))
    Eb, Bb = 2 * w * E0, 2 * w * B0
    g0, h0 = torch.empty_like(dpos), torch.empty_like(dpos)
    torch.cuda.synchronize()
    ctx.field_vjp_device(N, dpos.data_ptr(), dmom.data_ptr(), Eb.data_ptr(), Bb.data_ptr(),
                         g0.data_ptr(), h0.data_ptr(), *args)
    torch.cuda.synchronize()

    s = torch.cuda.Stream()
    slow = torch.ones(1 << 23, dtype=torch.float64, device=dev)
    try:
        ctx.set_stream(s.cuda_stream)
        with torch.cuda.stream(s):
            p2, m2 = torch.zeros_like(dpos), torch.zeros_like(dmom)
            Eb2, Bb2 = torch.zeros_like(Eb), torch.zeros_like(Bb)
            for _ in range(40):  # queued work ahead of the inputs on this stream
                slow.mul_(1.0000001)
            p2.copy_(dpos)
            m2.copy_(dmom)
            Eb2.copy_(Eb)
            Bb2.copy_(Bb)
            E1, B1 = torch.empty_like(dpos), torch.empty_like(dpos)
            ctx.field_device(N, p2.data_ptr(), m2.data_ptr(), E1.data_ptr(), B1.data_ptr(),
                             *args)
            g1, h1 = torch.empty_like(dpos), torch.empty_like(dpos)
            ctx.field_vjp_device(N, p2.data_ptr(), m2.data_ptr(), Eb2.data_ptr(),
                                 Bb2.data_ptr(), g1.data_ptr(), h1.data_ptr(), *args)
        s.synchronize()
    finally:
        ctx.set_stream(None)
    for a, b in ((E1, E0), (B1, B0), (g1, g0), (h1, h0)):
        assert torch.equal(a.cpu(), b.cpu())
    assert np.isfinite(E1.cpu().numpy()).all()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_serialized_helper_streams_are_bit_identical(ctx, prec):
    """fmm4rb_set_overlap(ctx, 0) runs the near field, the upward pass and the
    reverse-mode geometry in place on the call's stream: same results bit for bit."""
    pos, mom = W.bunch(1 << 15, seed=73)
    w = np.random.default_rng(74).random((len(pos), 1))
    args = (-1.0, LAM, 4, 125, 0.5, prec)
    out = []
    for on in (True, False):
        ctx.set_overlap(on)
        try:
            E, B, st = ctx.field(pos, mom, *args)
            g, h, _ = ctx.field_vjp(pos, mom, 2 * w * E, 2 * w * B, *args)
        finally:
            ctx.set_overlap(True)
        out.append((E, B, g, h, st.ms_m2l, st.ms_p2p))
    for a, b in zip(out[0][:4], out[1][:4]):
        assert np.array_equal(a, b)
    assert out[1][4] > 0 and out[1][5] > 0


def test_host_pointer_staging_matches_device(ctx):
    """The host-pointer entry points stage through pinned chunks of 64 MB copied by
    several host threads: at N = 2^22 (100 MB per array, two chunks) the fields equal
    the device-pointer call's bit for bit, and so do the gradients at N = 2^20."""
    import torch

    dev = torch.device("cuda:0")
    for N, vjp in ((1 << 22, False), (1 << 20, True)):
        pos, mom = W.bunch(N, seed=75)
        dp, dm = torch.from_numpy(pos).to(dev), torch.from_numpy(mom).to(dev)
        dE, dB = torch.empty_like(dp), torch.empty_like(dp)
        ctx.set_stream(None)
        ctx.field_device(N, dp.data_ptr(), dm.data_ptr(), dE.data_ptr(), dB.data_ptr(), -1.0,
                         LAM, 4, 125, 0.5, "mixed")
        torch.cuda.synchronize()
        E, B, _ = ctx.field(pos, mom, -1.0, LAM, 4, 125, 0.5, "mixed")
        assert np.array_equal(E, dE.cpu().numpy()) and np.array_equal(B, dB.cpu().numpy())
        if vjp:
            w = np.random.default_rng(76).random((N, 1))
            Eb, Bb = 2 * w * E, 2 * w * B
            gx, gp, _ = ctx.field_vjp(pos, mom, Eb, Bb, -1.0, LAM, 4, 125, 0.5, "mixed")
            tEb, tBb = torch.from_numpy(Eb).to(dev), torch.from_numpy(Bb).to(dev)
            dgx, dgp = torch.empty_like(dp), torch.empty_like(dp)
            ctx.field_vjp_device(N, dp.data_ptr(), dm.data_ptr(), tEb.data_ptr(), tBb.data_ptr(),
                                 dgx.data_ptr(), dgp.data_ptr(), -1.0, LAM, 4, 125, 0.5, "mixed")
            torch.cuda.synchronize()
            assert np.array_equal(gx, dgx.cpu().numpy()) and np.array_equal(gp, dgp.cpu().numpy())
