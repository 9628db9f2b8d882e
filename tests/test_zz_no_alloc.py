"""The allocation-free entry point fmm4rb_field_device_noalloc against
fmm4rb_field_device, through the C ABI: bit-identical fields on two calls in a
row (the second with a smaller input), a call larger than the static scratch
refused without writing the outputs, the context's own pipeline intact
afterwards, and no device memory taken by the call."""
import numpy as np
import pytest
import torch

import fmm4rb_hip as F
from fmm4rb_hip import workloads as W

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _inputs(N, seed):
    pos, mom = W.bunch(N, seed=seed)
    dpos = torch.from_numpy(pos).to(DEV)
    dmom = torch.from_numpy(mom).to(DEV)
    E = torch.full_like(dpos, float("nan"))
    B = torch.full_like(dpos, float("nan"))
    torch.cuda.synchronize()
    return dpos, dmom, E, B


def _call(fn, d, n=4, prec="fp64"):
    dpos, dmom, E, B = d
    E.fill_(float("nan"))
    B.fill_(float("nan"))
    torch.cuda.synchronize()
    print(f"  {fn.__name__} N={dpos.shape[0]} n={n} {prec}", flush=True)
    st = fn(dpos.shape[0], dpos.data_ptr(), dmom.data_ptr(), E.data_ptr(), B.data_ptr(), -1.0,
            1e-3, n, 125, 0.5, prec)
    return E.cpu().numpy(), B.cpu().numpy(), st


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_matches_field_device(prec):
    with F.Context(0) as ctx, F.Context(0) as ref:
        for N, seed in ((1 << 16, 11), (3000, 12)):  # the second call is smaller
            d = _inputs(N, seed)
            Ea, Ba, sa = _call(ctx.field_device_noalloc, d, prec=prec)
            Er, Br, sr = _call(ref.field_device, d, prec=prec)
            assert np.all(np.isfinite(Ea)) and np.all(np.isfinite(Ba))
            assert np.array_equal(Ea, Er) and np.array_equal(Ba, Br)
            assert (sa.np2p, sa.nm2l, sa.nleaf) == (sr.np2p, sr.nm2l, sr.nleaf)
            # the same context's own pipeline, right after its allocation-free call
            Ec, Bc, _ = _call(ctx.field_device, d, prec=prec)
            assert np.array_equal(Ec, Er) and np.array_equal(Bc, Br)


def test_noalloc_refuses_oversized_call():
    with F.Context(0) as ctx:
        d = _inputs(1 << 20, 13)
        with pytest.raises(F.FMMError) as ex:
            _call(ctx.field_device_noalloc, d)
        assert ex.value.code == 2 and "static scratch" in str(ex.value)
        assert np.all(np.isnan(d[2].cpu().numpy()))  # nothing written
        # a smaller call then works on the same context
        d2 = _inputs(5000, 14)
        E, B, _ = _call(ctx.field_device_noalloc, d2)
        Er, Br, _ = _call(ctx.field_device, d2)
        assert np.array_equal(E, Er) and np.array_equal(B, Br)


def test_noalloc_leaves_no_tree_to_refit():
    with F.Context(0) as ctx:
        d = _inputs(4000, 15)
        _call(ctx.field_device, d)
        _call(ctx.field_device_noalloc, d)
        with pytest.raises(F.FMMError) as ex:
            _call(ctx.field_device_refit, d)
        assert ex.value.code == 1


def test_noalloc_takes_no_device_memory():
    with F.Context(0) as ctx:
        big = _inputs(1 << 17, 16)
        small = _inputs(2000, 17)
        _call(ctx.field_device_noalloc, big, n=6)  # first use: code objects, queues
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        for d, n, prec in ((small, 4, "fp64"), (big, 4, "mixed"), (big, 6, "fp64")):
            _call(ctx.field_device_noalloc, d, n=n, prec=prec)
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        assert free1 == free0, (free0, free1)
