"""Python host mirror of FMM4RBGPU.jl's field-evaluation interface over the
gfx950 HIP library ``libfmm4rb_hip.so`` (C ABI in ``include/fmm4rb.h``).

Mirrors the reference API (src/particles.jl, src/update_particles_field.jl):

    beam = Beam(pos=pos, mom=mom)          # 3 x N arrays, charge=-1.0, mass=1.0
    update_particles_field(beam, FMMHIP(eta=0.5, N0=125, n=4), lambda_=1.0)
    beam.efields[i], beam.bfields[i]

``FMMHIP`` is the new algorithm type with the reference's ``(n, N0, eta)``
fields plus ``precision`` ("fp64" or "mixed").  ``BruteForceHIP`` runs the
reference's BruteForce on the GPU.  Every call goes through the HIP library;
there is no CPU fallback: if the library is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import os
import sys
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libfmm4rb_hip.so")

FP64, MIXED = 0, 1
_PREC = {"fp64": FP64, "mixed": MIXED}
AMP_RE = 2.8179403699772166e-15  # update_particles_field.jl:28

EXPORTS = (
    "fmm4rb_version", "fmm4rb_last_error", "fmm4rb_create", "fmm4rb_destroy",
    "fmm4rb_set_stream", "fmm4rb_set_overlap", "fmm4rb_field", "fmm4rb_field_device",
    "fmm4rb_bruteforce", "fmm4rb_tree", "fmm4rb_lists", "fmm4rb_op_m2l", "fmm4rb_op_p2p",
    "fmm4rb_set_debug", "fmm4rb_debug_fetch", "fmm4rb_field_vjp", "fmm4rb_field_vjp_device",
    "fmm4rb_bruteforce_vjp", "fmm4rb_field_refit", "fmm4rb_field_device_refit",
    "fmm4rb_field_pullback", "fmm4rb_field_pullback_device", "fmm4rb_field_device_noalloc",
)


class FMMError(RuntimeError):
    """A non-zero return code of the C ABI, with fmm4rb_last_error()."""

    def __init__(self, code, msg):
        super().__init__(f"fmm4rb error {code}: {msg}")
        self.code = code


class Stats(C.Structure):
    _fields_ = [
        ("N", C.c_int64), ("N0", C.c_int64), ("ncluster", C.c_int64), ("nleaf", C.c_int64),
        ("n", C.c_int32), ("precision", C.c_int32), ("maxlevel", C.c_int32),
        ("traversal_iterations", C.c_int32),
        ("np2p", C.c_int64), ("nm2l", C.c_int64), ("nm2l_groups", C.c_int64),
        ("p2p_particle_pairs", C.c_int64),
        ("p_avg", C.c_double * 3), ("g_avg", C.c_double),
        ("ms_total", C.c_float), ("ms_pavg", C.c_float), ("ms_tree", C.c_float),
        ("ms_lists", C.c_float), ("ms_upward", C.c_float), ("ms_m2l", C.c_float),
        ("ms_p2p", C.c_float), ("ms_downward", C.c_float),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["p_avg"] = list(self.p_avg)
        return d


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i64 = C.c_int64


def lib():
    """Load libfmm4rb_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with fmm4rbgpu.jl_amd/build.py")
        # PyTorch-ROCm ships its own libamdhip64 (same soname, loaded by RPATH).  If
        # this library is loaded first, /opt/rocm's HIP runtime comes in with it and a
        # later ``import torch`` brings a second runtime that finds no GPU; if torch is
        # loaded first, this library binds to torch's runtime and the two share
        # devices, memory and streams.  So load torch first whenever it is installed.
        if "torch" not in sys.modules and importlib.util.find_spec("torch") is not None:
            import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        L.fmm4rb_version.restype = C.c_char_p
        L.fmm4rb_last_error.restype = C.c_char_p
        L.fmm4rb_create.argtypes = [C.POINTER(_vp), C.c_int]
        L.fmm4rb_destroy.argtypes = [_vp]
        L.fmm4rb_set_stream.argtypes = [_vp, _vp]
        L.fmm4rb_set_debug.argtypes = [_vp, C.c_int]
        L.fmm4rb_set_overlap.argtypes = [_vp, C.c_int]
        L.fmm4rb_field.argtypes = [_vp, _i64, _dp, _dp, C.c_double, C.c_double, C.c_int, _i64,
                                   C.c_double, C.c_int, _dp, _dp, C.POINTER(Stats)]
        L.fmm4rb_field_device.argtypes = [_vp, _i64, _vp, _vp, C.c_double, C.c_double, C.c_int,
                                          _i64, C.c_double, C.c_int, _vp, _vp, C.POINTER(Stats)]
        L.fmm4rb_field_device_noalloc.argtypes = L.fmm4rb_field_device.argtypes
        L.fmm4rb_field_refit.argtypes = L.fmm4rb_field.argtypes
        L.fmm4rb_field_device_refit.argtypes = L.fmm4rb_field_device.argtypes
        L.fmm4rb_bruteforce.argtypes = [_vp, _i64, _dp, _dp, C.c_double, C.c_double, _i64, _vp,
                                        _dp, _dp]
        L.fmm4rb_tree.argtypes = [_vp, _i64, _dp, _dp, _i64, _ip, _ip, _ip, _dp, _ip, _dp]
        L.fmm4rb_lists.argtypes = [_vp, _i64, _dp, _dp, _i64, C.c_double, _i64, _vp, _i64, _vp,
                                   C.POINTER(_i64), C.POINTER(_i64)]
        L.fmm4rb_op_m2l.argtypes = [_vp, C.c_int, C.c_int, _i64, _dp, _dp, _dp, _i64, _vp, _dp,
                                    _dp, _dp]
        L.fmm4rb_op_p2p.argtypes = [_vp, C.c_int, _i64, _dp, _dp, _ip, _i64, _ip, _ip, _vp, _i64,
                                    _vp, _dp, _dp]
        L.fmm4rb_debug_fetch.argtypes = [_vp, C.c_int, _dp, _i64]
        L.fmm4rb_field_vjp.argtypes = [_vp, _i64, _dp, _dp, C.c_double, C.c_double, C.c_int, _i64,
                                       C.c_double, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(Stats)]
        L.fmm4rb_field_vjp_device.argtypes = [_vp, _i64, _vp, _vp, C.c_double, C.c_double, C.c_int,
                                              _i64, C.c_double, C.c_int, _vp, _vp, _vp, _vp,
                                              C.POINTER(Stats)]
        L.fmm4rb_field_pullback.argtypes = L.fmm4rb_field_vjp.argtypes
        L.fmm4rb_field_pullback_device.argtypes = L.fmm4rb_field_vjp_device.argtypes
        L.fmm4rb_bruteforce_vjp.argtypes = [_vp, _i64, _dp, _dp, C.c_double, C.c_double, _dp, _dp,
                                            _dp, _dp]
        for name in EXPORTS:
            getattr(L, name)
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise FMMError(rc, lib().fmm4rb_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64a(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def version():
    return lib().fmm4rb_version().decode()


class Context:
    """One fmm4rb context: a HIP stream plus device scratch kept between calls."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().fmm4rb_create(C.byref(h), device))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _check(lib().fmm4rb_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_stream(self, stream_ptr: int | None):
        _check(lib().fmm4rb_set_stream(self._h, stream_ptr or None))

    def set_overlap(self, on: bool):
        """False: run the near field and the upward pass in place on the call's
        stream (kernels one at a time; bit-identical results) for per-kernel timing."""
        _check(lib().fmm4rb_set_overlap(self._h, int(on)))

    def set_debug(self, on: bool):
        _check(lib().fmm4rb_set_debug(self._h, int(on)))

    # ------------------------------------------------------------ field
    def field(self, pos, mom, charge=-1.0, lam=1.0, n=4, N0=125, eta=0.5, precision="fp64"):
        """Host (N, 3) arrays -> (E, B, stats)."""
        pos, mom = _f64(pos), _f64(mom)
        N = pos.shape[0]
        E, B = np.empty((N, 3)), np.empty((N, 3))
        st = Stats()
        _check(lib().fmm4rb_field(self._h, N, pos, mom, charge, lam, n, N0, eta,
                                  _PREC[precision], E, B, C.byref(st)))
        return E, B, st

    def field_device(self, N, pos_ptr, mom_ptr, E_ptr, B_ptr, charge=-1.0, lam=1.0, n=4, N0=125,
                     eta=0.5, precision="fp64"):
        """Device pointers (e.g. torch tensors' data_ptr()) -> stats."""
        st = Stats()
        _check(lib().fmm4rb_field_device(self._h, N, pos_ptr, mom_ptr, charge, lam, n, N0, eta,
                                         _PREC[precision], E_ptr, B_ptr, C.byref(st)))
        return st

    def field_device_noalloc(self, N, pos_ptr, mom_ptr, E_ptr, B_ptr, charge=-1.0, lam=1.0,
                             n=4, N0=125, eta=0.5, precision="fp64"):
        """field_device() without device allocation: every buffer of the call comes
        from the library's static per-device scratch (N up to about 2^18 at n=4)."""
        st = Stats()
        _check(lib().fmm4rb_field_device_noalloc(self._h, N, pos_ptr, mom_ptr, charge, lam, n,
                                                 N0, eta, _PREC[precision], E_ptr, B_ptr,
                                                 C.byref(st)))
        return st

    def field_refit(self, pos, mom, charge=-1.0, lam=1.0, n=4, N0=125, eta=0.5,
                    precision="fp64"):
        """Like field() on the tree and lists of this context's last completed field or
        reverse-mode call with the same N, N0, eta (tracking steps): bounding boxes,
        p_avg and expansions are recomputed, the partition and lists are kept."""
        pos, mom = _f64(pos), _f64(mom)
        N = pos.shape[0]
        E, B = np.empty((N, 3)), np.empty((N, 3))
        st = Stats()
        _check(lib().fmm4rb_field_refit(self._h, N, pos, mom, charge, lam, n, N0, eta,
                                        _PREC[precision], E, B, C.byref(st)))
        return E, B, st

    def field_device_refit(self, N, pos_ptr, mom_ptr, E_ptr, B_ptr, charge=-1.0, lam=1.0, n=4,
                           N0=125, eta=0.5, precision="fp64"):
        """field_refit() with device pointers -> stats."""
        st = Stats()
        _check(lib().fmm4rb_field_device_refit(self._h, N, pos_ptr, mom_ptr, charge, lam, n, N0,
                                               eta, _PREC[precision], E_ptr, B_ptr,
                                               C.byref(st)))
        return st

    def field_vjp(self, pos, mom, Ebar, Bbar, charge=-1.0, lam=1.0, n=4, N0=125, eta=0.5,
                  precision="fp64"):
        """Cotangents of the fields -> (gpos, gmom, stats) (reverse mode, fixed tree)."""
        pos, mom = _f64(pos), _f64(mom)
        N = pos.shape[0]
        gx, gp = np.empty((N, 3)), np.empty((N, 3))
        st = Stats()
        _check(lib().fmm4rb_field_vjp(self._h, N, pos, mom, charge, lam, n, N0, eta,
                                      _PREC[precision], _f64(Ebar), _f64(Bbar), gx, gp,
                                      C.byref(st)))
        return gx, gp, st

    def field_vjp_device(self, N, pos_ptr, mom_ptr, Ebar_ptr, Bbar_ptr, gpos_ptr, gmom_ptr,
                         charge=-1.0, lam=1.0, n=4, N0=125, eta=0.5, precision="fp64"):
        st = Stats()
        _check(lib().fmm4rb_field_vjp_device(self._h, N, pos_ptr, mom_ptr, charge, lam, n, N0,
                                             eta, _PREC[precision], Ebar_ptr, Bbar_ptr,
                                             gpos_ptr, gmom_ptr, C.byref(st)))
        return st

    def field_pullback(self, pos, mom, Ebar, Bbar, charge=-1.0, lam=1.0, n=4, N0=125, eta=0.5,
                       precision="fp64"):
        """field_vjp() after a field() call on the same inputs and parameters, on that
        call's tree, lists, multipoles and final local expansions."""
        pos, mom = _f64(pos), _f64(mom)
        N = pos.shape[0]
        gx, gp = np.empty((N, 3)), np.empty((N, 3))
        st = Stats()
        _check(lib().fmm4rb_field_pullback(self._h, N, pos, mom, charge, lam, n, N0, eta,
                                           _PREC[precision], _f64(Ebar), _f64(Bbar), gx, gp,
                                           C.byref(st)))
        return gx, gp, st

    def field_pullback_device(self, N, pos_ptr, mom_ptr, Ebar_ptr, Bbar_ptr, gpos_ptr, gmom_ptr,
                              charge=-1.0, lam=1.0, n=4, N0=125, eta=0.5, precision="fp64"):
        st = Stats()
        _check(lib().fmm4rb_field_pullback_device(self._h, N, pos_ptr, mom_ptr, charge, lam, n,
                                                  N0, eta, _PREC[precision], Ebar_ptr, Bbar_ptr,
                                                  gpos_ptr, gmom_ptr, C.byref(st)))
        return st

    def bruteforce_vjp(self, pos, mom, Ebar, Bbar, charge=-1.0, lam=1.0):
        pos, mom = _f64(pos), _f64(mom)
        N = pos.shape[0]
        gx, gp = np.empty((N, 3)), np.empty((N, 3))
        _check(lib().fmm4rb_bruteforce_vjp(self._h, N, pos, mom, charge, lam, _f64(Ebar),
                                           _f64(Bbar), gx, gp))
        return gx, gp

    def bruteforce(self, pos, mom, charge=-1.0, lam=1.0, targets=None):
        pos, mom = _f64(pos), _f64(mom)
        N = pos.shape[0]
        if targets is None:
            nt, tp, keep = N, None, None
        else:
            keep = _i64a(targets)
            nt, tp = len(keep), keep.ctypes.data
        E, B = np.zeros((max(nt, 1), 3)), np.zeros((max(nt, 1), 3))
        _check(lib().fmm4rb_bruteforce(self._h, N, pos, mom, charge, lam, nt, tp, E, B))
        del keep
        return E[:nt], B[:nt]

    # ------------------------------------------------------------ parity hooks
    def tree(self, pos, mom, N0):
        pos, mom = _f64(pos), _f64(mom)
        N = pos.shape[0]
        ml = maxlevel(N, N0)
        nc = (1 << (ml + 1)) - 1
        t = dict(parindices=np.empty(N, np.int64), lo=np.empty(nc, np.int64),
                 hi=np.empty(nc, np.int64), bboxes=np.empty((nc, 6)),
                 children=np.empty((nc, 2), np.int64), p_avg=np.empty(3))
        _check(lib().fmm4rb_tree(self._h, N, pos, mom, N0, t["parindices"], t["lo"], t["hi"],
                                 t["bboxes"], t["children"], t["p_avg"]))
        t["nc"], t["ml"] = nc, ml
        return t

    def lists(self, pos, mom, N0, eta):
        pos, mom = _f64(pos), _f64(mom)
        N = pos.shape[0]
        n1, n2 = _i64(0), _i64(0)
        _check(lib().fmm4rb_lists(self._h, N, pos, mom, N0, eta, 0, None, 0, None,
                                  C.byref(n1), C.byref(n2)))
        p2p = np.zeros((max(n1.value, 1), 2), np.int64)
        m2l = np.zeros((max(n2.value, 1), 2), np.int64)
        _check(lib().fmm4rb_lists(self._h, N, pos, mom, N0, eta, n1.value, p2p.ctypes.data,
                                  n2.value, m2l.ctypes.data, C.byref(n1), C.byref(n2)))
        return p2p[: n1.value].copy(), m2l[: n2.value].copy()

    def op_m2l(self, n, bboxes, gam, mmom, pairs, p_avg, precision="fp64"):
        bboxes, gam, mmom = _f64(bboxes), _f64(gam), _f64(mmom)
        pairs = _i64a(pairs).reshape(-1, 2)
        nc = bboxes.shape[0]
        nn = (n + 1) ** 3
        le, lb = np.empty((nn * nc, 3)), np.empty((nn * nc, 3))
        _check(lib().fmm4rb_op_m2l(self._h, n, _PREC[precision], nc, bboxes, gam, mmom,
                                   len(pairs), pairs.ctypes.data if len(pairs) else None,
                                   _f64(p_avg), le, lb))
        return le, lb

    def op_p2p(self, pos, mom, parindices, lo, hi, pairs, bboxes=None, precision="fp64"):
        pos, mom = _f64(pos), _f64(mom)
        pairs = _i64a(pairs).reshape(-1, 2)
        lo, hi = _i64a(lo), _i64a(hi)
        N = pos.shape[0]
        bb = None if bboxes is None else _f64(bboxes)
        E, B = np.empty((N, 3)), np.empty((N, 3))
        _check(lib().fmm4rb_op_p2p(self._h, _PREC[precision], N, pos, mom, _i64a(parindices),
                                   len(lo), lo, hi, None if bb is None else bb.ctypes.data,
                                   len(pairs), pairs.ctypes.data if len(pairs) else None, E, B))
        return E, B

    def debug_fetch(self, what, n, nc, N):
        nn = (n + 1) ** 3
        size = {1: 4 * nn * nc, 2: 6 * nn * nc, 3: 6 * nn * nc, 4: 6 * N}[what]
        out = np.empty(size)
        _check(lib().fmm4rb_debug_fetch(self._h, what, out, size))
        if what == 1:
            return out[: nn * nc].copy(), out[nn * nc:].reshape(-1, 3).copy()
        half = size // 2
        return out[:half].reshape(-1, 3).copy(), out[half:].reshape(-1, 3).copy()


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def maxlevel(N, N0):
    """treeinfo.jl:6-8 (same floating-point formula as the library)."""
    import math
    return 0 if N <= N0 else int(math.ceil(math.log(N / N0) / math.log(2.0)))


# ---------------------------------------------------- reference-style interface
class Particles:
    """src/particles.jl:1-22: built from 3 x N matrices pos, mom."""

    def __init__(self, *, pos, mom, charge=-1.0, mass=1.0):
        pos = np.asarray(pos, dtype=np.float64)
        mom = np.asarray(mom, dtype=np.float64)
        assert pos.shape == mom.shape
        assert pos.shape[0] == 3
        self.npar = pos.shape[1]
        self.charge = float(charge)
        self.mass = float(mass)
        self.positions = np.ascontiguousarray(pos.T)
        self.momenta = np.ascontiguousarray(mom.T)
        self.efields = np.zeros((self.npar, 3))
        self.bfields = np.zeros((self.npar, 3))


Beam = Particles  # src/particles.jl:25


@dataclass(frozen=True)
class FMMHIP:
    """New algorithm type: the reference's FMM/FMMGPU fields (n, N0, eta),
    src/update_particles_field.jl:9-19, plus the arithmetic precision."""
    n: int
    N0: int
    eta: float
    precision: str = "fp64"


@dataclass(frozen=True)
class BruteForceHIP:
    """BruteForce (src/update_particles_field.jl:7) evaluated on the GPU."""


def update_particles_field(particles: Particles, alg, *, lambda_: float, ctx: Context | None = None):
    """update_particles_field!(particles, alg; lambda) (src/update_particles_field.jl:21-148).
    Overwrites particles.efields / bfields.  Returns the call's Stats for FMMHIP."""
    if not isinstance(alg, (FMMHIP, BruteForceHIP)):
        raise TypeError(f"no GPU method for {type(alg).__name__}")
    ctx = ctx or default_context()
    if isinstance(alg, FMMHIP):
        if alg.precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        E, B, st = ctx.field(particles.positions, particles.momenta, particles.charge, lambda_,
                             alg.n, alg.N0, alg.eta, alg.precision)
        particles.efields[...] = E
        particles.bfields[...] = B
        return st
    if isinstance(alg, BruteForceHIP):
        E, B = ctx.bruteforce(particles.positions, particles.momenta, particles.charge, lambda_)
        particles.efields[...] = E
        particles.bfields[...] = B
        return None
