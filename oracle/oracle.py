"""ctypes binding of the CPU oracle (fmm4rb_oracle.c).

TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library.  Only the
test suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module; the product never does.

Arrays follow the oracle's conventions: positions/momenta are ``(N, 3)``
float64 C-contiguous (the AoS layout of ``Vector{SVector{3,Float64}}``),
nodes are 0-based (node k is reference node k+1), ranges are inclusive.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libfmm4rb_oracle.so")
SINGULAR_ALL, SINGULAR_LAST = 0, 1

_i64 = C.c_int64
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_maxlevel.restype = _i64
        L.orc_maxlevel.argtypes = [_i64, _i64]
        L.orc_ncluster.restype = _i64
        L.orc_ncluster.argtypes = [_i64, _i64]
        L.orc_diameter.restype = C.c_double
        L.orc_diameter.argtypes = [_dp, _dp]
        L.orc_inside.restype = C.c_int
        L.orc_inside.argtypes = [_dp, _dp]
        L.orc_partition.restype = _i64
        L.orc_partition.argtypes = [_ip, _i64, _i64, _dp, _i64, _i64]
        L.orc_kpermute.restype = _i64
        L.orc_kpermute.argtypes = [_ip, _i64, _i64, _dp, _i64, _i64]
        L.orc_cheb2.restype = None
        L.orc_cheb2.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.orc_lgweight.restype = None
        L.orc_lgweight.argtypes = [C.c_double, _dp, C.c_int, C.c_int, _dp]
        L.orc_p_avg.restype = None
        L.orc_p_avg.argtypes = [_i64, _dp, _dp]
        L.orc_build_tree.restype = C.c_int
        L.orc_build_tree.argtypes = [_i64, _dp, _i64, _dp, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _up]
        L.orc_admissible.restype = C.c_int
        L.orc_admissible.argtypes = [_dp, _dp, _dp, C.c_double]
        L.orc_lists.restype = C.c_int
        L.orc_lists.argtypes = [_i64, _ip, _dp, _dp, C.c_double, _i64, _ip, _i64, _ip,
                                C.POINTER(_i64), C.POINTER(_i64)]
        L.orc_p2m.restype = None
        L.orc_p2m.argtypes = [C.c_int, C.c_int, _i64, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
        L.orc_m2m.restype = None
        L.orc_m2m.argtypes = [C.c_int, C.c_int, _i64, _ip, _ip, _dp, _dp, _dp]
        L.orc_m2l.restype = None
        L.orc_m2l.argtypes = [C.c_int, _i64, _ip, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_p2p.restype = None
        L.orc_p2p.argtypes = [_i64, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
        L.orc_l2l.restype = None
        L.orc_l2l.argtypes = [C.c_int, C.c_int, _i64, _ip, _ip, _dp, _dp, _dp]
        L.orc_l2p.restype = None
        L.orc_l2p.argtypes = [C.c_int, C.c_int, _i64, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp,
                              _dp, _dp]
        L.orc_fmm.restype = C.c_int
        L.orc_fmm.argtypes = [_i64, _dp, _dp, C.c_double, C.c_double, C.c_int, _i64, C.c_double,
                              C.c_void_p, C.c_int, _dp, _dp]
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_bruteforce.restype = None
        L.orc_bruteforce.argtypes = [_i64, _dp, _dp, C.c_double, C.c_double, _i64, C.c_void_p,
                                     C.c_int, _dp, _dp]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64a(a):
    return np.ascontiguousarray(a, dtype=np.int64)


# --------------------------------------------------------------- small helpers
def maxlevel(N, N0):
    return int(lib().orc_maxlevel(N, N0))


def ncluster(N, N0):
    return int(lib().orc_ncluster(N, N0))


def diameter(bbox, stretch=(1.0, 1.0, 1.0)):
    return float(lib().orc_diameter(_f64(bbox), _f64(stretch)))


def inside(child, parent):
    return bool(lib().orc_inside(_f64(child), _f64(parent)))


def partition(labels, lo, hi, data, pvtindex):
    """Lomuto partition on (data, label) order; labels/lo/hi/pvtindex 0-based."""
    lab = _i64a(labels)
    j = lib().orc_partition(lab, lo, hi, _f64(data), 1, pvtindex)
    labels[:] = lab
    return int(j)


def kpermute(labels, lo, hi, data, k):
    lab = _i64a(labels)
    j = lib().orc_kpermute(lab, lo, hi, _f64(data), 1, k)
    labels[:] = lab
    return int(j)


def cheb2(n, a=-1.0, b=1.0):
    out = np.empty(n + 1)
    lib().orc_cheb2(n, a, b, out)
    return out


def lgweight(x, nodes, rule=SINGULAR_LAST):
    nodes = _f64(nodes)
    out = np.empty(len(nodes))
    lib().orc_lgweight(float(x), nodes, len(nodes) - 1, rule, out)
    return out


def p_avg(mom):
    mom = _f64(mom)
    out = np.empty(3)
    lib().orc_p_avg(mom.shape[0], mom, out)
    return out


def stretch_of(pavg):
    g = np.sqrt(1.0 + ((pavg[0] * pavg[0] + pavg[1] * pavg[1]) + pavg[2] * pavg[2]))
    return np.array([1.0, 1.0, g])


# ------------------------------------------------------------------------ tree
def build_tree(pos, N0, stretch):
    pos = _f64(pos)
    N = pos.shape[0]
    nc = ncluster(N, N0)
    t = dict(
        parindices=np.empty(N, np.int64), lo=np.empty(nc, np.int64), hi=np.empty(nc, np.int64),
        levels=np.empty(nc, np.int64), parents=np.empty(nc, np.int64),
        children=np.empty((nc, 2), np.int64), bboxes=np.empty((nc, 6)), exists=np.empty(nc, np.uint8))
    rc = lib().orc_build_tree(N, pos, N0, _f64(stretch), t["parindices"], t["lo"], t["hi"],
                              t["levels"], t["parents"], t["children"], t["bboxes"], t["exists"])
    if rc:
        raise ValueError("orc_build_tree failed")
    t["nc"], t["ml"], t["N"], t["N0"] = nc, maxlevel(N, N0), N, N0
    return t


def admissible(b1, b2, stretch, eta):
    return bool(lib().orc_admissible(_f64(b1), _f64(b2), _f64(stretch), float(eta)))


def lists(tree, stretch, eta):
    """(p2p, m2l) as (k, 2) int64 arrays of (target, source) in DFS order."""
    ch, bb = _i64a(tree["children"]), _f64(tree["bboxes"])
    n1, n2 = _i64(0), _i64(0)
    dummy = np.zeros(2, np.int64)
    lib().orc_lists(tree["nc"], ch, bb, _f64(stretch), float(eta), 0, dummy, 0, dummy,
                    C.byref(n1), C.byref(n2))
    p2p = np.zeros((max(n1.value, 1), 2), np.int64)
    m2l = np.zeros((max(n2.value, 1), 2), np.int64)
    rc = lib().orc_lists(tree["nc"], ch, bb, _f64(stretch), float(eta), n1.value, p2p, n2.value,
                         m2l, C.byref(n1), C.byref(n2))
    assert rc == 0
    return p2p[: n1.value].copy(), m2l[: n2.value].copy()


def leaves(tree):
    ex, ch = tree["exists"], tree["children"]
    return np.nonzero((ex == 1) & (ch[:, 0] < 0))[0].astype(np.int64)


def level_nodes(tree, l):
    ks = np.arange((1 << l) - 1, (1 << (l + 1)) - 1, dtype=np.int64)
    return ks[tree["exists"][ks] == 1]


# ------------------------------------------------------------------- operators
def macro_zeros(n, nc):
    nn = (n + 1) ** 3
    return np.zeros(nn * nc), np.zeros((nn * nc, 3))


def p2m(n, tree, pos, mom, nodes=None, rule=SINGULAR_LAST, gam=None, mmom=None):
    if gam is None:
        gam, mmom = macro_zeros(n, tree["nc"])
    nodes = leaves(tree) if nodes is None else _i64a(nodes)
    lib().orc_p2m(n, rule, len(nodes), nodes, tree["parindices"], tree["lo"], tree["hi"],
                  _f64(tree["bboxes"]), _f64(pos), _f64(mom), gam, mmom)
    return gam, mmom


def m2m(n, tree, nodes, gam, mmom, rule=SINGULAR_LAST):
    nodes = _i64a(nodes)
    lib().orc_m2m(n, rule, len(nodes), nodes, tree["parents"], _f64(tree["bboxes"]), gam, mmom)
    return gam, mmom


def upward(n, tree, pos, mom, rule=SINGULAR_LAST):
    gam, mmom = p2m(n, tree, pos, mom, rule=rule)
    for l in range(tree["ml"], 0, -1):
        m2m(n, tree, level_nodes(tree, l), gam, mmom, rule=rule)
    return gam, mmom


def m2l(n, bboxes, pairs, pavg, gam, mmom, le=None, lb=None):
    nc = len(bboxes)
    if le is None:
        le, lb = macro_zeros(n, nc)[1], macro_zeros(n, nc)[1]
    pairs = _i64a(pairs).reshape(-1, 2)
    lib().orc_m2l(n, len(pairs), pairs, _f64(bboxes), _f64(pavg), _f64(gam), _f64(mmom), le, lb)
    return le, lb


def p2p(pairs, parindices, lo, hi, pos, mom, E=None, B=None):
    N = len(pos)
    if E is None:
        E, B = np.zeros((N, 3)), np.zeros((N, 3))
    pairs = _i64a(pairs).reshape(-1, 2)
    lib().orc_p2p(len(pairs), pairs, _i64a(parindices), _i64a(lo), _i64a(hi), _f64(pos),
                  _f64(mom), E, B)
    return E, B


def l2l(n, tree, nodes, le, lb, rule=SINGULAR_LAST):
    nodes = _i64a(nodes)
    lib().orc_l2l(n, rule, len(nodes), nodes, tree["parents"], _f64(tree["bboxes"]), le, lb)
    return le, lb


def l2p(n, tree, pos, le, lb, nodes=None, rule=SINGULAR_LAST, E=None, B=None):
    N = len(pos)
    if E is None:
        E, B = np.zeros((N, 3)), np.zeros((N, 3))
    nodes = leaves(tree) if nodes is None else _i64a(nodes)
    lib().orc_l2p(n, rule, len(nodes), nodes, tree["parindices"], tree["lo"], tree["hi"],
                  _f64(tree["bboxes"]), _f64(pos), _f64(le), _f64(lb), E, B)
    return E, B


# --------------------------------------------------------------------- drivers
def set_threads(n: int) -> None:
    """Threads for the operators (0 = all); results do not depend on it."""
    lib().orc_set_threads(int(n))


def fmm(pos, mom, charge=-1.0, lam=1.0, n=4, N0=125, eta=0.5, pavg=None, rule=SINGULAR_LAST):
    pos, mom = _f64(pos), _f64(mom)
    N = pos.shape[0]
    E, B = np.zeros((N, 3)), np.zeros((N, 3))
    pa = None if pavg is None else _f64(pavg)
    rc = lib().orc_fmm(N, pos, mom, charge, lam, n, N0, eta,
                       None if pa is None else pa.ctypes.data, rule, E, B)
    if rc:
        raise ValueError(f"orc_fmm failed rc={rc}")
    return E, B


def bruteforce(pos, mom, charge=-1.0, lam=1.0, targets=None, nthreads=1):
    pos, mom = _f64(pos), _f64(mom)
    N = pos.shape[0]
    if targets is None:
        nt, tp = N, None
    else:
        tg = _i64a(targets)
        nt, tp = len(tg), tg.ctypes.data
    E, B = np.zeros((nt, 3)), np.zeros((nt, 3))
    lib().orc_bruteforce(N, pos, mom, charge, lam, nt, tp, nthreads, E, B)
    return E, B
