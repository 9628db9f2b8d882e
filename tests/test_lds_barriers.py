"""Every kernel's barriers are reached with its LDS accesses completed (static check
of the gfx950 assembly hipcc produces from csrc/ with the library's build flags;
runs without a GPU)."""
import concurrent.futures as cf
import importlib.util
import os
import shutil
import subprocess
import tempfile

import pytest

from lds_barrier_check import unsafe_barriers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location(
    "fmm4rb_build", os.path.join(ROOT, "fmm4rbgpu.jl_amd", "build.py"))
libbuild = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(libbuild)

HIPCC = libbuild.HIPCC

LOOP_STORE_THEN_BARRIER = """
_Zk:
	s_mov_b32 s0, 0
.LBB0_1:
	s_barrier
	ds_read_b64 v[0:1], v2
	s_waitcnt lgkmcnt(0)
	v_add_f64 v[0:1], v[0:1], v[0:1]
	ds_write_b64 v2, v[0:1]
	s_cbranch_scc1 .LBB0_1
	s_endpgm
.Lfunc_end0:
"""


def test_checker_flags_store_before_loop_top_barrier():
    assert unsafe_barriers(LOOP_STORE_THEN_BARRIER) == {"_Zk": [".LBB0_1"]}


def test_checker_accepts_wait_before_barrier():
    fixed = LOOP_STORE_THEN_BARRIER.replace("\ts_barrier", "\ts_waitcnt lgkmcnt(0)\n\ts_barrier")
    assert unsafe_barriers(fixed) == {}


def test_checker_ignores_lane_permutes():
    perm = LOOP_STORE_THEN_BARRIER.replace("ds_write_b64 v2, v[0:1]", "ds_bpermute_b32 v3, v2, v0")
    assert unsafe_barriers(perm) == {}


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_every_kernel_waits_for_lds_before_barriers():
    srcs = sorted(os.path.join(libbuild.SRC, f) for f in os.listdir(libbuild.SRC)
                  if f.endswith(".hip"))
    flags = [f for f in libbuild.FLAGS if f not in ("-fPIC", "-Wall")]
    tmp = tempfile.mkdtemp(prefix="ldsbar")
    try:
        def asm(src):
            out = os.path.join(tmp, os.path.basename(src) + ".s")
            r = subprocess.run([HIPCC, *flags, "--cuda-device-only", "-S", "-o", out, src],
                               capture_output=True, text=True)
            assert r.returncode == 0, r.stderr[-2000:]
            return src, open(out).read()
        with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
            results = list(ex.map(asm, srcs))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    bad = {os.path.basename(src): unsafe_barriers(text) for src, text in results}
    bad = {k: v for k, v in bad.items() if v}
    assert not bad, bad
