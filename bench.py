"""Benchmark: field evaluation of a 2^22-particle relativistic Gaussian bunch on
one MI355X, mixed-precision FMM vs the FP64 HIP FMM.

A step is one full update_particles_field evaluation (mean momentum, tree,
lists, P2M/M2M, M2L, P2P, L2L/L2P) with positions and momenta already
resident in HBM and the fields written to HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4] [--N 4194304]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fmm4rbgpu.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "particles/s field eval N=2^22: mixed-prec vs FP64 HIP FMM at rel err <=1e-5"
PEAK = {"mixed": 157.3e12, "fp64": 78.6e12}  # MI355X FP32 / FP64 vector peaks (FLOP/s)
FLOPS_PER_PAIR = 40  # one kernel evaluation, P2P particle pair or M2L node pair
# HBM bytes per launch from two rocprofv3 PMC passes (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000,
# calibrated on a 1 GiB copy), N=2^22, n=4: profiles/m2l_p2p_bunch_2e22_n4_traffic.md
TRAFFIC = {(1 << 22, 4, "mixed", "m2l"): 3.523e9, (1 << 22, 4, "mixed", "p2p"): 3.000e9,
           (1 << 22, 4, "fp64", "m2l"): 1.266e10, (1 << 22, 4, "fp64", "p2p"): 2.045e9}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=1 << 22)
    ap.add_argument("--n", type=int, default=4)
    ap.add_argument("--N0", type=int, default=125)
    ap.add_argument("--eta", type=float, default=0.5)
    ap.add_argument("--workload", choices=["bunch", "ellipsoid"], default="bunch",
                    help="bunch: Gaussian, sigma=(1e-3,1e-3,1e-4), p_z~N(100,1); ellipsoid: "
                         "uniform, semi-axes (1e-3,1e-3,1e-4), p_z~N(200,2) (the 25.6M case)")
    ap.add_argument("--cpu-sample", type=int, default=1 << 17,
                    help="particles of the serial CPU FMM sample (0 = skip)")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--large-N", type=int, default=25_600_000,
                    help="particles of the large uniform-ellipsoid case reported beside the "
                         "headline (FP64 and mixed forward, one FP64 reverse pass; 0 = skip)")
    ap.add_argument("--adjoint-N", type=int, default=1 << 20,
                    help="the reverse-mode configuration (bunch, seed 4) also timed on its own; "
                         "0 skips it")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size "
                         "kernel that calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def timed(ctx, torch, N, bufs, n, N0, eta, prec, steps, warmup, lam):
    dpos, dmom, dE, dB = bufs
    call = lambda: ctx.field_device(N, dpos.data_ptr(), dmom.data_ptr(), dE.data_ptr(),
                                    dB.data_ptr(), -1.0, lam, n, N0, eta, prec)
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(call())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # after the timed region: the same call with the near field and the upward pass
    # in place on the call's stream, so that each kernel's event-timed duration is
    # its own (overlapped, P2P and M2L share the CUs and both durations stretch)
    ctx.set_overlap(False)
    try:
        alone = [call() for _ in range(ISOLATED_CALLS)]
    finally:
        ctx.set_overlap(True)
    return wall / steps, stats, alone


ISOLATED_CALLS = 2


def stage_avg(stats):
    keys = ["ms_total", "ms_pavg", "ms_tree", "ms_upward", "ms_lists", "ms_m2l", "ms_p2p",
            "ms_downward"]
    return {k[3:]: sum(getattr(s, k) for s in stats) / len(stats) for k in keys}


def roofline(st, stages, n, prec, N):
    nn = (n + 1) ** 3
    m2l_flops = FLOPS_PER_PAIR * st.nm2l * nn * nn
    p2p_flops = FLOPS_PER_PAIR * st.p2p_particle_pairs
    kern = {"m2l": (m2l_flops, stages["m2l"]), "p2p": (p2p_flops, stages["p2p"])}
    name = max(kern, key=lambda k: kern[k][1])
    flops, ms = kern[name]
    ach = flops / (ms * 1e-3)
    other = "p2p" if name == "m2l" else "m2l"
    of, oms = kern[other]
    return {"kernel": name, "bound": "valu", "achieved": ach / 1e12, "peak": PEAK[prec] / 1e12,
            "unit": "TFLOP/s", "frac": ach / PEAK[prec],
            "traffic": TRAFFIC.get((N, n, prec, name)),
            "flops_per_launch": flops, "avg_launch_ms": ms,
            "other": {"kernel": other, "achieved": of / (oms * 1e-3) / 1e12,
                      "frac": of / (oms * 1e-3) / PEAK[prec], "avg_launch_ms": oms}}


def host_cpu():
    """CPU model and the cores this process may use (the host's share)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail or 1, int(omp)) if omp.isdigit() and int(omp) > 0 else (avail or 1)
    return {"cpu_model": model, "nproc_available": avail, "os_cpu_count": os.cpu_count(),
            "OMP_NUM_THREADS": omp or None, "threads_used": threads}


def cpu_baseline(N, n, N0, eta, N_full):
    import numpy as np

    import oracle
    from fmm4rb_hip import workloads as W
    host = host_cpu()
    threads = host["threads_used"]
    pos, mom = W.bunch(N, seed=2)
    oracle.set_threads(1)
    t0 = time.perf_counter()
    oracle.fmm(pos, mom, -1.0, 1e-3, n, N0, eta)
    dt = time.perf_counter() - t0
    # the N = 8192 unit-cube case on 1 thread, and the N = 2^20 bunch on every core
    # (the oracle's loops over target groups on OpenMP threads, bit-identical to serial)
    pc, mc = W.uniform_cube(8192, seed=1)
    t0 = time.perf_counter()
    oracle.fmm(pc, mc, -1.0, 1.0, n, N0, eta)
    dt1 = time.perf_counter() - t0
    p2, m2 = W.bunch(1 << 20, seed=2)
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    oracle.fmm(p2, m2, -1.0, 1e-3, n, N0, eta)
    dt2 = time.perf_counter() - t0
    oracle.set_threads(0)
    # BruteForce (the accuracy oracle) on the host's cores: all N sources, 4096 targets
    tg = np.random.default_rng(7).choice(N, 4096, replace=False)
    t0 = time.perf_counter()
    oracle.bruteforce(pos, mom, -1.0, 1e-3, targets=tg, nthreads=threads)
    dtb = time.perf_counter() - t0
    pairs_s = N * len(tg) / dtb
    return {"value": N / dt, "unit": "particles/s", "cores": 1, "kind": "port",
            "sample": f"serial FMM (C restatement of src/fmm, 1 thread) on a {N}-particle bunch "
                      f"of the same distribution, n={n} N0={N0} eta={eta}: {dt:.1f} s",
            "host": host,
            "fmm_uniform8192_1thread_s": dt1,
            "fmm_bunch_2e20_all_threads": {"s": dt2, "particles_per_s": (1 << 20) / dt2,
                                           "cores": threads},
            "bruteforce": {"pair_evals_per_s": pairs_s, "cores": threads, "kind": "port",
                           "sample": f"BruteForce (C restatement of update_particles_field.jl:21-38, "
                                     f"OpenMP) on the same {N} sources x 4096 targets: {dtb:.2f} s",
                           "est_s_full_workload": float(N_full) * float(N_full) / pairs_s}}


def large_case(F, W, torch, N, n, N0, eta, lam):
    """The N = 25.6M uniform ellipsoid (semi-axes 1e-3, 1e-3, 1e-4; p_z ~ N(200, 2)),
    the reference README's problem size: FP64 and mixed forward (1 warm-up, 2 timed
    calls each), one FP64 reverse pass, mixed vs FP64 and FP64 vs GPU BruteForce on
    1024 targets.  Its own context, closed afterwards."""
    import numpy as np
    dev = torch.device("cuda:0")
    pos, mom = W.ellipsoid(N, seed=5)
    dpos, dmom = torch.from_numpy(pos).to(dev), torch.from_numpy(mom).to(dev)
    out = {"workload": f"uniform ellipsoid, N={N}, semi-axes (1e-3,1e-3,1e-4), p_z~N(200,2), "
                       "p_x,p_y~N(0,0.02), lambda=1e-3", "N": N, "n": n, "N0": N0, "eta": eta}
    fields = {}
    with F.Context(0) as ctx:
        for prec in ("fp64", "mixed"):
            E, B = torch.empty_like(dpos), torch.empty_like(dpos)
            call = lambda: ctx.field_device(N, dpos.data_ptr(), dmom.data_ptr(), E.data_ptr(),
                                            B.data_ptr(), -1.0, lam, n, N0, eta, prec)
            call()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = [call(), call()]
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / 2 * 1e3
            out[prec + "_forward_ms"] = ms
            out[prec + "_particles_per_s"] = N / (ms * 1e-3)
            out[prec + "_stages_ms"] = stage_avg(st)
            fields[prec] = (E, B)
        E, B = fields["fp64"]
        w = torch.rand(N, 1, dtype=torch.float64, device=dev,
                       generator=torch.Generator(device=dev).manual_seed(43))
        Eb, Bb = 2 * w * E, 2 * w * B
        gpos, gmom = torch.empty_like(dpos), torch.empty_like(dpos)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.field_vjp_device(N, dpos.data_ptr(), dmom.data_ptr(), Eb.data_ptr(), Bb.data_ptr(),
                             gpos.data_ptr(), gmom.data_ptr(), -1.0, lam, n, N0, eta, "fp64")
        torch.cuda.synchronize()
        out["fp64_reverse_ms"] = (time.perf_counter() - t0) * 1e3
        out["fp64_reverse_ratio_to_forward"] = out["fp64_reverse_ms"] / out["fp64_forward_ms"]
        del Eb, Bb, gpos, gmom, w
        Ed, Bd = E.cpu().numpy(), B.cpu().numpy()
        Em, Bm = fields["mixed"][0].cpu().numpy(), fields["mixed"][1].cpu().numpy()
        tg = np.random.default_rng(42).choice(N, 1024, replace=False)
        Ebf, Bbf = ctx.bruteforce(pos, mom, -1.0, lam, targets=tg)
    out["mixed_vs_fp64"] = {"E": W.rel_l2(Em, Ed), "B": W.rel_l2(Bm, Bd)}
    out["fp64_vs_bruteforce_1024"] = {"E": W.rel_l2(Ed[tg], Ebf), "B": W.rel_l2(Bd[tg], Bbf)}
    out["speedup_mixed_vs_fp64"] = out["fp64_forward_ms"] / out["mixed_forward_ms"]
    del fields, dpos, dmom
    torch.cuda.empty_cache()
    return out


def adjoint_case(F, W, torch, N, n, N0, eta, lam, steps=3, warmup=1):
    """The reverse-mode configuration: an N = 2^20 bunch (seed 4), L = sum w (|E|^2 +
    |B|^2) with w ~ U(0, 1) (seed 40).  Per precision: the forward evaluation, the reverse
    mode (tree and lists rebuilt) and the pullback after a forward call (which keeps that
    call's tree, lists and expansions), each 1 warm-up
    + `steps` timed calls, and each reverse mode's ratio to the forward of its own
    precision.  Its own context, closed afterwards."""
    dev = torch.device("cuda:0")
    pos, mom = W.bunch(N, seed=4)
    dpos, dmom = torch.from_numpy(pos).to(dev), torch.from_numpy(mom).to(dev)
    E, B = torch.empty_like(dpos), torch.empty_like(dpos)
    gpos, gmom = torch.empty_like(dpos), torch.empty_like(dpos)
    w = torch.rand(N, 1, dtype=torch.float64, device=dev,
                   generator=torch.Generator(device=dev).manual_seed(40))
    out = {"workload": f"relativistic Gaussian bunch, N={N} (seed 4), loss sum w(|E|^2+|B|^2)",
           "N": N, "n": n, "N0": N0, "eta": eta}

    def clock(call):
        for _ in range(warmup):
            call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            call()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3

    with F.Context(0) as ctx:
        fwd = lambda prec: ctx.field_device(N, dpos.data_ptr(), dmom.data_ptr(), E.data_ptr(),
                                            B.data_ptr(), -1.0, lam, n, N0, eta, prec)
        for prec in ("fp64", "mixed"):
            ms_f = clock(lambda: fwd(prec))
            Eb, Bb = 2 * w * E, 2 * w * B  # this precision's fields
            args = (N, dpos.data_ptr(), dmom.data_ptr(), Eb.data_ptr(), Bb.data_ptr(),
                    gpos.data_ptr(), gmom.data_ptr(), -1.0, lam, n, N0, eta, prec)
            ms_v = clock(lambda: ctx.field_vjp_device(*args))
            fwd(prec)
            ms_p = clock(lambda: ctx.field_pullback_device(*args))
            out[prec] = {"forward_ms": ms_f, "reverse_ms": ms_v, "pullback_ms": ms_p,
                         "reverse_ratio_to_forward": ms_v / ms_f,
                         "pullback_ratio_to_forward": ms_p / ms_f}
    del dpos, dmom, E, B, gpos, gmom, w
    torch.cuda.empty_cache()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU sharding design is documented, not built)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import fmm4rb_hip as F
    from fmm4rb_hip import workloads as W

    lam = 1e-3
    dev = torch.device("cuda:0")
    if a.workload == "bunch":
        pos, mom = W.bunch(a.N, seed=3)
        desc = (f"relativistic Gaussian bunch, N={a.N}, sigma=(1e-3,1e-3,1e-4), "
                "p_z~N(100,1), p_x,p_y~N(0,0.01), lambda=1e-3")
    else:
        pos, mom = W.ellipsoid(a.N, seed=5)
        desc = (f"uniform ellipsoid, N={a.N}, semi-axes (1e-3,1e-3,1e-4), p_z~N(200,2), "
                "p_x,p_y~N(0,0.02), lambda=1e-3")
    dpos = torch.from_numpy(pos).to(dev)
    dmom = torch.from_numpy(mom).to(dev)
    bufs_m = (dpos, dmom, torch.empty_like(dpos), torch.empty_like(dpos))
    bufs_d = (dpos, dmom, torch.empty_like(dpos), torch.empty_like(dpos))
    torch.cuda.synchronize()
    free0, res0 = torch.cuda.mem_get_info()[0], torch.cuda.memory_reserved()
    # device memory held by the library (not by torch's allocator)
    lib_bytes = lambda: (free0 - torch.cuda.mem_get_info()[0]) - (torch.cuda.memory_reserved() - res0)
    ctx = F.Context(0)
    ms_m, st_m, iso_m = timed(ctx, torch, a.N, bufs_m, a.n, a.N0, a.eta, "mixed", a.steps,
                              a.warmup, lam)
    res = {"library_device_bytes": {"forward_mixed": lib_bytes()}}
    # tracking steps: the same evaluation on the tree and lists of the last call
    # (fmm4rb_field_device_refit; bounding boxes, p_avg and expansions recomputed)
    dpm, dmm, dEm, dBm = bufs_m
    refit = lambda: ctx.field_device_refit(a.N, dpm.data_ptr(), dmm.data_ptr(), dEm.data_ptr(),
                                           dBm.data_ptr(), -1.0, lam, a.n, a.N0, a.eta, "mixed")
    ctx.field_device(a.N, dpm.data_ptr(), dmm.data_ptr(), dEm.data_ptr(), dBm.data_ptr(), -1.0,
                     lam, a.n, a.N0, a.eta, "mixed")
    refit()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st_r = [refit() for _ in range(a.steps)]
    torch.cuda.synchronize()
    ms_r = (time.perf_counter() - t0) / a.steps
    res["tracking_refit_mixed"] = {"ms_per_step": ms_r * 1e3, "particles_per_s": a.N / ms_r,
                                   "stages_ms": stage_avg(st_r)}
    if not a.skip_fp64:
        ms_d, st_d, iso_d = timed(ctx, torch, a.N, bufs_d, a.n, a.N0, a.eta, "fp64", a.steps,
                                  a.warmup, lam)
        res["fp64"] = {"value": a.N / ms_d, "ms_per_step": ms_d * 1e3,
                       "stages_ms": stage_avg(st_d), "stages_ms_isolated": stage_avg(iso_d),
                       "roofline": roofline(iso_d[-1], stage_avg(iso_d), a.n, "fp64", a.N)}
    if not a.skip_fp64 and not a.skip_adjoint:
        # reverse mode of L = sum w (|E|^2 + |B|^2): cotangents 2 w E, 2 w B (FP64 fields)
        w = torch.rand(a.N, 1, dtype=torch.float64, device=dev,
                       generator=torch.Generator(device=dev).manual_seed(40))
        Eb, Bb = 2 * w * bufs_d[2], 2 * w * bufs_d[3]
        grads = {}
        for prec in ("fp64", "mixed"):
            gpos, gmom = torch.empty_like(dpos), torch.empty_like(dpos)
            call = lambda: ctx.field_vjp_device(a.N, dpos.data_ptr(), dmom.data_ptr(),
                                                Eb.data_ptr(), Bb.data_ptr(), gpos.data_ptr(),
                                                gmom.data_ptr(), -1.0, lam, a.n, a.N0, a.eta, prec)
            for _ in range(a.warmup):
                call()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sa = [call() for _ in range(a.steps)]
            torch.cuda.synchronize()
            ms_a = (time.perf_counter() - t0) / a.steps
            fwd = ms_d if prec == "fp64" else ms_m
            res["adjoint_" + prec] = {"ms_per_step": ms_a * 1e3,
                                      "ratio_to_forward": ms_a / fwd,  # same precision
                                      "forward_ms": fwd * 1e3, "stages_ms": stage_avg(sa)}
            grads[prec] = (gpos.cpu().numpy(), gmom.cpu().numpy())
        res["adjoint_mixed"]["vs_fp64"] = {
            "gpos": W.rel_l2(grads["mixed"][0], grads["fp64"][0]),
            "gmom": W.rel_l2(grads["mixed"][1], grads["fp64"][1])}
        # the rrule's pullback: after a forward call on the same inputs, the reverse
        # mode reuses that call's tree, lists, multipoles and final local expansions
        # (repeated pullbacks of one forward call, as for several cotangents)
        for prec in ("fp64", "mixed"):
            gpos, gmom = torch.empty_like(dpos), torch.empty_like(dpos)
            Ef, Bf = torch.empty_like(dpos), torch.empty_like(dpos)
            ctx.field_device(a.N, dpos.data_ptr(), dmom.data_ptr(), Ef.data_ptr(), Bf.data_ptr(),
                             -1.0, lam, a.n, a.N0, a.eta, prec)
            call = lambda: ctx.field_pullback_device(a.N, dpos.data_ptr(), dmom.data_ptr(),
                                                     Eb.data_ptr(), Bb.data_ptr(),
                                                     gpos.data_ptr(), gmom.data_ptr(), -1.0, lam,
                                                     a.n, a.N0, a.eta, prec)
            for _ in range(a.warmup):
                call()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sp = [call() for _ in range(a.steps)]
            torch.cuda.synchronize()
            ms_p = (time.perf_counter() - t0) / a.steps
            g = (gpos.cpu().numpy(), gmom.cpu().numpy())
            fwd = ms_d if prec == "fp64" else ms_m
            res["pullback_" + prec] = {
                "ms_per_step": ms_p * 1e3, "ratio_to_forward": ms_p / fwd,  # same precision
                "forward_ms": fwd * 1e3,
                "stages_ms": stage_avg(sp),
                "vs_vjp": {"gpos": W.rel_l2(g[0], grads[prec][0]),
                           "gmom": W.rel_l2(g[1], grads[prec][1])}}
            del Ef, Bf
    res["library_device_bytes"]["after_reverse_mode"] = lib_bytes()
    # host-pointer entry point (fmm4rb_field): PCIe copies of pos/mom in and E/B out
    # included; reported beside the device-resident value, never as the value
    Eh, Bh, _ = ctx.field(pos, mom, -1.0, lam, a.n, a.N0, a.eta, "mixed")
    t0 = time.perf_counter()
    Eh, Bh, _ = ctx.field(pos, mom, -1.0, lam, a.n, a.N0, a.eta, "mixed")
    ms_h = time.perf_counter() - t0
    res["host_pointers_mixed"] = {"ms_per_call": ms_h * 1e3, "particles_per_s": a.N / ms_h}
    del Eh, Bh
    # accuracy, outside the timed region
    Em, Bm = bufs_m[2].cpu().numpy(), bufs_m[3].cpu().numpy()
    acc = {}
    if not a.skip_fp64:
        Ed, Bd = bufs_d[2].cpu().numpy(), bufs_d[3].cpu().numpy()
        acc["mixed_vs_fp64"] = {"E": W.rel_l2(Em, Ed), "B": W.rel_l2(Bm, Bd)}
    tg = np.random.default_rng(40).choice(a.N, 4096, replace=False)
    Eb, Bb = ctx.bruteforce(pos, mom, -1.0, lam, targets=tg)
    acc["mixed_vs_bruteforce_4096"] = {"E": W.rel_l2(Em[tg], Eb), "B": W.rel_l2(Bm[tg], Bb)}
    if not a.skip_fp64:
        acc["fp64_vs_bruteforce_4096"] = {"E": W.rel_l2(Ed[tg], Eb), "B": W.rel_l2(Bd[tg], Bb)}
    stages = stage_avg(st_m)
    s = st_m[-1]
    out = {
        "metric": METRIC, "value": a.N / ms_m, "unit": "particles/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_m * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64 (mixed)", "data": "synthetic",
        "config": {"workload": desc,
                   "N": a.N, "n": a.n, "N0": a.N0, "eta": a.eta, "precision": "mixed"},
        "accuracy": acc,
        "rel_err_ok": (acc.get("mixed_vs_fp64", {"E": 0, "B": 0})["E"] <= 1e-5
                       and acc.get("mixed_vs_fp64", {"E": 0, "B": 0})["B"] <= 1e-5),
        "counts": {"ncluster": s.ncluster, "nleaf": s.nleaf, "np2p": s.np2p, "nm2l": s.nm2l,
                   "p2p_particle_pairs": s.p2p_particle_pairs, "maxlevel": s.maxlevel},
        "stages_ms": stages,
        "stages_ms_isolated": stage_avg(iso_m),
        "roofline": roofline(s, stage_avg(iso_m), a.n, "mixed", a.N),
    }
    for k in ("adjoint_fp64", "adjoint_mixed", "pullback_fp64", "pullback_mixed",
              "tracking_refit_mixed", "host_pointers_mixed", "library_device_bytes"):
        if k in res:
            out[k] = res[k]
    if "fp64" in res:
        out["fp64"] = res["fp64"]
        out["speedup_vs_fp64"] = out["value"] / res["fp64"]["value"]
    ctx.close()
    if not a.skip_fp64 and not a.skip_adjoint and a.adjoint_N > 0 and a.workload == "bunch":
        out["adjoint_config"] = adjoint_case(F, W, torch, a.adjoint_N, a.n, a.N0, a.eta, lam)
    if a.large_N > 0:
        out["large_ellipsoid"] = large_case(F, W, torch, a.large_N, a.n, a.N0, a.eta, lam)
    if a.copy_probe_mib > 0:
        src = torch.ones(a.copy_probe_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = src.numel() * 4
    if a.cpu_sample > 0:
        out["cpu_baseline"] = cpu_baseline(a.cpu_sample, a.n, a.N0, a.eta, a.N)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
