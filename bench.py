"""Benchmark of the BN254 hot path on one MI355X.  Prints ONE JSON line.

Plain single-call path (the metric's first two numbers): forward NTT 2^24,
coset NTT 2^24 (shift = 7) and one G1 MSM of 2^22 affine bases k_i*G with
uniform scalars, all device resident; the MSM window c is swept over 12..20.
Batched vs plain (the third number): 64 uneven NTTs (2^18..2^23, forward /
inverse / coset-extended x8) and 40 commitment MSMs (2^16..2^22, shared 2^22
base table, 50% uniform / 30% selector / 20% small scalars), run as
sequential single calls with the untuned default configurations (plain) and
as batched launches replayed from a HIP graph with the cost model's picks
(and, for reference, the exhaustive tuner's picks); outputs compared bit for
bit.  Then the fused proving chunk at k = 22 under the same three schedules
(--no-chunk skips it).

Last, the polynomial library (poly_add / sub / mul, scan_mul,
batched_invert, poly_eval, kate_division, permute) at 2^24 elements, each
against its HBM and Montgomery-product rooflines, with sampled or
whole-array parity checks, the emitted fused gate against the hand-written
one, and the device-runtime layer's copy, file and page-mapping rates
(--no-poly skips these).

  python bench.py [--steps K] [--warmup W] [--window C] [--no-cpu] [--no-batch] [--no-chunk]
                  [--no-poly]
  python bench.py --profile ntt|msm|poly   (the rocprofv3 harness: a calibration
                                            copy, then a few calls of one path)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MSM 2^22 ms; NTT 2^24 HBM GB/s; batched-vs-plain speedup (bit-exact)"
HBM_PEAK_GBS = 8000.0
TUNED = os.path.join(ROOT, "impeller_amd", "tuned", "picks.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-ntt", type=int, default=24)
    ap.add_argument("--log-msm", type=int, default=22)
    ap.add_argument("--window", type=int, default=0, help="MSM window c (0: sweep 12..20)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched comparisons")
    ap.add_argument("--no-chunk", action="store_true", help="skip the k=22 proving chunk")
    ap.add_argument("--no-poly", action="store_true", help="skip the polynomial-library ops")
    ap.add_argument("--log-poly", type=int, default=24, help="log2 size of the poly-op leg")
    ap.add_argument("--profile", choices=("ntt", "msm", "poly"), default=None,
                    help="profiling harness for rocprofv3: one calibration copy of a known "
                         "size, then a few calls of the named path, nothing else")
    ap.add_argument("--plan", default="", help="--profile ntt: pass plan, e.g. 1,10,8,8,8")
    return ap.parse_args()


def summary_value(name, key):
    """An integer "key: value" line of a committed counter summary under
    profiles/, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(key + ":"):
                return int(line.split(":")[1])
    return None


def dit_traffic(log_n):
    """HBM bytes per k_ntt_dit launch (profiles/ntt<log>_dit_hbm_traffic.txt)."""
    return summary_value(f"ntt{log_n}_dit_hbm_traffic.txt", "traffic_bytes_per_launch")


def load_picks():
    if os.path.exists(TUNED):
        with open(TUNED) as f:
            return json.load(f)
    return {}


def profile_main(args):
    """Runs under rocprofv3 (e.g. `rocprofv3 --pmc FETCH_SIZE -- python3 bench.py
    --profile msm`): a device copy of a known byte count (the calibration of
    the HBM counters FETCH_SIZE / WRITE_SIZE), then the named path a few
    times -- NTT 2^log_ntt forward and coset with the tuned plans, MSM
    2^log_msm with precomputed tables (c = --window or 20, task 64), or every
    polynomial-library op at 2^log_poly.  Prints one JSON line of times."""
    import torch
    from impeller_amd import _lib
    from impeller_amd import msm as M
    from impeller_amd import ntt as N
    from impeller_amd import workloads as WL
    from impeller_amd.runtime import Stream, device_u32, to_device

    stream = Stream.current(0)

    def timed1(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        return round(a.elapsed_time(b), 4)

    out = {"profile": args.profile}
    if args.profile == "ntt":
        k = args.log_ntt
        n = 1 << k
        picks = load_picks()
        plan = [int(v) for v in args.plan.split(",")] if args.plan else \
            (picks.get(f"ntt_fwd:{k}") or {}).get("exhaustive")
        coset_plan = plan if args.plan else (picks.get(f"ntt_coset:{k}") or {}).get("exhaustive")
        x = device_u32(n * 8)
        _lib.call("impeller_gen_scalars", x.data_ptr(), n, 0, 2024, 0, stream.raw())
        tw = N.twiddles(WL.mont_limbs(WL.omega(k)), n // 2, stream)
        shift = to_device(WL.mont_limbs(WL.GENERATOR))
        y = torch.empty_like(x)
        stream.sync()
        out.update({"log": k, "plan": plan, "coset_plan": coset_plan,
                    "copy_bytes": 2 * n * 32, "copy_ms": timed1(lambda: y.copy_(x)),
                    "ntt_ms": [timed1(lambda: N.ntt(x, tw, k, stream, passes=plan))
                               for _ in range(3)],
                    "coset_ntt_ms": [timed1(lambda: N.ntt(x, tw, k, stream, coset_shift=shift,
                                                          passes=coset_plan))
                                     for _ in range(3)]})
        # the reference-ABI entry points on the same array: ssip_ntt (in
        # place, twiddle table) and recompute_ntt (no twiddle table: each
        # pass rebuilds its twiddles from pq / omegas)
        from impeller_amd.runtime import ScalarArray
        xs, tws = ScalarArray(n, x, "cuda"), ScalarArray(n // 2, tw, "cuda")
        g = N.GenPqOmegas()
        pq, om = ScalarArray(g.get_pq_len(k), None, "cpu"), ScalarArray(32, None, "cpu")
        g.get_fn()(pq, om, n, WL.mont_limbs(WL.omega(k)))
        pq_d, om_d = ScalarArray.from_numpy(pq.to_numpy()), ScalarArray.from_numpy(om.to_numpy())
        stream.sync()
        out["ssip_ntt_ms"] = [timed1(lambda: N.SsipNtt().get_fn()([xs], [tws, stream]))
                              for _ in range(3)]
        out["recompute_ntt_ms"] = [timed1(lambda: N.RecomputeNtt().get_fn()(
            [xs], [pq_d, om_d, stream])) for _ in range(3)]
    elif args.profile == "msm":
        nm, c, task = 1 << args.log_msm, args.window or 20, 64
        bases = device_u32(nm * 16)
        _lib.call("impeller_gen_bases", bases.data_ptr(), nm, 77, 0, stream.raw())
        scal = device_u32(nm * 8)
        _lib.call("impeller_gen_scalars", scal.data_ptr(), nm, 0, 78, 0, stream.raw())
        tables = device_u32(M.n_tables(c, 1) * nm * 16)
        tables[: nm * 16].copy_(bases)
        M.precompute_device(tables, nm, c, 1, stream)
        ws = device_u32((M.cfg_workspace_size(nm, c, 1, task) + 3) // 4)
        res = device_u32(16)
        y = torch.empty_like(bases)
        stream.sync()
        out.update({"log": args.log_msm, "window": c, "copy_bytes": 2 * nm * 64,
                    "copy_ms": timed1(lambda: y.copy_(bases)),
                    "msm_ms": [timed1(lambda: M.msm_device_cfg(
                        tables, scal, nm, c, 1, stream, task=task, workspace=ws, out=res,
                        sync=False)) for _ in range(2)]})
    else:
        info = poly_leg(args, stream, lambda fn, reps: timed1(fn), 1.0, False)
        out.update({"log": args.log_poly, "ms": {k: v["ms"] for k, v in info["ops"].items()}})
    print(json.dumps(out), flush=True)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def poly_leg(args, stream, timed, mul_peak, with_cpu):
    """Each op of the polynomial library at 2^log_poly elements: median of 3
    timed calls, algorithmic bytes (operands read once, results written
    once) against 8 TB/s and the Fr products its math needs against the
    live Montgomery-product peak; parity on the same data."""
    import ctypes

    import torch
    from impeller_amd import _lib
    from impeller_amd.runtime import GpuBuffer, ScalarArray, device_u32, to_host

    lg = args.log_poly
    n = 1 << lg

    def gen(seed):
        t = device_u32(n * 8)
        _lib.call("impeller_gen_scalars", t.data_ptr(), n, 0, seed, 0, stream.raw())
        return ScalarArray(n, t, "cuda")

    a, b, r = gen(501), gen(502), gen(503)
    pt = device_u32(8)
    _lib.call("impeller_gen_scalars", pt.data_ptr(), 1, 0, 504, 0, stream.raw())
    res = device_u32(8)
    sizes = {}

    def tmp(name, *args_):
        sz = ctypes.c_ulong(0)
        _lib.call(name, None, ctypes.byref(sz), *args_)
        sizes[name] = sz.value
        return GpuBuffer(sz.value)

    pv = a.poly_ptr()
    t_scan = tmp("scan_mul", pv, None)
    t_inv = tmp("batched_invert", pv, None)
    t_eval = tmp("poly_eval", pv, None, None, None)
    t_kate = tmp("kate_division", lg, pv, None, pv, None)
    usable = n - 25
    t_perm = tmp("permute", usable, pv, pv, pv, pv, None)
    r_u = ScalarArray(usable, device="cuda")
    t_u = ScalarArray(usable, device="cuda")
    src = a.values.clone()

    def restore():
        r.values.copy_(src)

    t_restore = timed(restore, 3)
    ops = {
        "poly_add": (lambda: _lib.call("poly_add", r.poly_ptr(), a.poly_ptr(), b.poly_ptr(),
                                       stream.raw()), 3, 1, 0),
        "poly_mul": (lambda: _lib.call("poly_mul", r.poly_ptr(), a.poly_ptr(), b.poly_ptr(),
                                       stream.raw()), 3, 1, 0),
        "poly_eval": (lambda: _lib.call("poly_eval", t_eval.data_ptr(), None, a.poly_ptr(),
                                        res.data_ptr(), pt.data_ptr(), stream.raw()), 1, 1, 0),
        "kate_division": (lambda: _lib.call("kate_division", t_kate.data_ptr(), None, lg,
                                            a.poly_ptr(), pt.data_ptr(), b.poly_ptr(),
                                            stream.raw()), 2, 1, 0),
        "scan_mul": (lambda: (restore(), _lib.call("scan_mul", t_scan.data_ptr(), None,
                                                   r.poly_ptr(), stream.raw())), 2, 1, 1),
        "batched_invert": (lambda: (restore(), _lib.call("batched_invert", t_inv.data_ptr(),
                                                         None, r.poly_ptr(), stream.raw())),
                           2, 3, 1),
        "permute": (lambda: _lib.call("permute", t_perm.data_ptr(), None, usable, a.poly_ptr(),
                                      b.poly_ptr(), r_u.poly_ptr(), t_u.poly_ptr(),
                                      stream.raw()), 4, 0, 0),
    }
    out = {}
    for name, (fn, nbytes, nmul, restores) in ops.items():
        ms = timed(fn, 3) - (t_restore if restores else 0.0)
        gbs = nbytes * n * 32 / (ms * 1e-3) / 1e9
        e = {"ms": round(ms, 4), "algorithmic_bytes": nbytes * n * 32, "gb_s": round(gbs, 1),
             "hbm_frac": round(gbs * 1e9 / (HBM_PEAK_GBS * 1e9), 4)}
        if nmul:
            rate = nmul * n / (ms * 1e-3)
            e.update({"fr_muls": nmul * n, "fr_mul_per_s": round(rate, 1),
                      "valu_frac": round(rate / mul_peak, 4)})
        out[name] = e
    info = {"log_n": lg, "ops": out, "temp_bytes": sizes, "mont_mul_peak": round(mul_peak, 1)}
    if with_cpu:
        # the checker leg: parity on the benchmarked data, and the oracle's
        # single-thread times at 2^20 beside the GPU ones
        from oracle import oracle as O
        ah = to_host(a.values)
        ph = to_host(pt)[0]
        parity = {}
        _lib.call("poly_eval", t_eval.data_ptr(), None, a.poly_ptr(), res.data_ptr(),
                  pt.data_ptr(), stream.raw())
        stream.sync()
        parity["poly_eval_full"] = bool((to_host(res)[0] == O.poly_eval(ah, ph)).all())
        _lib.call("kate_division", t_kate.data_ptr(), None, lg, a.poly_ptr(), pt.data_ptr(),
                  b.poly_ptr(), stream.raw())
        stream.sync()
        qh = to_host(b.values)
        ok = bool((qh[n - 1] == 0).all())
        for i in (0, n // 2, n - 2):
            ok &= bool((qh[i] == O.poly_eval(ah[i + 1:], ph)).all())
        parity["kate_sampled_quotients"] = ok
        restore()
        _lib.call("scan_mul", t_scan.data_ptr(), None, r.poly_ptr(), stream.raw())
        stream.sync()
        parity["scan_mul_full"] = bool((to_host(r.values) == O.scan_mul(ah)).all())
        restore()
        _lib.call("batched_invert", t_inv.data_ptr(), None, r.poly_ptr(), stream.raw())
        _lib.call("batched_invert", t_inv.data_ptr(), None, r.poly_ptr(), stream.raw())
        stream.sync()
        parity["batched_invert_involution"] = bool(torch.equal(r.values, src))
        info["parity"] = parity
        m = 1 << 20
        xs = O.gen_scalars(m, 9)
        bb = O.gen_scalars(1, 10)[0]
        cpu = {}
        for name, f in (("scan_mul", lambda: O.scan_mul(xs)),
                        ("batched_invert", lambda: O.batch_invert(xs)),
                        ("poly_eval", lambda: O.poly_eval(xs, bb)),
                        ("kate_division", lambda: O.kate_division(xs, bb))):
            t = time.perf_counter()
            f()
            cpu[name] = round((time.perf_counter() - t) * 1e3, 2)
        info["cpu_port_ms_2^20_1_thread"] = cpu
    del a, b, r, r_u, t_u, src, t_scan, t_inv, t_eval, t_kate, t_perm
    torch.cuda.empty_cache()
    return info


def runtime_leg():
    """The device-runtime layer (impeller_amd.hip_api over impeller_rt_*):
    pinned host <-> device copy rates on its own stream (256 MiB), the
    file -> device rate through 8 MiB of pinned staging (a 256 MiB file
    written just before, so mostly from the page cache), and the cost of
    mapping virtual-memory pages (allocate of 32 fresh 2 MiB-aligned pages
    into one address), each checked for content."""
    import ctypes
    import os
    import tempfile
    import time

    import numpy as np
    from impeller_amd import hip_api as H

    nbytes = 256 << 20
    s = H.HipStream(0)
    a, b = H.HipEvent(0, host_ordered=False), H.HipEvent(0, host_ordered=False)
    hp = H.alloc_pinned(nbytes)
    d = s.allocate(nbytes)
    src = np.frombuffer(os.urandom(1 << 20), np.uint8)
    host = np.ctypeslib.as_array((ctypes.c_uint8 * nbytes).from_address(hp))
    host[:] = np.tile(src, nbytes >> 20)

    def rate(kind):
        best = None
        for _ in range(3):
            a.record(s)
            if kind == "h2d":
                s.memcpy_h2d(d, hp, nbytes)
            else:
                s.memcpy_d2h(hp, d, nbytes)
            b.record(s)
            b.sync()
            ms = a.elapsed(b)
            best = ms if best is None else min(best, ms)
        return round(nbytes / best / 1e6, 1)

    h2d = rate("h2d")
    d2h = rate("d2h")
    ok = bool((host[: 1 << 20] == src).all())
    st_bytes = 8 << 20
    st = H.alloc_pinned(st_bytes)
    with tempfile.NamedTemporaryFile(dir="/tmp", delete=False) as f:
        f.write(np.tile(src, nbytes >> 20).tobytes())
        path = f.name
    fd = H.open_file(path)
    t0 = time.perf_counter()
    H.read_file_to_device(fd, 0, nbytes, d, st, st_bytes, s)
    file_s = time.perf_counter() - t0
    os.close(fd)
    os.unlink(path)
    back = np.zeros(1 << 20, np.uint8)
    s.memcpy_d2h(back.ctypes.data, d + nbytes - (1 << 20), 1 << 20)
    s.sync()
    ok &= bool((back == src).all())
    g = H.page_granularity(0)
    ps = max(g, 2 << 20) // g * g
    pa = H.PageAllocator(0, ps, 32, True)
    t0 = time.perf_counter()
    p = pa.allocate(32 * ps, list(range(32)), 0)
    map_ms = (time.perf_counter() - t0) * 1e3
    s.memcpy_h2d(p, hp, min(32 * ps, nbytes))
    s.memcpy_d2h(back.ctypes.data, p, 1 << 20)
    s.sync()
    ok &= bool((back == src).all())
    pa.deallocate(p, 0)
    pa.destroy()
    s.free(d)
    s.sync()
    H.free_pinned(st)
    H.free_pinned(hp)
    for e in (a, b):
        e.destroy()
    s.destroy()
    return {"pinned_h2d_gb_s": h2d, "pinned_d2h_gb_s": d2h,
            "file_to_device_gb_s": round(nbytes / file_s / 1e9, 2),
            "page_map_32_pages_ms": round(map_ms, 3), "page_size": ps, "granularity": g,
            "contents_equal": ok}


def fused_leg(stream, timed, log_n=24):
    """The C5 gate emitted by impeller_amd.kernel_gen (built beforehand into
    impeller_amd/_fused) against the hand-written gate kernel, through the
    same generated-kernel signature, at 2^log_n elements: 7 x 32 B per
    element of algorithmic traffic."""
    import ctypes

    import numpy as np
    import torch
    from impeller_amd import _lib
    from impeller_amd import kernel_gen as KG
    from impeller_amd import workloads as WL
    from impeller_amd.runtime import ScalarArray, device_u32

    op = WL.c5_gate_op()
    kern = KG.FusedKernel(op.meta(WL.FUSED_DIR))
    n = 1 << log_n
    arrs = []
    for i in range(7):
        t = device_u32(n * 8)
        _lib.call("impeller_gen_scalars", t.data_ptr(), n, 0, 900 + i, 0, stream.raw())
        arrs.append(ScalarArray(n, t, "cuda"))
    one = torch.from_numpy(WL.mont_limbs(1).view(np.int32).copy()).cuda()  # the scalar 1
    ptrs = [a.poly_ptr() for a in arrs[:6]] + [_lib.PolyPtr(one.data_ptr(), 1, 0, 0, 1),
                                               arrs[6].poly_ptr()]
    host = (_lib.PolyPtr * 8)(*ptrs)
    argb = torch.from_numpy(np.frombuffer(bytes(host), dtype=np.uint8).copy()).cuda()
    sl = _lib.PolySlice(0, n, n)
    base = argb.data_ptr()
    emitted = lambda: _lib.check(op.name, kern.fn(base, base + 7 * 40, n, ctypes.byref(sl),  # noqa
                                                  None, stream.raw()))
    # hand-written: vars = the six arrays, mut_vars = the output
    host2 = (_lib.PolyPtr * 7)(*([a.poly_ptr() for a in arrs[:6]] + [arrs[6].poly_ptr()]))
    argb2 = torch.from_numpy(np.frombuffer(bytes(host2), dtype=np.uint8).copy()).cuda()
    b2 = argb2.data_ptr()
    hand = lambda: _lib.call("fused_gate_abc_d_qe", b2, b2 + 6 * 40, n, ctypes.byref(sl),  # noqa
                             None, stream.raw())
    t_e = timed(emitted, 5)
    out_e = arrs[6].values.clone()
    t_h = timed(hand, 5)
    same = bool(torch.equal(out_e, arrs[6].values))
    nbytes = 7 * n * 32
    res = {"log_n": log_n, "emitted_ms": round(t_e, 4), "handwritten_ms": round(t_h, 4),
           "emitted_gb_s": round(nbytes / (t_e * 1e-3) / 1e9, 1),
           "emitted_hbm_frac": round(nbytes / (t_e * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "handwritten_gb_s": round(nbytes / (t_h * 1e-3) / 1e9, 1),
           "outputs_equal": same}
    del arrs, argb, argb2
    torch.cuda.empty_cache()
    return res


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.profile:
        profile_main(args)
        return

    import numpy as np
    import torch
    from impeller_amd import _lib
    from impeller_amd import msm as M
    from impeller_amd import ntt as N
    from impeller_amd import workloads as WL
    from impeller_amd.runtime import Stream, device_u32, to_device, to_host

    # graphs cannot be captured on the legacy default stream
    side = torch.cuda.Stream()
    torch.cuda.set_stream(side)
    stream = Stream.current(0)
    r = WL.R_MOD
    mont = WL.mont_limbs
    k = args.log_ntt
    n = 1 << k
    nm = 1 << args.log_msm
    w = WL.omega(k)
    picks = load_picks()

    def ev():
        return torch.cuda.Event(enable_timing=True)

    def timed(fn, reps):
        fn()
        best = None
        for _ in range(reps):
            a, b = ev(), ev()
            a.record()
            fn()
            b.record()
            b.synchronize()
            t = a.elapsed_time(b)
            best = t if best is None else min(best, t)
        return best

    # ---------------------------------------------------------------- inputs
    x = device_u32(n * 8)
    _lib.call("impeller_gen_scalars", x.data_ptr(), n, 0, 2024, 0, stream.raw())
    tw = N.twiddles(mont(w), n // 2, stream)
    shift = to_device(mont(WL.GENERATOR))
    bases = device_u32(nm * 16)
    _lib.call("impeller_gen_bases", bases.data_ptr(), nm, 77, 0, stream.raw())
    scal = device_u32(nm * 8)
    _lib.call("impeller_gen_scalars", scal.data_ptr(), nm, 0, 78, 0, stream.raw())
    stream.sync()
    x_orig = x.clone()
    # plans / MSM task from the autotuner's exhaustive picks when present
    ntt_plan = (picks.get(f"ntt_fwd:{k}") or {}).get("exhaustive")
    coset_plan = (picks.get(f"ntt_coset:{k}") or {}).get("exhaustive")
    msm_task = ((picks.get(f"msm:{args.log_msm}:0") or {}).get("exhaustive") or [0, 0, 64])[2]

    # ---------------------------------------- MSM 2^22: window sweep + tables
    windows = [args.window] if args.window else list(range(12, 21))

    def time_msm(tables, c, target, task, reps=2):
        ws = device_u32((M.cfg_workspace_size(nm, c, target, task) + 3) // 4)
        out = device_u32(16)
        t = timed(lambda: M.msm_device_cfg(tables, scal, nm, c, target, stream, task=task,
                                           workspace=ws, out=out, sync=False), reps)
        del ws
        return t

    sweep, sweep_plain, setup_ms = {}, {}, {}
    for c in windows:
        tables = device_u32(M.n_tables(c, 1) * nm * 16)
        tables[: nm * 16].copy_(bases)
        a, b = ev(), ev()
        a.record()
        M.precompute_device(tables, nm, c, 1, stream)
        b.record()
        b.synchronize()
        setup_ms[c] = a.elapsed_time(b)
        sweep[c] = time_msm(tables, c, 1, msm_task)
        del tables
        torch.cuda.empty_cache()
    for c in (windows if args.window else (15, 16, 17)):
        sweep_plain[c] = time_msm(bases, c, 0, 64)

    # end to end through the reference's host-buffer entry point `msm`
    # (msm.cu:9-41 contract; c = 16, one bucket set per window): the points
    # (256 MiB) and scalars (128 MiB) cross PCIe from pinned host memory and
    # the affine result comes back, each call
    def msm_end_to_end(parts=1, stages=1, reps=3):
        import ctypes
        pts_h = torch.empty(nm * 16, dtype=torch.int32, pin_memory=True)
        scal_h = torch.empty(nm * 8, dtype=torch.int32, pin_memory=True)
        res_h = torch.zeros(16, dtype=torch.int32, pin_memory=True)
        pts_h.copy_(bases)
        scal_h.copy_(scal)
        sizes = (ctypes.c_ulong * 1)()
        cards = (ctypes.c_uint * 1)(0)
        _lib.call("msm", None, sizes, nm, 1, parts, stages, stages, 1, cards, None, 0, None,
                  None)
        buf = device_u32((sizes[0] + 3) // 4)
        arr = lambda *p: (ctypes.c_void_p * len(p))(*p)
        times = []
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t = time.perf_counter()
            _lib.call("msm", arr(buf.data_ptr()), None, nm, 1, parts, stages, stages, 1, cards,
                      arr(pts_h.data_ptr()), 1, arr(scal_h.data_ptr()), arr(res_h.data_ptr()))
            times.append((time.perf_counter() - t) * 1e3)
        del buf
        torch.cuda.empty_cache()
        return sorted(times[1:])[reps // 2], res_h.numpy().view(np.uint32).copy(), sizes[0]

    # parts x stage buffers: points and scalars of part p + 1 cross PCIe on
    # a copy stream while part p computes (msm_impl.cuh:473-501)
    e2e = {}
    for parts, stages in ((1, 1), (4, 2), (8, 2)):
        ms_, res_, dev_bytes = msm_end_to_end(parts, stages)
        e2e[(parts, stages)] = (ms_, res_, dev_bytes)
    e2e_key = min(e2e, key=lambda kk: e2e[kk][0])
    e2e_ms, e2e_res = e2e[e2e_key][0], e2e[e2e_key][1]
    e2e_all_equal = all(bool((v[1] == e2e_res).all()) for v in e2e.values())
    best_c = min(sweep, key=sweep.get)
    tables = device_u32(M.n_tables(best_c, 1) * nm * 16)
    tables[: nm * 16].copy_(bases)
    M.precompute_device(tables, nm, best_c, 1, stream)
    ws = device_u32((M.cfg_workspace_size(nm, best_c, 1, msm_task) + 3) // 4)
    res_buf = device_u32(16)

    def run_msm():
        return M.msm_device_cfg(tables, scal, nm, best_c, 1, stream, task=msm_task,
                                workspace=ws, out=res_buf, sync=False)

    # ------------------------------------------------------ plain timed steps
    for _ in range(args.warmup):
        N.ntt(x, tw, k, stream, passes=ntt_plan)
        N.ntt(x, tw, k, stream, coset_shift=shift, passes=coset_plan)
        run_msm()
    torch.cuda.synchronize()
    evs = [[ev() for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        e = evs[i]
        e[0].record()
        N.ntt(x, tw, k, stream, passes=ntt_plan)
        e[1].record()
        N.ntt(x, tw, k, stream, coset_shift=shift, passes=coset_plan)
        e[2].record()
        res = run_msm()
        e[3].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    med = args.steps // 2
    ntt_ms = sorted(e[0].elapsed_time(e[1]) for e in evs)[med]
    coset_ms = sorted(e[1].elapsed_time(e[2]) for e in evs)[med]
    msm_ms = sorted(e[2].elapsed_time(e[3]) for e in evs)[med]
    # per-kernel times from the library's timing registry, in extra steps
    # after the timed ones (its events around every launch would otherwise
    # sit inside the timed region)
    _lib.timing_reset()
    _lib.timing_enable(True)
    for _ in range(2):
        N.ntt(x, tw, k, stream, passes=ntt_plan)
        N.ntt(x, tw, k, stream, coset_shift=shift, passes=coset_plan)
        run_msm()
    torch.cuda.synchronize()
    _lib.timing_enable(False)
    kt = {name: _lib.timing_read(name) for name in _lib.KERNEL_IDS}
    bytes_ntt = 2 * n * 32
    ntt_gbs = bytes_ntt / (ntt_ms * 1e-3) / 1e9
    dit_ms, dit_cnt = kt["ntt_dit"]
    dit_avg_ms = dit_ms / max(dit_cnt, 1)
    dit_gbs = bytes_ntt / (dit_avg_ms * 1e-3) / 1e9
    msm_res_host = to_host(res, 16)[0].copy()

    # VALU side: Montgomery-product throughput measured live (two independent
    # chains per thread, 4096 x 256 threads) against the products the NTT and
    # the MSM bucket kernel perform
    def mul_rate(form, iters=200, blocks=4096):
        out = device_u32(blocks * 256)
        launch = lambda: _lib.call("impeller_mont_mul_rate", form, blocks, iters,  # noqa: E731
                                   out.data_ptr(), stream.raw())
        return blocks * 256 * 2 * iters / (timed(launch, 3) * 1e-3)

    rate_fr32, rate_fq29 = mul_rate(0), mul_rate(1)
    ntt_muls = (n // 2) * k
    tasks_ms, tasks_cnt = kt["msm_buckets"]
    tasks_avg_ms = tasks_ms / max(tasks_cnt, 1)
    msm_madds = (254 // best_c + 1) * nm          # one bucket set: W entries per point
    valu = {
        "ntt": {"what": f"forward NTT 2^{k}: (n/2) log n Fr Montgomery products",
                "achieved": round(ntt_muls / (ntt_ms * 1e-3), 1), "peak": round(rate_fr32, 1),
                "unit": "Fr mont-mul/s (8 x 32-bit form)",
                "frac": round(ntt_muls / (ntt_ms * 1e-3) / rate_fr32, 4)},
        "msm": {"what": f"kb_tasks: {msm_madds} mixed additions x 10 Fq products",
                "achieved": round(10 * msm_madds / (tasks_avg_ms * 1e-3), 1),
                "peak": round(rate_fq29, 1), "unit": "Fq mont-mul/s (9 x 29-bit form)",
                "frac": round(10 * msm_madds / (tasks_avg_ms * 1e-3) / rate_fq29, 4),
                "kernel_avg_ms": round(tasks_avg_ms, 4),
                # fetch + write bytes per launch at 2^22, c = 20 (counter summary)
                "traffic": summary_value("msm22_hbm_traffic.txt", "kb_tasks_traffic_bytes_per_launch")
                if (nm, best_c) == (1 << 22, 20) else None},
    }

    # ---------------------------------------------------- batched vs plain
    # Configurations come from this run: the reduced candidate space of
    # every benched problem is timed, and each problem's model pick comes
    # from a cost model trained (on the GPU) without that problem's rows.
    # Legs: plain / batched x default / model / exhaustive configurations,
    # so that the batching gain and the configuration gain show apart.
    from impeller_amd import autotune as A

    run_picks, cm = {}, None
    if not (args.no_batch and args.no_chunk):
        ntt_probs = {("ntt_inv", 22), ("ntt_ext", 25), ("ntt_inv", 25)}
        for kind, lg in WL.ntt_batch_spec():
            ntt_probs.add({"fwd": ("ntt_fwd", lg), "inv": ("ntt_inv", lg),
                           "ext": ("ntt_ext", lg + 3)}[kind])
        msm_probs = set(WL.msm_batch_spec()) | {(22, 0)}
        del ws
        torch.cuda.empty_cache()
        _, run_picks, cm = A.inrun_tune(stream, sorted(ntt_probs), sorted(msm_probs),
                                        log=lambda m: print(m, file=sys.stderr, flush=True))
        torch.cuda.empty_cache()

    def picked(prob, key):
        pk = run_picks.get(prob)
        return pk.get(key) if pk else None

    def ntt_plans(key):
        out = {}
        for kind, lg in WL.ntt_batch_spec():
            L = lg + 3 if kind == "ext" else lg
            op = {"fwd": "ntt_fwd", "inv": "ntt_inv", "ext": "ntt_ext"}[kind]
            pl = picked(f"{op}:{L}", key)
            if pl:
                out[(kind, L)] = pl
        return out

    def msm_configs(key):
        out = []
        for lg, kind in WL.msm_batch_spec():
            cfg = picked(f"msm:{lg}:{kind}", key)
            out.append(tuple(cfg) if cfg else A.default_msm(lg))
        return out

    batch = None
    batch_parity = {}
    if not args.no_batch:
        # 64 uneven NTTs; every schedule restores its in-place inputs first and
        # the restore time is measured on its own and subtracted
        spec = WL.ntt_batch_spec()
        t_ntt, t_msm = {}, {}
        ref_out = None
        ntt_equal = True
        for key in ("default", "model", "exhaustive"):
            wl = WL.NttWorkload(spec, stream, plans=ntt_plans(key))
            t_restore = timed(wl.restore, 2)
            if key != "exhaustive":
                t_ntt[("plain", key)] = timed(lambda: (wl.restore(), wl.run_plain()), 2) - t_restore
                if ref_out is None:
                    ref_out = [it["data"].clone() for it in wl.items]
                ntt_equal &= all(bool(torch.equal(o, it["data"]))
                                 for o, it in zip(ref_out, wl.items))
            nb_ = wl.batch()
            g = N.Graph(stream, lambda: nb_.run(stream))
            t_ntt[("batched", key)] = timed(lambda: (wl.restore(), g.launch(stream)), 2) - t_restore
            ntt_equal &= all(bool(torch.equal(o, it["data"])) for o, it in zip(ref_out, wl.items))
            del g, nb_, wl
            torch.cuda.empty_cache()
        batch_parity["ntt_batched_equals_plain"] = ntt_equal
        del ref_out
        torch.cuda.empty_cache()

        # 40 commitment MSMs on a shared 2^22 base table
        mspec = WL.msm_batch_spec()
        msm_wl = WL.MsmWorkload(mspec, stream)
        ref_res = None
        msm_equal = True
        for key in ("default", "model", "exhaustive"):
            cfgs = msm_configs(key)
            if key != "exhaustive":
                wss = msm_wl.workspaces(cfgs)
                res_p = device_u32(16 * len(mspec))
                msm_wl.run_plain(cfgs, res_p, wss)     # builds the table sets
                t_msm[("plain", key)] = timed(lambda: msm_wl.run_plain(cfgs, res_p, wss), 2)
                if ref_res is None:
                    ref_res = res_p.clone()
                msm_equal &= bool(torch.equal(ref_res, res_p))
                del wss, res_p
                torch.cuda.empty_cache()
            mb = M.MsmBatch(msm_wl.items(cfgs), stream)
            g = N.Graph(stream, lambda: mb.run(stream))
            t_msm[("batched", key)] = timed(lambda: g.launch(stream), 2)
            msm_equal &= bool(torch.equal(ref_res, mb.results))
            del g, mb
            torch.cuda.empty_cache()
        batch_parity["msm_batched_equals_plain"] = msm_equal

        def tot(leg, key):
            return t_ntt[(leg, key)] + t_msm[(leg, key)]

        batch = {"legs_ms": {f"{leg}_{key}": {"ntt64": round(t_ntt[(leg, key)], 3),
                                              "msm40": round(t_msm[(leg, key)], 3),
                                              "total": round(tot(leg, key), 3)}
                             for (leg, key) in t_ntt},
                 "plain_configs": "default (no tuner)",
                 "batched_configs": "cost model picks (in-run, held out)"}
        batch["ntt64_plain_ms"] = round(t_ntt[("plain", "default")], 3)
        batch["ntt64_batched_graph_ms"] = round(t_ntt[("batched", "model")], 3)
        batch["msm40_plain_ms"] = round(t_msm[("plain", "default")], 3)
        batch["msm40_batched_graph_ms"] = round(t_msm[("batched", "model")], 3)
        batch["ntt64_speedup"] = round(t_ntt[("plain", "default")] / t_ntt[("batched", "model")], 3)
        batch["msm40_speedup"] = round(t_msm[("plain", "default")] / t_msm[("batched", "model")], 3)
        # plain-default -> batched-model, and its two factors
        batch["speedup"] = round(tot("plain", "default") / tot("batched", "model"), 3)
        batch["batching_gain"] = round(tot("plain", "model") / tot("batched", "model"), 3)
        batch["config_gain"] = round(tot("batched", "default") / tot("batched", "model"), 3)
        batch["model_vs_exhaustive"] = round(tot("batched", "model")
                                             / tot("batched", "exhaustive"), 3)
        batch["model_training"] = cm["model_training"] if cm else None
        del msm_wl
        torch.cuda.empty_cache()
        ws = None

    chunk = None
    if not args.no_chunk:
        # k = 22 proving chunk: plain single calls with the default
        # configuration vs the batched graph with the cost model's picks
        pc = WL.ProvingChunk(stream)
        K = pc.k + pc.EXT

        def chunk_plans(key):
            out = {}
            for kind, L, op in (("inv", pc.k, "ntt_inv"), ("ext", K, "ntt_ext"),
                                ("inv", K, "ntt_inv")):
                pl = picked(f"{op}:{L}", key)
                if pl:
                    out[(kind, L)] = pl
            return out

        def chunk_tables(cfg):
            if M.n_tables(cfg[0], cfg[1]) == 1:
                return pc.bases, 0
            t = device_u32(M.n_tables(cfg[0], cfg[1]) * (1 << pc.k) * 16)
            t[: (1 << pc.k) * 16].copy_(pc.bases)
            M.precompute_device(t, 1 << pc.k, cfg[0], cfg[1], stream)
            return t, 1 << pc.k

        cfg_d = A.default_msm(pc.k)
        tab_d, _ = chunk_tables(cfg_d)
        wsc = device_u32((M.cfg_workspace_size(1 << pc.k, *cfg_d) + 3) // 4)
        t_plain = timed(lambda: pc.run_plain(cfg_d, msm_tables=tab_d, ws=wsc), 1)
        plain_commit = pc.results.clone()
        del wsc, tab_d
        torch.cuda.empty_cache()
        chunk = {"plain_ms": round(t_plain, 2), "plain_config": list(cfg_d)}
        equal = True
        for key in ("model", "exhaustive"):
            cfg = tuple(picked(f"msm:{pc.k}:0", key) or cfg_d)  # in-run picks
            tab, stride = chunk_tables(cfg)
            nbs = pc.batches(chunk_plans(key))
            mbc = pc.msm_batch(cfg, tables=tab, table_stride=stride)
            g = N.Graph(stream, lambda: pc.run_batched(nbs, mbc))
            t_b = timed(lambda: g.launch(stream), 1)
            equal &= bool(torch.equal(plain_commit, mbc.results))
            chunk[f"batched_{key}_ms"] = round(t_b, 2)
            chunk[f"{key}_config"] = list(cfg)
            del g, nbs, mbc, tab
            torch.cuda.empty_cache()
        chunk["batched_graph_ms"] = chunk["batched_model_ms"]
        chunk["speedup"] = round(t_plain / chunk["batched_model_ms"], 3)
        chunk["commitments_equal"] = equal
        del pc

    # ------------------------------------------------- polynomial library
    poly = None
    fused = None
    rt = None
    if not args.no_poly:
        poly = poly_leg(args, stream, timed, rate_fr32, not args.no_cpu)
        fused = fused_leg(stream, timed)
        rt = runtime_leg()

    # ------------------------------------------------ parity + CPU baseline
    parity = dict(batch_parity)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        kk = O.gen_scalars(nm, 77)
        expect = O.msm_closed_form(to_host(scal), kk)
        parity["msm_closed_form"] = bool((msm_res_host == expect).all())
        parity["msm_end_to_end_closed_form"] = bool((e2e_res == expect.reshape(-1)).all())
        x.copy_(x_orig)
        N.ntt(x, tw, k, stream, passes=ntt_plan)
        stream.sync()
        xh, yh = to_host(x_orig), to_host(x)
        ok = True
        for j in (0, 1, n // 3, n - 1):
            ok &= bool((O.poly_eval(xh, mont(pow(w, j, r))) == yh[j]).all())
        parity["ntt_sampled_outputs"] = ok
        # the checker's CPU path at the plain-path sizes: best_fft 2^k forward
        # and coset (distribute_powers then best_fft), best_multiexp 2^log_msm
        # on the same device-generated bases, all host threads
        threads = O.host_threads()
        xc = to_host(x_orig)
        t = time.perf_counter()
        cpu_fft = O.best_fft(xc, mont(w), k, threads)
        t_fft = time.perf_counter() - t
        parity["ntt_full_vs_cpu"] = bool((cpu_fft == yh).all())
        gpow = O.powers(mont(WL.GENERATOR), n)      # a precomputed table, untimed
        t = time.perf_counter()
        cpu_coset = O.best_fft(O.distribute_powers(xc, gpow), mont(w), k, threads)
        t_coset = time.perf_counter() - t
        x.copy_(x_orig)
        N.ntt(x, tw, k, stream, coset_shift=shift, passes=coset_plan)
        stream.sync()
        parity["coset_ntt_full_vs_cpu"] = bool((cpu_coset == to_host(x)).all())
        bh = to_host(bases, 16)
        sh = to_host(scal)
        t = time.perf_counter()
        cpu_msm = O.best_multiexp(sh, bh, threads)
        t_msm = time.perf_counter() - t
        parity["msm_cpu_best_multiexp"] = bool((cpu_msm.reshape(-1) == expect.reshape(-1)).all())
        cpu = {
            "value": round(2 * n * 32 / t_fft / 1e9, 4),
            "unit": "GB/s",
            "cores": threads,
            "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"the full plain-path workload on the host: best_fft 2^{k} "
                      f"({t_fft * 1e3:.0f} ms), coset best_fft 2^{k} ({t_coset * 1e3:.0f} ms), "
                      f"best_multiexp 2^{args.log_msm} ({t_msm * 1e3:.0f} ms); C restatement "
                      f"of halo2's CPU path, {threads} threads (the coset's g^i multiply "
                      f"is one thread)",
            "ntt_ms": round(t_fft * 1e3, 2), "coset_ntt_ms": round(t_coset * 1e3, 2),
            "ntt_log": k, "msm_ms": round(t_msm * 1e3, 2), "msm_log": args.log_msm,
            "gpu_speedup": {"ntt": round(t_fft * 1e3 / ntt_ms, 1),
                            "coset_ntt": round(t_coset * 1e3 / coset_ms, 1),
                            "msm": round(t_msm * 1e3 / msm_ms, 1)},
        }

    line = {
        "metric": METRIC,
        "value": round(ntt_gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (BN254 Fr/Fq Montgomery, 8x32-bit limbs)",
        "data": "synthetic (seeded SplitMix64 uniform Fr; bases k_i*G)",
        "config": {
            "workload": f"NTT 2^{k} + coset NTT 2^{k} + G1 MSM 2^{args.log_msm} "
                        "(plain single-call path, device resident)",
            "ntt_log": k, "msm_log": args.log_msm, "msm_window": best_c,
            "ntt_plan": ntt_plan or "default",
            "msm_layout": f"precomputed window tables (one bucket set), task {msm_task}",
            "parallelism": "single GPU",
        },
        "ntt_2^24_ms": round(ntt_ms, 3),
        "coset_ntt_2^24_ms": round(coset_ms, 3),
        "msm_2^22_ms": round(msm_ms, 3),
        "msm_window_sweep_ms": {str(c): round(v, 3) for c, v in sweep.items()},
        "msm_table_setup_ms": {str(c): round(v, 2) for c, v in setup_ms.items()},
        "msm_no_tables_ms": {str(c): round(v, 3) for c, v in sweep_plain.items()},
        "msm_2^22_end_to_end_ms": {
            "value": round(e2e_ms, 3), "parts": e2e_key[0], "stage_buffers": e2e_key[1],
            "vs_device_resident_c16": round(e2e_ms / sweep_plain[16], 3)
            if 16 in sweep_plain else None,
            "by_parts": {f"{p_}x{s_}": {"ms": round(v[0], 3), "device_bytes": v[2]}
                         for (p_, s_), v in e2e.items()},
            "results_equal": e2e_all_equal,
            "what": "reference host-buffer entry point msm (c = 16, per-window buckets): "
                    "pinned host points and scalars in, affine result out, PCIe included; "
                    "parts x stage buffers overlap the copies with the computation"},
        "batched_vs_plain_speedup": batch["speedup"] if batch else None,
        "batched": batch,
        "cost_model": cm,
        "proving_chunk": chunk,
        "poly_ops": poly,
        "fused_emitter": fused,
        "runtime_api": rt,
        "kernels_ms": {name: {"total": round(v[0], 3), "launches": v[1]}
                       for name, v in kt.items() if v[1]},
        "roofline": {
            "kernel": "k_ntt_dit (NTT LDS pass)",
            "bound": "hbm",
            "achieved": round(dit_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dit_gbs / HBM_PEAK_GBS, 5),
            "traffic": dit_traffic(k),
            "algorithmic_bytes_per_launch": bytes_ntt,
            "avg_launch_ms": round(dit_avg_ms, 4),
        },
        "valu_roofline": valu,
        "parity": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
