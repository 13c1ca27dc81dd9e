"""Profiling harness for the NTT kernels: a few forward and coset NTTs of one
size on device-resident data, plus one plain device copy of the same array
(a known byte count, used to calibrate the HBM counters FETCH_SIZE and
WRITE_SIZE).  Run it under rocprofv3, e.g.

  rocprofv3 --pmc FETCH_SIZE -- python3 -m impeller_amd.ntt_profile --log 24 --reps 3

Prints one JSON line with the per-call times.
"""
import argparse
import json
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--log", type=int, default=24)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--plan", default="", help="pass plan, e.g. 5,10,7,6,6 (default: the tuned picks)")
    args = ap.parse_args()

    import torch
    from . import _lib
    from . import ntt as N
    from . import workloads as WL
    from .runtime import Stream, device_u32, to_device

    stream = Stream.current(0)
    k = args.log
    n = 1 << k
    plan = [int(v) for v in args.plan.split(",")] if args.plan else None
    coset_plan = plan
    if not args.plan:
        # the plans the benchmark uses: the autotuner's exhaustive picks
        import os
        picks_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned", "picks.json")
        if os.path.exists(picks_path):
            with open(picks_path) as f:
                picks = json.load(f)
            plan = (picks.get(f"ntt_fwd:{k}") or {}).get("exhaustive")
            coset_plan = (picks.get(f"ntt_coset:{k}") or {}).get("exhaustive")
    x = device_u32(n * 8)
    _lib.call("impeller_gen_scalars", x.data_ptr(), n, 0, 2024, 0, stream.raw())
    tw = N.twiddles(WL.mont_limbs(WL.omega(k)), n // 2, stream)
    shift = to_device(WL.mont_limbs(WL.GENERATOR))
    y = torch.empty_like(x)
    stream.sync()

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b)

    copy_ms = timed(lambda: y.copy_(x))
    fwd = [timed(lambda: N.ntt(x, tw, k, stream, passes=plan)) for _ in range(args.reps)]
    coset = [timed(lambda: N.ntt(x, tw, k, stream, coset_shift=shift, passes=coset_plan))
             for _ in range(args.reps)]
    print(json.dumps({"log": k, "plan": plan, "coset_plan": coset_plan, "copy_bytes": 2 * n * 32, "copy_ms": round(copy_ms, 4),
                      "ntt_ms": [round(t, 4) for t in fwd],
                      "coset_ntt_ms": [round(t, 4) for t in coset]}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
