"""Profiling harness for the MSM kernels: one device copy of a known size
(to calibrate the HBM counters FETCH_SIZE and WRITE_SIZE), then MSM 2^22 with
precomputed window tables (c = 20, one bucket set, task 64) twice.  Run it
under rocprofv3, e.g.

  rocprofv3 --pmc FETCH_SIZE -- python3 impeller_amd/msm_profile.py

Prints one JSON line with the per-call times.
"""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    import torch
    from impeller_amd import _lib
    from impeller_amd import msm as M
    from impeller_amd.runtime import Stream, device_u32

    stream = Stream.current(0)
    nm, c, task = 1 << 22, 20, 64
    bases = device_u32(nm * 16)
    _lib.call("impeller_gen_bases", bases.data_ptr(), nm, 77, 0, stream.raw())
    scal = device_u32(nm * 8)
    _lib.call("impeller_gen_scalars", scal.data_ptr(), nm, 0, 78, 0, stream.raw())
    tables = device_u32(M.n_tables(c, 1) * nm * 16)
    tables[: nm * 16].copy_(bases)
    M.precompute_device(tables, nm, c, 1, stream)
    ws = device_u32((M.cfg_workspace_size(nm, c, 1, task) + 3) // 4)
    out = device_u32(16)
    stream.sync()
    y = torch.empty_like(bases)
    times = []
    ev = lambda: torch.cuda.Event(enable_timing=True)
    a, b = ev(), ev()
    a.record()
    y.copy_(bases)
    b.record()
    b.synchronize()
    copy_ms = a.elapsed_time(b)
    for _ in range(2):
        a, b = ev(), ev()
        a.record()
        M.msm_device_cfg(tables, scal, nm, c, 1, stream, task=task, workspace=ws, out=out,
                         sync=False)
        b.record()
        b.synchronize()
        times.append(round(a.elapsed_time(b), 3))
    print(json.dumps({"copy_bytes": 2 * nm * 64, "copy_ms": round(copy_ms, 4), "msm_ms": times}),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
