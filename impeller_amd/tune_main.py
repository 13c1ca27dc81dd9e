"""Run the exhaustive autotuner on the GPU, train the cost model on its
timings and evaluate the model's picks.

  python -m impeller_amd.tune_main [--out DIR] [--ntt-logs 18-25] [--msm-logs 16-22]
                                   [--reuse TIMINGS.json] [--sweep all|ntt|msm|none]

Writes DIR/timings.json (every measured configuration), DIR/picks.json
(exhaustive best and the model's leave-one-problem-out pick per problem)
and DIR/report.json (the regret of those picks vs the exhaustive best and
of the library default).  --reuse starts from the rows
of an earlier timing file; --sweep re-measures only the named operation
(rows of that operation in the reused file are dropped).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_range(s):
    a, b = s.split("-")
    return list(range(int(a), int(b) + 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuned"))
    ap.add_argument("--ntt-logs", default="18-25")
    ap.add_argument("--msm-logs", default="16-22")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--reuse", default="")
    ap.add_argument("--sweep", default="all", choices=("all", "ntt", "msm", "none"))
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)

    from . import _lib
    from . import autotune as A
    from .runtime import Stream, device_u32

    stream = Stream.current(0)
    db = A.TimingDB()
    t0 = time.time()
    picks = {}
    sweep_ntt = args.sweep in ("all", "ntt")
    sweep_msm = args.sweep in ("all", "msm")
    if args.reuse:
        old = A.TimingDB.load(args.reuse)
        for r in old.rows:
            is_msm = r["problem"].startswith("msm")
            if (is_msm and sweep_msm) or (not is_msm and sweep_ntt):
                continue
            db.rows.append(r)
        for prob in db.problems():
            best = min(db.for_problem(prob), key=lambda r: r["ms"])
            picks[prob] = {"exhaustive": best["config"], "exhaustive_ms": best["ms"]}
        print(f"[tune] reused {len(db.rows)} rows from {args.reuse}", flush=True)
    for L in (parse_range(args.ntt_logs) if sweep_ntt else []):
        best = A.tune_ntt(db, L, stream, ops=("ntt_fwd", "ntt_inv", "ntt_coset", "ntt_ext"),
                          reps=args.reps)
        for prob, (cfg, ms) in best.items():
            picks[prob] = {"exhaustive": cfg, "exhaustive_ms": ms}
        print(f"[tune] ntt 2^{L}: {len(A.ntt_plans(L))} plans x 4 ops, {time.time() - t0:.1f}s",
              flush=True)
    msm_logs = parse_range(args.msm_logs)
    if sweep_msm:
        nb = 1 << max(msm_logs)
        bases = device_u32(nb * 16)
        _lib.call("impeller_gen_bases", bases.data_ptr(), nb, 77, 0, stream.raw())
        stream.sync()
    for lg in (msm_logs if sweep_msm else []):
        best = A.tune_msm(db, lg, bases, stream, kinds=(0, 1, 2), reps=max(1, args.reps - 1))
        for prob, (cfg, ms) in best.items():
            picks[prob] = {"exhaustive": cfg, "exhaustive_ms": ms}
        print(f"[tune] msm 2^{lg}: {len(A.msm_candidates())} configs x 3 kinds, "
              f"{time.time() - t0:.1f}s", flush=True)
    db.save(os.path.join(args.out, "timings.json"))

    # the model's pick for each problem comes from a model trained without
    # that problem's rows (leave-one-problem-out): a pick without a sweep
    t1 = time.time()
    report = A.regret_report(db, stream, steps=args.steps, holdout=True,
                             log=lambda m: print(m, flush=True))
    print(f"[tune] leave-one-problem-out picks, {time.time() - t1:.1f}s", flush=True)
    for prob, rep in report.items():
        picks[prob]["model"] = rep["model_config"]
        picks[prob]["model_ms"] = rep["model_ms"]
        picks[prob]["model_training"] = "leave-one-problem-out"
    with open(os.path.join(args.out, "picks.json"), "w") as f:
        json.dump(picks, f, indent=1)
    regrets = [v["model_regret"] for v in report.values()]
    dregrets = [v["default_regret"] for v in report.values() if v["default_regret"] is not None]
    summary = {
        "problems": len(report),
        "rows": len(db.rows),
        "model_regret_mean": sum(regrets) / len(regrets),
        "model_regret_max": max(regrets),
        "model_within_5pct": sum(1 for r in regrets if r <= 0.05) / len(regrets),
        "default_regret_mean": sum(dregrets) / len(dregrets) if dregrets else None,
        "default_regret_max": max(dregrets) if dregrets else None,
        "seconds": time.time() - t0,
    }
    with open(os.path.join(args.out, "report.json"), "w") as f:
        json.dump({"summary": summary, "per_problem": report}, f, indent=1)
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
