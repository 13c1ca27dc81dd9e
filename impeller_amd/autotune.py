"""Exhaustive autotuner and learned configuration picker.

The tuner times every candidate configuration of an operation on the GPU
(HIP events, median of a few warm runs) and records (features, ms) rows in a
timing database.  The cost model (costmodel.CostModel, an FP32 MLP trained
on those rows by the library's HIP kernels) then ranks candidates for a new
problem without a sweep; the exhaustive result stays the fallback and the
yardstick of the pick's regret.

Configuration spaces:
  NTT  plan = [fuse, lds_log, d0, d1, ...]  (fuse in {0, 4, 5}: DIT levels
       done in the bit-reversal kernel, or 1: bit reversal fused into the
       first pass; lds_log in {10, 11}; passes of the remaining levels)
  MSM  (window c, target windows (0 = per-window buckets, 1 = precomputed
       window tables), task size T)
"""
import itertools
import json
import os

import numpy as np

OPS = ["ntt_fwd", "ntt_inv", "ntt_coset", "msm"]
# ntt_ext: the coset extension of 2^(L-3) coefficients to 2^L points (a batch
# item with a zero-extended source), featured as ntt_coset plus its levels
EXT_LEVELS = 3
N_FEATURES = 24


# ------------------------------------------------------------------ spaces
def ntt_plans(L, max_passes=4):
    """Candidate plans for a 2^L transform."""
    out = []
    for fuse in ((0, 1, 4, 5) if L >= 10 else ((0, 1, 4) if L >= 8 else (0, 1))):
        rest = L - fused_levels(fuse)
        for lds_log in (10, 11):
            for k in range(2 if fuse == 1 else 1, max_passes + 1):
                lo = max(1, rest // k - 2)
                for parts in itertools.product(range(lo, lds_log + 1), repeat=k):
                    if sum(parts) != rest:
                        continue
                    # keep non-increasing splits only (order matters little),
                    # except the first pass of a fused reversal: it reads runs
                    # of 2^(lds_log - d0) elements, so it is kept at least 4
                    # columns wide and may be shallower than the rest
                    tail = parts[1:] if fuse == 1 else parts
                    if list(tail) != sorted(tail, reverse=True):
                        continue
                    if fuse == 1 and lds_log - parts[0] < min(2, rest - parts[0]):
                        continue
                    out.append([fuse, lds_log, *parts])
    return out


def fused_levels(fuse):
    """DIT levels the plan's first entry takes out of the LDS passes."""
    return fuse if fuse in (4, 5) else 0


def default_ntt_plan(L):
    """The library's default plan (default_plan in csrc/ntt.hip)."""
    if L >= 20:
        rest = L - 8
        np_ = (rest + 8) // 9
        return [1, 10, 8] + [rest // np_ + (1 if i < rest % np_ else 0) for i in range(np_)]
    fuse = 4 if L >= 8 else 0
    rest = L - fuse
    np_ = max(1, (rest + 7) // 8)
    return [fuse, 10] + [rest // np_ + (1 if i < rest % np_ else 0) for i in range(np_)]


def msm_candidates(windows=range(10, 21), targets=(0, 1), tasks=(32, 64, 128)):
    return [(c, t, T) for c in windows for t in targets for T in tasks]


DEFAULT_MSM = (16, 0, 64)  # the plain `msm` symbol's configuration


def default_msm(log_n):
    """The configuration the batched workloads use without a tuned pick:
    heuristic window, precomputed tables, task 64."""
    from .workloads import heuristic_window
    return (heuristic_window(log_n), 1, 64)


# --------------------------------------------------------------- features
# Slots 0-3: op one-hot; 4: log2 size; 15: log2 batch.  NTT: 5 fused levels
# / 5, 6 LDS tile log - 10, 7 passes, 8/9 deepest/shallowest pass, 16-19
# pass depths (sorted, zero padded), 20 levels after the fused ones, 21
# passes deeper than the LDS tile, 22 bit reversal fused into the first
# pass, 23 log2 column run of that pass, 10 zero-extension levels (ntt_ext,
# whose fused first pass skips them).  MSM: 10 window c, 11 precomputed tables,
# 12 log2 task, 13/14/7 fraction of zero / one / small scalars, then the
# derived shape: 16 log2 entries per bucket set, 17 log2 entries per bucket,
# 18 sort passes, 19 bucket sets, 20 log2 tasks per bucket, 21 log2 buckets.
def ntt_features(op, L, plan, batch=1):
    f = np.zeros(N_FEATURES, np.float32)
    if op == "ntt_ext":
        f[OPS.index("ntt_coset")] = 1
        f[10] = EXT_LEVELS
    else:
        f[OPS.index(op)] = 1
    f[4] = L
    fl = fused_levels(plan[0])
    f[5] = fl / 5.0
    f[6] = plan[1] - 10
    d = list(plan[2:])
    f[7] = len(d)
    f[8] = max(d)
    f[9] = min(d)
    dd = sorted(d, reverse=True)[:4]
    f[16:16 + len(dd)] = dd
    f[20] = L - fl
    f[21] = sum(1 for x in d if x > plan[1])
    if plan[0] == 1:
        f[22] = 1
        f[23] = min(plan[1] - d[0], L - d[0])
    f[15] = np.log2(batch)
    return f


def msm_features(logN, cfg, sparsity=(0.0, 0.0, 0.0), batch=1):
    c, target, task = cfg
    W = 254 // c + 1
    nw = W if target == 0 else min(target, W)
    npre = -(-W // nw)
    f = np.zeros(N_FEATURES, np.float32)
    f[3] = 1
    f[4] = logN
    f[10] = c
    f[11] = 1.0 if target == 1 else 0.0
    f[12] = np.log2(task)
    f[13], f[14] = sparsity[0], sparsity[1]
    f[7] = sparsity[2]
    f[16] = logN + np.log2(npre)
    f[17] = f[16] - (c - 1)
    f[18] = -(-c // 8)
    f[19] = nw
    f[20] = f[17] - f[12]
    f[21] = np.log2(nw) + c - 1
    f[15] = np.log2(batch)
    return f


def row_features(row):
    """Features of a timing row, recomputed from its problem and config (so
    rows recorded with an older feature layout stay usable)."""
    parts = row["problem"].split(":")
    if parts[0] == "msm":
        return msm_features(int(parts[1]), row["config"], scalar_sparsity(int(parts[2])))
    return ntt_features(parts[0], int(parts[1]), row["config"])


def scalar_sparsity(kind):
    """(fraction zero, fraction one, fraction small) of the seeded generators."""
    return {0: (0.0, 0.0, 0.0), 1: (0.9, 0.1, 0.0), 2: (0.0, 0.0, 1.0)}[kind]


# ---------------------------------------------------------------- database
class TimingDB:
    def __init__(self, rows=None):
        self.rows = list(rows or [])

    def add(self, problem, config, feats, ms):
        self.rows.append({"problem": problem, "config": config,
                          "features": [float(x) for x in feats], "ms": float(ms)})

    def save(self, path):
        with open(path, "w") as f:
            json.dump(self.rows, f)

    @classmethod
    def load(cls, path):
        if not os.path.exists(path):
            return cls()
        with open(path) as f:
            return cls(json.load(f))

    def problems(self):
        seen = []
        for r in self.rows:
            if r["problem"] not in seen:
                seen.append(r["problem"])
        return seen

    def for_problem(self, problem):
        return [r for r in self.rows if r["problem"] == problem]


# ----------------------------------------------------------------- timing
def _median_ms(fn, stream, reps):
    import torch
    fn()
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    return float(np.median(times))


def tune_ntt(db, L, stream, ops=("ntt_fwd",), reps=3, plans=None):
    """Time every plan of a 2^L transform for each op kind; returns best per op."""
    from . import _lib
    from . import ntt as N
    from .runtime import device_u32
    n = 1 << L
    x = device_u32(n * 8)
    _lib.call("impeller_gen_scalars", x.data_ptr(), n, 0, 1000 + L, 0, stream.raw())
    w = device_u32(8)
    _lib.call("impeller_gen_scalars", w.data_ptr(), 1, 0, 7, 0, stream.raw())
    stream.sync()
    omega = w.cpu().numpy().view(np.uint32)
    tw = N.twiddles(omega, max(1, n // 2), stream)
    best = {}
    for op in ops:
        problem = f"{op}:{L}"
        for plan in (plans or ntt_plans(L)):
            kw = {}
            if op == "ntt_inv":
                kw["inverse"] = True
            if op == "ntt_coset":
                kw["coset_shift"] = w

            if op == "ntt_ext":
                # one batch item: the first 2^(L-3) elements of x as its source
                b = N.NttBatch([dict(data=x, src=x, log_n=L, log_src=L - EXT_LEVELS,
                                     twiddle=tw, coset_shift=w, plan=plan)], stream)

                def run(b=b):
                    b.run(stream)
            else:
                def run(plan=plan, kw=kw):
                    N.ntt(x, tw, L, stream, passes=plan, **kw)

            ms = _median_ms(run, stream, reps)
            db.add(problem, plan, ntt_features(op, L, plan), ms)
            if problem not in best or ms < best[problem][1]:
                best[problem] = (plan, ms)
    return best


def tune_msm(db, logN, bases, stream, kinds=(0,), cands=None, reps=2):
    """Time every MSM configuration on the first 2^logN bases for each scalar kind."""
    import torch
    from . import _lib
    from . import msm as M
    from .runtime import device_u32
    n = 1 << logN
    cands = cands or msm_candidates()
    tables = {}
    best = {}
    scal = {}
    for kind in kinds:
        s = device_u32(n * 8)
        _lib.call("impeller_gen_scalars", s.data_ptr(), n, kind, 500 + kind, 0, stream.raw())
        scal[kind] = s
    for (c, target, task) in cands:
        key = (c, target)
        if key not in tables:
            tables.clear()
            torch.cuda.empty_cache()
            nt = M.n_tables(c, target)
            t = device_u32(nt * n * 16)
            t[: n * 16].copy_(bases[: n * 16])
            M.precompute_device(t, n, c, target, stream)
            tables[key] = t
        t = tables[key]
        ws = device_u32((M.cfg_workspace_size(n, c, target, task) + 3) // 4)
        out = device_u32(16)
        for kind in kinds:
            problem = f"msm:{logN}:{kind}"

            def run(c=c, target=target, task=task, kind=kind):
                M.msm_device_cfg(t, scal[kind], n, c, target, stream, task=task, workspace=ws,
                                 out=out, sync=False)

            ms = _median_ms(run, stream, reps)
            cfg = [c, target, task]
            db.add(problem, cfg, msm_features(logN, cfg, scalar_sparsity(kind)), ms)
            if problem not in best or ms < best[problem][1]:
                best[problem] = (cfg, ms)
        del ws
    return best


# ------------------------------------------------------------ model picks
def fit_model(db, stream, rows=None, steps=1500, seed=0):
    """Cost model on the rows' features; the targets are log-times centred
    per problem, so the network learns how configurations rank within a
    problem rather than how problem sizes compare (a pick only needs the
    ranking)."""
    from .costmodel import CostModel
    rows = rows if rows is not None else db.rows
    feats = np.array([row_features(r) for r in rows], np.float32)
    ms = np.array([r["ms"] for r in rows])
    model = CostModel(N_FEATURES, 64, seed)
    model.fit(feats, ms, stream, steps=steps, groups=[r["problem"] for r in rows])
    return model


def pick(model, rows, stream):
    """The row (configuration) the model predicts fastest among candidates."""
    feats = np.array([row_features(r) for r in rows], np.float32)
    pred = model.predict_ms(feats, stream)
    return rows[int(np.argmin(pred))]


def regret_report(db, stream, steps=1500, holdout=True, log=None):
    """Per problem: measured ms of the model's pick vs the exhaustive best and
    vs the configuration used without the model (default_ntt_plan,
    default_msm).  With holdout, the model for a problem is trained
    without that problem's rows (leave-one-problem-out), i.e. a pick without
    a sweep of that problem."""
    report = {}
    probs = db.problems()
    full = None if holdout else fit_model(db, stream, steps=steps)
    for p in probs:
        rows = db.for_problem(p)
        best = min(rows, key=lambda r: r["ms"])
        model = fit_model(db, stream, [r for r in db.rows if r["problem"] != p], steps) \
            if holdout else full
        choice = pick(model, rows, stream)
        op = p.split(":")[0]
        if op == "msm":
            dcfg = default_msm(int(p.split(":")[1]))
            default = [r for r in rows if tuple(r["config"]) == dcfg]
        else:
            L = int(p.split(":")[1])
            default = [r for r in rows if r["config"] == default_ntt_plan(L)]
        d_ms = default[0]["ms"] if default else None
        report[p] = {
            "best_ms": best["ms"], "best_config": best["config"],
            "model_ms": choice["ms"], "model_config": choice["config"],
            "model_regret": choice["ms"] / best["ms"] - 1.0,
            "default_ms": d_ms,
            "default_regret": (d_ms / best["ms"] - 1.0) if d_ms else None,
        }
        if log:
            log(f"[regret] {p}: model {report[p]['model_regret']:.3f} "
                f"default {report[p]['default_regret'] if d_ms else float('nan'):.3f}")
    return report


# ------------------------------------------------- in-run, held-out tuning
def reduced_ntt_plans(L):
    """The in-run candidate plans of a 2^L transform: 1024-element LDS
    tiles, bit reversal fused into the first pass or four levels fused into
    the bit-reversal kernel, at most three passes, plus the default plan
    (the shape of every exhaustive best of the full sweep at 2^18..2^26,
    apart from one four-pass coset plan)."""
    out = [p for p in ntt_plans(L, max_passes=3) if p[1] == 10 and p[0] in (1, 4)]
    d = default_ntt_plan(L)
    if d not in out:
        out.append(d)
    return out


def reduced_msm_candidates():
    """Precomputed window tables, windows 10..20, tasks 32 / 64."""
    return msm_candidates(windows=range(10, 21), targets=(1,), tasks=(32, 64))


def default_config(problem):
    op, lg = problem.split(":")[0], int(problem.split(":")[1])
    return list(default_msm(lg)) if op == "msm" else default_ntt_plan(lg)


def inrun_tune(stream, ntt_problems, msm_problems, folds=5, steps=1500, reps=2, log=None):
    """Times every reduced candidate of each problem in this run, then picks
    each problem's configuration with a cost model trained (on the GPU) only
    on the other problems' rows: the problems are split into `folds` groups
    and the model for a group never sees that group's timings (a pick
    without a sweep of the problem).  Returns (db, picks, summary); picks[p]
    holds the model's, the in-run exhaustive best's and the default's
    configuration and measured ms."""
    import time

    from . import _lib
    from .runtime import device_u32
    t0 = time.time()
    db = TimingDB()
    by_L = {}
    for op, L in ntt_problems:
        by_L.setdefault(L, []).append(op)
    for L in sorted(by_L):
        tune_ntt(db, L, stream, ops=tuple(sorted(set(by_L[L]))), reps=reps,
                 plans=reduced_ntt_plans(L))
        if log:
            log(f"[inrun] ntt 2^{L} {sorted(set(by_L[L]))}, {time.time() - t0:.1f}s")
    by_lg = {}
    for lg, kind in msm_problems:
        by_lg.setdefault(lg, set()).add(kind)
    if by_lg:
        nb = 1 << max(by_lg)
        bases = device_u32(nb * 16)
        _lib.call("impeller_gen_bases", bases.data_ptr(), nb, 77, 0, stream.raw())
        stream.sync()
        for lg in sorted(by_lg):
            tune_msm(db, lg, bases, stream, kinds=tuple(sorted(by_lg[lg])),
                     cands=reduced_msm_candidates(), reps=reps)
            if log:
                log(f"[inrun] msm 2^{lg}, {time.time() - t0:.1f}s")
        del bases
    t_sweep = time.time() - t0
    probs = db.problems()
    picks = {}
    for f in range(folds):
        held = [p for i, p in enumerate(probs) if i % folds == f]
        if not held:
            continue
        model = fit_model(db, stream, [r for r in db.rows if r["problem"] not in held], steps)
        for p in held:
            rows = db.for_problem(p)
            best = min(rows, key=lambda r: r["ms"])
            choice = pick(model, rows, stream)
            dflt = [r for r in rows if list(r["config"]) == default_config(p)]
            picks[p] = {"model": list(choice["config"]), "model_ms": choice["ms"],
                        "exhaustive": list(best["config"]), "exhaustive_ms": best["ms"],
                        "default": default_config(p),
                        "default_ms": dflt[0]["ms"] if dflt else None}
        if log:
            log(f"[inrun] fold {f}: {len(held)} problems held out, {time.time() - t0:.1f}s")
    regrets = [v["model_ms"] / v["exhaustive_ms"] - 1.0 for v in picks.values()]
    dregrets = [v["default_ms"] / v["exhaustive_ms"] - 1.0 for v in picks.values()
                if v["default_ms"]]
    summary = {
        "model_training": f"in-run, held-out ({folds}-fold by problem)",
        "rows": len(db.rows), "problems": len(probs),
        "sweep_s": round(t_sweep, 1), "fit_and_pick_s": round(time.time() - t0 - t_sweep, 1),
        "model_regret_mean": round(float(np.mean(regrets)), 4),
        "model_regret_max": round(float(np.max(regrets)), 4),
        "model_within_5pct": round(sum(1 for r in regrets if r <= 0.05) / len(regrets), 3),
        "default_regret_mean": round(float(np.mean(dregrets)), 4) if dregrets else None,
        "default_regret_max": round(float(np.max(dregrets)), 4) if dregrets else None,
    }
    return db, picks, summary
