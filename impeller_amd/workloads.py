"""Synthetic proving workloads used by the benchmark and the tests.

batched NTT: 64 polynomials with seeded log-sizes in 18..23 and seeded kinds
  (forward, inverse, coset-extended 2^k -> 2^(k+3) with shift 7), run either
  as 64 sequential single calls or as one batched launch per stage (optionally
  replayed from a HIP graph).
"""
import os
import random

import numpy as np

from . import _lib
from . import ntt as N
from .runtime import Stream, device_u32

R_MOD = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
MONT_R = 1 << 256
ROOT_OF_UNITY = pow(7, (R_MOD - 1) >> 28, R_MOD)
GENERATOR = 7


def mont_limbs(x):
    v = x * MONT_R % R_MOD
    return np.array([(v >> (32 * i)) & 0xFFFFFFFF for i in range(8)], np.uint32)


def omega(k, inverse=False):
    w = pow(ROOT_OF_UNITY, 1 << (28 - k), R_MOD)
    return pow(w, -1, R_MOD) if inverse else w


class TwiddleCache:
    """Device twiddle tables omega_k^i (i < 2^(k-1)) per (k, inverse)."""

    def __init__(self, stream: Stream):
        self.stream = stream
        self.tabs = {}

    def get(self, k, inverse=False):
        key = (k, inverse)
        if key not in self.tabs:
            self.tabs[key] = N.twiddles(mont_limbs(omega(k, inverse)), max(1, (1 << k) // 2),
                                        self.stream)
        return self.tabs[key]


def ntt_batch_spec(count=64, seed=2024, lo=18, hi=23):
    rng = random.Random(seed)
    return [(rng.choice(["fwd", "inv", "ext"]), rng.randint(lo, hi)) for _ in range(count)]


class NttWorkload:
    """Device buffers for a batch spec; inputs are seeded and restorable."""

    def __init__(self, spec, stream: Stream, seed=7, plans=None):
        import torch
        self.spec = spec
        self.stream = stream
        self.tw = TwiddleCache(stream)
        self.shift = torch.from_numpy(mont_limbs(GENERATOR).view(np.int32).copy()).cuda()
        self.ginv = torch.from_numpy(
            mont_limbs(pow(GENERATOR, -1, R_MOD)).view(np.int32).copy()).cuda()
        self.items = []
        self.inputs = []
        for i, (kind, k) in enumerate(spec):
            n_in = 1 << k
            src = device_u32(n_in * 8)
            _lib.call("impeller_gen_scalars", src.data_ptr(), n_in, 0, seed * 1000 + i, 0,
                      stream.raw())
            if kind == "ext":
                L = k + 3
                data = device_u32((1 << L) * 8)
                item = dict(data=data, src=src, log_n=L, log_src=k, twiddle=self.tw.get(L),
                            coset_shift=self.shift)
            else:
                data = src.clone()
                item = dict(data=data, log_n=k, twiddle=self.tw.get(k, kind == "inv"),
                            inverse=kind == "inv")
            if plans:
                p = plans.get((kind, item["log_n"]))
                if p:
                    item["plan"] = p
            self.items.append(item)
            self.inputs.append(src)
        stream.sync()

    def restore(self):
        for (kind, _), it, src in zip(self.spec, self.items, self.inputs):
            if kind != "ext":
                it["data"].copy_(src)

    def run_plain(self):
        """64 sequential single calls (the plain per-call path)."""
        s = self.stream
        for (kind, k), it in zip(self.spec, self.items):
            if kind == "ext":
                n_src = 1 << k
                d = it["data"]
                d[: n_src * 8].copy_(it["src"])
                d[n_src * 8:].zero_()
                N.ntt(d, it["twiddle"], it["log_n"], s, coset_shift=self.shift,
                      passes=it.get("plan"))
            else:
                N.ntt(it["data"], it["twiddle"], k, s, inverse=(kind == "inv"),
                      passes=it.get("plan"))

    def batch(self):
        return N.NttBatch(self.items, self.stream)


# --------------------------------------------------------------- batched MSM
def msm_batch_spec(count=40, seed=2025, lo=16, hi=22):
    """(log size, scalar kind): 50% uniform, 30% selector-like, 20% small."""
    rng = random.Random(seed)
    kinds = [0] * 20 + [1] * 12 + [2] * 8
    rng.shuffle(kinds)
    kinds = (kinds * ((count + 39) // 40))[:count]
    return [(rng.randint(lo, hi), kinds[i]) for i in range(count)]


def heuristic_window(log_n):
    """Default window when no tuned configuration is known."""
    return max(10, min(16, log_n - 6))


class MsmWorkload:
    """40 commitment MSMs sharing one 2^22 base table (seeded k_i * G)."""

    def __init__(self, spec, stream: Stream, log_bases=22, seed=77):
        from . import msm as M
        self.M = M
        self.spec = spec
        self.stream = stream
        self.nb = 1 << log_bases
        self.bases = device_u32(self.nb * 16)
        _lib.call("impeller_gen_bases", self.bases.data_ptr(), self.nb, seed, 0, stream.raw())
        self.scalars = []
        for i, (lg, kind) in enumerate(spec):
            s = device_u32((1 << lg) * 8)
            _lib.call("impeller_gen_scalars", s.data_ptr(), 1 << lg, kind, 9000 + i, 0,
                      stream.raw())
            self.scalars.append(s)
        self.tables = {}
        stream.sync()

    def table(self, c, target):
        if target == 0:
            return self.bases
        key = (c, target)
        if key not in self.tables:
            M = self.M
            t = device_u32(M.n_tables(c, target) * self.nb * 16)
            t[: self.nb * 16].copy_(self.bases)
            M.precompute_device(t, self.nb, c, target, self.stream)
            self.tables[key] = t
        return self.tables[key]

    def items(self, configs):
        """configs[i] = (c, target, task) for MSM i."""
        out = []
        for (lg, _), s, (c, target, task) in zip(self.spec, self.scalars, configs):
            out.append(dict(tables=self.table(c, target), scalars=s, len=1 << lg, window=c,
                            target=target, task=task, table_stride=self.nb if target else 0))
        return out

    def workspaces(self, configs):
        M = self.M
        return [device_u32((M.cfg_workspace_size(1 << lg, c, t, T) + 3) // 4)
                for (lg, _), (c, t, T) in zip(self.spec, configs)]

    def run_plain(self, configs, results, workspaces):
        """Per-MSM single calls (the plain path); results: device tensor 16 x count."""
        for i, it in enumerate(self.items(configs)):
            ws = workspaces[i]
            _lib.call("impeller_msm_cfg", it["tables"].data_ptr(), it["scalars"].data_ptr(),
                      it["len"], it["table_stride"], it["window"], it["target"], it["task"],
                      results.data_ptr() + 64 * i, ws.data_ptr(), ws.numel() * 4,
                      self.stream.raw())


# ----------------------------------------------------- fused proving chunk
class ProvingChunk:
    """One proving chunk at k: INTT of `cols` columns (2^k), coset NTT of each
    to the 2^(k+3) extended domain, the gate a*b*c - d + q*e*(1 - q) on
    columns 0..5 over the extended domain, inverse coset NTT of the gate
    output back to coefficients, and one MSM commitment per 2^k chunk of it
    (8 MSMs against the first 2^k bases)."""

    EXT = 3

    def __init__(self, stream: Stream, k=22, cols=20, seed=5, bases=None):
        import torch
        self.stream = stream
        self.k, self.cols = k, cols
        self.K = k + self.EXT
        n, N = 1 << k, 1 << self.K
        self.tw = TwiddleCache(stream)
        self.shift = torch.from_numpy(mont_limbs(GENERATOR).view(np.int32).copy()).cuda()
        self.ginv = torch.from_numpy(
            mont_limbs(pow(GENERATOR, -1, R_MOD)).view(np.int32).copy()).cuda()
        self.lagrange = []
        for i in range(cols):
            t = device_u32(n * 8)
            _lib.call("impeller_gen_scalars", t.data_ptr(), n, 0, seed * 100 + i, 0,
                      stream.raw())
            self.lagrange.append(t)
        self.coeffs = [device_u32(n * 8) for _ in range(cols)]
        self.ext = [device_u32(N * 8) for _ in range(cols)]
        self.h = device_u32(N * 8)
        self.arg = device_u32(7 * 10)
        if bases is None:
            bases = device_u32(n * 16)
            _lib.call("impeller_gen_bases", bases.data_ptr(), n, 77, 0, stream.raw())
        self.bases = bases
        self.results = device_u32(16 * (1 << self.EXT))
        stream.sync()

    # ------------------------------------------------------ plain schedule
    def run_plain(self, msm_cfg, plans=None, msm_tables=None, ws=None):
        """Sequential single calls: ssip-style NTTs, the gate ABI, single MSMs."""
        from . import msm as M
        from .fused import FusedGate
        from .runtime import ScalarArray
        s = self.stream
        k, K = self.k, self.K
        plans = plans or {}
        for i in range(self.cols):
            self.coeffs[i].copy_(self.lagrange[i])
            N.ntt(self.coeffs[i], self.tw.get(k, True), k, s, inverse=True,
                  passes=plans.get(("inv", k)))
        for i in range(self.cols):
            e = self.ext[i]
            e[: (1 << k) * 8].copy_(self.coeffs[i])
            e[(1 << k) * 8:].zero_()
            N.ntt(e, self.tw.get(K), K, s, coset_shift=self.shift,
                  passes=plans.get(("ext", K)))
        self._gate()
        N.ntt(self.h, self.tw.get(K, True), K, s, inverse=True, coset_shift=self.ginv,
              passes=plans.get(("inv", K)))
        c, target, task = msm_cfg
        tables = msm_tables if msm_tables is not None else self.bases
        for j in range(1 << self.EXT):
            chunk = self.h[j * (1 << k) * 8:(j + 1) * (1 << k) * 8]
            M.msm_device_cfg(tables, chunk, 1 << k, c, target, s, task=task, workspace=ws,
                             out=self.results[16 * j:16 * (j + 1)], sync=False)

    def _gate(self):
        from .fused import FusedGate
        from .runtime import ScalarArray
        N_ = 1 << self.K
        ins = []
        for i in range(6):
            a = ScalarArray(N_, self.ext[i], "cuda")
            ins.append(a)
        out = ScalarArray(N_, self.h, "cuda")
        FusedGate().get_fn()([self.arg, None, out], [self.stream] + ins)

    # ---------------------------------------------------- batched schedule
    def batches(self, plans=None):
        """Three NTT batches (INTT of all columns, coset-extension of all
        columns, inverse coset of the gate output)."""
        plans = plans or {}
        k, K = self.k, self.K
        intt = [dict(data=self.coeffs[i], src=self.lagrange[i], log_n=k, log_src=k,
                     twiddle=self.tw.get(k, True), inverse=True, plan=plans.get(("inv", k)))
                for i in range(self.cols)]
        ext = [dict(data=self.ext[i], src=self.coeffs[i], log_n=K, log_src=k,
                    twiddle=self.tw.get(K), coset_shift=self.shift, plan=plans.get(("ext", K)))
               for i in range(self.cols)]
        back = [dict(data=self.h, log_n=K, twiddle=self.tw.get(K, True), inverse=True,
                     coset_shift=self.ginv, plan=plans.get(("inv", K)))]
        return (N.NttBatch(intt, self.stream), N.NttBatch(ext, self.stream),
                N.NttBatch(back, self.stream))

    def msm_batch(self, msm_cfg, tables=None, table_stride=0):
        from . import msm as M
        c, target, task = msm_cfg
        n = 1 << self.k
        tables = tables if tables is not None else self.bases
        items = [dict(tables=tables, scalars=self.h[j * n * 8:(j + 1) * n * 8], len=n,
                      window=c, target=target, task=task, table_stride=table_stride)
                 for j in range(1 << self.EXT)]
        return M.MsmBatch(items, self.stream)

    def run_batched(self, nb, mb):
        """The whole chunk as batched launches (graph-capturable)."""
        intt, ext, back = nb
        intt.run(self.stream)
        ext.run(self.stream)
        self._gate_batched()
        back.run(self.stream)
        mb.run(self.stream)

    def _gate_batched(self):
        from .fused import gate_device
        gate_device(self.ext[:6], self.h, 1 << self.K, self.stream)


# ----------------------------------------------------- emitted C5 gate
FUSED_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_fused")


def c5_gate_graph():
    """The proving chunk's gate a*b*c - d + q*e*(1 - q) as an ArithGraph
    (Lagrange form, the constant one a scalar input), for the emitter."""
    from . import kernel_gen as KG
    g = KG.ArithGraph(KG.LAGRANGE)
    a, b, c, d, q, e = (g.input(KG.ARRAY) for _ in range(6))
    one = g.input(KG.SCALAR)
    t = g.binop("pp", "mul", g.binop("pp", "mul", a, b), c)
    t = g.binop("pp", "sub", t, d)
    u = g.binop("pp", "mul", g.binop("pp", "mul", q, e), g.binop("sp", "sub", one, q))
    g.output(KG.ARRAY, g.binop("pp", "add", t, u))
    return g


def c5_gate_op():
    from . import kernel_gen as KG
    return KG.FusedOp(c5_gate_graph(), "emitted_gate_abc_d_qe")


def build_fused():
    """Emits and builds the workloads' fused kernels (run beforehand by the
    library build; the bench loads the result)."""
    from . import kernel_gen as KG
    return KG.build([c5_gate_op()], FUSED_DIR)
