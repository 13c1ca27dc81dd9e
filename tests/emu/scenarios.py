"""Kernel scenarios on the CPU-emulated library, checked against the oracle.
The same shapes and entry points as the GPU tests (batched NTT and MSM, graph
replay, single-call NTT/MSM, gate), at sizes the emulation finishes quickly.

  python tests/emu/scenarios.py [--asan] [name ...]
"""
import ctypes
import random
import sys
import time

import numpy as np

from tests.emu.harness import Emu, garbage_u32, ptr, u32
from impeller_amd import _lib
from oracle import oracle as O

r = O.R_MOD
L = O.int_to_limbs
ROOT = pow(7, (r - 1) >> 28, r)


def mont(x):
    return L(O.to_mont_int(x))


def omega(k):
    return pow(ROOT, 1 << (28 - k), r)


def flat(a):
    return np.ascontiguousarray(a, dtype=np.uint32).reshape(-1)


def ntt_items(specs, seed):
    """(items, expects); kinds fwd / inv / ext / fwd_random as in the GPU tests."""
    rng = random.Random(seed)
    items, expects = [], []
    for i, (kind, k) in enumerate(specs):
        n = 1 << k
        if kind == "ext":
            ks = k - 3
            coeffs = O.gen_scalars(1 << ks, seed * 100 + i)
            w, g = mont(omega(k)), mont(7)
            ext = np.zeros((n, 8), np.uint32)
            ext[: 1 << ks] = coeffs
            expects.append(O.best_fft(O.distribute_powers(ext, O.powers(g, n)), w, k))
            items.append(dict(data=garbage_u32(n * 8), src=flat(coeffs), log_n=k, log_src=ks,
                              twiddle=flat(O.powers(w, n // 2)), coset_shift=flat(g)))
        elif kind == "inv":
            x = O.gen_scalars(n, seed * 100 + i)
            winv = mont(pow(omega(k), -1, r))
            expects.append(O.vec_scale(O.best_fft(x, winv, k), mont(pow(n, -1, r))))
            items.append(dict(data=flat(x).copy(), log_n=k, inverse=True,
                              twiddle=flat(O.powers(winv, max(1, n // 2)))))
        else:
            x = O.gen_scalars(n, seed * 100 + i)
            w = mont(rng.randrange(r)) if kind == "fwd_random" else mont(omega(k))
            expects.append(O.best_fft(x, w, k))
            items.append(dict(data=flat(x).copy(), log_n=k,
                              twiddle=flat(O.powers(w, max(1, n // 2)))))
    return items, expects


class NttBatch:
    def __init__(self, E, items, second_buffers=True):
        self.E = E
        n = len(items)
        arr = (_lib.ImpNttItem * n)()
        for i, it in enumerate(items):
            a = arr[i]
            a.data = ptr(it["data"])
            a.src = ptr(it.get("src"))
            a.log_n = it["log_n"]
            a.log_src = it.get("log_src", it["log_n"])
            a.twiddle = ptr(it["twiddle"])
            a.inverse = 1 if it.get("inverse") else 0
            a.coset_shift = ptr(it.get("coset_shift"))
            plan = it.get("plan") or []
            a.plan_len = len(plan)
            for k, v in enumerate(plan):
                a.plan[k] = v
        # second_buffers=False: the minimal buffer, fused-reversal items then
        # run in place after the bit-reversal stage
        nbytes = int(E.call("impeller_ntt_batch_bytes_items", arr, n) if second_buffers
                     else E.call("impeller_ntt_batch_bytes", n))
        self.dev = garbage_u32((nbytes + 3) // 4)
        self.info = _lib.ImpNttBatch()
        E.call("impeller_ntt_batch_prepare", arr, n, ptr(self.dev), nbytes,
               ctypes.byref(self.info), None)

    def run(self, max_blocks=0):
        self.E.call("impeller_ntt_batch_run", ctypes.byref(self.info), max_blocks, None)


class Graph:
    def __init__(self, E, fn):
        self.E = E
        E.call("impeller_graph_begin", None)
        fn()
        h = ctypes.c_void_p()
        E.call("impeller_graph_end", None, ctypes.byref(h))
        self.h = h

    def launch(self):
        self.E.call("impeller_graph_launch", self.h, None)


def check_items(items, expects, tag):
    for it, e in zip(items, expects):
        got = it["data"].reshape(-1, 8)
        assert (got == e).all(), f"{tag}: log_n {it['log_n']}"


SPECS = [("fwd", 12), ("inv", 14), ("ext", 13), ("fwd_random", 10), ("fwd", 3), ("inv", 9),
         ("ext", 15), ("fwd_random", 16), ("fwd", 1)]


def batched_ntt_mixed(E):
    for max_blocks in (0, 7, 256):
        items, expects = ntt_items(SPECS, 11 + max_blocks)
        b = NttBatch(E, items)
        b.run(max_blocks)
        check_items(items, expects, f"mixed max_blocks={max_blocks}")


def batched_ntt_plans(E):
    for second in (True, False):
        items, expects = ntt_items([("fwd_random", 14), ("inv", 12), ("ext", 14), ("fwd", 9),
                                    ("ext", 13), ("inv", 11), ("ext", 12)], 5)
        items[0]["plan"] = [5, 11, 9]
        items[1]["plan"] = [0, 10, 2, 3, 7]
        items[2]["plan"] = [5, 10, 4, 5]
        items[3]["plan"] = [1, 10, 4, 5]
        items[4]["plan"] = [1, 10, 6, 7]  # coset extension with the reversal fused
        items[5]["plan"] = [1, 11, 2, 9]
        items[6]["plan"] = [1, 10, 2, 10]  # zero-extension levels beyond the first pass
        NttBatch(E, items, second_buffers=second).run(7 if second else 0)
        check_items(items, expects, f"plans second_buffers={second}")


def ntt_graph_replay(E):
    items, expects = ntt_items([("fwd", 12), ("ext", 13), ("inv", 11)], 9)
    restore = [(it["data"], it["data"].copy()) for it in items if it.get("src") is None]
    b = NttBatch(E, items)
    g = Graph(E, lambda: b.run())
    for rep in range(3):
        for dst, src in restore:
            dst[:] = src
        g.launch()
        check_items(items, expects, f"graph rep {rep}")


def single_ntt(E):
    for k in (1, 5, 8, 9, 10, 13, 16):
        n = 1 << k
        x = O.gen_scalars(n, 300 + k)
        w = mont(omega(k))
        tw = flat(O.powers(w, max(1, n // 2)))
        for inverse, shift in ((0, None), (0, flat(mont(7))), (1, None), (1, flat(mont(5)))):
            d = flat(x).copy()
            E.call("impeller_ntt", ptr(d), ptr(tw), k, inverse, ptr(shift), None, 0, None)
            if inverse:
                e = O.best_fft(x, w, k)
                e = O.vec_scale(e, mont(pow(n, -1, r)))
                if shift is not None:
                    e = O.vec_mul(e, O.powers(mont(5), n))
            else:
                xin = x if shift is None else O.vec_mul(x, O.powers(mont(7), n))
                e = O.best_fft(xin, w, k)
            assert (d.reshape(-1, 8) == e).all(), (k, inverse, shift is not None)


def ntt_fused_reversal(E):
    """Forward, coset, inverse and inverse-coset transforms with plans whose
    first pass reads the source bit-reversed, against best_fft."""
    for k, plan in ((2, [1, 10, 1, 1]), (5, [1, 10, 3, 2]), (9, [1, 11, 5, 4]),
                    (13, [1, 10, 7, 6]), (14, [1, 10, 4, 5, 5])):
        n = 1 << k
        x = O.gen_scalars(n, 400 + k)
        w = mont(omega(k))
        tw = flat(O.powers(w, max(1, n // 2)))
        arr = (ctypes.c_uint * len(plan))(*plan)
        for inverse, shift in ((0, None), (0, flat(mont(7))), (1, None), (1, flat(mont(5)))):
            d = flat(x).copy()
            E.call("impeller_ntt", ptr(d), ptr(tw), k, inverse, ptr(shift),
                   ctypes.cast(arr, ctypes.c_void_p), len(plan), None)
            if inverse:
                e = O.best_fft(x, w, k)
                e = O.vec_scale(e, mont(pow(n, -1, r)))
                if shift is not None:
                    e = O.vec_mul(e, O.powers(mont(5), n))
            else:
                xin = x if shift is None else O.vec_mul(x, O.powers(mont(7), n))
                e = O.best_fft(xin, w, k)
            assert (d.reshape(-1, 8) == e).all(), (k, plan, inverse, shift is not None)


PLANS = [[0, 10, 4, 4, 4, 4], [0, 10, 8, 8], [5, 10, 3, 8], [0, 11, 10, 6], [0, 10] + [2] * 8,
         [5, 11, 11], [5, 10, 1, 1, 9], [5, 11, 5, 6], [5, 10] + [1] * 3 + [2] * 4,
         [4, 10, 6, 6], [4, 11, 11, 1], [4, 10, 2, 5, 5],
         # bit reversal fused into the first pass (column runs of 2^(lds_log - d0))
         [1, 10, 8, 8], [1, 10, 7, 9], [1, 11, 6, 10], [1, 10, 3, 3, 10], [1, 10, 10, 6],
         [1, 11, 11, 5], [1, 10, 1, 5, 10], [1, 10, 2, 2, 2, 2, 2, 2, 2, 2]]


def ntt_plans(E):
    k = 16
    n = 1 << k
    x = O.gen_scalars(n, 7000)
    w = mont(random.Random(70).randrange(r))
    expect = O.best_fft(x, w, k)
    tw = flat(O.powers(w, n // 2))
    for plan in PLANS:
        d = flat(x).copy()
        arr = (ctypes.c_uint * len(plan))(*plan)
        E.call("impeller_ntt", ptr(d), ptr(tw), k, 0, None, ctypes.cast(arr, ctypes.c_void_p),
               len(plan), None)
        assert (d.reshape(-1, 8) == expect).all(), plan


def msm_items(E, nb, seed=91):
    bases = u32(nb * 16)
    E.call("impeller_gen_bases", ptr(bases), nb, seed, 0, None)
    k = O.gen_scalars(nb, seed)
    tabs = {}
    for c, target in ((13, 1), (11, 3)):
        nt = int(E.call("impeller_msm_tables", c, target))
        t = u32(nt * nb * 16)
        t[: nb * 16] = bases
        E.call("impeller_msm_precompute", ptr(t), nb, c, target, None)
        tabs[(c, target)] = t
    return bases, k, tabs


class MsmBatch:
    def __init__(self, E, items):
        self.E = E
        n = len(items)
        self.results = u32(16 * n)
        arr = (_lib.ImpMsmItem * n)()
        for i, it in enumerate(items):
            a = arr[i]
            a.tables = ptr(it["tables"])
            a.scalars = ptr(it["scalars"])
            a.len = it["len"]
            a.table_stride = it.get("table_stride", 0)
            a.window = it["window"]
            a.target_windows = it.get("target", 0)
            a.task = it.get("task", 64)
            a.result = self.results.ctypes.data + 64 * i
        nbytes = int(E.call("impeller_msm_batch_bytes", arr, n))
        assert nbytes > 0
        self.dev = garbage_u32((nbytes + 3) // 4)
        self.info = _lib.ImpMsmBatch()
        E.call("impeller_msm_batch_prepare", arr, n, ptr(self.dev), nbytes,
               ctypes.byref(self.info), None)

    def run(self, max_blocks=0):
        self.E.call("impeller_msm_batch_run", ctypes.byref(self.info), max_blocks, None)


def batched_msm_shared_tables(E, nb=1 << 12):
    bases, k, tabs = msm_items(E, nb)
    items, expects = [], []
    rng = random.Random(3)
    for i in range(9):
        n = rng.choice([1, 37, 1 << 8, 1 << 10, nb])
        kind = i % 3
        s = O.gen_scalars(n, 700 + i, kind=kind)
        c, target = [(13, 1), (11, 3), (12, 0)][i % 3]
        t = bases if target == 0 else tabs[(c, target)]
        items.append(dict(tables=t, scalars=flat(s), len=n, window=c, target=target,
                          table_stride=nb if target else 0, task=rng.choice([16, 64])))
        expects.append(O.msm_closed_form(s, k[:n]))
    b = MsmBatch(E, items)
    for max_blocks in (0, 64):
        b.results[:] = 0
        b.run(max_blocks)
        got = b.results.reshape(-1, 16)
        for i, e in enumerate(expects):
            assert (got[i] == e).all(), (i, items[i]["len"], max_blocks)


def degenerate_msm_items(bases, tabs, k, nb):
    """Batch items whose sums are the identity or trivial: an empty MSM, all
    zero scalars, r - 1 and 1 on one repeated base, beside a normal item."""
    zeros = np.zeros((300, 8), np.uint32)
    cancel = np.zeros((64, 8), np.uint32)
    cancel[0] = O.int_to_limbs(O.to_mont_int(O.R_MOD - 1))
    cancel[1] = O.one()
    bb = np.ascontiguousarray(bases[: 64 * 16]).reshape(64, 16).copy()
    bb[1] = bb[0]
    normal = O.gen_scalars(257, 808)
    specs = [  # (scalars, bases, tables, len, c, target)
        (np.zeros((1, 8), np.uint32), None, None, 0, 12, 0),
        (zeros, None, (13, 1), 300, 13, 1),
        (cancel, bb, None, 64, 10, 0),
        (normal, None, (11, 3), 257, 11, 3),
    ]
    items, expects = [], []
    for s, b, tab, n, c, target in specs:
        if b is not None:
            t = flat(b)
            expects.append(O.best_multiexp(s, b))
        else:
            t = bases if tab is None else tabs[tab]
            expects.append(O.msm_closed_form(s[:n], k[:n]) if n else np.zeros(16, np.uint32))
        items.append(dict(tables=t, scalars=flat(s), len=n, window=c, target=target,
                          table_stride=nb if target else 0))
    return items, expects


def batched_msm_degenerate(E, nb=1 << 10):
    bases, k, tabs = msm_items(E, nb)
    items, expects = degenerate_msm_items(bases, tabs, k, nb)
    assert not expects[0].any() and not expects[1].any() and not expects[2].any()
    b = MsmBatch(E, items)
    for max_blocks in (0, 16):
        b.results[:] = 0xFFFFFFFF
        b.run(max_blocks)
        got = b.results.reshape(-1, 16)
        for i, e in enumerate(expects):
            assert (got[i] == e).all(), (i, items[i]["len"], max_blocks)


def msm_graph_replay(E):
    n = 1 << 9
    bases = u32(n * 16)
    E.call("impeller_gen_bases", ptr(bases), n, 5, 0, None)
    k = O.gen_scalars(n, 5)
    ss = [O.gen_scalars(n, 40 + i) for i in range(3)]
    items = [dict(tables=bases, scalars=flat(s), len=n, window=w, target=0)
             for s, w in zip(ss, (10, 12, 14))]
    b = MsmBatch(E, items)
    g = Graph(E, lambda: b.run())
    for rep in range(2):
        b.results[:] = 0
        g.launch()
        got = b.results.reshape(-1, 16)
        for i, s in enumerate(ss):
            assert (got[i] == O.msm_closed_form(s, k)).all(), (rep, i)


def single_msm(E):
    # (16384, 8, 0): 32 windows of 4 sort tiles each
    for n, c, target, task, kind in ((1, 8, 0, 64, 0), (7, 13, 0, 64, 0), (1000, 10, 1, 16, 1),
                                     (4096, 16, 0, 64, 2), (4096, 12, 2, 32, 0),
                                     (1024, 17, 1, 32, 0), (512, 20, 1, 32, 2),
                                     (16384, 8, 0, 64, 0)):
        bases = u32(n * 16)
        E.call("impeller_gen_bases", ptr(bases), n, 77, 0, None)
        nt = int(E.call("impeller_msm_tables", c, target))
        tables = u32(nt * n * 16)
        tables[: n * 16] = bases
        E.call("impeller_msm_precompute", ptr(tables), n, c, target, None)
        s = O.gen_scalars(n, 12 + n, kind=kind)
        ws_bytes = int(E.call("impeller_msm_cfg_workspace_size", n, c, target, task))
        ws = garbage_u32((ws_bytes + 3) // 4)
        out = garbage_u32(16)
        E.call("impeller_msm_cfg", ptr(tables), ptr(flat(s)), n, 0, c, target, task, ptr(out),
               ptr(ws), ws_bytes, None)
        assert (out == O.msm_closed_form(s, O.gen_scalars(n, 77))).all(), (n, c, target, task)


def msm_hot_bucket(E):
    """Thousands of entries in one bucket: multi-level combine trees (G = 16)."""
    n = 4096
    bases = u32(n * 16)
    E.call("impeller_gen_bases", ptr(bases), n, 77, 0, None)
    k = O.gen_scalars(n, 77)
    one = O.ints_to_array([O.to_mont_int(1)])[0]
    three = O.ints_to_array([O.to_mont_int(3)])[0]
    cases = []
    s1 = np.tile(one, (n, 1))                                  # every entry in bucket 1
    s2 = np.tile(one, (n, 1)); s2[::3] = three                 # two hot buckets
    s3 = O.gen_scalars(n, 5, kind=1)                           # selector-like
    for s, c, task in ((s1, 8, 4), (s2, 10, 1), (s3, 12, 2), (s1, 16, 7)):
        ws_bytes = int(E.call("impeller_msm_cfg_workspace_size", n, c, 0, task))
        ws = garbage_u32((ws_bytes + 3) // 4)
        out = garbage_u32(16)
        table = bases.copy()                                  # one table: the bases, converted
        E.call("impeller_msm_precompute", ptr(table), n, c, 0, None)
        E.call("impeller_msm_cfg", ptr(table), ptr(flat(s)), n, 0, c, 0, task, ptr(out), ptr(ws),
               ws_bytes, None)
        assert (out == O.msm_closed_form(s, k)).all(), (c, task)


def field_inverse(E):
    """Inversion in Fr and Fq (binary extended Euclid) against the oracle,
    including 0, 1, m - 1 and powers of two, then affine conversion."""
    for f, m in ((O.FR, O.R_MOD), (O.FQ, O.Q_MOD)):
        vals = [0, 1, 2, m - 1, m - 2, (m + 1) // 2, 1 << 200, (1 << 253) % m]
        rng = random.Random(5 + f)
        vals += [rng.randrange(m) for _ in range(56)]
        # every bit length, long runs of trailing zeros, values next to m
        vals += [1 << k for k in range(254)] + [(1 << k) - 1 for k in range(1, 254)]
        vals += [m - (1 << k) for k in range(0, 253, 5)]
        vals += [rng.randrange(1, 1 << 40) << rng.randrange(0, 200) for _ in range(300)]
        vals += [rng.randrange(1 << rng.randrange(1, 255)) % m for _ in range(300)]
        a = O.ints_to_array([v % m for v in vals])
        out = u32(a.size)
        E.call("impeller_field_op", f, 7, ptr(out), ptr(flat(a)), None, len(vals), None)
        exp = np.stack([O.inv(a[i], f) for i in range(len(vals))])
        assert (out.reshape(-1, 8) == exp).all(), f


def _g1(E, op, a, b, cols, n):
    out = u32(n * cols)
    E.call("impeller_g1_op", op, ptr(out), ptr(flat(a)), ptr(flat(b)) if b is not None else None,
           n, None)
    return out.reshape(n, cols)


def _same_xyzz(got, exp):
    """Bit-identical XYZZ rows; identity rows (zz == 0) only need zz == 0,
    since the other coordinates of an identity carry no value."""
    for i in range(len(exp)):
        if not exp[i][16:24].any():
            assert not got[i][16:24].any(), i
        else:
            assert (got[i] == exp[i]).all(), i


def arith29(E):
    """The 9 x 29-bit Montgomery forms (product, padded subtractions, XYZZ
    formulas, table format) agree bit for bit with the 8 x 32-bit forms and
    the oracle, including the doubling and identity branches."""
    for f, m in ((O.FR, O.R_MOD), (O.FQ, O.Q_MOD)):
        rng = random.Random(29 + f)
        vals = [0, 1, m - 1, m - 2, (1 << 253) % m] + [rng.randrange(m) for _ in range(59)]
        a = O.ints_to_array([O.to_mont_int(v) if f == O.FR else v for v in vals])
        b = np.roll(a, 3, axis=0).copy()
        n = len(vals)
        for op, ref in ((8, lambda x, y: O.mul(x, y, f)), (9, lambda x, y: O.sub(x, y, f)),
                        (10, lambda x, y: O.sub(O.sub(O.sub(x, y, f), y, f), y, f)),
                        (11, lambda x, y: O.mul(x, x, f))):
            out = u32(n * 8)
            E.call("impeller_field_op", f, op, ptr(out), ptr(flat(a)), ptr(flat(b)), n, None)
            exp = np.stack([ref(a[i], b[i]) for i in range(n)])
            assert (out.reshape(n, 8) == exp).all(), (f, op)
    n = 48
    bases = O.gen_bases(n + 1, 77)
    x = np.stack([O.xyzz_op("orc_affine_to_xyzz", bases[i]) for i in range(n)])
    x = np.stack([O.xyzz_op("orc_xyzz_dbl", x[i]) for i in range(n)])
    x[5] = 0
    x[5][24:] = O.one(O.FQ)                                       # identity
    q = bases[1:n + 1].copy()
    q[11] = 0                                                     # affine identity
    q[12] = O.xyzz_op("orc_xyzz_to_affine", x[12], out_len=16)    # same point: doubling
    q[13] = O.xyzz_op("orc_xyzz_to_affine", x[13], out_len=16)
    q[13][8:] = O.neg(q[13][8:], O.FQ)                            # opposite point: identity
    _same_xyzz(_g1(E, 5, x, q, 32, n), _g1(E, 0, x, q, 32, n))
    y = np.roll(x, 1, axis=0).copy()
    y[20] = x[20]                                                 # doubling branch of add
    y[21] = x[21]
    y[21][8:16] = O.neg(x[21][8:16], O.FQ)                        # P + (-P)
    _same_xyzz(_g1(E, 6, x, y, 32, n), _g1(E, 1, x, y, 32, n))
    _same_xyzz(_g1(E, 7, x, None, 32, n), _g1(E, 2, x, None, 32, n))
    # the group-of-four forms: the same coordinates as the one-lane forms
    _same_xyzz(_g1(E, 9, x, y, 32, n), _g1(E, 1, x, y, 32, n))
    _same_xyzz(_g1(E, 10, x, None, 32, n), _g1(E, 2, x, None, 32, n))
    neg = bases[:n].copy()
    neg[:, 8:] = np.stack([O.neg(bases[i][8:], O.FQ) for i in range(n)])
    neg[0] = 0
    src = bases[:n].copy()
    src[0] = 0
    assert (_g1(E, 8, src, None, 16, n) == neg).all()


def mlp_grad(E):
    """Cost-model MLP forward and gradient kernels against the float64
    restatement, at a batch larger than one work-group's stride."""
    from oracle import mlp_ref
    rng = np.random.default_rng(7)
    F, H = 9, 64
    for B in (37, 1000):
        n = int(E.call("impeller_mlp_param_count", F, H))
        p = rng.normal(0, 0.3, n).astype(np.float32)
        X = rng.normal(0, 1, (B, F)).astype(np.float32)
        T = rng.normal(0, 1, B).astype(np.float32)
        ws = np.zeros(int(E.call("impeller_mlp_workspace_floats", F, H, B)), np.float32)
        Y = np.zeros(B, np.float32)
        E.call("impeller_mlp_forward", ptr(p), ptr(X), B, F, H, ptr(Y), ptr(ws), None)
        Yref, _ = mlp_ref.forward(p.astype(np.float64), X, F, H)
        assert np.allclose(Y, Yref, rtol=1e-4, atol=1e-5), B
        g = np.zeros(n, np.float32)
        loss = np.zeros(1, np.float32)
        E.call("impeller_mlp_grad", ptr(p), ptr(X), ptr(T), B, F, H, ptr(g), ptr(loss), ptr(ws),
               None)
        gref = mlp_ref.grad(p.astype(np.float64), X, T, F, H)
        assert np.allclose(g, gref, rtol=1e-3, atol=1e-5), B
        assert np.isclose(loss[0] / B, mlp_ref.loss(p.astype(np.float64), X, T, F, H),
                          rtol=1e-4), B


def _view(arr, n, rotate=0, offset=0, whole=None):
    return _lib.PolyPtr(arr.ctypes.data, n, rotate, offset, whole if whole is not None else n)


def _logical(phys, n, rotate=0, offset=0, whole=None):
    wl = whole if whole is not None else n
    off = (n - rotate) if rotate else offset
    return (off + np.arange(n)) % wl


def ref_ntt_abi(E):
    """The reference-ABI transforms: ssip_ntt (in place, on rotated views) and
    recompute_ntt (twiddles from pq / omegas, every pq_deg the compiler can
    pass), for true roots and a random omega, against best_fft."""
    rng = random.Random(77)
    for k in (1, 2, 5, 9, 13, 16):
        n = 1 << k
        x = O.gen_scalars(n, 900 + k)
        for w in (mont(omega(k)), mont(rng.randrange(r))):
            expect = O.best_fft(x, w, k)
            tw = flat(O.powers(w, max(1, n // 2)))
            rot = (3 * k) % n
            d = flat(np.roll(x, -rot, axis=0)).copy()  # logical x under rotate = rot
            E.call("ssip_ntt", _view(d, n, rotate=rot), ptr(tw), k, None, 8, 8)
            got = d.reshape(-1, 8)[_logical(d, n, rotate=rot)]
            assert (got == expect).all(), ("ssip_ntt", k)
            for pq_deg in (0, 1, 3, min(k, 8), min(k, 9)):
                npq = max(1, (1 << pq_deg) >> 1)
                pq = garbage_u32(npq * 8)
                om = garbage_u32(32 * 8)
                E.L.gen_pq_omegas(ptr(pq), ptr(om), pq_deg, n, ptr(flat(w)))
                d = flat(x).copy()
                E.call("recompute_ntt", _view(d, n), ptr(pq) if pq_deg else None, pq_deg,
                       ptr(om), k, None, 8, 8)
                assert (d.reshape(-1, 8) == expect).all(), ("recompute_ntt", k, pq_deg)
    # caller-owned scratch: fused-reversal plan with and without room for its
    # second buffer, coset and inverse factors
    k, plan = 13, [1, 10, 7, 6]
    n = 1 << k
    x = O.gen_scalars(n, 950)
    w = mont(omega(k))
    tw = flat(O.powers(w, n // 2))
    arr = (ctypes.c_uint * len(plan))(*plan)
    pl = ctypes.cast(arr, ctypes.c_void_p)
    full = int(E.call("impeller_ntt_workspace_size", k, pl, len(plan)))
    part = int(E.call("impeller_ntt_workspace_size", k, None, 0))
    assert full > part > 0
    for bytes_ in (full, part):
        ws = garbage_u32((bytes_ + 3) // 4)
        for inverse, shift in ((0, None), (0, flat(mont(7))), (1, None), (1, flat(mont(5)))):
            d = flat(x).copy()
            E.call("impeller_ntt_ws", ptr(d), ptr(tw), k, inverse, ptr(shift), pl, len(plan),
                   ptr(ws), bytes_, None)
            if inverse:
                e = O.vec_scale(O.best_fft(x, w, k), mont(pow(n, -1, r)))
                if shift is not None:
                    e = O.vec_mul(e, O.powers(mont(5), n))
            else:
                e = O.best_fft(x if shift is None else O.vec_mul(x, O.powers(mont(7), n)), w, k)
            assert (d.reshape(-1, 8) == e).all(), (bytes_, inverse, shift is not None)


def ntt_noalloc(E):
    """impeller_ntt_noalloc (factor tables in the library's static scratch)
    against impeller_ntt, two calls in a row, the second smaller."""
    for k, inverse, shift, plan in ((13, 0, 7, [1, 10, 7, 6]), (9, 1, 5, None), (6, 1, None, None),
                                    (11, 0, None, [4, 10, 7])):
        n = 1 << k
        x = O.gen_scalars(n, 960 + k)
        w = mont(omega(k))
        tw = flat(O.powers(w, n // 2))
        sh = flat(mont(shift)) if shift else None
        arr = (ctypes.c_uint * len(plan))(*plan) if plan else None
        pl = ctypes.cast(arr, ctypes.c_void_p) if plan else None
        a, b = flat(x).copy(), flat(x).copy()
        E.call("impeller_ntt", ptr(a), ptr(tw), k, inverse, ptr(sh), pl, len(plan or []), None)
        E.call("impeller_ntt_noalloc", ptr(b), ptr(tw), k, inverse, ptr(sh), pl,
               len(plan or []), None)
        assert (a == b).all(), (k, inverse, shift, plan)


def gate_ref_abi(E):
    """fused_gate_abc_d_qe through the generated-kernel convention: the views
    (rotated, sliced) sit in a device array that the kernel reads."""
    n = 1000
    ins = [O.gen_scalars(n, 60 + i) for i in range(6)]
    whole = 1500
    views, bufs = [], []
    for i, a in enumerate(ins):
        if i % 3 == 0:
            rot = 17 * (i + 1)
            b = flat(np.roll(a, -rot, axis=0)).copy()
            views.append(_view(b, n, rotate=rot))
        elif i % 3 == 1:
            off = 1300  # wraps around the end of the backing array
            b = garbage_u32(whole * 8).reshape(-1, 8)
            b[(off + np.arange(n)) % whole] = a
            b = flat(b)
            views.append(_view(b, n, offset=off, whole=whole))
        else:
            b = flat(a).copy()
            views.append(_view(b, n))
        bufs.append(b)
    out = garbage_u32(n * 8)
    orot = 5
    arr = (_lib.PolyPtr * 6)(*views)
    marr = (_lib.PolyPtr * 1)(_view(out, n, rotate=orot))
    sl = _lib.PolySlice(3, n, n)
    E.call("fused_gate_abc_d_qe", ctypes.addressof(arr), ctypes.addressof(marr), n,
           ctypes.byref(sl), None, None)
    got = out.reshape(-1, 8)[_logical(out, n, rotate=orot)]
    assert (got == O.gate(*ins)).all()


def msm_host_pipeline(E):
    """The reference's host-buffer msm_cfg with parts, point / scalar stages
    and batches per run (one table and precomputed tables), each batch
    against the closed form."""
    n = 600
    kk = O.gen_scalars(n, 77)
    bases = u32(n * 16)
    E.call("impeller_gen_bases", ptr(bases), n, 77, 0, None)
    cards = (ctypes.c_uint * 1)(0)
    for c, target, parts, sp, ss, bpr, batches in ((8, 0, 1, 1, 1, 1, 2), (8, 0, 3, 2, 2, 2, 3),
                                                  (10, 2, 3, 2, 3, 2, 2), (9, 1, 4, 3, 1, 0, 2)):
        nt = int(E.call("impeller_msm_tables", c, target))
        tabs = [bases.copy()]
        if nt > 1:
            t = u32(nt * n * 16)
            t[: n * 16] = bases
            E.call("msm_precompute_cfg", c, target, n,
                   (ctypes.c_void_p * nt)(*[t[i * n * 16:].ctypes.data for i in range(nt)]), 1)
            tabs = [t[i * n * 16:(i + 1) * n * 16].copy() for i in range(nt)]
        sizes = (ctypes.c_ulong * 1)()
        E.call("msm_cfg", c, target, None, sizes, n, bpr, parts, ss, sp, 1, cards, None, 0,
               None, None)
        buf = garbage_u32((sizes[0] + 3) // 4)
        scal = [flat(O.gen_scalars(n, 700 + b, kind=b % 3)) for b in range(batches)]
        res = [garbage_u32(16) for _ in range(batches)]
        P = lambda xs: (ctypes.c_void_p * len(xs))(*[x.ctypes.data for x in xs])  # noqa: E731
        E.call("msm_cfg", c, target, P([buf]), None, n, bpr, parts, ss, sp, 1, cards, P(tabs),
               batches, P(scal), P(res))
        for b in range(batches):
            exp = O.msm_closed_form(scal[b].reshape(-1, 8), kk)
            assert (res[b] == exp).all(), (c, target, parts, sp, ss, bpr, b)


def gate(E):
    for n in (1, 777, 4096):
        ins = [O.gen_scalars(n, 40 + i) for i in range(6)]
        arrs = [flat(a) for a in ins]
        ptrs = (ctypes.c_void_p * 6)(*[a.ctypes.data for a in arrs])
        out = u32(n * 8)
        E.call("impeller_gate", ptrs, ptr(out), n, None)
        assert (out.reshape(-1, 8) == O.gate(*ins)).all(), n


def _temp(E, name, *args):
    size = ctypes.c_ulong(0)
    E.call(name, None, ctypes.byref(size), *args)
    return garbage_u32((size.value + 3) // 4 + 4)


def _placed(x, rotate=0):
    """Physical buffer holding logical x under a rotation (iter.cuh:309-317)."""
    n = x.shape[0]
    phys = np.zeros_like(x)
    phys[_logical(None, n, rotate)] = x
    return flat(phys)


POLY_SIZES = (1, 2, 9, 2048, 2049, 5000)


def poly_scans(E):
    """scan_mul, batched_invert and poly_eval on rotated views, one and two
    levels of tiles, against the oracle."""
    for n in POLY_SIZES:
        x = O.gen_scalars(n, 700 + n)
        rot = 3 % n
        # scan_mul in place
        d = _placed(x, rot)
        tmp = _temp(E, "scan_mul", _view(d, n), None)
        E.call("scan_mul", ptr(tmp), None, _view(d, n, rot), None)
        got = d.reshape(-1, 8)[_logical(None, n, rot)]
        assert (got == O.scan_mul(x)).all(), ("scan_mul", n)
        # batched_invert with zeros (core/tests/poly.rs:434-440)
        z = x.copy()
        z[::17] = 0
        d = _placed(z, rot)
        tmp = _temp(E, "batched_invert", _view(d, n), None)
        E.call("batched_invert", ptr(tmp), None, _view(d, n, rot), None)
        got = d.reshape(-1, 8)[_logical(None, n, rot)]
        assert (got == O.batch_invert(z)).all(), ("batched_invert", n)
        # poly_eval
        pt = O.gen_scalars(1, 900 + n)[0]
        d = _placed(x, rot)
        res = garbage_u32(8)
        tmp = _temp(E, "poly_eval", _view(d, n), None, None, None)
        E.call("poly_eval", ptr(tmp), None, _view(d, n, rot), ptr(res), ptr(flat(pt)), None)
        assert (res == O.poly_eval(x, pt)).all(), ("poly_eval", n)
    # all zeros and all ones for the inverse
    for n in (5, 3000):
        for fill in ("zero", "one"):
            z = np.zeros((n, 8), np.uint32) if fill == "zero" else np.tile(O.one(), (n, 1))
            d = flat(z).copy()
            tmp = _temp(E, "batched_invert", _view(d, n), None)
            E.call("batched_invert", ptr(tmp), None, _view(d, n), None)
            assert (d.reshape(-1, 8) == z).all(), ("batched_invert", fill, n)


def poly_kate(E):
    """kate_division on rotated views (core/tests/poly.rs:170-246 rotates p by
    2 and q by -2) against the oracle."""
    for k in (0, 1, 3, 11, 12, 13):
        n = 1 << k
        p = O.gen_scalars(n, 800 + k)
        b = O.gen_scalars(1, 850 + k)[0]
        rp, rq = 2 % n, (n - 2) % n
        dp = _placed(p, rp)
        dq = garbage_u32(n * 8)
        tmp = _temp(E, "kate_division", k, _view(dp, n), None, _view(dq, n), None)
        E.call("kate_division", ptr(tmp), None, k, _view(dp, n, rp), ptr(flat(b)),
               _view(dq, n, rq), None)
        got = dq.reshape(-1, 8)[_logical(None, n, rq)]
        assert (got == O.kate_division(p, b)).all(), ("kate", k)
    # a buffer smaller than the query is rejected when its size is given
    n = 4096
    size = ctypes.c_ulong(0)
    E.call("kate_division", None, ctypes.byref(size), 12, _view(u32(8), n), None,
           _view(u32(8), n), None)
    small = ctypes.c_ulong(size.value - 32)
    try:
        E.call("kate_division", ptr(garbage_u32(size.value // 4)), ctypes.byref(small), 12,
               _view(u32(n * 8), n), ptr(u32(8)), _view(u32(n * 8), n), None)
        raise AssertionError("small buffer accepted")
    except _lib.HipError:
        pass


def poly_permute(E):
    """permute on lookup-shaped data (core/tests/poly.rs:556-640: table values
    with duplicates, inputs drawn from the usable table rows), a rotated
    input view, one and two sort tiles, and a missing input value.  The
    sort's schedule takes every branch: small integers with a few spread
    values (different keys share their top 8 bytes: the full sort runs),
    spread values only (the top-8-byte order is final), small integers only
    (the prefix passes see constant bytes and skip), and a table whose keys
    agree in the top 8 bytes and differ lower down."""
    rng = np.random.default_rng(5)
    cases = ((40, 37, 30, 5), (3000, 2975, 5000, 5), (2100, 2100, 7, 5), (3000, 2990, 0, 1),
             (2500, 2500, 100000, 0), (600, 600, -1, 0))
    for n, usable, spread, every in cases:
        if spread > 0:
            table_ints = rng.integers(0, spread, n)
            tab = O.ints_to_array([O.to_mont_int(int(v)) for v in table_ints])
        elif spread == 0:
            tab = O.gen_scalars(n, 78)
        else:
            # one shared top 8 bytes, the low words drawn from a few values
            hi = (int(rng.integers(1, 1 << 60)) << 192)
            table_ints = [hi + (int(v) << 64) + int(w) for v, w in
                          zip(rng.integers(0, 9, n), rng.integers(0, 5, n))]
            tab = O.ints_to_array([O.to_mont_int(v % O.R_MOD) for v in table_ints])
        if every:
            # large canonical values too: the top limbs take part in the sort
            tab[::every] = O.gen_scalars(n, 77)[::every]
        # inputs drawn from the usable table rows
        inp = tab[rng.integers(0, usable, n)]
        ei, et, missing = O.permute(inp, tab, usable)
        assert missing == 0
        rot = 9 % n
        di = _placed(inp, rot)
        dt = flat(tab).copy()
        ri, rt = garbage_u32(usable * 8), garbage_u32(usable * 8)
        tmp = _temp(E, "permute", usable, _view(di, n), _view(dt, n), _view(ri, usable),
                    _view(rt, usable), None)
        E.call("permute", ptr(tmp), None, usable, _view(di, n, rot), _view(dt, n),
               _view(ri, usable), _view(rt, usable), None)
        assert (ri.reshape(-1, 8) == ei).all(), ("permute input", n)
        assert (rt.reshape(-1, 8) == et).all(), ("permute table", n)
    inp = O.ints_to_array([O.to_mont_int(v) for v in (9, 2, 2, 9, 4, 4)])
    tab = O.ints_to_array([O.to_mont_int(v) for v in (1, 2, 3, 4, 5, 6)])
    ei, et, missing = O.permute(inp, tab, 6)
    assert missing == 1
    ri, rt = garbage_u32(48), garbage_u32(48)
    tmp = _temp(E, "permute", 6, _view(u32(48), 6), _view(u32(48), 6), _view(ri, 6),
                _view(rt, 6), None)
    E.call("permute", ptr(tmp), None, 6, _view(flat(inp), 6), _view(flat(tab), 6),
           _view(ri, 6), _view(rt, 6), None)
    assert (ri.reshape(-1, 8) == ei).all() and (rt.reshape(-1, 8) == et).all()


def poly_elementwise(E):
    """Mixed-length add / sub / mul on rotated and sliced views, the three
    constant fills and the scalar inverse / power."""
    for na, nb in ((1000, 1000), (700, 1000), (1000, 300), (5, 1), (1, 3)):
        a, b = O.gen_scalars(na, 11 + na), O.gen_scalars(nb, 12 + nb)
        ra, rb = 10 % na, (nb - 10 % nb) % nb
        up = max(na, nb)
        for op in ("add", "sub", "mul"):
            da, db = _placed(a, ra), _placed(b, rb)
            whole = up + 7
            dr = garbage_u32(whole * 8)
            E.call("poly_" + op, _lib.PolyPtr(dr.ctypes.data + 32 * 0, up, 0, 5, whole),
                   _view(da, na, ra), _view(db, nb, rb), None)
            got = dr.reshape(-1, 8)[(5 + np.arange(up)) % whole]
            assert (got == O.poly_binary(op, a, b)).all(), (op, na, nb)
    n = 300
    d = garbage_u32(n * 8)
    E.call("poly_zero", _view(d, n), None)
    assert (d == 0).all()
    E.call("poly_one_lagrange", _view(d, n, 7), None)
    assert (d.reshape(-1, 8) == O.one()).all()
    for begin, slen in ((0, n), (37, n), (0, 1000), (650, 1000)):
        d = garbage_u32(n * 8)
        sl = _lib.PolySlice(begin, n, slen)
        E.call("poly_one_coef", _view(d, n), None, ctypes.byref(sl))
        exp = np.zeros((n, 8), np.uint32)
        for i in range(n):
            if (begin + i) % slen == 0:
                exp[i] = O.one()
        assert (d.reshape(-1, 8) == exp).all(), (begin, slen)
    for v in (O.gen_scalars(3, 5)):
        t = flat(v).copy()
        E.call("inv_scalar", ptr(t), None)
        assert (t == O.inv(v)).all()
        for e in (0, 1, 2, 0xFFFFFFFFFFFFFFFF, 0x123456789ABCDEF):
            t = flat(v).copy()
            E.call("scalar_pow", ptr(t), e, None)
            assert (t == O.pow_(v, L(e))).all(), e
    t = u32(8)
    E.call("inv_scalar", ptr(t), None)
    assert (t == 0).all()


ALL = [("single_ntt", single_ntt), ("ntt_plans", ntt_plans),
       ("ntt_fused_reversal", ntt_fused_reversal), ("batched_ntt_mixed", batched_ntt_mixed),
       ("batched_ntt_plans", batched_ntt_plans), ("ntt_graph_replay", ntt_graph_replay),
       ("gate", gate), ("field_inverse", field_inverse), ("arith29", arith29), ("mlp_grad", mlp_grad), ("single_msm", single_msm), ("msm_hot_bucket", msm_hot_bucket), ("msm_graph_replay", msm_graph_replay),
       ("batched_msm_shared_tables", batched_msm_shared_tables),
       ("batched_msm_degenerate", batched_msm_degenerate), ("ref_ntt_abi", ref_ntt_abi),
       ("gate_ref_abi", gate_ref_abi), ("msm_host_pipeline", msm_host_pipeline),
       ("ntt_noalloc", ntt_noalloc), ("poly_scans", poly_scans), ("poly_kate", poly_kate),
       ("poly_elementwise", poly_elementwise), ("poly_permute", poly_permute)]


def main(argv):
    asan = "--asan" in argv
    names = [a for a in argv if not a.startswith("--")]
    E = Emu(asan)
    for name, fn in ALL:
        if names and name not in names:
            continue
        t = time.time()
        print(f"[emu] {name} ...", flush=True)
        fn(E)
        print(f"[emu] {name} ok ({time.time() - t:.1f}s)", flush=True)


if __name__ == "__main__":
    import os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    main(sys.argv[1:])
