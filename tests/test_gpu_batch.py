"""Batched NTT (one launch per stage over all transforms) and HIP-graph
replay, bit-exact against the oracle and against the single-call path."""
import random

import numpy as np
import pytest

from impeller_amd import ntt as N
from impeller_amd.runtime import Stream, device_u32, to_device, to_host
from oracle import oracle as O

pytestmark = pytest.mark.gpu
r = O.R_MOD
L = O.int_to_limbs
ROOT = pow(7, (r - 1) >> 28, r)


def mont(x):
    return L(O.to_mont_int(x))


def omega(k):
    return pow(ROOT, 1 << (28 - k), r)


def build_items(specs, seed):
    """specs: (kind, log_n) with kind in fwd / inv / ext / fwd_random."""
    rng = random.Random(seed)
    items, expects = [], []
    for i, (kind, k) in enumerate(specs):
        n = 1 << k
        if kind == "ext":
            ks = k - 3
            coeffs = O.gen_scalars(1 << ks, seed * 100 + i)
            w = mont(omega(k))
            g = mont(7)
            ext = np.zeros((n, 8), np.uint32)
            ext[: 1 << ks] = coeffs
            expects.append(O.best_fft(O.distribute_powers(ext, O.powers(g, n)), w, k))
            items.append(dict(data=device_u32(n * 8), src=to_device(coeffs), log_n=k, log_src=ks,
                              twiddle=to_device(O.powers(w, n // 2)), coset_shift=to_device(g)))
        elif kind == "inv":
            x = O.gen_scalars(n, seed * 100 + i)
            winv = mont(pow(omega(k), -1, r))
            ninv = mont(pow(n, -1, r))
            expects.append(O.vec_scale(O.best_fft(x, winv, k), ninv))
            items.append(dict(data=to_device(x), log_n=k, inverse=True,
                              twiddle=to_device(O.powers(winv, max(1, n // 2)))))
        else:
            x = O.gen_scalars(n, seed * 100 + i)
            w = mont(rng.randrange(r)) if kind == "fwd_random" else mont(omega(k))
            expects.append(O.best_fft(x, w, k))
            items.append(dict(data=to_device(x), log_n=k,
                              twiddle=to_device(O.powers(w, max(1, n // 2)))))
    return items, expects


SPECS = [("fwd", 12), ("inv", 14), ("ext", 13), ("fwd_random", 10), ("fwd", 3), ("inv", 9),
         ("ext", 15), ("fwd_random", 16), ("fwd", 1)]


@pytest.mark.parametrize("max_blocks", [0, 7, 256])
def test_batched_ntt_mixed(stream, max_blocks):
    items, expects = build_items(SPECS, 11 + max_blocks)
    b = N.NttBatch(items, stream)
    b.run(stream, max_blocks)
    stream.sync()
    for it, e in zip(items, expects):
        assert (to_host(it["data"]) == e).all(), it["log_n"]


def test_batched_ntt_plans(stream):
    specs = [("fwd_random", 14), ("inv", 12), ("ext", 14), ("fwd", 9)]
    items, expects = build_items(specs, 5)
    items[0]["plan"] = [5, 11, 9]
    items[1]["plan"] = [0, 10, 2, 3, 7]
    items[2]["plan"] = [5, 10, 4, 5]
    items[3]["plan"] = [4, 10, 5]
    b = N.NttBatch(items, stream)
    b.run(stream)
    stream.sync()
    for it, e in zip(items, expects):
        assert (to_host(it["data"]) == e).all()


def test_batched_ntt_fused_reversal():
    # fused-reversal plans in a batch (second buffers in the batch buffer),
    # including the library default at 2^20 and 2^21, eager and replayed
    import torch
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        stream = Stream.current(0)
        specs = [("fwd", 20), ("ext", 21), ("inv", 14), ("fwd_random", 13), ("fwd", 9)]
        items, expects = build_items(specs, 21)
        items[2]["plan"] = [1, 10, 6, 8]
        items[3]["plan"] = [1, 11, 3, 10]
        items[4]["plan"] = [4, 10, 5]
        restore = [(it["data"], it["data"].clone()) for it in items if it.get("src") is None]
        b = N.NttBatch(items, stream)
        g = N.Graph(stream, lambda: b.run(stream))
        for rep in range(3):
            for dst, src in restore:
                dst.copy_(src)
            torch.cuda.synchronize()
            if rep == 0:
                b.run(stream)
            else:
                g.launch(stream)
            stream.sync()
            for it, e in zip(items, expects):
                assert (to_host(it["data"]) == e).all(), (rep, it["log_n"])


def test_graph_replay_matches_eager():
    import torch
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        stream = Stream.current(0)
        specs = [("fwd", 12), ("ext", 13), ("inv", 11)]
        items, expects = build_items(specs, 9)
        # inputs restored before every replay (the batch runs in place)
        restore = [(it["data"], it["data"].clone()) for it in items if it.get("src") is None]
        b = N.NttBatch(items, stream)
        g = N.Graph(stream, lambda: b.run(stream))
        for rep in range(3):
            for dst, src in restore:
                dst.copy_(src)
            torch.cuda.synchronize()
            g.launch(stream)
            stream.sync()
            for it, e in zip(items, expects):
                assert (to_host(it["data"]) == e).all(), (rep, it["log_n"])


def test_batched_msm_uneven_shared_tables(stream):
    from impeller_amd import _lib
    from impeller_amd import msm as M
    # one 2^14 base set with precomputed tables for c = 13 (target 1), shared
    # by MSMs of different lengths through the table stride
    nb = 1 << 14
    bases = device_u32(nb * 16)
    _lib.call("impeller_gen_bases", bases.data_ptr(), nb, 91, 0, stream.raw())
    k = O.gen_scalars(nb, 91)
    tabs = {}
    for c, target in ((13, 1), (11, 3)):
        t = device_u32(M.n_tables(c, target) * nb * 16)
        t[: nb * 16].copy_(bases)
        M.precompute_device(t, nb, c, target, stream)
        tabs[(c, target)] = t
    items, expects = [], []
    rng = random.Random(3)
    for i in range(9):
        n = rng.choice([1, 37, 1 << 10, 1 << 12, nb])
        kind = i % 3
        s = O.gen_scalars(n, 700 + i, kind=kind)
        c, target = [(13, 1), (11, 3), (12, 0)][i % 3]
        if target == 0:
            items.append(dict(tables=bases, scalars=to_device(s), len=n, window=c, target=0,
                              task=rng.choice([16, 64])))
        else:
            items.append(dict(tables=tabs[(c, target)], scalars=to_device(s), len=n, window=c,
                              target=target, table_stride=nb, task=rng.choice([16, 64])))
        expects.append(O.msm_closed_form(s, k[:n]))
    b = M.MsmBatch(items, stream)
    for max_blocks in (0, 64):
        b.results.zero_()
        b.run(stream, max_blocks)
        stream.sync()
        got = b.results_host()
        for i, e in enumerate(expects):
            assert (got[i] == e).all(), (i, items[i]["len"], max_blocks)


def test_batched_msm_graph_replay():
    import torch
    from impeller_amd import _lib
    from impeller_amd import msm as M
    s_t = torch.cuda.Stream()
    with torch.cuda.stream(s_t):
        stream = Stream.current(0)
        n = 1 << 11
        bases = device_u32(n * 16)
        _lib.call("impeller_gen_bases", bases.data_ptr(), n, 5, 0, stream.raw())
        k = O.gen_scalars(n, 5)
        ss = [O.gen_scalars(n, 40 + i) for i in range(3)]
        items = [dict(tables=bases, scalars=to_device(s), len=n, window=w, target=0)
                 for s, w in zip(ss, (10, 12, 14))]
        b = M.MsmBatch(items, stream)
        g = N.Graph(stream, lambda: b.run(stream))
        for rep in range(3):
            b.results.zero_()
            torch.cuda.synchronize()
            g.launch(stream)
            stream.sync()
            got = b.results_host()
            for i, s in enumerate(ss):
                assert (got[i] == O.msm_closed_form(s, k)).all(), (rep, i)


def test_batched_msm_degenerate_items(stream):
    """An empty MSM, all-zero scalars and a cancelling pair (identity (0, 0))
    batched beside a normal item; results start as all-ones words."""
    from impeller_amd import _lib
    from impeller_amd import msm as M
    nb = 1 << 12
    bases = device_u32(nb * 16)
    _lib.call("impeller_gen_bases", bases.data_ptr(), nb, 93, 0, stream.raw())
    k = O.gen_scalars(nb, 93)
    tabs = {}
    for c, target in ((13, 1), (11, 3)):
        t = device_u32(M.n_tables(c, target) * nb * 16)
        t[: nb * 16].copy_(bases)
        M.precompute_device(t, nb, c, target, stream)
        tabs[(c, target)] = t
    cancel = np.zeros((64, 8), np.uint32)
    cancel[0] = O.int_to_limbs(O.to_mont_int(O.R_MOD - 1))
    cancel[1] = O.one()
    bb = to_host(bases, 16)[:64].copy()
    bb[1] = bb[0]
    normal = O.gen_scalars(1000, 808)
    items = [
        dict(tables=bases, scalars=to_device(np.zeros((1, 8), np.uint32)), len=0, window=12,
             target=0),
        dict(tables=tabs[(13, 1)], scalars=to_device(np.zeros((300, 8), np.uint32)), len=300,
             window=13, target=1, table_stride=nb),
        dict(tables=to_device(bb), scalars=to_device(cancel), len=64, window=10, target=0),
        dict(tables=tabs[(11, 3)], scalars=to_device(normal), len=1000, window=11, target=3,
             table_stride=nb),
    ]
    expects = [np.zeros(16, np.uint32)] * 2 + [O.best_multiexp(cancel, bb),
                                               O.msm_closed_form(normal, k[:1000])]
    assert not expects[2].any()
    b = M.MsmBatch(items, stream)
    for max_blocks in (0, 16):
        b.results.fill_(-1)
        b.run(stream, max_blocks)
        stream.sync()
        got = b.results_host()
        for i, e in enumerate(expects):
            assert (got[i] == e).all(), (i, items[i]["len"], max_blocks)
