"""Parity at the benchmark's own sizes, not only at reduced ones:
  * NTT 2^24 forward and coset (shift 7) with the tuned plans and through
    ssip_ntt's in-place plan: all 2^24 outputs against best_fft;
  * the headline MSM 2^22 (c = 20, one bucket set, task 64) against the
    closed form (bases k_i G: sum_i s_i k_i G);
  * all 40 commitment MSMs of the batched workload (sizes 2^16..2^22 on one
    shared 2^22 table, mixed scalar kinds) as one batch, each against the
    closed form;
  * the 64-transform batch: every item of log size <= 23 against best_fft,
    the larger coset extensions by Horner samples and an inverse round trip;
  * the proving chunk end to end at k = 18 (20 columns) against the oracle,
    and at k = 22: its commitments against the closed form of the GPU's h and
    h itself at sampled points of the extended coset."""
import json
import os
import random

import numpy as np
import pytest
import torch

from impeller_amd import _lib
from impeller_amd import autotune as A
from impeller_amd import msm as M
from impeller_amd import ntt as N
from impeller_amd import workloads as WL
from impeller_amd.runtime import Stream, device_u32, to_device, to_host
from oracle import oracle as O

pytestmark = pytest.mark.gpu
mont = WL.mont_limbs
R = WL.R_MOD
PICKS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     "impeller_amd", "tuned", "picks.json")


def picks():
    with open(PICKS) as f:
        return json.load(f)


def picked(prob, key="exhaustive"):
    p = picks().get(prob)
    if p is None and prob.startswith("ntt_ext:"):
        p = picks().get("ntt_coset:" + prob.split(":")[1])
    return p.get(key) if p else None


# ------------------------------------------------------------- NTT 2^24
@pytest.fixture(scope="module")
def ntt24():
    k = 24
    n = 1 << k
    x = O.gen_scalars(n, 2024)
    w = mont(WL.omega(k))
    return k, x, w, O.best_fft(x, w, k)


def test_ntt24_forward_tuned_plan(stream, ntt24):
    k, x, w, expect = ntt24
    plan = picked("ntt_fwd:24") or [1, 10, 8, 8, 8]
    xd = to_device(x)
    tw = N.twiddles(w, (1 << k) // 2, stream)
    N.ntt(xd, tw, k, stream, passes=plan)
    stream.sync()
    assert (to_host(xd) == expect).all(), plan


def test_ntt24_ssip_inplace(stream, ntt24):
    k, x, w, expect = ntt24
    xd = to_device(x)
    tw = N.twiddles(w, (1 << k) // 2, stream)
    _lib.call("ssip_ntt", _lib.PolyPtr(xd.data_ptr(), 1 << k, 0, 0, 1 << k), tw.data_ptr(), k,
              stream.raw(), 8, 8)
    stream.sync()
    assert (to_host(xd) == expect).all()


def test_ntt24_coset_tuned_plan(stream, ntt24):
    k, x, w, _ = ntt24
    n = 1 << k
    expect = O.best_fft(O.vec_mul(x, O.powers(mont(WL.GENERATOR), n)), w, k)
    plan = picked("ntt_coset:24") or [1, 10, 8, 8, 8]
    xd = to_device(x)
    tw = N.twiddles(w, n // 2, stream)
    N.ntt(xd, tw, k, stream, coset_shift=to_device(mont(WL.GENERATOR)), passes=plan)
    stream.sync()
    assert (to_host(xd) == expect).all(), plan


# ------------------------------------------------------------- MSM 2^22
def test_msm22_headline_config(stream):
    """c = 20, one bucket set (13 precomputed tables), task 64."""
    nm, c, target, task = 1 << 22, 20, 1, 64
    tables = device_u32(M.n_tables(c, target) * nm * 16)
    _lib.call("impeller_gen_bases", tables.data_ptr(), nm, 77, 0, stream.raw())
    M.precompute_device(tables, nm, c, target, stream)
    scal = device_u32(nm * 8)
    _lib.call("impeller_gen_scalars", scal.data_ptr(), nm, 0, 78, 0, stream.raw())
    got = M.msm_device_cfg(tables, scal, nm, c, target, stream, task=task)
    expect = O.msm_closed_form(to_host(scal), O.gen_scalars(nm, 77))
    assert (got == expect).all()


@pytest.mark.parametrize("key", ["exhaustive", "model"])
def test_msm40_batched_all_items(stream, key):
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        st = Stream.current(0)
        spec = WL.msm_batch_spec()
        wl = WL.MsmWorkload(spec, st)
        cfgs = []
        for lg, kind in spec:
            cfg = picked(f"msm:{lg}:{kind}", key)
            cfgs.append(tuple(cfg) if cfg else A.default_msm(lg))
        mb = M.MsmBatch(wl.items(cfgs), st)
        g = N.Graph(st, lambda: mb.run(st))
        g.launch(st)
        st.sync()
        got = mb.results_host()
    kk = O.gen_scalars(1 << 22, 77)
    for i, ((lg, kind), s) in enumerate(zip(spec, wl.scalars)):
        n = 1 << lg
        exp = O.msm_closed_form(to_host(s), kk[:n])
        assert (got[i] == exp).all(), (i, lg, kind, cfgs[i])


# ------------------------------------------------------ 64-NTT batch
@pytest.fixture(scope="module")
def ntt64():
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        st = Stream.current(0)
        spec = WL.ntt_batch_spec()
        plans = {}
        for kind, lg in spec:
            L = lg + 3 if kind == "ext" else lg
            op = {"fwd": "ntt_fwd", "inv": "ntt_inv", "ext": "ntt_ext"}[kind]
            pl = picked(f"{op}:{L}")
            if pl:
                plans[(kind, L)] = pl
        wl = WL.NttWorkload(spec, st, plans=plans)
        nb = wl.batch()
        g = N.Graph(st, lambda: nb.run(st))
        g.launch(st)
        st.sync()
        srcs = [to_host(s) for s in wl.inputs]
        outs = [to_host(it["data"]) for it in wl.items]
        # inverse coset round trip of the large extensions, on the GPU
        back = {}
        for i, ((kind, lg), it) in enumerate(zip(spec, wl.items)):
            if kind == "ext" and lg + 3 > 23:
                L = lg + 3
                d = it["data"].clone()
                N.ntt(d, wl.tw.get(L, True), L, st, inverse=True, coset_shift=wl.ginv)
                st.sync()
                back[i] = to_host(d)
                del d
    return spec, srcs, outs, back


@pytest.mark.parametrize("part", range(4))
def test_ntt64_batch_items(ntt64, part):
    spec, srcs, outs, back = ntt64
    g = mont(WL.GENERATOR)
    rng = random.Random(part)
    for i in range(part * 16, part * 16 + 16):
        kind, lg = spec[i]
        src, out = srcs[i], outs[i]
        if kind == "fwd":
            exp = O.best_fft(src, mont(WL.omega(lg)), lg)
        elif kind == "inv":
            exp = O.vec_scale(O.best_fft(src, mont(WL.omega(lg, True)), lg),
                              mont(pow(1 << lg, -1, R)))
        else:
            L = lg + 3
            if L <= 23:
                e = np.zeros((1 << L, 8), np.uint32)
                e[: 1 << lg] = src
                exp = O.best_fft(O.vec_mul(e, O.powers(g, 1 << L)), mont(WL.omega(L)), L)
            else:
                # out[j] = P(g w^j) for the source polynomial P, at sampled j,
                # and the inverse coset transform gives P back exactly
                wL = WL.omega(L)
                for j in [0, 1, (1 << L) - 1] + [rng.randrange(1 << L) for _ in range(3)]:
                    pt = mont(WL.GENERATOR * pow(wL, j, R) % R)
                    assert (O.poly_eval(src, pt) == out[j]).all(), (i, j)
                b = back[i]
                assert (b[: 1 << lg] == src).all(), i
                assert not b[1 << lg:].any(), i
                continue
        assert (out == exp).all(), (i, kind, lg)


# ------------------------------------------------------ proving chunk
def _oracle_chunk(k, cols, seed, ext=3):
    n, big, K = 1 << k, 1 << (k + ext), k + ext
    lag = [O.gen_scalars(n, seed * 100 + i) for i in range(cols)]
    ninv = mont(pow(n, -1, R))
    coeffs = [O.vec_scale(O.best_fft(a, mont(WL.omega(k, True)), k), ninv) for a in lag]
    gpow = O.powers(mont(WL.GENERATOR), big)
    evals = []
    for c in coeffs[:6]:
        e = np.zeros((big, 8), np.uint32)
        e[:n] = c
        evals.append(O.best_fft(O.vec_mul(e, gpow), mont(WL.omega(K)), K))
    h = O.gate(*evals)
    hc = O.vec_scale(O.best_fft(h, mont(WL.omega(K, True)), K), mont(pow(big, -1, R)))
    hc = O.vec_mul(hc, O.powers(mont(pow(WL.GENERATOR, -1, R)), big))
    kk = O.gen_scalars(n, 77)
    commits = [O.msm_closed_form(hc[j * n:(j + 1) * n], kk) for j in range(1 << ext)]
    return coeffs, hc, np.stack(commits)


def _chunk_tables(pc, cfg, stream):
    if M.n_tables(cfg[0], cfg[1]) == 1:
        return None, 0
    n = 1 << pc.k
    t = device_u32(M.n_tables(cfg[0], cfg[1]) * n * 16)
    t[: n * 16].copy_(pc.bases)
    M.precompute_device(t, n, cfg[0], cfg[1], stream)
    return t, n


def test_chunk18_end_to_end():
    k, cols = 18, 20
    coeffs, hc, commits = _oracle_chunk(k, cols, seed=5)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        st = Stream.current(0)
        pc = WL.ProvingChunk(st, k=k, cols=cols, seed=5)
        cfg = tuple(picked(f"msm:{k}:0") or A.default_msm(k))
        tab, stride = _chunk_tables(pc, cfg, st)
        pc.run_plain(cfg, msm_tables=tab)
        st.sync()
        for i in range(cols):
            assert (to_host(pc.coeffs[i]) == coeffs[i]).all(), i
        assert (to_host(pc.h) == hc).all()
        assert (to_host(pc.results, 16) == commits).all()
        nbs = pc.batches()
        mb = pc.msm_batch(cfg, tables=tab, table_stride=stride)
        g = N.Graph(st, lambda: pc.run_batched(nbs, mb))
        pc.h.zero_()
        g.launch(st)
        st.sync()
        assert (to_host(pc.h) == hc).all()
        assert (mb.results_host() == commits).all()


def test_chunk22_commitments_and_samples():
    k, cols = 22, 20
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        st = Stream.current(0)
        pc = WL.ProvingChunk(st, k=k, cols=cols, seed=5)
        cfg = tuple(picked(f"msm:{k}:0") or A.default_msm(k))
        tab, stride = _chunk_tables(pc, cfg, st)
        nbs = pc.batches()
        mb = pc.msm_batch(cfg, tables=tab, table_stride=stride)
        g = N.Graph(st, lambda: pc.run_batched(nbs, mb))
        g.launch(st)
        st.sync()
        h = to_host(pc.h)
        got = mb.results_host()
        coeffs6 = [to_host(pc.coeffs[i]) for i in range(6)]
    n, K = 1 << k, k + 3
    kk = O.gen_scalars(n, 77)
    for j in range(8):
        assert (got[j] == O.msm_closed_form(h[j * n:(j + 1) * n], kk)).all(), j
    # the six gate columns' coefficients against the oracle's INTT
    for i in range(6):
        lag = O.gen_scalars(n, 5 * 100 + i)
        exp = O.vec_scale(O.best_fft(lag, mont(WL.omega(k, True)), k), mont(pow(n, -1, R)))
        assert (coeffs6[i] == exp).all(), i
    # h is the coefficient form of gate(P_0..P_5) on the coset g<w_K>: sample it
    rng = random.Random(22)
    wK = WL.omega(K)
    for j in [0, 1, (1 << K) - 1] + [rng.randrange(1 << K) for _ in range(3)]:
        pt = mont(WL.GENERATOR * pow(wK, j, R) % R)
        vals = [O.poly_eval(c, pt).reshape(1, 8) for c in coeffs6]
        assert (O.poly_eval(h, pt) == O.gate(*vals)[0]).all(), j
