"""The proving chunk (INTT of the columns, coset extension x8, gate, inverse
coset, one MSM commitment per 2^k chunk) against the oracle run end to end,
and the batched/graph schedule against the plain single-call schedule."""
import numpy as np
import pytest
import torch

from impeller_amd import ntt as N
from impeller_amd import workloads as WL
from impeller_amd.fused import gate_device
from impeller_amd.runtime import device_u32, to_device, to_host
from oracle import oracle as O

pytestmark = pytest.mark.gpu

mont = WL.mont_limbs
R = WL.R_MOD


def _coset_factors(g, n):
    return O.powers(mont(g), n)


def oracle_chunk(k, cols, seed, ext=3):
    n, big = 1 << k, 1 << (k + ext)
    K = k + ext
    lag = [O.gen_scalars(n, seed * 100 + i) for i in range(cols)]
    ninv = mont(pow(n, -1, R))
    coeffs = [O.vec_scale(O.best_fft(a, mont(WL.omega(k, True)), k), ninv) for a in lag]
    gpow = _coset_factors(WL.GENERATOR, big)
    evals = []
    for c in coeffs[:6]:
        e = np.zeros((big, 8), np.uint32)
        e[:n] = c
        evals.append(O.best_fft(O.vec_mul(e, gpow), mont(WL.omega(K)), K))
    h = O.gate(*evals)
    hc = O.vec_scale(O.best_fft(h, mont(WL.omega(K, True)), K), mont(pow(big, -1, R)))
    hc = O.vec_mul(hc, _coset_factors(pow(WL.GENERATOR, -1, R), big))
    kk = O.gen_scalars(n, 77)
    commits = [O.msm_closed_form(hc[j * n:(j + 1) * n], kk) for j in range(1 << ext)]
    return coeffs, hc, np.stack(commits)


@pytest.mark.parametrize("msm_cfg", [(10, 0, 64), (8, 1, 32)])
def test_chunk_plain_and_batched_match_oracle(msm_cfg):
    from impeller_amd import msm as M
    from impeller_amd.runtime import Stream
    k, cols = 10, 7
    coeffs, hc, commits = oracle_chunk(k, cols, seed=5)
    side = torch.cuda.Stream()  # graphs are not captured on the legacy stream
    with torch.cuda.stream(side):
        stream = Stream.current(0)
        pc = WL.ProvingChunk(stream, k=k, cols=cols, seed=5)
        c, target, task = msm_cfg
        tables = None
        if target:
            tables = device_u32(M.n_tables(c, target) * (1 << k) * 16)
            tables[: (1 << k) * 16].copy_(pc.bases)
            M.precompute_device(tables, 1 << k, c, target, stream)
        pc.run_plain(msm_cfg, msm_tables=tables)
        stream.sync()
        for i in range(cols):
            assert (to_host(pc.coeffs[i]) == coeffs[i]).all(), f"column {i} coefficients"
        assert (to_host(pc.h) == hc).all()
        assert (to_host(pc.results, 16) == commits).all()

        # batched schedule: the extension reads the coefficients through src
        # (zero-extended in the kernel), so stale extended buffers must not matter
        for e in pc.ext:
            e.fill_(-1)
        nbs = pc.batches()
        mb = pc.msm_batch(msm_cfg, tables=tables, table_stride=(1 << k) if target else 0)
        g = N.Graph(stream, lambda: pc.run_batched(nbs, mb))
        for rep in range(2):
            pc.h.zero_()
            mb.results.zero_()
            torch.cuda.synchronize()
            g.launch(stream)
            stream.sync()
            assert (to_host(pc.h) == hc).all(), rep
            assert (mb.results_host() == commits).all(), rep


@pytest.mark.parametrize("n", [1, 777, 1 << 15])
def test_gate_device(stream, n):
    ins = [O.gen_scalars(n, 40 + i) for i in range(6)]
    expect = O.gate(*ins)
    xs = [to_device(a) for a in ins]
    out = device_u32(n * 8)
    gate_device(xs, out, n, stream)
    stream.sync()
    assert (to_host(out) == expect).all()
