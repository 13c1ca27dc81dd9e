"""GPU MSM parity with the oracle (halo2 best_multiexp restatement and the
closed form (sum s_i k_i) * G for bases k_i * G), affine results bit-exact.
Mirrors core/tests/msm.rs: several batches through the host-memory C ABI with
the buffer-size query, plus the device-resident entry point."""
import ctypes

import numpy as np
import pytest

from impeller_amd import _lib
from impeller_amd import msm as M
from impeller_amd.runtime import PointArray, ScalarArray, device_u32, to_device, to_host
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bases_small():
    n = 1 << 12
    return O.gen_bases(n, 555), O.gen_scalars(n, 555)


@pytest.mark.parametrize("n", [1, 2, 7, 100, 1 << 12])
@pytest.mark.parametrize("c", [8, 13, 16])
def test_msm_device_uniform(stream, bases_small, n, c):
    print(f"msm n={n} c={c}", flush=True)
    bases, _ = bases_small
    s = O.gen_scalars(n, 600 + n)
    expect = O.best_multiexp(s, bases[:n])
    got = M.msm_device(to_device(bases[:n]), to_device(s), n, c, stream)
    assert (got == expect).all()


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("c", [10, 16])
def test_msm_device_skewed(stream, bases_small, kind, c):
    bases, k = bases_small
    n = bases.shape[0]
    s = O.gen_scalars(n, 700 + kind, kind=kind)
    expect = O.msm_closed_form(s, k)
    assert (expect == O.best_multiexp(s, bases)).all()
    got = M.msm_device(to_device(bases), to_device(s), n, c, stream)
    assert (got == expect).all()


def test_msm_edge_cases(stream, bases_small):
    bases, _ = bases_small
    n = 64
    b = to_device(bases[:n])
    # all-zero scalars -> identity (0, 0)
    got = M.msm_device(b, to_device(np.zeros((n, 8), np.uint32)), n, 12, stream)
    assert not got.any()
    # r - 1 and 1 on the same base cancel; repeated base doubles
    s = np.zeros((n, 8), np.uint32)
    s[0] = O.int_to_limbs(O.to_mont_int(O.R_MOD - 1))
    s[1] = O.one()
    bb = bases[:n].copy()
    bb[1] = bb[0]
    bb[2] = bb[3]
    s[2] = O.one()
    s[3] = O.one()
    expect = O.best_multiexp(s, bb)
    assert (expect == O.g1_mul(bases[3], 2)).all()
    got = M.msm_device(to_device(bb), to_device(s), n, 12, stream)
    assert (got == expect).all()


def test_msm_host_abi_batches(stream, bases_small):
    # core/tests/msm.rs:132-277 call pattern: buffer query, BATCHES = 4
    bases, k = bases_small
    n = bases.shape[0]
    cfg = M.MsmConfig(window_size=12, target_window=32)
    sizes = M.MSM(cfg).get_buffer_size(n)
    buf = device_u32((sizes[0] + 3) // 4)
    pts = PointArray.from_numpy(bases, device="cpu")
    scal = [ScalarArray.from_numpy(O.gen_scalars(n, 800 + b), device="cpu") for b in range(4)]
    res = [np.zeros(16, np.uint32) for _ in range(4)]
    M.MSM(cfg).get_fn()([buf] + res, [pts] + scal)
    for b in range(4):
        assert (res[b] == O.msm_closed_form(scal[b].to_numpy(), k)).all()


@pytest.mark.parametrize("window,target", [(16, 1), (12, 4), (10, 26)])
def test_msm_precompute_tables(stream, bases_small, window, target):
    # msm_precompute fills tables P * 2^(j * n_windows * c); the MSM then uses
    # n_precompute point tables (msm.cu:43-49, msm_impl.cuh:56-67)
    bases, k = bases_small
    n = 1 << 10
    cfg = M.MsmConfig(window_size=window, target_window=target)
    npre = cfg.get_precompute()
    tables = [PointArray.from_numpy(bases[:n], device="cpu")] + [
        PointArray.from_numpy(np.zeros((n, 16), np.uint32), device="cpu")
        for _ in range(npre - 1)]
    M.MSMPrecompute(cfg).get_fn()(tables, n)
    step = cfg.n_windows() * window
    for j in (1, npre - 1):
        if j >= 1 and j < npre:
            assert (tables[j].to_numpy()[3] == O.g1_mul(bases[3], 1 << (j * step))).all()
    sizes = M.MSM(cfg).get_buffer_size(n)
    buf = device_u32((sizes[0] + 3) // 4)
    s = ScalarArray.from_numpy(O.gen_scalars(n, 900 + window), device="cpu")
    res = np.zeros(16, np.uint32)
    M.MSM(cfg).get_fn()([buf, res], tables + [s])
    assert (res == O.msm_closed_form(s.to_numpy(), k[:n])).all()


def test_msm_2_20_closed_form(stream):
    n = 1 << 20
    bases = device_u32(n * 16)
    _lib.call("impeller_gen_bases", bases.data_ptr(), n, 4242, 0, stream.raw())
    scal = device_u32(n * 8)
    _lib.call("impeller_gen_scalars", scal.data_ptr(), n, 0, 4343, 0, stream.raw())
    stream.sync()
    k = O.gen_scalars(n, 4242)
    s = to_host(scal)
    expect = O.msm_closed_form(s, k)
    for c in (14, 16):
        assert (M.msm_device(bases, scal, n, c, stream) == expect).all()


@pytest.mark.parametrize("c", [12, 14, 18])
def test_msm_top_window_hot_buckets(stream, bases_small, c):
    # with these windows the top window holds 2 bits of the scalar, so a few
    # buckets receive every point: the task split + partial-sum tree must
    # keep this exact (and fast)
    bases, k = bases_small
    n = bases.shape[0]
    s = O.gen_scalars(n, 1200 + c)
    got = M.msm_device(to_device(bases), to_device(s), n, c, stream)
    assert (got == O.msm_closed_form(s, k)).all()


@pytest.mark.parametrize("window,target,task", [(16, 1, 64), (13, 1, 7), (12, 3, 1),
                                                (15, 2, 4096), (16, 0, 33)])
def test_msm_device_precomputed_tables(stream, bases_small, window, target, task):
    bases, k = bases_small
    n = bases.shape[0]
    nt = M.n_tables(window, target)
    tables = device_u32(nt * n * 16)
    tables[: n * 16].copy_(to_device(bases))
    M.precompute_device(tables, n, window, target, stream)
    stream.sync()
    if nt > 1:
        step = M.MsmConfig(window_size=window, target_window=target or 10 ** 6).n_windows() * window
        # the tables are stored in the table format (Montgomery R' = 2^261)
        t1 = to_host(tables[n * 16: 2 * n * 16], 16)
        assert (t1[5] == O.to_table_format([O.g1_mul(bases[5], 1 << step)])[0]).all()
        t0 = to_host(tables[: n * 16], 16)
        assert (t0[:8] == O.to_table_format(bases[:8])).all()
    for kind in (0, 1, 2):
        s = O.gen_scalars(n, 1300 + kind + window, kind=kind)
        got = M.msm_device_cfg(tables, to_device(s), n, window, target, stream, task=task)
        assert (got == O.msm_closed_form(s, k)).all(), (kind, window, target, task)


def test_msm_selectors_2_18(stream):
    n = 1 << 18
    bases = device_u32(n * 16)
    _lib.call("impeller_gen_bases", bases.data_ptr(), n, 31, 0, stream.raw())
    k = O.gen_scalars(n, 31)
    for kind in (1, 2):
        scal = device_u32(n * 8)
        _lib.call("impeller_gen_scalars", scal.data_ptr(), n, kind, 32 + kind, 0, stream.raw())
        stream.sync()
        expect = O.msm_closed_form(to_host(scal), k)
        assert (M.msm_device(bases, scal, n, 15, stream) == expect).all()


@pytest.mark.parametrize("c,target,task", [(20, 1, 32), (16, 0, 64)])
def test_msm_2_24_closed_form(stream, c, target, task):
    # the largest size the benchmarks do not reach: 2^24 bases k_i * G,
    # uniform scalars, checked through (sum s_i k_i) * G
    n = 1 << 24
    nt = M.n_tables(c, target)
    tables = device_u32(nt * n * 16)
    _lib.call("impeller_gen_bases", tables.data_ptr(), n, 41, 0, stream.raw())
    M.precompute_device(tables, n, c, target, stream)
    scal = device_u32(n * 8)
    _lib.call("impeller_gen_scalars", scal.data_ptr(), n, 0, 42, 0, stream.raw())
    stream.sync()
    got = M.msm_device_cfg(tables, scal, n, c, target, stream, task=task)
    assert (got == O.msm_closed_form(to_host(scal), O.gen_scalars(n, 41))).all()
