"""GPU NTT parity with the oracle's best_fft restatement (bit-exact).

Mirrors core/tests/ntt.rs: random omega that is not a root of unity (:63),
rotated and sliced device views (:92-97, :225), recompute tables (:212-250),
chunked distribute_powers with a 3-power zeta table (:335-410)."""
import ctypes
import random

import numpy as np
import pytest

from impeller_amd import _lib
from impeller_amd import ntt as N
from impeller_amd.runtime import ScalarArray, device_u32, to_device, to_host
from oracle import oracle as O

pytestmark = pytest.mark.gpu
r = O.R_MOD
L = O.int_to_limbs
ROOT = pow(7, (r - 1) >> 28, r)


def omega(k):
    return pow(ROOT, 1 << (28 - k), r)


def mont(x):
    return L(O.to_mont_int(x))


def random_omega(seed):
    return mont(random.Random(seed).randrange(r))


def host_twiddles(w_mont, k):
    return O.powers(w_mont, max(1, (1 << k) // 2))


# 20: the default plan fuses the bit reversal into the first pass from 2^20
@pytest.mark.parametrize("k", list(range(1, 15)) + [16, 18, 20])
def test_ssip_ntt_random_omega(stream, k):
    print(f"ssip_ntt k={k}", flush=True)
    x = O.gen_scalars(1 << k, 1000 + k)
    w = random_omega(k)
    expect = O.best_fft(x, w, k)
    xa = ScalarArray.from_numpy(x)
    tw = ScalarArray.from_numpy(host_twiddles(w, k))
    N.SsipNtt().get_fn()([xa], [tw, stream])
    stream.sync()
    assert (xa.to_numpy() == expect).all()


@pytest.mark.parametrize("k", [10, 13, 20])
def test_ssip_ntt_rotated_slice(stream, k):
    # core/tests/ntt.rs:92-97: buffer of 2^(k+1), rotate(3), slice(2^k, 2^(k+1))
    x = O.gen_scalars(1 << k, 2000 + k)
    w = mont(omega(k))
    expect = O.best_fft(x, w, k)
    buf = ScalarArray.from_numpy(O.gen_scalars(1 << (k + 1), 3000 + k))
    buf.rotate(3)
    view = buf.slice(1 << k, 1 << (k + 1))
    whole = buf.to_numpy().copy()
    idx = (view.slice_info.offset + np.arange(1 << k)) % (1 << (k + 1))
    whole[idx] = x
    buf.values.copy_(to_device(whole))
    tw = ScalarArray.from_numpy(host_twiddles(w, k))
    N.SsipNtt().get_fn()([view], [tw, stream])
    stream.sync()
    after = buf.to_numpy()
    assert (after[idx] == expect).all()
    untouched = np.setdiff1d(np.arange(1 << (k + 1)), idx)
    assert (after[untouched] == whole[untouched]).all()


def test_ssip_precompute_matches_powers():
    w = random_omega(5)
    tw = ScalarArray(1 << 11, None, "cpu")
    N.SsipPrecompute().get_fn()(tw, w)
    assert (tw.to_numpy() == O.powers(w, 1 << 11)).all()


@pytest.mark.parametrize("k", [2, 9, 12, 15, 20])
def test_recompute_ntt(stream, k):
    # core/tests/ntt.rs:157-281: rotate(-2), pq/omegas from gen_pq_omegas
    g = N.GenPqOmegas()
    x = O.gen_scalars(1 << k, 4000 + k)
    w = random_omega(40 + k)
    expect = O.best_fft(x, w, k)
    pq = ScalarArray(g.get_pq_len(k), None, "cpu")
    om = ScalarArray(32, None, "cpu")
    g.get_fn()(pq, om, 1 << k, w)
    # tables follow precompute.cuh:89-112
    assert (om.to_numpy()[0] == w).all()
    assert (om.to_numpy()[3] == O.pow_(w, L(8))).all()
    pq_deg = g.get_pq_len(k).bit_length()
    step = (1 << k) >> pq_deg
    if pq.len > 1:
        assert (pq.to_numpy()[1] == O.pow_(w, L(step))).all()
    buf = ScalarArray.from_numpy(np.zeros((1 << k, 8), np.uint32))
    buf.rotate(-2)
    phys = ((1 << k) - buf.get_rotation() + np.arange(1 << k)) % (1 << k)
    whole = np.zeros((1 << k, 8), np.uint32)
    whole[phys] = x
    buf.values.copy_(to_device(whole))
    pq_d = ScalarArray.from_numpy(pq.to_numpy())
    om_d = ScalarArray.from_numpy(om.to_numpy())
    N.RecomputeNtt().get_fn()([buf], [pq_d, om_d, stream])
    stream.sync()
    assert (buf.to_numpy()[phys] == expect).all()


@pytest.mark.parametrize("k", [10, 14])
def test_distribute_zeta_chunked(stream, k):
    # core/tests/ntt.rs:330-410: extended domain (x 4 -> ext k+2), powers
    # [1, g, g^2] applied in 4 slices, then NTT of the whole rotated array
    ext = k + 2
    n = 1 << ext
    coeffs = np.zeros((n, 8), np.uint32)
    coeffs[: 1 << k] = O.gen_scalars(1 << k, 5000 + k)
    zeta = mont(pow(7, (r - 1) // 3, r))
    powers = np.stack([O.one(), zeta, O.mul(zeta, zeta)])
    w_ext = mont(omega(ext))
    shifted = O.distribute_powers(coeffs, powers)
    expect = O.best_fft(shifted, w_ext, ext)

    buf = ScalarArray.from_numpy(np.zeros((n, 8), np.uint32))
    buf.rotate(-2)
    phys = (n - buf.get_rotation() + np.arange(n)) % n
    whole = np.zeros((n, 8), np.uint32)
    whole[phys] = coeffs
    buf.values.copy_(to_device(whole))
    pw = ScalarArray.from_numpy(powers)
    chunk = n // 4
    fn = N.DistributePowers().get_fn()
    for i in range(4):
        sl = buf.slice(i * chunk, (i + 1) * chunk)
        fn([sl], [pw, stream], None, _lib.PolySlice(i * chunk, chunk, n))
    tw = ScalarArray.from_numpy(host_twiddles(w_ext, ext))
    N.SsipNtt().get_fn()([buf], [tw, stream])
    stream.sync()
    assert (buf.to_numpy()[phys] == expect).all()


@pytest.mark.parametrize("k", [3, 11, 17])
def test_impeller_ntt_inverse_and_coset(stream, k):
    n = 1 << k
    x = O.gen_scalars(n, 6000 + k)
    w = mont(omega(k))
    winv = mont(pow(omega(k), -1, r))
    g = mont(7)
    ginv = mont(pow(7, -1, r))
    tw = to_device(host_twiddles(w, k))
    twi = to_device(host_twiddles(winv, k))
    # forward coset: NTT(x_i * g^i)
    d = to_device(x)
    N.ntt(d, tw, k, stream, coset_shift=to_device(g))
    stream.sync()
    coset_expect = O.best_fft(O.distribute_powers(x, O.powers(g, n)), w, k)
    assert (to_host(d) == coset_expect).all()
    # inverse coset: back to x
    N.ntt(d, twi, k, stream, inverse=True, coset_shift=to_device(ginv))
    stream.sync()
    assert (to_host(d) == x).all()
    # plain inverse == n^-1 * NTT(omega^-1) (intt_mending.rs:29-62)
    d = to_device(x)
    N.ntt(d, twi, k, stream, inverse=True)
    stream.sync()
    ninv = mont(pow(n, -1, r))
    assert (to_host(d) == O.vec_scale(O.best_fft(x, winv, k), ninv)).all()


@pytest.mark.parametrize("passes", [[0, 10, 4, 4, 4, 4], [0, 10, 8, 8], [5, 10, 3, 8],
                                    [0, 11, 10, 6], [0, 10] + [2] * 8, [5, 11, 11],
                                    [5, 10, 1, 1, 9], [5, 11, 5, 6], [5, 10] + [1] * 3 + [2] * 4,
                                    [4, 10, 6, 6], [4, 11, 11, 1], [4, 10, 2, 5, 5],
                                    # bit reversal fused into the first pass
                                    [1, 10, 8, 8], [1, 10, 7, 9], [1, 11, 6, 10],
                                    [1, 10, 3, 3, 10], [1, 10, 10, 6], [1, 10, 1, 5, 10]])
def test_pass_splits_agree(stream, passes):
    k = 16
    x = O.gen_scalars(1 << k, 7000)
    w = random_omega(70)
    expect = O.best_fft(x, w, k)
    d = to_device(x)
    N.ntt(d, to_device(host_twiddles(w, k)), k, stream, passes=passes)
    stream.sync()
    assert (to_host(d) == expect).all()


@pytest.mark.parametrize("passes", [[0, 10] + [1] * 16, [0, 12, 8, 8], [5, 10, 11],
                                    [3, 10, 13], [0, 10, 0, 16], [5, 10, 4, 4]])
def test_invalid_plans_rejected(stream, passes):
    # at most 8 LDS passes of <= lds_log levels, fuse 0 or 5, levels summing to log n
    k = 16
    d = to_device(O.gen_scalars(1 << k, 7001))
    with pytest.raises(_lib.HipError):
        N.ntt(d, to_device(host_twiddles(random_omega(71), k)), k, stream, passes=passes)


def test_fused_reversal_needs_two_passes(stream):
    k = 8
    d = to_device(O.gen_scalars(1 << k, 7002))
    with pytest.raises(_lib.HipError):
        N.ntt(d, to_device(host_twiddles(random_omega(72), k)), k, stream, passes=[1, 10, 8])


@pytest.mark.parametrize("k,passes", [(17, [1, 10, 7, 10]), (12, [1, 11, 5, 7])])
def test_fused_reversal_inverse_and_coset(stream, k, passes):
    n = 1 << k
    x = O.gen_scalars(n, 6100 + k)
    w = mont(omega(k))
    winv = mont(pow(omega(k), -1, r))
    g = mont(7)
    ginv = mont(pow(7, -1, r))
    tw = to_device(host_twiddles(w, k))
    twi = to_device(host_twiddles(winv, k))
    d = to_device(x)
    N.ntt(d, tw, k, stream, coset_shift=to_device(g), passes=passes)
    stream.sync()
    assert (to_host(d) == O.best_fft(O.distribute_powers(x, O.powers(g, n)), w, k)).all()
    N.ntt(d, twi, k, stream, inverse=True, coset_shift=to_device(ginv), passes=passes)
    stream.sync()
    assert (to_host(d) == x).all()


def test_fused_reversal_2_24_matches_in_place(stream):
    # the fused-reversal plan of the benchmark (also the library default at
    # this size) against an in-place plan with the bit-reversal kernel, bit
    # for bit, forward and coset at 2^24
    k = 24
    n = 1 << k
    x = device_u32(n * 8)
    _lib.call("impeller_gen_scalars", x.data_ptr(), n, 0, 9100, 0, stream.raw())
    tw = N.twiddles(mont(omega(k)), n // 2, stream)
    g = to_device(mont(7))
    for shift in (None, g):
        a, b = x.clone(), x.clone()
        N.ntt(a, tw, k, stream, coset_shift=shift, passes=[4, 10, 7, 7, 6])
        N.ntt(b, tw, k, stream, coset_shift=shift, passes=[1, 10, 8, 8, 8])
        stream.sync()
        assert bool((a == b).all()), shift is not None


def test_full_size_2_20_vs_oracle(stream):
    k = 20
    x = O.gen_scalars(1 << k, 8000)
    w = random_omega(80)
    expect = O.best_fft(x, w, k)
    d = to_device(x)
    N.ntt(d, to_device(host_twiddles(w, k)), k, stream)
    stream.sync()
    assert (to_host(d) == expect).all()


def test_full_size_2_24_properties(stream):
    # size-independent checks at the benchmark size: y_j = p(omega^j) for
    # sampled j (Horner on the CPU), and INTT(NTT(x)) == x
    k = 24
    n = 1 << k
    w = omega(k)
    x = device_u32(n * 8)
    _lib.call("impeller_gen_scalars", x.data_ptr(), n, 0, 9000, 0, stream.raw())
    tw = N.twiddles(mont(w), n // 2, stream)
    twi = N.twiddles(mont(pow(w, -1, r)), n // 2, stream)
    orig = x.clone()
    N.ntt(x, tw, k, stream)
    stream.sync()
    host_x = to_host(orig)
    y = to_host(x)
    for j in (0, 1, 12345, n - 1):
        assert (O.poly_eval(host_x, mont(pow(w, j, r))) == y[j]).all()
    N.ntt(x, twi, k, stream, inverse=True)
    stream.sync()
    assert bool((x == orig).all())


@pytest.mark.parametrize("k", [24, 26, 27])
def test_coset_sparse_inputs_closed_form(stream, k):
    # x has a few nonzero coefficients, so y_j = sum_q x_q (g w^j)^q is a short
    # closed form: checks the coset factor shift^q at large q (all factor-table
    # digits) and the transform at sizes past the CPU oracle's reach; then the
    # inverse coset transform must give x back
    n = 1 << k
    g = 7
    w = omega(k)
    rng = random.Random(k)
    pos = [0, 1, 12345, (n >> 1) + 77, n - 1]
    vals = [rng.randrange(r) for _ in pos]
    x = device_u32(n * 8)
    x.zero_()
    host = np.zeros((len(pos), 8), np.uint32)
    for i, (q, v) in enumerate(zip(pos, vals)):
        host[i] = mont(v)
        x[q * 8:(q + 1) * 8] = to_device(host[i])
    tw = N.twiddles(mont(w), n // 2, stream)
    shift = to_device(mont(g))
    N.ntt(x, tw, k, stream, coset_shift=shift)
    stream.sync()
    for j in (0, 1, 999, n - 2):
        t = g * pow(w, j, r) % r
        expect = sum(v * pow(t, q, r) for q, v in zip(pos, vals)) % r
        got = O.limbs_to_int(to_host(x[j * 8:(j + 1) * 8])[0])
        assert got == O.to_mont_int(expect), j
    twi = N.twiddles(mont(pow(w, -1, r)), n // 2, stream)
    N.ntt(x, twi, k, stream, inverse=True, coset_shift=to_device(mont(pow(g, -1, r))))
    stream.sync()
    back = to_host(x)
    nz = np.flatnonzero(back.any(axis=1))
    assert list(nz) == pos
    assert (back[pos] == host).all()
