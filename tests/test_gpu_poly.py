"""GPU parity of the polynomial library (core/src/poly/src/poly.h:6-30)
through the reference-shaped operators of impeller_amd.poly, against the
oracle.  The cases follow core/tests/poly.rs (K = 20, the same rotations of
operands and results, zeros every 17th element for the inverse, random
exponents for the power), plus ragged and multi-level sizes (2^22 + 5
elements make three tile levels), mixed lengths, empty and one-element
inputs, and HIP graph capture and replay.  Bit-exact."""
import ctypes
import random

import numpy as np
import pytest

from impeller_amd import _lib
from impeller_amd import poly as P
from impeller_amd.runtime import GpuBuffer, Scalar, ScalarArray
from oracle import oracle as O

pytestmark = pytest.mark.gpu

K = 20


def dev(x, rotate=0):
    a = ScalarArray.from_numpy(x)
    if rotate:
        # keep the logical contents: place them where the rotated view reads
        phys = np.zeros_like(x)
        n = x.shape[0]
        rr = rotate % n
        phys[(n - rr + np.arange(n)) % n if rr else np.arange(n)] = x
        a = ScalarArray.from_numpy(phys)
        a.rotate(rotate)
    return a


def empty(n, rotate=0):
    a = ScalarArray.from_numpy(np.full((n, 8), 0xA5A5A5A5, np.uint32))
    if rotate:
        a.rotate(rotate)
    return a


@pytest.mark.parametrize("op", ["add", "sub", "mul"])
def test_binary_k20(stream, op):
    """core/tests/poly.rs:29-118: a rotated 10, b rotated -10, res rotated 3."""
    n = 1 << K
    a, b = O.gen_scalars(n, 1), O.gen_scalars(n, 2)
    ad, bd, rd = dev(a, 10), dev(b, -10), empty(n, 3)
    cls = {"add": P.PolyAdd, "sub": P.PolySub, "mul": P.PolyMul}[op]
    cls().get_fn()([rd], [ad, bd, stream])
    stream.sync()
    assert (rd.logical_numpy() == O.poly_binary(op, a, b)).all()


@pytest.mark.parametrize("na,nb", [(1 << 16, 3000), (5, 1 << 12), (1, 1), (777, 778)])
def test_binary_mixed_lengths(stream, na, nb):
    a, b = O.gen_scalars(na, 3), O.gen_scalars(nb, 4)
    for op, cls in (("add", P.PolyAdd), ("sub", P.PolySub), ("mul", P.PolyMul)):
        rd = empty(max(na, nb), 1)
        cls().get_fn()([rd], [dev(a, 2), dev(b, 1), stream])
        stream.sync()
        assert (rd.logical_numpy() == O.poly_binary(op, a, b)).all(), (op, na, nb)


def test_zero_one(stream):
    """core/tests/poly.rs:251-360, plus PolyOneCoef on a slice whose origin
    falls inside it and outside it."""
    n = 1 << K
    d = empty(n)
    P.PolyZero().get_fn()([d], [stream])
    stream.sync()
    assert (d.to_numpy() == 0).all()
    P.PolyOneLagrange().get_fn()([d], [stream])
    stream.sync()
    assert (d.to_numpy() == O.one()).all()
    P.PolyOneCoef().get_fn()([d], [stream])
    stream.sync()
    got = d.to_numpy()
    assert (got[0] == O.one()).all() and (got[1:] == 0).all()
    m = 4096
    for begin, total in ((m, 4 * m), (3 * m + 5, 4 * m)):
        d = empty(m)
        P.PolyOneCoef().get_fn()([d], [stream], None, _lib.PolySlice(begin, m, total))
        stream.sync()
        exp = np.zeros((m, 8), np.uint32)
        for i in range(m):
            if (begin + i) % total == 0:
                exp[i] = O.one()
        assert (d.to_numpy() == exp).all(), begin


def run_eval(stream, x, pt, rotate=0):
    f = P.PolyEval()
    tmp = GpuBuffer(f.get_buffer_size(x.shape[0]))
    res = Scalar()
    f.get_fn()([tmp, res], [dev(x, rotate), Scalar.from_numpy(pt), stream])
    stream.sync()
    return res.to_numpy()


@pytest.mark.parametrize("n", [1 << K, 1, 2, 2047, 2048, 2049, (1 << 22) + 5])
def test_eval(stream, n):
    """core/tests/poly.rs:120-168 (poly rotated 10) and ragged sizes."""
    x = O.gen_scalars(n, 10 + n % 97)
    pt = O.gen_scalars(1, 11)[0]
    assert (run_eval(stream, x, pt, 10 % n) == O.poly_eval(x, pt)).all()


def test_eval_at_roots_and_zero(stream):
    """p(1) = sum of coefficients; p(0) = p[0]."""
    n = 5000
    x = O.gen_scalars(n, 12)
    one, zero = O.one(), np.zeros(8, np.uint32)
    s = x[0]
    for v in x[1:]:
        s = O.add(s, v)
    assert (run_eval(stream, x, one) == s).all()
    assert (run_eval(stream, x, zero) == x[0]).all()


def run_kate(stream, p, b, rp=0, rq=0):
    f = P.KateDivision()
    n = p.shape[0]
    tmp = GpuBuffer(f.get_buffer_size(n.bit_length() - 1))
    q = empty(n, rq)
    f.get_fn()([tmp, q], [dev(p, rp), Scalar.from_numpy(b), stream])
    stream.sync()
    return q.logical_numpy()


@pytest.mark.parametrize("k", [K, 0, 1, 5, 11, 12, 22])
def test_kate(stream, k):
    """core/tests/poly.rs:170-249: p rotated 2, q rotated -2; the quotient
    of halo2 kate_division in q[0 .. 2^k - 1), q[2^k - 1] = 0."""
    n = 1 << k
    p = O.gen_scalars(n, 20 + k)
    b = O.gen_scalars(1, 21 + k)[0]
    assert (run_kate(stream, p, b, 2 % n, -2 % n if n > 1 else 0) == O.kate_division(p, b)).all()


def run_scan(stream, p, x0, rp=0, rt=0):
    f = P.PolyScan()
    n = p.shape[0]
    tmp = GpuBuffer(f.get_buffer_size(n))
    t = empty(n, rt)
    f.get_fn()([tmp, t], [dev(p, rp), Scalar.from_numpy(x0), stream])
    stream.sync()
    return t.logical_numpy()


@pytest.mark.parametrize("n", [1 << K, 1, 2, 2049, (1 << 22) + 5])
def test_scan(stream, n):
    """core/tests/poly.rs:362-416: poly rotated 3, result rotated -5;
    res[0] = x0, res[i] = res[i-1] * poly[i-1]."""
    p = O.gen_scalars(n, 30 + n % 89)
    x0 = O.gen_scalars(1, 31)[0]
    got = run_scan(stream, p, x0, 3 % n, -5 % n)
    exp = O.scan_mul(np.concatenate([x0[None], p[:-1]]))
    assert (got == exp).all()


def run_invert(stream, x):
    f = P.PolyInvert()
    tmp = GpuBuffer(f.get_buffer_size(x.shape[0]))
    d = dev(x)
    f.get_fn()([tmp, d], [stream])
    stream.sync()
    return d.logical_numpy()


@pytest.mark.parametrize("n", [1 << K, 1, 3, 2048, 6000, (1 << 22) + 5])
def test_invert(stream, n):
    """core/tests/poly.rs:418-470: zeros at every 17th element kept."""
    x = O.gen_scalars(n, 40 + n % 83)
    x[::17] = 0
    assert (run_invert(stream, x) == O.batch_invert(x)).all()


def test_invert_degenerate(stream):
    for x in (np.zeros((5000, 8), np.uint32), np.tile(O.one(), (3000, 1)),
              np.tile(O.neg(O.one()), (2049, 1))):
        assert (run_invert(stream, x) == O.batch_invert(x)).all()


def test_scalar_inv_and_pow(stream):
    """core/tests/poly.rs:472-580 (random exponent), plus 0^-1 = 0 and x^0 = 1."""
    rng = random.Random(5)
    for v in O.gen_scalars(4, 50):
        s = Scalar.from_numpy(v)
        P.ScalarInv().get_fn()([s], [stream])
        stream.sync()
        assert (s.to_numpy() == O.inv(v)).all()
        for e in (rng.getrandbits(64), 0, 1, (1 << 64) - 1):
            s = Scalar.from_numpy(v)
            P.ScalarPow(e).get_fn()([s], [stream])
            stream.sync()
            assert (s.to_numpy() == O.pow_(v, O.int_to_limbs(e))).all(), e
    s = Scalar.from_numpy(np.zeros(8, np.uint32))
    P.ScalarInv().get_fn()([s], [stream])
    stream.sync()
    assert (s.to_numpy() == 0).all()


def test_size_query_and_small_buffer(stream):
    n = 1 << 16
    need = P.PolyScan().get_buffer_size(n)
    assert need > 0
    d = dev(O.gen_scalars(n, 60))
    small = ctypes.c_ulong(need - 32)
    tmp = GpuBuffer(need)
    with pytest.raises(_lib.HipError):
        _lib.call("scan_mul", tmp.data_ptr(), ctypes.byref(small), d.poly_ptr(), stream.raw())
    ok = ctypes.c_ulong(need)
    _lib.call("scan_mul", tmp.data_ptr(), ctypes.byref(ok), d.poly_ptr(), stream.raw())
    stream.sync()


def test_graph_capture_replay():
    """kate_division, batched_invert and poly_eval launch kernels only: one
    captured graph replays them, checked on every replay, with the inputs
    rewritten by eager copies between replays."""
    import torch
    from impeller_amd.ntt import Graph
    from impeller_amd.runtime import Stream
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        stream = Stream.current(0)
        n, kp = 5000, 1 << 12
        pk, b = O.gen_scalars(kp, 70), O.gen_scalars(1, 71)[0]
        xi = O.gen_scalars(n, 72)
        xi[::13] = 0
        pd, qd = dev(pk), empty(kp)
        bd = Scalar.from_numpy(b)
        idd = dev(xi)
        pv = O.gen_scalars(1, 73)[0]
        ed, pt = Scalar(), Scalar.from_numpy(pv)
        tk = GpuBuffer(P.KateDivision().get_buffer_size(12))
        ti = GpuBuffer(P.PolyInvert().get_buffer_size(n))
        te = GpuBuffer(P.PolyEval().get_buffer_size(n))
        src = idd.values.clone()
        torch.cuda.synchronize()

        def body():
            P.KateDivision().get_fn()([tk, qd], [pd, bd, stream])
            P.PolyInvert().get_fn()([ti, idd], [stream])
            P.PolyEval().get_fn()([te, ed], [pd, pt, stream])

        g = Graph(stream, body)
        for rep in range(3):
            idd.values.copy_(src)
            qd.values.fill_(rep)
            ed.values.fill_(-1)
            torch.cuda.synchronize()
            g.launch(stream)
            stream.sync()
            assert (qd.logical_numpy() == O.kate_division(pk, b)).all(), rep
            assert (idd.logical_numpy() == O.batch_invert(xi)).all(), rep
            assert (ed.to_numpy() == O.poly_eval(pk, pv)).all(), rep


def lookup_data(n, usable, seed):
    """core/tests/poly.rs:556-600: table values rng % (2 len) (duplicates),
    inputs sampled from the distinct values of the usable table rows; here
    every fifth table value is a full-width field element as well."""
    rng = np.random.default_rng(seed)
    tab = O.ints_to_array([O.to_mont_int(int(v)) for v in rng.integers(0, 2 * n, n)])
    tab[::5] = O.gen_scalars(n, seed)[::5]
    inp = tab[rng.integers(0, usable, n)]
    return inp, tab


@pytest.mark.parametrize("k,blinding", [(K, 24), (12, 0), (16, 5)])
def test_permute(stream, k, blinding):
    """core/tests/poly.rs:556-680 (K = 20, 24 blinding rows): sorted inputs;
    each run start matched by one table copy; the other rows take the
    leftover table values in ascending order."""
    n = 1 << k
    usable = n - (blinding + 1) if blinding else n
    inp, tab = lookup_data(n, usable, k)
    ei, et, missing = O.permute(inp, tab, usable)
    assert missing == 0
    f = P.PolyPermute()
    tmp = GpuBuffer(f.get_buffer_size(usable))
    ri, rt = empty(usable), empty(usable)
    f.get_fn()([ri, rt, tmp], [dev(inp, 3), dev(tab), stream])
    stream.sync()
    assert (ri.logical_numpy() == ei).all()
    assert (rt.logical_numpy() == et).all()


@pytest.mark.parametrize("kind", ["spread", "small", "shared_top"])
def test_permute_sort_schedules(stream, kind):
    """The radix sort's device-side schedule on each kind of key set: values
    spread over the field (the top-8-byte order is final, the remaining
    passes skip), small integers only (constant high bytes skip, the low
    bytes sort), and keys sharing their top 8 bytes but differing lower
    down (the full sort runs)."""
    n = 1 << 16
    rng = np.random.default_rng(len(kind))
    if kind == "spread":
        tab = O.gen_scalars(n, 91)
    elif kind == "small":
        tab = O.ints_to_array([O.to_mont_int(int(v)) for v in rng.integers(0, 3 * n, n)])
    else:
        hi = int(rng.integers(1, 1 << 60)) << 192
        tab = O.ints_to_array([O.to_mont_int(hi + (int(v) << 64) + int(w)) for v, w in
                               zip(rng.integers(0, 300, n), rng.integers(0, 300, n))])
    inp = tab[rng.integers(0, n, n)]
    ei, et, missing = O.permute(inp, tab, n)
    assert missing == 0
    f = P.PolyPermute()
    tmp = GpuBuffer(f.get_buffer_size(n))
    ri, rt = empty(n), empty(n)
    f.get_fn()([ri, rt, tmp], [dev(inp, 5), dev(tab), stream])
    stream.sync()
    assert (ri.logical_numpy() == ei).all()
    assert (rt.logical_numpy() == et).all()


def test_permute_missing_value(stream):
    inp = O.ints_to_array([O.to_mont_int(v) for v in (9, 2, 2, 9, 4, 4)])
    tab = O.ints_to_array([O.to_mont_int(v) for v in (1, 2, 3, 4, 5, 6)])
    ei, et, missing = O.permute(inp, tab, 6)
    assert missing == 1
    f = P.PolyPermute()
    tmp = GpuBuffer(f.get_buffer_size(6))
    ri, rt = empty(6), empty(6)
    f.get_fn()([ri, rt, tmp], [dev(inp), dev(tab), stream])
    stream.sync()
    assert (ri.logical_numpy() == ei).all() and (rt.logical_numpy() == et).all()
