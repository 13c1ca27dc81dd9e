"""Device field and curve arithmetic vs the CPU oracle (bit-exact)."""
import json
import os
import random

import numpy as np
import pytest

from impeller_amd import _lib
from impeller_amd.runtime import device_u32, to_device, to_host
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
L = O.int_to_limbs


def field_op(stream, f, op, a, b=None):
    n = a.shape[0]
    ra = device_u32(n * 8)
    da = to_device(a)
    db = to_device(b) if b is not None else None
    _lib.call("impeller_field_op", f, op, ra.data_ptr(), da.data_ptr(),
              db.data_ptr() if db is not None else None, n, stream.raw())
    stream.sync()
    return to_host(ra)


@pytest.mark.parametrize("f", [O.FR, O.FQ])
def test_field_ops_vs_oracle(stream, f):
    m = O.modulus(f)
    rng = random.Random(7 + f)
    specials = [0, 1, 2, m - 1, m - 2, (m - 1) // 2, (1 << 256) % m]
    vals = specials + [rng.randrange(m) for _ in range(2041)]
    a = O.ints_to_array(vals)
    b = O.ints_to_array(vals[3:] + vals[:3])
    n = a.shape[0]
    exp_mul = np.stack([O.mul(a[i], b[i], f) for i in range(n)])
    assert (field_op(stream, f, 0, a, b) == exp_mul).all()
    assert (field_op(stream, f, 1, a, b) == np.stack([O.add(a[i], b[i], f) for i in range(n)])).all()
    assert (field_op(stream, f, 2, a, b) == np.stack([O.sub(a[i], b[i], f) for i in range(n)])).all()
    assert (field_op(stream, f, 3, a) == np.stack([O.neg(a[i], f) for i in range(n)])).all()
    assert (field_op(stream, f, 4, a) == np.stack([O.from_number(a[i], f) for i in range(n)])).all()
    assert (field_op(stream, f, 5, a) == np.stack([O.to_number(a[i], f) for i in range(n)])).all()
    k = 64
    e = O.ints_to_array([rng.randrange(1 << 256) for _ in range(k)])
    assert (field_op(stream, f, 6, a[:k], e)
            == np.stack([O.pow_(a[i], e[i], f) for i in range(k)])).all()
    assert (field_op(stream, f, 7, a[:k]) == np.stack([O.inv(a[i], f) for i in range(k)])).all()


@pytest.mark.parametrize("f", [O.FR, O.FQ])
def test_inverse_many_values(stream, f):
    """The single-lane inversion (optimized binary GCD, 17 rounds of 31
    steps on 64-bit approximations) on 12 000 inputs: random values, values
    of every bit length, powers of two and their neighbours, values next to
    m and values with long runs of trailing zero bits.  The input A is the
    Montgomery representative of a (A = a R); the output is a^-1 R, i.e.
    A^-1 R^2 mod m (and 0 for 0)."""
    m = O.modulus(f)
    R = 1 << 256
    rng = random.Random(31 + f)
    vals = [0, 1, 2, 3, m - 1, m - 2, (m + 1) // 2, (1 << 64) - 1, 1 << 64, 1 << 63, (1 << 253)]
    vals += [1 << k for k in range(254)] + [(1 << k) - 1 for k in range(1, 254)]
    vals += [m - (1 << k) for k in range(0, 253, 3)]
    vals += [rng.randrange(1, 1 << 40) << rng.randrange(0, 200) for _ in range(1000)]
    vals += [rng.randrange(1 << rng.randrange(1, 255)) for _ in range(3000)]
    vals += [rng.randrange(m) for _ in range(12000 - len(vals))]
    vals = [v % m for v in vals]
    got = field_op(stream, f, 7, O.ints_to_array(vals))
    for i, v in enumerate(vals):
        exp = 0 if v == 0 else pow(v, -1, m) * R * R % m
        assert O.limbs_to_int(got[i]) == exp, (f, hex(v))


def test_field_kat_on_device(stream):
    kat = json.load(open(os.path.join(GOLD, "field_kat.json")))["instances"]
    for inst in ("instance1", "instance2"):
        d = {k: int(v, 16) for k, v in kat[inst].items()}
        a, b = L(d["a"])[None], L(d["b"])[None]
        assert O.limbs_to_int(field_op(stream, O.FR, 0, a, b)) == d["prod_mont"]
        assert O.limbs_to_int(field_op(stream, O.FR, 1, a, b)) == d["sum_mont"]
        assert O.limbs_to_int(field_op(stream, O.FR, 2, a, b)) == d["sub_mont"]
        ea = field_op(stream, O.FR, 4, a)
        assert O.limbs_to_int(field_op(stream, O.FR, 6, ea, b)) == d["pow_mont"]
        assert O.limbs_to_int(field_op(stream, O.FR, 7, ea)) == d["a_inv_mont"]


def g1_op(stream, op, a, b, out_cols, n):
    r = device_u32(n * out_cols)
    da = to_device(a)
    db = to_device(b) if b is not None else None
    _lib.call("impeller_g1_op", op, r.data_ptr(), da.data_ptr(),
              db.data_ptr() if db is not None else None, n, stream.raw())
    stream.sync()
    return to_host(r, out_cols)


def test_curve_ops_vs_oracle(stream):
    n = 96
    bases = O.gen_bases(n + 1, 77)
    x = np.stack([O.xyzz_op("orc_affine_to_xyzz", bases[i]) for i in range(n)])
    # make the XYZZ inputs non-trivial (zz != 1)
    x = np.stack([O.xyzz_op("orc_xyzz_dbl", x[i]) for i in range(n)])
    x[5] = 0
    x[5][24:] = O.one(O.FQ)                       # identity
    q = bases[1:n + 1].copy()
    q[9] = bases[9 // 1]                          # generic
    q[11] = 0                                     # affine identity
    got = g1_op(stream, 0, x, q, 32, n)
    exp = np.stack([O.xyzz_op("orc_xyzz_madd", x[i], q[i]) for i in range(n)])
    assert (got == exp).all()
    y = np.roll(x, 1, axis=0)
    got = g1_op(stream, 1, x, y, 32, n)
    exp = np.stack([O.xyzz_op("orc_xyzz_add", x[i], y[i]) for i in range(n)])
    assert (got == exp).all()
    got = g1_op(stream, 2, x, None, 32, n)
    assert (got == np.stack([O.xyzz_op("orc_xyzz_dbl", x[i]) for i in range(n)])).all()
    got = g1_op(stream, 3, x, None, 16, n)
    assert (got == np.stack([O.xyzz_op("orc_xyzz_to_affine", x[i], out_len=16)
                             for i in range(n)])).all()
    ks = [random.Random(i).randrange(1 << 254) for i in range(8)]
    got = g1_op(stream, 4, bases[:8], O.ints_to_array(ks), 16, 8)
    assert (got == np.stack([O.g1_mul(bases[i], ks[i]) for i in range(8)])).all()


def test_quad_point_ops_vs_oracle(stream):
    """The group-of-four-lanes 29-bit addition and doubling (the MSM's final
    window sums run on them) against the oracle's XYZZ formulas, including
    the identity, doubling and P + (-P) branches of the addition."""
    n = 96
    bases = O.gen_bases(n + 1, 91)
    x = np.stack([O.xyzz_op("orc_affine_to_xyzz", bases[i]) for i in range(n)])
    x = np.stack([O.xyzz_op("orc_xyzz_dbl", x[i]) for i in range(n)])
    x[5] = 0
    x[5][24:] = O.one(O.FQ)                       # identity
    y = np.roll(x, 1, axis=0).copy()
    y[20] = x[20]                                 # doubling branch
    y[21] = x[21]
    y[21][8:16] = O.neg(x[21][8:16], O.FQ)        # P + (-P)
    got = g1_op(stream, 9, x, y, 32, n)
    exp = np.stack([O.xyzz_op("orc_xyzz_add", x[i], y[i]) for i in range(n)])
    for i in range(n):
        if not exp[i][16:24].any():               # identity: only zz == 0 carries meaning
            assert not got[i][16:24].any(), i
        else:
            assert (got[i] == exp[i]).all(), i
    got = g1_op(stream, 10, x, None, 32, n)
    exp = np.stack([O.xyzz_op("orc_xyzz_dbl", x[i]) for i in range(n)])
    for i in range(n):
        if not exp[i][16:24].any():
            assert not got[i][16:24].any(), i
        else:
            assert (got[i] == exp[i]).all(), i


def test_generators_match_oracle(stream):
    n = 4096
    for kind in (0, 1, 2):
        t = device_u32(n * 8)
        _lib.call("impeller_gen_scalars", t.data_ptr(), n, kind, 1234, 77, stream.raw())
        stream.sync()
        assert (to_host(t) == O.gen_scalars(n, 1234, kind=kind, start=77)).all()
    nb = 512
    t = device_u32(nb * 16)
    _lib.call("impeller_gen_bases", t.data_ptr(), nb, 99, 5, stream.raw())
    stream.sync()
    assert (to_host(t, 16) == O.gen_bases(nb, 99, start=5)).all()
