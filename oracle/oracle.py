"""ctypes wrapper over oracle/liboracle.so -- the CPU checker for the HIP path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Field vectors are numpy uint32 arrays of shape (n, 8): 8 little-endian u32
limbs per element, Montgomery form (R = 2^256), the reference's layout
(core/src/common/mont/src/field.cuh:685-743).  G1 affine points are (n, 16)
(x || y), XYZZ points (n, 32).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FR, FQ = 0, 1

R_MOD = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
Q_MOD = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
MONT_R = 1 << 256

_lib = None


def host_threads():
    """Worker threads for the multi-threaded checker routines: the
    OMP_NUM_THREADS share when the environment sets one, else every core."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return os.cpu_count() or 1


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_INT = ctypes.c_int


def _declare(L):
    sig = {
        "orc_mul": [_INT, _P, _P, _P],
        "orc_sqr": [_INT, _P, _P],
        "orc_add": [_INT, _P, _P, _P],
        "orc_sub": [_INT, _P, _P, _P],
        "orc_neg": [_INT, _P, _P],
        "orc_pow": [_INT, _P, _P, _P],
        "orc_inv": [_INT, _P, _P],
        "orc_from_number": [_INT, _P, _P],
        "orc_to_number": [_INT, _P, _P],
        "orc_one": [_INT, _P],
        "orc_modulus": [_INT, _P],
        "orc_vec_mul": [_INT, _P, _P, _P, _U64],
        "orc_vec_scale": [_INT, _P, _U64, _P],
        "orc_gen_scalars": [_INT, _U64, _U64, _U64, _P],
        "orc_best_fft": [_P, _P, ctypes.c_uint32, _INT],
        "orc_powers": [_P, _U64, _P],
        "orc_poly_eval": [_P, _U64, _P, _P],
        "orc_poly_binary": [_INT, _P, _P, _U64, _P, _U64],
        "orc_kate_division": [_P, _U64, _P, _P],
        "orc_batch_invert": [_P, _U64],
        "orc_scan_mul": [_P, _U64],
        "orc_permute": [_P, _P, _U64, _P, _P],
        "orc_distribute_powers": [_P, _U64, _U64, _U64, _U64, _P, _U64, _U64, _U64, _U64],
        "orc_xyzz_madd": [_P, _P, _P],
        "orc_xyzz_add": [_P, _P, _P],
        "orc_xyzz_dbl": [_P, _P],
        "orc_xyzz_to_affine": [_P, _P],
        "orc_affine_to_xyzz": [_P, _P],
        "orc_affine_on_curve": [_P],
        "orc_xyzz_on_curve": [_P],
        "orc_g1_generator": [_P],
        "orc_g1_mul": [_P, _P, _P],
        "orc_gen_bases": [_U64, _U64, _U64, _P, _INT],
        "orc_best_multiexp": [_P, _P, _U64, _INT, _P],
        "orc_msm_closed_form": [_P, _P, _U64, _P],
        "orc_gate": [_P, _P, _P, _P, _P, _P, _P, _U64],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = _INT if name.endswith("on_curve") else None
    L.orc_permute.restype = _U64


def _ptr(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


# ----------------------------------------------------------------- conversions
def int_to_limbs(x, nlimbs=8):
    return np.array([(x >> (32 * i)) & 0xFFFFFFFF for i in range(nlimbs)], dtype=np.uint32)


def limbs_to_int(a):
    a = np.asarray(a, dtype=np.uint32).reshape(-1)
    return sum(int(v) << (32 * i) for i, v in enumerate(a))


def ints_to_array(xs):
    out = np.zeros((len(xs), 8), dtype=np.uint32)
    for i, x in enumerate(xs):
        out[i] = int_to_limbs(x)
    return out


def array_to_ints(a):
    return [limbs_to_int(row) for row in np.asarray(a).reshape(-1, 8)]


def modulus(f):
    return R_MOD if f == FR else Q_MOD


def to_mont_int(x, f=FR):
    """Python-int Montgomery representation (independent of the C code)."""
    m = modulus(f)
    return (x * MONT_R) % m


def from_mont_int(x, f=FR):
    m = modulus(f)
    return (x * pow(MONT_R, -1, m)) % m


# ------------------------------------------------------------------- scalars
def _un(f, name, a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    r = np.zeros(8, dtype=np.uint32)
    getattr(lib(), name)(f, _ptr(r), _ptr(a))
    return r


def _bin(f, name, a, b):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    r = np.zeros(8, dtype=np.uint32)
    getattr(lib(), name)(f, _ptr(r), _ptr(a), _ptr(b))
    return r


def mul(a, b, f=FR):
    return _bin(f, "orc_mul", a, b)


def add(a, b, f=FR):
    return _bin(f, "orc_add", a, b)


def sub(a, b, f=FR):
    return _bin(f, "orc_sub", a, b)


def neg(a, f=FR):
    return _un(f, "orc_neg", a)


def pow_(a, e_int_limbs, f=FR):
    return _bin(f, "orc_pow", a, e_int_limbs)


def inv(a, f=FR):
    return _un(f, "orc_inv", a)


def from_number(n, f=FR):
    return _un(f, "orc_from_number", n)


def to_number(a, f=FR):
    return _un(f, "orc_to_number", a)


def one(f=FR):
    r = np.zeros(8, dtype=np.uint32)
    lib().orc_one(f, _ptr(r))
    return r


# ------------------------------------------------------------------- vectors
def gen_scalars(n, seed, kind=0, start=0):
    out = np.zeros((n, 8), dtype=np.uint32)
    lib().orc_gen_scalars(kind, seed, start, n, _ptr(out))
    return out


def vec_mul(a, b, f=FR):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    out = np.zeros_like(a)
    lib().orc_vec_mul(f, _ptr(out), _ptr(a), _ptr(b), a.shape[0])
    return out


def vec_scale(x, s, f=FR):
    x = np.ascontiguousarray(x, dtype=np.uint32).copy()
    s = np.ascontiguousarray(s, dtype=np.uint32)
    lib().orc_vec_scale(f, _ptr(x), x.shape[0], _ptr(s))
    return x


def best_fft(a, omega, log_n, threads=None):
    a = np.ascontiguousarray(a, dtype=np.uint32).copy()
    assert a.shape == (1 << log_n, 8)
    omega = np.ascontiguousarray(omega, dtype=np.uint32)
    lib().orc_best_fft(_ptr(a), _ptr(omega), log_n, threads or host_threads())
    return a


def poly_eval(coeffs, point):
    c = np.ascontiguousarray(coeffs, dtype=np.uint32)
    p = np.ascontiguousarray(point, dtype=np.uint32)
    r = np.zeros(8, dtype=np.uint32)
    lib().orc_poly_eval(_ptr(c), c.shape[0], _ptr(p), _ptr(r))
    return r


def poly_binary(op, a, b):
    """op "add" / "sub" / "mul" over mixed lengths (poly_basic.cu:4-54)."""
    a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, 8)
    b = np.ascontiguousarray(b, dtype=np.uint32).reshape(-1, 8)
    r = np.zeros((max(a.shape[0], b.shape[0]), 8), dtype=np.uint32)
    lib().orc_poly_binary({"add": 0, "sub": 1, "mul": 2}[op], _ptr(r), _ptr(a), a.shape[0],
                          _ptr(b), b.shape[0])
    return r


def kate_division(p, b):
    p = np.ascontiguousarray(p, dtype=np.uint32).reshape(-1, 8)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    q = np.zeros_like(p)
    lib().orc_kate_division(_ptr(p), p.shape[0], _ptr(b), _ptr(q))
    return q


def batch_invert(x):
    x = np.ascontiguousarray(x, dtype=np.uint32).reshape(-1, 8).copy()
    lib().orc_batch_invert(_ptr(x), x.shape[0])
    return x


def scan_mul(x):
    x = np.ascontiguousarray(x, dtype=np.uint32).reshape(-1, 8).copy()
    lib().orc_scan_mul(_ptr(x), x.shape[0])
    return x


def permute(inp, table, usable):
    """(res_input, res_table, missing) of the lookup permutation."""
    inp = np.ascontiguousarray(inp, dtype=np.uint32).reshape(-1, 8)
    table = np.ascontiguousarray(table, dtype=np.uint32).reshape(-1, 8)
    ri = np.zeros((usable, 8), dtype=np.uint32)
    rt = np.zeros((usable, 8), dtype=np.uint32)
    missing = lib().orc_permute(_ptr(inp), _ptr(table), usable, _ptr(ri), _ptr(rt))
    return ri, rt, missing


def powers(base, n):
    out = np.zeros((n, 8), dtype=np.uint32)
    base = np.ascontiguousarray(base, dtype=np.uint32)
    lib().orc_powers(_ptr(out), n, _ptr(base))
    return out


def distribute_powers(x, powers_arr, rotate=0, offset=0, whole_len=None, length=None,
                      s_begin=0, s_slice_len=None, s_len=None):
    """x is the *whole* physical array; (length, rotate, offset, whole_len) the view."""
    x = np.ascontiguousarray(x, dtype=np.uint32).copy()
    whole_len = whole_len or x.shape[0]
    length = length or whole_len
    s_slice_len = s_slice_len or length
    s_len = s_len or length
    p = np.ascontiguousarray(powers_arr, dtype=np.uint32)
    lib().orc_distribute_powers(_ptr(x), length, rotate, offset, whole_len, _ptr(p), p.shape[0],
                                s_begin, s_slice_len, s_len)
    return x


# --------------------------------------------------------------------- curve
def g1_generator():
    r = np.zeros(16, dtype=np.uint32)
    lib().orc_g1_generator(_ptr(r))
    return r


def g1_mul(p_aff, k_int):
    r = np.zeros(16, dtype=np.uint32)
    p = np.ascontiguousarray(p_aff, dtype=np.uint32)
    k = np.ascontiguousarray(int_to_limbs(k_int) if isinstance(k_int, int) else k_int,
                             dtype=np.uint32)
    lib().orc_g1_mul(_ptr(r), _ptr(p), _ptr(k))
    return r


def xyzz_op(name, *args, out_len=32):
    r = np.zeros(out_len, dtype=np.uint32)
    args = [np.ascontiguousarray(a, dtype=np.uint32) for a in args]
    getattr(lib(), name)(_ptr(r), *[_ptr(a) for a in args])
    return r


def on_curve_affine(p):
    return bool(lib().orc_affine_on_curve(_ptr(np.ascontiguousarray(p, dtype=np.uint32))))


def on_curve_xyzz(p):
    return bool(lib().orc_xyzz_on_curve(_ptr(np.ascontiguousarray(p, dtype=np.uint32))))


def gen_bases(n, seed, start=0, threads=None):
    out = np.zeros((n, 16), dtype=np.uint32)
    lib().orc_gen_bases(seed, start, n, _ptr(out), threads or host_threads())
    return out


def best_multiexp(scalars, bases, threads=None):
    s = np.ascontiguousarray(scalars, dtype=np.uint32)
    b = np.ascontiguousarray(bases, dtype=np.uint32)
    assert s.shape[0] == b.shape[0]
    r = np.zeros(16, dtype=np.uint32)
    lib().orc_best_multiexp(_ptr(s), _ptr(b), s.shape[0], threads or host_threads(), _ptr(r))
    return r


def msm_closed_form(scalars, k):
    s = np.ascontiguousarray(scalars, dtype=np.uint32)
    k = np.ascontiguousarray(k, dtype=np.uint32)
    r = np.zeros(16, dtype=np.uint32)
    lib().orc_msm_closed_form(_ptr(s), _ptr(k), s.shape[0], _ptr(r))
    return r


def gate(a, b, c, d, q, e):
    arrs = [np.ascontiguousarray(x, dtype=np.uint32) for x in (a, b, c, d, q, e)]
    out = np.zeros_like(arrs[0])
    lib().orc_gate(_ptr(out), *[_ptr(x) for x in arrs], arrs[0].shape[0])
    return out


def to_table_format(points):
    """Affine points (rows of 16 limbs, Montgomery R = 2^256) in the MSM table
    format of the HIP library: each coordinate in Montgomery form with
    R' = 2^261 (x R' = (x R) 2^5 mod q), canonical, as 8 x 32-bit limbs."""
    out = np.array(points, dtype=np.uint32, copy=True)
    for row in out:
        for c in (0, 8):
            v = limbs_to_int(row[c:c + 8])
            row[c:c + 8] = int_to_limbs((v << 5) % Q_MOD)
    return out
