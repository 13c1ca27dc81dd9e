"""Python-integer evaluation of a fused element-wise graph -- the checker for
the HIP fused-kernel emitter (impeller_amd/kernel_gen.py).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker, never by the
product.  It takes the graph as plain data (poly_repr, vertices), so it
shares no code with the emitter.  Semantics restated from
core/src/fused_kernels/kernel_gen.rs:460-600 (and the ArithGraph of
common/src/arith.rs):

  ("input", typ, mutable)             scalar: one element; array: x[idx]
  ("bin", "pp"|"ss", op, l, r)        add / sub / mul
  ("bin", "sp", op, s, p)             add: s + p, sub: s - p, subby: p - s,
                                      mul: s * p; in coefficient form add /
                                      sub / subby apply the scalar only where
                                      slice.to_origin(idx) == 0, elsewhere the
                                      result is p, -p, p
  ("unr", "p"|"s", "neg", a, _)       -a
  ("unr", "p", "pow", a, e)           a^e (e < 2^64)
  ("output", typ, store, in_node)     the stored value

Values are field elements in Montgomery form, shape (n, 8) / (8,) uint32,
converted to integers mod r for the arithmetic.
"""
import numpy as np

from . import oracle as O

R = O.R_MOD


def _to_ints(a):
    return [O.from_mont_int(v) for v in O.array_to_ints(np.asarray(a).reshape(-1, 8))]


def _to_arr(vals):
    return O.ints_to_array([O.to_mont_int(v % R) for v in vals])


def evaluate(poly_repr, vertices, inputs, n, slice_begin=0, slice_len=None):
    """inputs: {input vertex id: array (n, 8) or scalar (8,)}.  Returns
    {output vertex id: array (n, 8) or scalar (8,)}."""
    slice_len = n if slice_len is None else slice_len
    vals = {}
    for i, v in enumerate(vertices):
        if v[0] == "input":
            xs = _to_ints(inputs[i])
            vals[i] = xs if v[1] == "array" else xs[0]
    outs = {}
    for idx in range(n):
        origin = (slice_begin + idx) % slice_len == 0 if slice_len else idx == 0
        cur = {}

        def get(j):
            x = vals[j] if j in vals else cur[j]
            return x[idx] if isinstance(x, list) else x

        for i, v in enumerate(vertices):
            kind = v[0]
            if kind == "bin":
                _, k, op, lhs, rhs = v
                a, b = get(lhs), get(rhs)
                if op == "mul":
                    r = a * b
                elif k == "sp" and poly_repr == "coef" and not origin:
                    r = {"add": b, "sub": -b, "subby": b}[op]
                else:
                    r = {"add": a + b, "sub": a - b, "subby": b - a}[op]
                cur[i] = r % R
            elif kind == "unr":
                _, k, op, arg, power = v
                a = get(arg)
                cur[i] = (-a) % R if op == "neg" else pow(a, power, R)
            elif kind == "output":
                _, typ, store, _ = v
                outs.setdefault(i, []).append(get(store))
    result = {}
    for i, v in enumerate(vertices):
        if v[0] == "output":
            vs = outs.get(i, [])
            result[i] = _to_arr(vs) if v[1] == "array" else _to_arr(vs[:1])[0]
    return result
