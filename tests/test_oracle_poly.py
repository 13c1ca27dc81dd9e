"""Pins the oracle's polynomial-library restatements (oracle/oracle.c,
orc_poly_binary / orc_kate_division / orc_batch_invert / orc_scan_mul /
orc_permute / orc_poly_eval) against Python big-integer arithmetic and the
identities that define each result, independently of the C code:

  kate_division  q(X) (X - b) + p(b) = p(X) coefficient by coefficient
  poly_eval      sum_i a_i x^i with Python pow
  batch_invert   x * x^-1 = 1, zeros kept
  scan_mul       products of prefixes
  poly_binary    the element-wise definition over mixed lengths
  permute        the truth construction of core/tests/poly.rs:556-640 (a
                 BTreeMap of table counts) plus the lookup invariants
"""
import random
from collections import Counter

import numpy as np
import pytest

from oracle import oracle as O

r = O.R_MOD


def mont_arr(vals):
    return O.ints_to_array([O.to_mont_int(v % r) for v in vals])


def plain(arr):
    return [O.from_mont_int(v) for v in O.array_to_ints(arr)]


@pytest.mark.parametrize("n", [1, 2, 3, 8, 64, 257])
def test_kate_division_is_the_quotient(n):
    rng = random.Random(n)
    p = [rng.randrange(r) for _ in range(n)]
    b = rng.randrange(r)
    q = plain(O.kate_division(mont_arr(p), mont_arr([b])[0]))
    pb = sum(c * pow(b, i, r) for i, c in enumerate(p)) % r
    assert q[n - 1] == 0
    for i in range(n):
        prev = q[i - 1] if i > 0 else 0
        assert (prev - b * q[i] + (pb if i == 0 else 0)) % r == p[i], i


def test_kate_division_small_closed_form():
    # p = X^2 - (1 + b) X + b = (X - 1)(X - b) -> q = X - 1, p(b) = 0
    b = 12345
    p = [b, -(1 + b), 1]
    q = plain(O.kate_division(mont_arr(p), mont_arr([b])[0]))
    assert q == [r - 1, 1, 0]


@pytest.mark.parametrize("n", [0, 1, 5, 300])
def test_poly_eval_vs_python(n):
    rng = random.Random(7 + n)
    a = [rng.randrange(r) for _ in range(n)]
    x = rng.randrange(r)
    got = plain(O.poly_eval(mont_arr(a) if n else np.zeros((0, 8), np.uint32),
                            mont_arr([x])[0]).reshape(1, 8))[0]
    assert got == sum(c * pow(x, i, r) for i, c in enumerate(a)) % r


@pytest.mark.parametrize("n,zero_every", [(1, 0), (50, 17), (300, 3), (64, 1)])
def test_batch_invert_vs_python(n, zero_every):
    rng = random.Random(n)
    vals = [0 if zero_every and i % zero_every == 0 else rng.randrange(1, r) for i in range(n)]
    got = plain(O.batch_invert(mont_arr(vals)))
    for v, g in zip(vals, got):
        assert g == (0 if v == 0 else pow(v, -1, r))


def test_scan_mul_vs_python():
    rng = random.Random(3)
    vals = [rng.randrange(r) for _ in range(200)]
    vals[50] = 0
    got = plain(O.scan_mul(mont_arr(vals)))
    acc = 1
    for v, g in zip(vals, got):
        acc = acc * v % r
        assert g == acc


@pytest.mark.parametrize("na,nb", [(40, 40), (25, 40), (40, 9), (1, 3)])
def test_poly_binary_vs_python(na, nb):
    rng = random.Random(na * 100 + nb)
    a = [rng.randrange(r) for _ in range(na)]
    b = [rng.randrange(r) for _ in range(nb)]
    for op in ("add", "sub", "mul"):
        got = plain(O.poly_binary(op, mont_arr(a), mont_arr(b)))
        for i, g in enumerate(got):
            if i < min(na, nb):
                e = {"add": a[i] + b[i], "sub": a[i] - b[i], "mul": a[i] * b[i]}[op]
            elif na > nb:
                e = a[i]
            else:
                e = -b[i] if op == "sub" else b[i]
            assert g == e % r, (op, i)


def permute_truth(inp, table):
    """core/tests/poly.rs:556-640 in Python: sorted input; run starts take
    their value out of the table's multiset; the repeated rows, in order,
    take the leftovers in ascending order."""
    pin = sorted(inp)
    left = Counter(table)
    ptab = [None] * len(pin)
    repeated = []
    for row, v in enumerate(pin):
        if row == 0 or v != pin[row - 1]:
            ptab[row] = v
            assert left[v] > 0
            left[v] -= 1
        else:
            repeated.append(row)
    leftovers = sorted(v for v, c in left.items() for _ in range(c))
    for row, v in zip(repeated, leftovers):
        ptab[row] = v
    return pin, ptab


@pytest.mark.parametrize("n,usable,spread", [(64, 60, 40), (300, 275, 500), (16, 16, 3),
                                             (1, 1, 1)])
def test_permute_vs_reference_truth(n, usable, spread):
    rng = random.Random(n + spread)
    table = [rng.randrange(spread) for _ in range(n)]
    distinct = sorted(set(table[:usable]))
    inp = [distinct[rng.randrange(len(distinct))] for _ in range(n)]
    ri, rt, missing = O.permute(mont_arr(inp), mont_arr(table), usable)
    assert missing == 0
    pin, ptab = permute_truth(inp[:usable], table[:usable])
    assert plain(ri) == pin
    assert plain(rt) == ptab
    # lookup invariants: permutations of the originals; each row either
    # matches the table or repeats the previous input
    assert Counter(plain(ri)) == Counter(inp[:usable])
    assert Counter(plain(rt)) == Counter(table[:usable])
    for i, (a, t) in enumerate(zip(plain(ri), plain(rt))):
        assert a == t or (i > 0 and a == plain(ri)[i - 1])


def test_permute_counts_missing_values():
    table = [1, 2, 3, 4, 5, 6]
    inp = [9, 2, 2, 9, 4, 4]
    ri, rt, missing = O.permute(mont_arr(inp), mont_arr(table), 6)
    assert missing == 1
    assert plain(ri) == [2, 2, 4, 4, 9, 9]
    # run starts take their value, the rest the unused table values ascending
    assert plain(rt) == [2, 1, 4, 3, 9, 5]
