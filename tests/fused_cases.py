"""Fused-kernel graphs the emitter tests use (seeded, so the library built
beforehand and the tests see the same graphs), and the build of their HIP
library into tests/_fused/ (`python -m tests.fused_cases` or
__graft_entry__.build()).

  gate_c5        a*b*c - d + q*e*(1 - q) with the constant 1 as a scalar
                 input, the gate of the proving chunk (Lagrange form)
  coef_mix       coefficient form: Sp add / sub / subby at the slice origin,
                 Sp mul, pow, neg, a scalar output and an in-place output
  rand_<form><k> seeded random graphs over 3 arrays and 2 scalars
"""
import os
import random

from impeller_amd import kernel_gen as KG

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_fused")


def gate_c5():
    g = KG.ArithGraph(KG.LAGRANGE)
    a, b, c, d, q, e = (g.input(KG.ARRAY) for _ in range(6))
    one = g.input(KG.SCALAR)
    t = g.binop("pp", "mul", a, b)
    t = g.binop("pp", "mul", t, c)
    t = g.binop("pp", "sub", t, d)
    u = g.binop("pp", "mul", q, e)
    w = g.binop("sp", "sub", one, q)
    u = g.binop("pp", "mul", u, w)
    g.output(KG.ARRAY, g.binop("pp", "add", t, u))
    return g


def coef_mix():
    g = KG.ArithGraph(KG.COEF)
    x, y = g.input(KG.ARRAY), g.input(KG.ARRAY)
    acc = g.input(KG.ARRAY, mutable=True)
    s, t = g.input(KG.SCALAR), g.input(KG.SCALAR)
    st = g.binop("ss", "mul", s, t)
    st2 = g.binop("ss", "sub", st, s)
    sn = g.unop("s", "neg", st2)
    a1 = g.binop("sp", "add", s, x)
    a2 = g.binop("sp", "sub", t, y)
    a3 = g.binop("sp", "subby", st, a1)
    a4 = g.binop("sp", "mul", sn, a2)
    a5 = g.unop("p", "pow", a4, 0xD1B54A32D192ED03)
    a6 = g.binop("pp", "add", a5, g.unop("p", "neg", a3))
    a7 = g.binop("pp", "mul", a6, acc)
    g.output(KG.ARRAY, a6)
    g.output(KG.ARRAY, a7, in_node=acc)
    g.output(KG.SCALAR, st2)
    return g


def random_graph(seed, form, n_ops=40):
    rng = random.Random(seed)
    g = KG.ArithGraph(form)
    arrays = [g.input(KG.ARRAY) for _ in range(3)]
    scalars = [g.input(KG.SCALAR) for _ in range(2)]
    for _ in range(n_ops):
        r = rng.random()
        if r < 0.5:
            op = rng.choice(("add", "sub", "mul"))
            node = g.binop("pp", op, rng.choice(arrays), rng.choice(arrays))
            arrays.append(node)
        elif r < 0.75:
            op = rng.choice(("add", "sub", "subby", "mul"))
            node = g.binop("sp", op, rng.choice(scalars), rng.choice(arrays))
            arrays.append(node)
        elif r < 0.85:
            op = rng.choice(("add", "sub", "mul"))
            node = g.binop("ss", op, rng.choice(scalars), rng.choice(scalars))
            scalars.append(node)
        elif r < 0.95:
            node = g.unop("p", "neg", rng.choice(arrays))
            arrays.append(node)
        else:
            node = g.unop("p", "pow", rng.choice(arrays), rng.randrange(1 << 64))
            arrays.append(node)
    g.output(KG.ARRAY, arrays[-1])
    g.output(KG.ARRAY, rng.choice(arrays[3:-1]))
    return g


def cases():
    out = [("gate_c5", gate_c5()), ("coef_mix", coef_mix())]
    for k in range(3):
        out.append((f"rand_lagrange{k}", random_graph(100 + k, KG.LAGRANGE)))
        out.append((f"rand_coef{k}", random_graph(200 + k, KG.COEF)))
    return out


def ops():
    return [KG.FusedOp(g, name) for name, g in cases()]


def build():
    return KG.build(ops(), OUT)


if __name__ == "__main__":
    print(build())
