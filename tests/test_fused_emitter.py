"""The HIP fused-kernel emitter on the CPU: its checker pinned against
independent restatements, the emitted sources built for gfx950 with every
symbol exported, and the emitted kernels run through the CPU launch
emulation (tests/emu) against the checker on the cases of
tests/fused_cases.py (the C5 gate, a coefficient-form mix with in-place and
scalar outputs, seeded random graphs), with rotated views and slices."""
import ctypes
import os
import shutil

import numpy as np
import pytest

from impeller_amd import _lib
from impeller_amd import kernel_gen as KG
from oracle import fused_ref as FR
from oracle import oracle as O
from tests import fused_cases as F

r = O.R_MOD


def test_checker_matches_the_gate():
    """fused_ref on the gate graph equals the C checker's gate."""
    g = F.gate_c5()
    n = 64
    ins = [O.gen_scalars(n, 10 + i) for i in range(6)]
    inputs = {g.inputs[i]: ins[i] for i in range(6)}
    inputs[g.inputs[6]] = O.one()
    out = FR.evaluate(*g.plain(), inputs, n)
    assert (out[g.outputs[0]] == O.gate(*ins)).all()


def test_checker_coef_form_closed_form():
    """Coefficient form: s + p adds s only at the slice origin, s - p is -p
    elsewhere, p - s (subby) is p elsewhere (kernel_gen.rs:525-555)."""
    g = KG.ArithGraph(KG.COEF)
    p, s = g.input(KG.ARRAY), g.input(KG.SCALAR)
    o1 = g.output(KG.ARRAY, g.binop("sp", "add", s, p))
    o2 = g.output(KG.ARRAY, g.binop("sp", "sub", s, p))
    o3 = g.output(KG.ARRAY, g.binop("sp", "subby", s, p))
    n, begin, total = 6, 10, 12   # origin (begin + i) % total == 0 at i = 2
    vals = [5, 6, 7, 8, 9, 10]
    pv = O.ints_to_array([O.to_mont_int(v) for v in vals])
    sv = O.ints_to_array([O.to_mont_int(100)])[0]
    out = FR.evaluate(*g.plain(), {p: pv, s: sv}, n, begin, total)

    def ints(a):
        return [O.from_mont_int(v) for v in O.array_to_ints(a)]

    assert ints(out[o1]) == [5, 6, 107, 8, 9, 10]
    assert ints(out[o2]) == [(-v) % r if i != 2 else 100 - 7 for i, v in enumerate(vals)]
    assert ints(out[o3]) == [v if i != 2 else (7 - 100) % r for i, v in enumerate(vals)]


def test_var_lists_follow_gen_var_lists():
    g = F.coef_mix()
    op = KG.FusedOp(g, "coef_mix")
    # every output first (mut_vars), then the inputs no output stores back to
    assert [v for _, v in op.mut_vars] == g.outputs
    inplace = g.vertices[g.outputs[1]][3]
    assert inplace not in [v for _, v in op.vars]
    assert op.mut_var_mapping[inplace] == 1
    assert op.arg_buffer_bytes() == (op.num_vars() + op.num_mut_vars()) * 40


@pytest.mark.skipif(not KG.hipcc_available(), reason="no hipcc")
def test_emitted_library_builds_for_gfx950(tmp_path):
    path = KG.build(F.ops(), str(tmp_path))
    L = ctypes.CDLL(path)
    for op in F.ops():
        assert hasattr(L, op.name), op.name


def _emu_lib():
    from tests.emu import build as EB
    srcs = KG.gen(F.ops(), os.path.join(F.OUT, "emu_src"))
    out = os.path.join(F.OUT, "emu", "libfused_emu.so")
    newest = max(os.path.getmtime(s) for s in srcs + [os.path.join(KG.CSRC, "field.h"),
                                                      os.path.join(KG.CSRC, "common.h")])
    if not os.path.exists(out) or os.path.getmtime(out) < newest:
        EB.build_extra(srcs, out)
    return ctypes.CDLL(out)


def random_inputs(g, n, seed):
    rng = np.random.default_rng(seed)
    vals = {}
    for k, i in enumerate(g.inputs):
        if g.vertices[i][1] == KG.ARRAY:
            vals[i] = O.gen_scalars(n, 1000 * seed + k)
            vals[i][rng.integers(0, n, 3)] = 0
        else:
            vals[i] = O.gen_scalars(1, 1000 * seed + k)[0]
    return vals


def run_emulated(L, op, inputs, n, rotate=3, slice_=None):
    """Host buffers stand in for device memory; arrays are placed under a
    rotation and read back in logical order."""
    g = op.graph
    keep = []

    def view_of(arr):
        phys = np.zeros_like(arr)
        rr = rotate % n
        phys[(n - rr + np.arange(n)) % n if rr else np.arange(n)] = arr
        buf = np.ascontiguousarray(phys.reshape(-1))
        keep.append(buf)
        return buf, _lib.PolyPtr(buf.ctypes.data, n, rr, 0, n)

    vars_, muts, out_bufs = [], [], {}
    for _, vid in op.vars:
        if g.vertices[vid][1] == KG.ARRAY:
            vars_.append(view_of(inputs[vid])[1])
        else:
            b = np.ascontiguousarray(inputs[vid].reshape(-1))
            keep.append(b)
            vars_.append(_lib.PolyPtr(b.ctypes.data, 1, 0, 0, 1))
    for _, vid in op.mut_vars:
        v = g.vertices[vid]
        if v[1] == KG.ARRAY:
            init = inputs[v[3]] if v[3] is not None else np.full((n, 8), 0xA5A5A5A5, np.uint32)
            buf, p = view_of(init)
            out_bufs[vid] = (buf, True)
        else:
            buf = np.full(8, 0xA5A5A5A5, np.uint32)
            keep.append(buf)
            p = _lib.PolyPtr(buf.ctypes.data, 1, 0, 0, 1)
            out_bufs[vid] = (buf, False)
        muts.append(p)
    va = (_lib.PolyPtr * max(1, len(vars_)))(*vars_)
    ma = (_lib.PolyPtr * len(muts))(*muts)
    sl = slice_ or _lib.PolySlice(0, n, n)
    fn = getattr(L, op.name)
    fn.restype = ctypes.c_int
    rc = fn(ctypes.cast(va, ctypes.c_void_p), ctypes.cast(ma, ctypes.c_void_p),
            ctypes.c_ulonglong(n), ctypes.byref(sl), None, None)
    assert rc == 0
    res = {}
    rr = rotate % n
    idx = (n - rr + np.arange(n)) % n if rr else np.arange(n)
    for vid, (buf, is_arr) in out_bufs.items():
        res[vid] = buf.reshape(-1, 8)[idx] if is_arr else buf.reshape(8)
    return res


@pytest.mark.skipif(shutil.which("/opt/rocm/llvm/bin/clang++") is None
                    and shutil.which("clang++") is None, reason="no host clang++")
@pytest.mark.parametrize("case", [name for name, _ in F.cases()])
def test_emitted_kernels_on_cpu_emulation(case):
    L = _emu_lib()
    op = {o.name: o for o in F.ops()}[case]
    n = 300
    inputs = random_inputs(op.graph, n, len(case))
    # coefficient form: a slice whose origin falls at index 100
    sl = (_lib.PolySlice(4 * n - 100, n, 4 * n) if op.graph.poly_repr == KG.COEF
          else _lib.PolySlice(0, n, n))
    got = run_emulated(L, op, inputs, n, rotate=7, slice_=sl)
    exp = FR.evaluate(*op.graph.plain(), inputs, n, sl.begin, sl.len)
    for o in op.graph.outputs:
        assert (got[o] == exp[o]).all(), (case, o)
