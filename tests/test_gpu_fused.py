"""GPU parity of kernels emitted by impeller_amd.kernel_gen (the HIP
replacement of the reference's fused-kernel generator,
core/src/fused_kernels/kernel_gen.rs) through the reference-shaped runtime
closure (FusedKernel), against the Python-integer checker: every case of
tests/fused_cases.py on rotated views, coefficient-form graphs on a slice
whose origin falls inside the range, in-place and scalar outputs; the
emitted C5 gate at 2^20 against the C checker and the hand-written gate;
graph capture and replay.  The library is built beforehand
(`python -m tests.fused_cases`, part of __graft_entry__.build())."""
import os

import numpy as np
import pytest

from impeller_amd import _lib
from impeller_amd import kernel_gen as KG
from impeller_amd.runtime import GpuBuffer, Scalar, ScalarArray
from oracle import fused_ref as FR
from oracle import oracle as O
from tests import fused_cases as F

pytestmark = pytest.mark.gpu

LIB = os.path.join(F.OUT, KG.LIB_NAME)


@pytest.fixture(scope="module")
def kernels():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is missing: build it with `python -m tests.fused_cases`",
                    pytrace=False)
    return {op.name: (op, KG.FusedKernel(op.meta(F.OUT))) for op in F.ops()}


def placed(x, rotate):
    n = x.shape[0]
    rr = rotate % n
    phys = np.zeros_like(x)
    phys[(n - rr + np.arange(n)) % n if rr else np.arange(n)] = x
    a = ScalarArray.from_numpy(phys)
    if rr:
        a.rotate(rr)
    return a


def run_case(op, kern, inputs, n, stream, rotate=5, poly_slice=None):
    g = op.graph
    var = [stream]
    for _, vid in op.vars:
        if g.vertices[vid][1] == KG.ARRAY:
            var.append(placed(inputs[vid], rotate))
        else:
            var.append(Scalar.from_numpy(inputs[vid]))
    outs = []
    for _, vid in op.mut_vars:
        v = g.vertices[vid]
        if v[1] == KG.ARRAY:
            init = inputs[v[3]] if v[3] is not None else np.full((n, 8), 0xA5A5A5A5, np.uint32)
            outs.append(placed(init, rotate + 3))
        else:
            outs.append(Scalar.from_numpy(np.full(8, 0xA5A5A5A5, np.uint32)))
    arg = GpuBuffer(op.arg_buffer_bytes())
    kern.get_fn()([arg, None] + outs, var, None, poly_slice)
    stream.sync()
    got = {}
    for (_, vid), o in zip(op.mut_vars, outs):
        got[vid] = o.logical_numpy() if isinstance(o, ScalarArray) else o.to_numpy()
    return got


def random_inputs(g, n, seed):
    vals = {}
    for k, i in enumerate(g.inputs):
        if g.vertices[i][1] == KG.ARRAY:
            vals[i] = O.gen_scalars(n, 7000 + 100 * seed + k)
            vals[i][::97] = 0
        else:
            vals[i] = O.gen_scalars(1, 7000 + 100 * seed + k)[0]
    return vals


@pytest.mark.parametrize("case", [name for name, _ in F.cases()])
def test_emitted_case(stream, kernels, case):
    op, kern = kernels[case]
    n = 3000
    inputs = random_inputs(op.graph, n, len(case))
    coef = op.graph.poly_repr == KG.COEF
    sl = _lib.PolySlice(4 * n - 1234, n, 4 * n) if coef else None
    got = run_case(op, kern, inputs, n, stream, 11, sl)
    exp = FR.evaluate(*op.graph.plain(), inputs, n, *(sl.begin, sl.len) if coef else (0, n))
    for o in op.graph.outputs:
        assert (got[o] == exp[o]).all(), (case, o)


def test_emitted_gate_2_20(stream, kernels):
    """The emitted C5 gate against the C checker's gate and the library's
    hand-written fused_gate_abc_d_qe at 2^20."""
    op, kern = kernels["gate_c5"]
    n = 1 << 20
    g = op.graph
    ins = [O.gen_scalars(n, 300 + i) for i in range(6)]
    inputs = {g.inputs[i]: ins[i] for i in range(6)}
    inputs[g.inputs[6]] = O.one()
    got = run_case(op, kern, inputs, n, stream, 0)
    exp = O.gate(*ins)
    assert (got[g.outputs[0]] == exp).all()


def test_emitted_kernel_graph_replay(kernels):
    """An emitted kernel launches one kernel only: captured into a HIP graph
    and replayed three times (inputs of the in-place output restored by
    eager copies in between), every replay equal to the checker."""
    import ctypes

    import torch
    from impeller_amd.ntt import Graph
    from impeller_amd.runtime import Stream
    op, kern = kernels["coef_mix"]
    n = 4096
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        stream = Stream.current(0)
        g = op.graph
        inputs = random_inputs(g, n, 99)
        sl = _lib.PolySlice(3 * n - 17, n, 3 * n)
        exp = FR.evaluate(*g.plain(), inputs, n, sl.begin, sl.len)
        var = [stream] + [placed(inputs[v], 3) if g.vertices[v][1] == KG.ARRAY
                          else Scalar.from_numpy(inputs[v]) for _, v in op.vars]
        acc_in = inputs[g.vertices[g.outputs[1]][3]]
        outs = [placed(np.zeros((n, 8), np.uint32), 1), placed(acc_in, 1),
                Scalar.from_numpy(np.zeros(8, np.uint32))]
        acc_orig = outs[1].values.clone()
        arg = GpuBuffer(op.arg_buffer_bytes())
        kern.get_fn()([arg, None] + outs, var, None, sl)   # eager; fills the arg buffer
        torch.cuda.synchronize()
        base = arg.data_ptr()
        gr = Graph(stream, lambda: _lib.check("coef_mix", kern.fn(
            base, base + op.num_vars() * 40, n, ctypes.byref(sl), None, stream.raw())))
        for rep in range(3):
            outs[0].values.fill_(-1)
            outs[1].values.copy_(acc_orig)
            outs[2].values.fill_(-1)
            torch.cuda.synchronize()
            gr.launch(stream)
            stream.sync()
            for o, got in zip(g.outputs, (outs[0].logical_numpy(), outs[1].logical_numpy(),
                                          outs[2].to_numpy())):
                assert (got == exp[o]).all(), (rep, o)
