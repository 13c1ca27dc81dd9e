"""HIP emitter for fused element-wise kernels over BN254 Fr.

The reference's compiler turns an ArithGraph (common/src/arith.rs: inputs,
arithmetic vertices, outputs; Lagrange or coefficient form) into CUDA text
and builds it with xmake (core/src/fused_kernels/kernel_gen.rs:53-600,
core/src/fused_kernels.rs:37-200).  This module takes the same graph,
emits one HIP kernel per FusedOp over the library's field arithmetic
(csrc/field.h, csrc/common.h), builds them with hipcc for gfx950 into one
shared object, and exposes each symbol with the generated-kernel signature

  hipError_t name(const ConstPolyPtr* vars, const PolyPtr* mut_vars,
                  unsigned long long len, const PolySlice* slice_info,
                  void* local_buffer, hipStream_t stream)

through FusedKernel, which mirrors the reference's runtime closure (argument
lists [stream, inputs...] / [arg_buffer, local_buffer, outputs...]).

Semantics kept from kernel_gen.rs:460-600:
  * Pp / Ss add, sub, mul; Sp (scalar lhs, polynomial rhs) add, sub, subby,
    mul; in coefficient form the scalar of an Sp add / sub / subby applies
    only at the index whose slice origin is 0 (elsewhere: rhs, -rhs, rhs);
    P / S neg, P pow(u64);
  * scalar inputs are read from one element, array inputs through their
    PolyPtr views at the loop index; a mutable input that an output stores
    back to (the output's in_node) is read from and written to that output's
    mut_vars slot, and is not listed among vars (gen_var_lists,
    common/src/arith.rs:501-527).
Differences: straight-line code in one kernel, registers left to the
compiler (no Belady spill schedule, so local_buffer is accepted and unused),
a grid of 8 work-groups of 256 per CU, and scalar outputs written once (by
index 0).
"""
import ctypes
import hashlib
import os
import shutil
import subprocess

from . import _lib
from . import build as _build

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")

SCALAR, ARRAY = "scalar", "array"
LAGRANGE, COEF = "lagrange", "coef"
BIN_KINDS = {"pp": ("add", "sub", "mul"), "ss": ("add", "sub", "mul"),
             "sp": ("add", "sub", "subby", "mul")}
UNR_KINDS = {"p": ("neg", "pow"), "s": ("neg",)}


class ArithGraph:
    """Vertices in creation order, each a tuple:
      ("input", typ, mutable)
      ("bin", kind, op, lhs, rhs)        kind pp / ss / sp, lhs the scalar for sp
      ("unr", kind, op, arg, power)      kind p / s, power for pow
      ("output", typ, store_node, in_node)
    poly_repr: "lagrange" or "coef" (common/src/typ.rs PolyType)."""

    def __init__(self, poly_repr=LAGRANGE):
        assert poly_repr in (LAGRANGE, COEF)
        self.poly_repr = poly_repr
        self.vertices = []
        self.inputs = []
        self.outputs = []

    def _add(self, v):
        self.vertices.append(v)
        return len(self.vertices) - 1

    def input(self, typ=ARRAY, mutable=False):
        assert typ in (SCALAR, ARRAY)
        i = self._add(("input", typ, bool(mutable)))
        self.inputs.append(i)
        return i

    def binop(self, kind, op, lhs, rhs):
        assert op in BIN_KINDS[kind], (kind, op)
        assert lhs < len(self.vertices) and rhs < len(self.vertices)
        return self._add(("bin", kind, op, lhs, rhs))

    def unop(self, kind, op, arg, power=None):
        assert op in UNR_KINDS[kind], (kind, op)
        if op == "pow":
            assert power is not None and 0 <= power < (1 << 64)
        return self._add(("unr", kind, op, arg, power))

    def output(self, typ, store_node, in_node=None):
        if in_node is not None:
            v = self.vertices[in_node]
            assert v[0] == "input" and v[2], "an in-place output stores to a mutable input"
        o = self._add(("output", typ, store_node, in_node))
        self.outputs.append(o)
        return o

    def typ(self, i):
        """ScalarArray or Scalar: the value type of a vertex."""
        v = self.vertices[i]
        if v[0] == "input":
            return v[1]
        if v[0] == "bin":
            return SCALAR if v[1] == "ss" else ARRAY
        if v[0] == "unr":
            return SCALAR if v[1] == "s" else ARRAY
        return v[1]

    def var_lists(self):
        """(vars, mut_vars) as (type, vertex) lists, gen_var_lists order:
        every output, then the inputs that no output stores back to."""
        mut_vars = [(self.vertices[o][1], o) for o in self.outputs]
        inplace = {self.vertices[o][3] for o in self.outputs if self.vertices[o][3] is not None}
        vars_ = [(self.vertices[i][1], i) for i in self.inputs if i not in inplace]
        return vars_, mut_vars

    def plain(self):
        """The graph as plain data (poly_repr, vertices), for checkers."""
        return self.poly_repr, list(self.vertices)


def topo_order(graph):
    """Kahn's order of the vertices (inputs first, outputs after what they
    store)."""
    n = len(graph.vertices)
    deps = [[] for _ in range(n)]
    for i, v in enumerate(graph.vertices):
        if v[0] == "bin":
            deps[i] = [v[3], v[4]]
        elif v[0] == "unr":
            deps[i] = [v[3]]
        elif v[0] == "output":
            deps[i] = [v[2]]
    indeg = [len(d) for d in deps]
    users = [[] for _ in range(n)]
    for i, d in enumerate(deps):
        for j in d:
            users[j].append(i)
    ready = [i for i in range(n) if indeg[i] == 0]
    order = []
    while ready:
        i = ready.pop(0)
        order.append(i)
        for u in users[i]:
            indeg[u] -= 1
            if indeg[u] == 0:
                ready.append(u)
    assert len(order) == n, "the graph has a cycle"
    return order


class FusedOp:
    """core/src/fused_kernels/kernel_gen.rs FusedOp: one graph, one symbol."""

    def __init__(self, graph: ArithGraph, name: str):
        assert name.isidentifier()
        self.graph = graph
        self.name = name
        self.vars, self.mut_vars = graph.var_lists()
        self.var_mapping = {vid: k for k, (_, vid) in enumerate(self.vars)}
        self.mut_var_mapping = {vid: k for k, (_, vid) in enumerate(self.mut_vars)}
        for o in graph.outputs:
            inn = graph.vertices[o][3]
            if inn is not None:
                self.mut_var_mapping[inn] = self.mut_var_mapping[o]

    def num_vars(self):
        return len(self.vars)

    def num_mut_vars(self):
        return len(self.mut_vars)

    def arg_buffer_bytes(self):
        return (self.num_vars() + self.num_mut_vars()) * ctypes.sizeof(_lib.PolyPtr)

    def temp_buffer_bytes(self):
        """local_buffer bytes: registers are left to the compiler."""
        return 0

    def source(self):
        g = self.graph
        order = topo_order(g)
        lines = []
        A = lines.append
        A(f"// generated by impeller_amd.kernel_gen: fused kernel {self.name}")
        A(f"// {g.poly_repr} form, {len(g.vertices)} vertices, "
          f"{self.num_vars()} vars, {self.num_mut_vars()} mut_vars")
        A('#include "common.h"')
        A("namespace imp_fused {")
        A("using namespace imp;")
        A("__device__ __forceinline__ Fr pw(const Fr& x, unsigned long long e) {")
        A("  Fr acc = Fr::one();")
        A("  for (int i = 63; i >= 0; i--) {")
        A("    acc = mul<FrParams>(acc, acc);")
        A("    if ((e >> i) & 1) acc = mul<FrParams>(acc, x);")
        A("  }")
        A("  return acc;")
        A("}")
        A(f"__global__ __launch_bounds__(256) void k_{self.name}(const ConstPolyPtr* vars, "
          "const PolyPtr* mut_vars, u64 len, PolySlice slice) {")

        def slot(vid):
            if vid in self.var_mapping:
                return f"reinterpret_cast<const PolyPtr*>(vars) + {self.var_mapping[vid]}"
            return f"mut_vars + {self.mut_var_mapping[vid]}"

        # views and scalar inputs are read once per thread
        for i in g.inputs:
            if g.vertices[i][1] == ARRAY:
                A(f"  const View v{i} = dev_view({slot(i)});")
            else:
                A(f"  const Fr t{i} = load_fe<Fr>(gld({slot(i)}).ptr);")
        for o in g.outputs:
            if g.vertices[o][1] == ARRAY:
                A(f"  const View w{o} = dev_view(mut_vars + {self.mut_var_mapping[o]});")
        need_origin = g.poly_repr == COEF and any(
            v[0] == "bin" and v[1] == "sp" and v[2] in ("add", "sub", "subby")
            for v in g.vertices)
        A("  const u64 stride = (u64)gridDim.x * blockDim.x;")
        A("  for (u64 idx = (u64)blockIdx.x * blockDim.x + threadIdx.x; idx < len; "
          "idx += stride) {")
        if need_origin:
            A("    const bool origin = slice.len == 0 ? idx == 0 : "
              "(slice.begin + idx) % slice.len == 0;")
        for i in order:
            v = g.vertices[i]
            if v[0] == "input":
                if v[1] == ARRAY:
                    A(f"    const Fr t{i} = load_fe<Fr>(v{i}.base + 8 * view_phys(v{i}, idx));")
            elif v[0] == "bin":
                _, kind, op, lhs, rhs = v
                a, b = f"t{lhs}", f"t{rhs}"
                if op == "mul":
                    e = f"mul<FrParams>({a}, {b})"
                elif kind == "sp" and g.poly_repr == COEF:
                    # kernel_gen.rs:525-555: the scalar only at the origin
                    e = {"add": f"origin ? add<FrParams>({a}, {b}) : {b}",
                         "sub": f"origin ? sub<FrParams>({a}, {b}) : neg<FrParams>({b})",
                         "subby": f"origin ? sub<FrParams>({b}, {a}) : {b}"}[op]
                else:
                    e = {"add": f"add<FrParams>({a}, {b})", "sub": f"sub<FrParams>({a}, {b})",
                         "subby": f"sub<FrParams>({b}, {a})"}[op]
                A(f"    const Fr t{i} = {e};")
            elif v[0] == "unr":
                _, kind, op, arg, power = v
                e = f"neg<FrParams>(t{arg})" if op == "neg" else f"pw(t{arg}, {power}ull)"
                A(f"    const Fr t{i} = {e};")
            else:
                _, typ, store, _ = v
                if typ == ARRAY:
                    A(f"    store_fe<Fr>(w{i}.base + 8 * view_phys(w{i}, idx), t{store});")
                else:
                    A(f"    if (idx == 0) store_fe<Fr>(mut_vars[{self.mut_var_mapping[i]}].ptr, "
                      f"t{store});")
        A("  }")
        A("}")
        A("}  // namespace imp_fused")
        A(f'extern "C" hipError_t {self.name}(const ConstPolyPtr* vars, const PolyPtr* mut_vars,')
        A("                              unsigned long long len, const PolySlice* slice_info,")
        A("                              void* local_buffer, hipStream_t stream) {")
        A("  (void)local_buffer;")
        A("  if (len == 0) return hipSuccess;")
        A("  if (vars == nullptr && " + str(self.num_vars()) + " > 0) return hipErrorInvalidValue;")
        A("  if (mut_vars == nullptr) return hipErrorInvalidValue;")
        A("  PolySlice slice;")
        A("  slice.begin = slice_info ? slice_info->begin : 0;")
        A("  slice.slice_len = slice_info ? slice_info->slice_len : len;")
        A("  slice.len = slice_info ? slice_info->len : len;")
        A("  int dev = 0, cus = 256;")
        A("  if (hipGetDevice(&dev) != hipSuccess ||")
        A("      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) "
          "!= hipSuccess)")
        A("    cus = 256;")
        A("  const unsigned long long want = (len + 255) / 256, cap = (unsigned long long)cus * 8;")
        A("  const unsigned grid = (unsigned)(want < cap ? want : cap);")
        A(f"  hipLaunchKernelGGL(imp_fused::k_{self.name}, dim3(grid), dim3(256), 0, stream, vars,")
        A("                     mut_vars, (imp::u64)len, slice);")
        A("  return hipGetLastError();")
        A("}")
        return "\n".join(lines) + "\n"

    def meta(self, lib_path):
        """runtime/src/functions.rs FusedKernelMeta: name, counts, library."""
        return {"name": self.name, "num_vars": self.num_vars(),
                "num_mut_vars": self.num_mut_vars(), "lib_path": lib_path}


LIB_NAME = "libfused_kernels.so"


def gen(ops, target_dir):
    """Writes one .hip source per op into target_dir; returns their paths."""
    os.makedirs(target_dir, exist_ok=True)
    paths = []
    for op in ops:
        path = os.path.join(target_dir, f"{op.name}.hip")
        text = op.source()
        if not (os.path.exists(path) and open(path).read() == text):
            with open(path, "w") as f:
                f.write(text)
        paths.append(path)
    return paths


def build(ops, target_dir, lib_name=LIB_NAME):
    """Emits and builds every op into target_dir/lib_name (gfx950).  A
    library whose recorded source digest matches is not rebuilt."""
    srcs = gen(ops, target_dir)
    digest = hashlib.sha256()
    for p in srcs + [os.path.join(CSRC, f) for f in ("common.h", "field.h")]:
        digest.update(open(p, "rb").read())
    out = os.path.join(target_dir, lib_name)
    stamp = out + ".sha256"
    if os.path.exists(out) and os.path.exists(stamp) and open(stamp).read() == digest.hexdigest():
        return out
    hipcc = _build.HIPCC

    def one(src):
        obj = src + ".o"
        cmd = [hipcc, *_build.COMMON, "-I", CSRC, "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
        # every device body inlined (the library's own rule, csrc build check)
        bad = _build._device_functions(src, include_dirs=(CSRC,))
        if bad:
            raise RuntimeError(f"out-of-line device functions in {src}: {bad}")
        return obj

    import concurrent.futures
    workers = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "8"))))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(one, srcs))
    tmp = out + ".tmp"
    r = subprocess.run([hipcc, "-shared", "-fPIC", f"--offload-arch={_build.ARCH}", *objs,
                        "-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, out)
    with open(stamp, "w") as f:
        f.write(digest.hexdigest())
    return out


class FusedKernel:
    """core/src/fused_kernels.rs:95-200: mut_var = [arg_buffer, local_buffer,
    outputs...], var = [stream, inputs...]; inputs and outputs in the op's
    vars / mut_vars order; len is the arrays' common length (1 when every
    argument is a scalar)."""

    def __init__(self, meta):
        self.meta = meta
        self.lib = ctypes.CDLL(os.path.join(meta["lib_path"], LIB_NAME)
                               if os.path.isdir(meta["lib_path"]) else meta["lib_path"])
        fn = getattr(self.lib, meta["name"])
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong,
                       ctypes.POINTER(_lib.PolySlice), ctypes.c_void_p, ctypes.c_void_p]
        self.fn = fn

    @staticmethod
    def _ptr(v):
        from .runtime import Scalar
        if isinstance(v, Scalar):
            return _lib.PolyPtr(v.data_ptr(), 1, 0, 0, 1), 0
        return v.poly_ptr(), v.len

    def get_fn(self):
        meta = self.meta

        def f(mut_var, var, gpu_mapping=None, poly_slice=None):
            import numpy as np
            import torch
            assert len(var) - 1 == meta["num_vars"]
            assert len(mut_var) - 2 == meta["num_mut_vars"]
            stream = var[0]
            arg_buffer, local_buffer = mut_var[0], mut_var[1]
            length = 0
            ptrs = []
            for v in list(var[1:]) + list(mut_var[2:]):
                p, n = self._ptr(v)
                if n:
                    assert length in (0, 1) or length == n, "arrays of different lengths"
                    length = n
                elif length == 0:
                    length = 1
                ptrs.append(p)
            assert length > 0
            sl = poly_slice or _lib.PolySlice(0, length, length)
            host = (_lib.PolyPtr * len(ptrs))(*ptrs)
            raw = np.frombuffer(bytes(host), dtype=np.uint8)
            dst = arg_buffer.values if hasattr(arg_buffer, "values") else arg_buffer
            assert dst.numel() * dst.element_size() >= raw.nbytes
            with torch.cuda.stream(stream.torch_stream):
                dst.view(torch.uint8)[:raw.nbytes].copy_(torch.from_numpy(raw.copy()),
                                                        non_blocking=False)
            base = dst.data_ptr()
            lb = None
            if local_buffer is not None:
                lb = local_buffer.data_ptr()
            _lib.check(meta["name"], self.fn(
                base, base + meta["num_vars"] * ctypes.sizeof(_lib.PolyPtr), length,
                ctypes.byref(sl), lb, stream.raw()))
        return f


def hipcc_available():
    return shutil.which(_build.HIPCC) is not None or os.path.exists(_build.HIPCC)
