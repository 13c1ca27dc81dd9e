// misc.hip -- fused gate kernel, seeded input generators and the small
// device-arithmetic entry points used to test the field and curve code.
#include "common.h"
#include "field29.h"
#include "timing.h"

namespace imp {

// ---------------------------------------------------------------- fused gate
// out = a*b*c - d + q*e*(1 - q) over Fr, one element per index, grid-stride
// (the generated kernels' loop, core/src/fused_kernels/kernel_gen.rs:281-304).
// Operands are PolyPtr views (make_slice_iter semantics, iter.cuh:309-327).
// The gate is a Lagrange-form expression (its constant 1 applies at every
// index), so the slice, which only places Coef-form constants
// (kernel_gen.rs:525-555), does not change its values.
struct GateViews {
  View in[6];
  View out;
};

__device__ __forceinline__ void gate_at(const GateViews& g, u64 i) {
  Fr a = load_fe<Fr>(g.in[0].base + 8 * view_phys(g.in[0], i));
  Fr b = load_fe<Fr>(g.in[1].base + 8 * view_phys(g.in[1], i));
  Fr c = load_fe<Fr>(g.in[2].base + 8 * view_phys(g.in[2], i));
  Fr d = load_fe<Fr>(g.in[3].base + 8 * view_phys(g.in[3], i));
  Fr q = load_fe<Fr>(g.in[4].base + 8 * view_phys(g.in[4], i));
  Fr e = load_fe<Fr>(g.in[5].base + 8 * view_phys(g.in[5], i));
  Fr t1 = sub<FrParams>(mul<FrParams>(mul<FrParams>(a, b), c), d);
  Fr t2 = mul<FrParams>(mul<FrParams>(q, e), sub<FrParams>(Fr::one(), q));
  store_fe<Fr>(g.out.base + 8 * view_phys(g.out, i), add<FrParams>(t1, t2));
}

__global__ __launch_bounds__(256) void k_gate(GateViews g, u64 len) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) gate_at(g, i);
}

// The reference calling convention: the views live in device memory
// (fused_kernels.rs:172-188 copies them there), so each thread reads them.
__global__ __launch_bounds__(256) void k_gate_dev(const ConstPolyPtr* vars, const PolyPtr* mut_vars,
                                                  u64 len) {
  GateViews g;
#pragma unroll
  for (int k = 0; k < 6; k++) g.in[k] = dev_view(reinterpret_cast<const PolyPtr*>(vars) + k);
  g.out = dev_view(mut_vars);
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) gate_at(g, i);
}

static hipError_t gate_grid(unsigned long long len, u32* blocks) {
  int dev = 0, cus = 256;
  IMP_CHECK(hipGetDevice(&dev));
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  u64 b = (len + 255) / 256, cap = (u64)cus * 8;
  *blocks = (u32)(b > cap ? cap : b);
  return hipSuccess;
}

// ---------------------------------------------------------------- generators
__global__ void k_gen_scalars(u32* out, u64 n, int kind, u64 seed, u64 start) {
  u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  u64 i = start + t;
  Fr v;
  if (kind == 0) {
    v = gen_uniform_fr(seed, i);
  } else {
    u64 h = smix(smix(seed) ^ (i * 0xD1B54A32D192ED03ull) ^ 0x5DEECE66Dull);
    if (kind == 1) {
      v = (h % 10) == 0 ? Fr::one() : Fr::zero();
    } else {
      v = Fr::zero();
      v.v[0] = (u32)(h & 0xFFFFu);
      v = from_number<FrParams>(v);
    }
  }
  store_fe<Fr>(out + 8 * t, v);
}

// P_i = to_number(k_i) * G, k_i = uniform Fr; G = (1, 2)
__global__ void k_gen_bases(u32* out, u64 n, u64 seed, u64 start) {
  u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  Fr k = to_number<FrParams>(gen_uniform_fr(seed, start + t));
  G1Aff g;
  g.x = Fq::zero();
  g.x.v[0] = 1;
  g.x = from_number<FqParams>(g.x);
  g.y = Fq::zero();
  g.y.v[0] = 2;
  g.y = from_number<FqParams>(g.y);
  store_aff(out + 16 * t, xyzz_to_aff(scalar_mul(g, k.v), true));
}

// ------------------------------------------------------- arithmetic probes
template <class P, class P29>
__global__ void k_field_op(int op, u32* r, const u32* a, const u32* b, u64 n) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typedef Fp<P> F;
  F x = load_fe<F>(a + 8 * i);
  F y = b ? load_fe<F>(b + 8 * i) : F::zero();
  F z;
  switch (op) {
    case 0: z = mul<P>(x, y); break;
    case 1: z = add<P>(x, y); break;
    case 2: z = sub<P>(x, y); break;
    case 3: z = neg<P>(x); break;
    case 4: z = from_number<P>(x); break;
    case 5: z = to_number<P>(x); break;
    case 6: z = pow_big<P>(x, y.v); break;
    case 7: z = inv<P>(x); break;
    // the 9 x 29-bit forms, entered and left through the R-form
    case 8: z = from_r29<P29>(mul29<P29>(to_r29<P29>(x), to_r29<P29>(y))); break;
    case 9: z = from_r29<P29>(sub29<P29, 2>(to_r29<P29>(x), to_r29<P29>(y))); break;
    case 11: z = from_r29<P29>(sqr29<P29>(to_r29<P29>(x))); break;
    default: {
      F29<P29> ty = to_r29<P29>(y);
      z = from_r29<P29>(sub3_29<P29, 6>(to_r29<P29>(x), ty, ty, ty));
    }
  }
  store_fe<F>(r + 8 * i, z);
}

__global__ void k_g1_op(int op, u32* r, const u32* a, const u32* b, u64 n) {
  const u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (op >= 9) {
    // the group-of-four forms: lanes 4i .. 4i + 3 work on item i together
    const u64 i = g >> 2;
    const u32 q = (u32)(g & 3);
    if (i >= n) return;
    const G1Xyzz29 x = xyzz_to_29(load_xyzz(a + 32 * i));
    const G1Xyzz29 res = op == 9 ? xyzz29_add_quad(x, xyzz_to_29(load_xyzz(b + 32 * i)), q)
                                 : xyzz29_dbl_quad(x, q);
    if (q == 0) store_xyzz(r + 32 * i, xyzz_from_29(res));
    return;
  }
  const u64 i = g;
  if (i >= n) return;
  switch (op) {
    case 0: store_xyzz(r + 32 * i, xyzz_madd(load_xyzz(a + 32 * i), load_aff(b + 16 * i))); break;
    case 1: store_xyzz(r + 32 * i, xyzz_add(load_xyzz(a + 32 * i), load_xyzz(b + 32 * i))); break;
    case 2: store_xyzz(r + 32 * i, xyzz_dbl(load_xyzz(a + 32 * i))); break;
    case 3: store_aff(r + 16 * i, xyzz_to_aff(load_xyzz(a + 32 * i))); break;
    // the 29-bit forms, entered and left through the R-form XYZZ
    case 5: {
      G1Aff q = load_aff(b + 16 * i);
      G1Aff29 q29;
      q29.x = to_r29<Q29>(q.x);
      q29.y = to_r29<Q29>(q.y);
      store_xyzz(r + 32 * i, xyzz_from_29(xyzz29_madd(xyzz_to_29(load_xyzz(a + 32 * i)), q29)));
      break;
    }
    case 6:
      store_xyzz(r + 32 * i, xyzz_from_29(xyzz29_add(xyzz_to_29(load_xyzz(a + 32 * i)),
                                                     xyzz_to_29(load_xyzz(b + 32 * i)))));
      break;
    case 7: store_xyzz(r + 32 * i, xyzz_from_29(xyzz29_dbl(xyzz_to_29(load_xyzz(a + 32 * i))))); break;
    case 8: {
      // table format round trip, then -P through the table path
      u32 tab[16];
      aff_to_table(tab, load_aff(a + 16 * i));
      G1Aff29 t = aff29_neg(load_aff_table(tab));
      G1Aff o;
      o.x = from_r29<Q29>(t.x);
      o.y = from_r29<Q29>(t.y);
      store_aff(r + 16 * i, o);
      break;
    }
    default: {
      Fr k = load_fe<Fr>(b + 8 * i);
      store_aff(r + 16 * i, xyzz_to_aff(scalar_mul(load_aff(a + 16 * i), k.v)));
    }
  }
}

// two independent product chains per thread (their results are consumed so
// the loop is not removed)
template <int FORM>
__global__ __launch_bounds__(256) void k_mul_rate(u32* out, u32 iters) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  u32 acc = 0;
  if (FORM == 0) {
    Fr a = Fr::one(), b = Fr::one(), c = Fr::r2();
    a.v[0] ^= t;
    b.v[1] ^= t;
    for (u32 i = 0; i < iters; i++) {
      a = mul<FrParams>(a, c);
      b = mul<FrParams>(b, c);
    }
    for (int j = 0; j < 8; j++) acc ^= a.v[j] ^ b.v[j];
  } else {
    Fq29 a = Fq29::one(), b = Fq29::one(), c = consts29<Fq29Params>(Fq29Params::kin);
    a.v[0] ^= t & 0xffffu;
    b.v[1] ^= t & 0xffffu;
    for (u32 i = 0; i < iters; i++) {
      a = mul29<Fq29Params>(a, c);
      b = mul29<Fq29Params>(b, c);
    }
    for (int j = 0; j < 9; j++) acc ^= a.v[j] ^ b.v[j];
  }
  out[t] = acc;
}

}  // namespace imp

using namespace imp;

extern "C" hipError_t impeller_mont_mul_rate(int form, unsigned int blocks, unsigned int iters,
                                             uint32_t* out_d, hipStream_t stream) {
  IMP_REQUIRE(out_d && blocks > 0 && (form == 0 || form == 1), "mul rate args");
  if (form == 0)
    hipLaunchKernelGGL(k_mul_rate<0>, dim3(blocks), dim3(256), 0, stream, out_d, iters);
  else
    hipLaunchKernelGGL(k_mul_rate<1>, dim3(blocks), dim3(256), 0, stream, out_d, iters);
  return hipGetLastError();
}

extern "C" hipError_t fused_gate_abc_d_qe(const ConstPolyPtr* vars, const PolyPtr* mut_vars,
                                          unsigned long long len, const PolySlice* slice_info,
                                          void* local_buffer, hipStream_t stream) {
  (void)slice_info;  // Lagrange-form gate: no Coef-form constants to place
  (void)local_buffer;  // no spills
  IMP_REQUIRE(vars && mut_vars, "null pointer");
  if (len == 0) return hipSuccess;
  u32 blocks = 0;
  IMP_CHECK(gate_grid(len, &blocks));
  {
    TimedScope ts(K_GATE, stream);
    hipLaunchKernelGGL(k_gate_dev, dim3(blocks), dim3(256), 0, stream, vars, mut_vars, (u64)len);
  }
  return hipGetLastError();
}

extern "C" hipError_t impeller_gate(const uint32_t* const* inputs, uint32_t* out,
                                    unsigned long long len, hipStream_t stream) {
  IMP_REQUIRE(inputs && out, "null pointer");
  GateViews g;
  for (int k = 0; k < 6; k++) {
    IMP_REQUIRE(inputs[k], "gate input");
    g.in[k] = View{const_cast<u32*>(inputs[k]), 0, len ? len : 1};
  }
  g.out = View{out, 0, len ? len : 1};
  if (len == 0) return hipSuccess;
  u32 blocks = 0;
  IMP_CHECK(gate_grid(len, &blocks));
  {
    TimedScope ts(K_GATE, stream);
    hipLaunchKernelGGL(k_gate, dim3(blocks), dim3(256), 0, stream, g, (u64)len);
  }
  return hipGetLastError();
}

extern "C" hipError_t impeller_gen_scalars(uint32_t* out_d, unsigned long long n, int kind,
                                           unsigned long long seed, unsigned long long start,
                                           hipStream_t stream) {
  IMP_REQUIRE(out_d || n == 0, "null pointer");
  IMP_REQUIRE(kind >= 0 && kind <= 2, "kind");
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gen_scalars, dim3((u32)((n + 255) / 256)), dim3(256), 0, stream, out_d,
                     (u64)n, kind, (u64)seed, (u64)start);
  return hipGetLastError();
}

extern "C" hipError_t impeller_gen_bases(uint32_t* out_d, unsigned long long n,
                                         unsigned long long seed, unsigned long long start,
                                         hipStream_t stream) {
  IMP_REQUIRE(out_d || n == 0, "null pointer");
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gen_bases, dim3((u32)((n + 127) / 128)), dim3(128), 0, stream, out_d,
                     (u64)n, (u64)seed, (u64)start);
  return hipGetLastError();
}

extern "C" hipError_t impeller_field_op(int field, int op, uint32_t* r_d, const uint32_t* a_d,
                                        const uint32_t* b_d, unsigned long long n,
                                        hipStream_t stream) {
  IMP_REQUIRE(r_d && a_d, "null pointer");
  IMP_REQUIRE(op >= 0 && op <= 11 && (field == 0 || field == 1), "op/field");
  IMP_REQUIRE(b_d || op == 3 || op == 4 || op == 5 || op == 7 || op == 11, "b operand");
  if (n == 0) return hipSuccess;
  dim3 grid((u32)((n + 127) / 128));
  if (field == 0)
    hipLaunchKernelGGL((k_field_op<FrParams, Fr29Params>), grid, dim3(128), 0, stream, op, r_d,
                       a_d, b_d, (u64)n);
  else
    hipLaunchKernelGGL((k_field_op<FqParams, Fq29Params>), grid, dim3(128), 0, stream, op, r_d,
                       a_d, b_d, (u64)n);
  return hipGetLastError();
}

extern "C" hipError_t impeller_g1_op(int op, uint32_t* r_d, const uint32_t* a_d,
                                     const uint32_t* b_d, unsigned long long n,
                                     hipStream_t stream) {
  IMP_REQUIRE(r_d && a_d, "null pointer");
  IMP_REQUIRE(op >= 0 && op <= 10, "op");
  IMP_REQUIRE(b_d || op == 2 || op == 3 || op == 7 || op == 8 || op == 10, "b operand");
  IMP_REQUIRE(n < (1ull << 36), "n");
  if (n == 0) return hipSuccess;
  const u64 lanes = op >= 9 ? 4 * n : n;
  hipLaunchKernelGGL(k_g1_op, dim3((u32)((lanes + 63) / 64)), dim3(64), 0, stream, op, r_d, a_d,
                     b_d, (u64)n);
  return hipGetLastError();
}

extern "C" hipError_t impeller_graph_begin(hipStream_t stream) {
  return hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
}

extern "C" hipError_t impeller_graph_end(hipStream_t stream, void** graph_exec) {
  IMP_REQUIRE(graph_exec, "null pointer");
  hipGraph_t g = nullptr;
  IMP_CHECK(hipStreamEndCapture(stream, &g));
  hipGraphExec_t e = nullptr;
  hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  IMP_CHECK(r);
  *graph_exec = (void*)e;
  return hipSuccess;
}

extern "C" hipError_t impeller_graph_launch(void* graph_exec, hipStream_t stream) {
  IMP_REQUIRE(graph_exec, "null graph");
  return hipGraphLaunch((hipGraphExec_t)graph_exec, stream);
}

extern "C" hipError_t impeller_graph_destroy(void* graph_exec) {
  if (!graph_exec) return hipSuccess;
  return hipGraphExecDestroy((hipGraphExec_t)graph_exec);
}

extern "C" const char* impeller_version(void) { return "impeller_hip 0.1 (gfx950)"; }
