// Learned closure of the cascaded collision on gfx950 (no reference
// counterpart: the reference fixes the third- and fourth-order rates at 1,
// clbm_moving.py:188-196).
//
// A per-cell network 19 -> 32 -> 32 -> 4 (GELU, bf16 weights and
// activations, fp32 accumulation) maps the cell's normalised central moments
// to corrections of four relaxation rates: s = 1 + y for (third order in the
// xy plane, third order with z, fourth order k220, fourth order k202/k022).
// A cell whose rates leave [lo, hi] takes the plain cascade (all ones).
//
// MFMA layout.  Every product is one __builtin_amdgcn_mfma_f32_16x16x32_bf16
// (A 16x32, B 32x16, D 16x16; lane l = 16 g + r holds A[r][8g..8g+7],
// B[8g..8g+7][r] and D[4g..4g+3][r]).  The network runs transposed,
// Y^T = W^T X^T, with the 64 cells of a wavefront as four 16-column blocks:
// a layer's D (hidden units 4g+i and 16+4g+i of column r in this lane) is
// directly the next layer's B operand once the next weight matrix's K index
// is permuted to match (perm32), so activations never leave registers.
// Only the input (one LDS transpose) and the output (lane shuffles) move.
//
// Training (closure_grad) runs the same forward, backward through the
// transposed weights (dh2^T = W3 dy^T, dh1^T = W2 dz2^T, again register to
// register), and forms the weight gradients as MFMA products over the
// samples from bf16 transposes staged in LDS.  Each workgroup writes its
// partial gradient; closure_reduce sums the partials in a fixed order.
#include "clbm_step.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace clbm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4v;

constexpr int CP = 19 * 32 + 32 + 32 * 32 + 32 + 32 * 4 + 4;  // 1828 parameters
constexpr int OW1 = 0, OB1 = 19 * 32, OW2 = OB1 + 32, OB2 = OW2 + 32 * 32, OW3 = OB2 + 32,
              OB3 = OW3 + 32 * 4;
static_assert(OB3 + 4 == CP, "parameter layout");
// packed bf16 operands, 64 lanes x 8 values each
enum { P_W1A0, P_W1A1, P_W2A0, P_W2A1, P_W3A, P_W3B0, P_W3B1, P_W2B0, P_W2B1, P_COUNT };
constexpr int P_INFER = 5;  // the forward operands come first

// MFMA K slot -> hidden unit whose D value this lane group holds (g = slot/8:
// slots 8g..8g+3 are units 4g..4g+3, slots 8g+4..8g+7 are 16+4g..16+4g+3).
__host__ __device__ constexpr int perm32(int slot) {
  return (slot & 7) < 4 ? 4 * (slot >> 3) + (slot & 7) : 16 + 4 * (slot >> 3) + (slot & 7) - 4;
}

// GELU, tanh form: x/2 (1 + tanh(y)), y = sqrt(2/pi) (x + 0.044715 x^3).
// Since (1 + tanh(y)) / 2 = 1 / (1 + exp(-2y)), it is x s with
// s = 1 / (1 + 2^(x (c1 + c3 x^2))), c1 = -2 log2(e) sqrt(2/pi), c3 =
// 0.044715 c1: seven instructions with the hardware base-2 exponential and
// reciprocal, within a few 1e-7 relative of the tanh form (far below the bf16
// rounding the hidden activations take next).  Large |x| saturate: 2^+big is
// infinity (s = 0), 2^-big is 0 (s = 1).
constexpr float GELU_C1 = -2.302208198144325f;   // -2 log2(e) sqrt(2/pi)
constexpr float GELU_C3 = -0.1029432395800235f;  // 0.044715 GELU_C1
__device__ __forceinline__ float gelu_sig(float x) {
  const float e = __builtin_amdgcn_exp2f(x * __builtin_fmaf(GELU_C3, x * x, GELU_C1));
  return __builtin_amdgcn_rcpf(e + 1.0f);
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sig(x); }
// d/dx x s = s + x s (1 - s) 2 y'(x), y' = sqrt(2/pi) (1 + 3 0.044715 x^2)
__device__ __forceinline__ float gelu_grad(float x) {
  const float s = gelu_sig(x);
  return s + 2.0f * x * s * (1.0f - s) * 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
}

__device__ __forceinline__ f4v mfma(bf16x8 a, bf16x8 b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// All lanes of the wavefront see each other's LDS writes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// operand packing (fp32 parameters -> bf16 MFMA A operands)
// ---------------------------------------------------------------------------
__global__ void k_closure_pack(const float* __restrict__ p, __bf16* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P_COUNT * 512) return;
  const int op = t / 512, l = (t / 8) % 64, j = t % 8, r = l % 16;
  const int k = 8 * (l / 16) + j;
  float v = 0.0f;
  switch (op) {
    case P_W1A0: case P_W1A1:  // W1^T: row = hidden, K = input (19 of 32)
      v = k < 19 ? p[OW1 + k * 32 + 16 * (op - P_W1A0) + r] : 0.0f;
      break;
    case P_W2A0: case P_W2A1:  // W2^T: row = output unit, K = permuted input unit
      v = p[OW2 + perm32(k) * 32 + 16 * (op - P_W2A0) + r];
      break;
    case P_W3A:                // W3^T: row = output (4 of 16), K = permuted hidden
      v = r < 4 ? p[OW3 + perm32(k) * 4 + r] : 0.0f;
      break;
    case P_W3B0: case P_W3B1:  // W3: row = hidden, K = output (4 of 32)
      v = k < 4 ? p[OW3 + (16 * (op - P_W3B0) + r) * 4 + k] : 0.0f;
      break;
    case P_W2B0: case P_W2B1:  // W2: row = input unit, K = permuted output unit
      v = p[OW2 + (16 * (op - P_W2B0) + r) * 32 + perm32(k)];
      break;
  }
  out[t] = (__bf16)v;
}

// ---------------------------------------------------------------------------
// forward of one 16-column block: B operand x (this lane's 8 inputs of column r)
// ---------------------------------------------------------------------------
struct Fwd {
  f4v z1[2], z2[2], y;
  bf16x8 h1, h2;
};
__device__ __forceinline__ void forward_block(const bf16x8* __restrict__ A, const float* __restrict__ bias,
                                              bf16x8 x, int lane, Fwd& o) {
  const int g = lane >> 4;
  const f4v zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    o.z1[rb] = mfma(A[(P_W1A0 + rb) * 64 + lane], x, zero);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      o.z1[rb][i] += bias[16 * rb + 4 * g + i];
      o.h1[4 * rb + i] = (__bf16)gelu(o.z1[rb][i]);
    }
  }
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    o.z2[rb] = mfma(A[(P_W2A0 + rb) * 64 + lane], o.h1, zero);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      o.z2[rb][i] += bias[32 + 16 * rb + 4 * g + i];
      o.h2[4 * rb + i] = (__bf16)gelu(o.z2[rb][i]);
    }
  }
  o.y = mfma(A[P_W3A * 64 + lane], o.h2, zero);
#pragma unroll
  for (int i = 0; i < 4; ++i) o.y[i] += g == 0 ? bias[64 + i] : 0.0f;
}

// The network on the 64 cells of a wavefront: x[19] of this lane's cell in,
// y[4] of this lane's cell out.  sx: this wavefront's 64 x 32 bf16 LDS tile.
// Every lane of the wavefront must call it.
__device__ __forceinline__ void mlp_wave(const float* x, float* y, const bf16x8* __restrict__ A,
                                         const float* __restrict__ bias, __bf16* __restrict__ sx) {
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int k = 0; k < 32; ++k) sx[lane * 32 + k] = (__bf16)(k < 19 ? x[k] : 0.0f);
  wave_sync();
  float out[4][4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const bf16x8 xb = *(const bf16x8*)&sx[(16 * cb + r) * 32 + 8 * g];
    Fwd o;
    forward_block(A, bias, xb, lane, o);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[cb][i] = o.y[i];
  }
  wave_sync();  // sx may be reused by the caller
  // the result of cell t sits in lane t % 16 (lane group 0) of block t / 16
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float v = 0.0f;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const float w = __shfl(out[cb][i], r);
      v = (lane >> 4) == cb ? w : v;
    }
    y[i] = v;
  }
}

// ---------------------------------------------------------------------------
// collision with per-cell third/fourth-order rates (clbm_moving.py:120-249
// with s_q, s_pi replaced by the four rates)
// ---------------------------------------------------------------------------
struct Moments {
  float rho, u, v, w;
  float k[18];  // central moments in the raw-moment order, from index 1 (k100 ... k022)
};
__device__ __forceinline__ void central_moments(const float* f, float Fx, float Fy, float Fz,
                                                Moments& m) {
  float rho = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    rho += f[q];
    mx += float(ex(q)) * f[q];
    my += float(ey(q)) * f[q];
    mz += float(ez(q)) * f[q];
  }
  const float inv = 1.0f / rho;
  m.rho = rho;
  m.u = (mx + 0.5f * Fx) * inv;
  m.v = (my + 0.5f * Fy) * inv;
  m.w = (mz + 0.5f * Fz) * inv;
#pragma unroll
  for (int i = 0; i < 18; ++i) m.k[i] = 0.0f;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const float a = float(ex(q)) - m.u, b = float(ey(q)) - m.v, c = float(ez(q)) - m.w;
    const float fq = f[q];
    const float fa = fq * a, fb = fq * b, fc = fq * c;
    m.k[0] += fa;
    m.k[1] += fb;
    m.k[2] += fc;
    m.k[3] += fa * b;
    m.k[4] += fa * c;
    m.k[5] += fb * c;
    m.k[6] += fa * a;
    m.k[7] += fb * b;
    m.k[8] += fc * c;
    m.k[9] += fa * a * b;
    m.k[10] += fa * b * b;
    m.k[11] += fa * a * c;
    m.k[12] += fb * b * c;
    m.k[13] += fa * c * c;
    m.k[14] += fb * c * c;
    m.k[15] += fa * a * b * b;
    m.k[16] += fa * a * c * c;
    m.k[17] += fb * b * c * c;
  }
}

// The 19 features: moments / rho minus their equilibrium values.
__device__ __forceinline__ void features(const Moments& m, float* x) {
  const float inv = 1.0f / m.rho;
  x[0] = m.rho - 1.0f;
#pragma unroll
  for (int i = 0; i < 18; ++i) x[1 + i] = m.k[i] * inv;
  x[7] -= 1.0f / 3.0f;
  x[8] -= 1.0f / 3.0f;
  x[9] -= 1.0f / 3.0f;
  x[16] -= 1.0f / 9.0f;
  x[17] -= 1.0f / 9.0f;
  x[18] -= 1.0f / 9.0f;
}

__device__ __forceinline__ void relax_rebuild(const Moments& m, const float* s, float Fx, float Fy,
                                              float Fz, float s_v, float s_e, float* f) {
  const float rho = m.rho, u = m.u, v = m.v, w = m.w;
  const float* k = m.k;
  const float om = 1.0f - s_v;
  const float a110 = om * k[3], a101 = om * k[4], a011 = om * k[5];
  const float kA = k[6] + k[7] + k[8];
  const float kAs = kA - s_e * (kA - rho);
  const float kDs = om * (k[6] - k[7]), kEs = om * (k[6] - k[8]);
  const float third = 1.0f / 3.0f;
  const float a200 = (kAs + kDs + kEs) * third;
  const float a020 = (kAs - 2.0f * kDs + kEs) * third;
  const float a002 = (kAs + kDs - 2.0f * kEs) * third;
  const float a210 = (1.0f - s[0]) * k[9], a120 = (1.0f - s[0]) * k[10];
  const float a201 = (1.0f - s[1]) * k[11], a021 = (1.0f - s[1]) * k[12];
  const float a102 = (1.0f - s[1]) * k[13], a012 = (1.0f - s[1]) * k[14];
  const float r9 = rho * (1.0f / 9.0f);
  const float a220 = r9 + (1.0f - s[2]) * (k[15] - r9);
  const float a202 = r9 + (1.0f - s[3]) * (k[16] - r9);
  const float a022 = r9 + (1.0f - s[3]) * (k[17] - r9);
  const float ru = rho * u, rv = rho * v, rw = rho * w, uu = u * u, vv = v * v, ww = w * w;
  const float T110 = a110 + ru * v, T101 = a101 + ru * w, T011 = a011 + rv * w;
  const float T200 = a200 + ru * u, T020 = a020 + rv * v, T002 = a002 + rw * w;
  const float T210 = a210 + 2.0f * u * a110 + v * a200 + ru * u * v;
  const float T120 = a120 + 2.0f * v * a110 + u * a020 + ru * vv;
  const float T201 = a201 + 2.0f * u * a101 + w * a200 + ru * u * w;
  const float T021 = a021 + 2.0f * v * a011 + w * a020 + rv * v * w;
  const float T102 = a102 + 2.0f * w * a101 + u * a002 + ru * ww;
  const float T012 = a012 + 2.0f * w * a011 + v * a002 + rv * ww;
  const float T220 = a220 + 2.0f * u * a120 + 2.0f * v * a210 + 4.0f * u * v * a110 + vv * a200 +
                     uu * a020 + ru * u * vv;
  const float T202 = a202 + 2.0f * u * a102 + 2.0f * w * a201 + 4.0f * u * w * a101 + ww * a200 +
                     uu * a002 + ru * u * ww;
  const float T022 = a022 + 2.0f * v * a012 + 2.0f * w * a021 + 4.0f * v * w * a011 + ww * a020 +
                     vv * a002 + rv * v * ww;
  const float half = 0.5f, quarter = 0.25f;
  f[0] = rho - (T200 + T020 + T002) + (T220 + T202 + T022);
  {
    const float a = T200 - T220 - T202, b = ru - T120 - T102;
    f[1] = half * (a + b);
    f[2] = half * (a - b);
  }
  {
    const float a = T020 - T220 - T022, b = rv - T210 - T012;
    f[3] = half * (a + b);
    f[4] = half * (a - b);
  }
  {
    const float a = T002 - T202 - T022, b = rw - T201 - T021;
    f[5] = half * (a + b);
    f[6] = half * (a - b);
  }
  {
    const float p = T220 + T210, mm = T220 - T210, sp = T120 + T110, d = T120 - T110;
    f[7] = quarter * (p + sp);
    f[8] = quarter * (p - sp);
    f[9] = quarter * (mm + d);
    f[10] = quarter * (mm - d);
  }
  {
    const float p = T202 + T201, mm = T202 - T201, sp = T102 + T101, d = T102 - T101;
    f[11] = quarter * (p + sp);
    f[12] = quarter * (p - sp);
    f[13] = quarter * (mm + d);
    f[14] = quarter * (mm - d);
  }
  {
    const float p = T022 + T021, mm = T022 - T021, sp = T012 + T011, d = T012 - T011;
    f[15] = quarter * (p + sp);
    f[16] = quarter * (p - sp);
    f[17] = quarter * (mm + d);
    f[18] = quarter * (mm - d);
  }
  const float c = 1.0f - half * s_v;
  const float uF = u * Fx + v * Fy + w * Fz;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const float eF = float(ex(q)) * Fx + float(ey(q)) * Fy + float(ez(q)) * Fz;
    const float eu = float(ex(q)) * u + float(ey(q)) * v + float(ez(q)) * w;
    f[q] += (wq<float>(q) * c) * (3.0f * (eF - uF) + 9.0f * eu * eF);
  }
}

// Rates from the network outputs, or all ones outside [lo, hi].
__device__ __forceinline__ void rates_from(const float* y, float lo, float hi, float* s) {
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    s[i] = 1.0f + y[i];
    ok = ok && s[i] >= lo && s[i] <= hi;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) s[i] = ok ? s[i] : 1.0f;
}

struct ClosureArgs {
  const bf16x8* A;      // packed operands (global)
  const float* params;  // fp32 parameter vector (the biases are read from it)
  float lo, hi;         // rates outside [lo, hi] fall back to the plain cascade
};

constexpr int CBLK = 256;
// Stage the forward operands and the biases b1 (32), b2 (32), b3 (4) of the
// network in LDS (whole block).
__device__ __forceinline__ void stage_net(const ClosureArgs& n, bf16x8* sA, float* sb) {
  for (int i = threadIdx.x; i < P_INFER * 64; i += blockDim.x) sA[i] = n.A[i];
  for (int i = threadIdx.x; i < 68; i += blockDim.x)
    sb[i] = n.params[i < 32 ? OB1 + i : (i < 64 ? OB2 + i - 32 : OB3 + i - 64)];
  __syncthreads();
}

// ---------------------------------------------------------------------------
// the step with the learned closure (every cell of the lattice)
// ---------------------------------------------------------------------------
template <int S, int PAR, bool FORCE, bool BOX>
__global__ __launch_bounds__(CBLK) void k_step_closure(StepArgs a, ClosureArgs n) {
  __shared__ bf16x8 sA[P_INFER * 64];
  __shared__ float sb[68];
  __shared__ __bf16 sx[CBLK / 64][64 * 32];
  stage_net(n, sA, sb);
  using R = typename Store<S>::R;
  const Geo& g = a.g;
  const Planes& pl = a.pl;
  const int c = sweep_block<PAR>(blockIdx.x, gridDim.x) * CBLK + threadIdx.x;
  int x = 0, y = 0, z = 0;
  bool act = c < g.ncell;
  if (act) cell_xyz(g, c, x, y, z);
  act = act && !(pl.on && x == g.nx - 1);
  const Pop<S> pop((typename Store<S>::T*)a.f, g);
  R fq[Q];
  float Fx = 0.0f, Fy = 0.0f, Fz = 0.0f;
  float Gx = 0.0f, Gy = 0.0f, Gz = 0.0f;
  Nbr nb{};
  const bool outlet = act && pl.on && x == g.nx - 2;
  bool twin = true;
  Moments m;
  float feat[Q];
  if (act) {
    nb = make_nbr(g, x, y, z);
    pop_load<PAR>(pop, c, nb, fq);
    R rx, ry, rz;
    force_at<R, FORCE, BOX>(c, x, y, z, a.fx, a.fy, a.fz, a.box, rx, ry, rz);
    Fx = float(rx);
    Fy = float(ry);
    Fz = float(rz);
    if (outlet) {
      force_at<R, FORCE, BOX>(c + 1, x + 1, y, z, a.fx, a.fy, a.fz, a.box, rx, ry, rz);
      Gx = float(rx);
      Gy = float(ry);
      Gz = float(rz);
      twin = (Gx == Fx) & (Gy == Fy) & (Gz == Fz);
      if (!twin) pop_store_in<PAR>(pop, c + 1, make_nbr(g, x + 1, y, z), fq);
    }
    if (pl.on && x == 0) inlet_rule<R>(fq, R(pl.ux_in), R(pl.uy_in), pl.rule);
    float ff[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) ff[q] = float(fq[q]);
    central_moments(ff, Fx, Fy, Fz, m);
    features(m, feat);
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) feat[q] = 0.0f;
  }
  float yv[4], s[4];
  mlp_wave(feat, yv, sA, sb, sx[threadIdx.x >> 6]);
  if (!act) return;
  rates_from(yv, n.lo, n.hi, s);
  float out[Q];
  relax_rebuild(m, s, Fx, Fy, Fz, a.s_v.f, a.s_e.f, out);
#pragma unroll
  for (int q = 0; q < Q; ++q) fq[q] = R(out[q]);
  pop_store<PAR>(pop, c, nb, fq);
  if (outlet) {
    const Nbr nb1 = make_nbr(g, x + 1, y, z);
    if (!twin) {
      // the outlet cell of a row whose force differs from its neighbour's takes
      // the plain cascade (the network runs on the cells of the wavefront only)
      pop_load<PAR>(pop, c + 1, nb1, fq);
      collide_cell<R>(fq, R(Gx), R(Gy), R(Gz), a.s_v.template as<R>(), a.s_e.template as<R>());
    }
    pop_store<PAR>(pop, c + 1, nb1, fq);
  }
}

// ---------------------------------------------------------------------------
// the network on arbitrary inputs (inference check, evaluation)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(CBLK) void k_closure_infer(ClosureArgs n, int count,
                                                        const float* __restrict__ X,
                                                        float* __restrict__ Y) {
  __shared__ bf16x8 sA[P_INFER * 64];
  __shared__ float sb[68];
  __shared__ __bf16 sx[CBLK / 64][64 * 32];
  stage_net(n, sA, sb);
  const int i = blockIdx.x * CBLK + threadIdx.x;
  float x[Q], y[4];
#pragma unroll
  for (int k = 0; k < Q; ++k) x[k] = i < count ? X[(size_t)i * Q + k] : 0.0f;
  mlp_wave(x, y, sA, sb, sx[threadIdx.x >> 6]);
  if (i < count)
#pragma unroll
    for (int k = 0; k < 4; ++k) Y[(size_t)i * 4 + k] = y[k];
}

// ---------------------------------------------------------------------------
// training: partial gradients of L = sum_n w_n |y_n - t_n|^2 / wsum, or with
// one weight per output (wdim 4) L = sum_nj w_nj (y_nj - t_nj)^2 / wsum
// ---------------------------------------------------------------------------
constexpr int GPART = CP + 1;  // gradient + loss per workgroup
__global__ __launch_bounds__(64) void k_closure_grad(const bf16x8* __restrict__ Ag,
                                                     const float* __restrict__ params, int count,
                                                     const float* __restrict__ X,
                                                     const float* __restrict__ Tg,
                                                     const float* __restrict__ Wg, int wdim,
                                                     float wsum, float* __restrict__ partial) {
  __shared__ bf16x8 sA[P_COUNT * 64];
  __shared__ float sb[68];
  __shared__ __bf16 sx[64 * 32];
  // transposes [unit][sample] for the weight gradients
  __shared__ __bf16 tX[32 * 64], tH1[32 * 64], tH2[32 * 64], tD1[32 * 64], tD2[32 * 64],
      tDY[16 * 64];
  const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
  for (int i = lane; i < P_COUNT * 64; i += 64) sA[i] = Ag[i];
  for (int i = lane; i < 32; i += 64) {
    sb[i] = params[OB1 + i];
    sb[32 + i] = params[OB2 + i];
  }
  if (lane < 4) sb[64 + lane] = params[OB3 + lane];
  for (int i = lane; i < 16 * 64; i += 64) tDY[i] = (__bf16)0.0f;
  __syncthreads();
  const f4v zero = {0.0f, 0.0f, 0.0f, 0.0f};
  f4v gw1[2][2] = {{zero, zero}, {zero, zero}}, gw2[2][2] = {{zero, zero}, {zero, zero}};
  f4v gw3[2] = {zero, zero};
  float gb1[2][4] = {}, gb2[2][4] = {}, gb3[4] = {}, loss = 0.0f;
  const float inv = 1.0f / wsum;
  for (int base = blockIdx.x * 64; base < count; base += gridDim.x * 64) {
    const int smp = base + lane;
    const bool live = smp < count;
    float t[4], wl[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      t[k] = live ? Tg[(size_t)smp * 4 + k] : 0.0f;
      wl[k] = live ? Wg[wdim == 4 ? (size_t)smp * 4 + k : (size_t)smp] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      const __bf16 v = (__bf16)(live && k < 19 ? X[(size_t)smp * 19 + k] : 0.0f);
      sx[lane * 32 + k] = v;
      tX[k * 64 + lane] = v;
    }
    wave_sync();
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const int col = 16 * cb + r;  // the sample (column) this lane holds in this block
      const bf16x8 xb = *(const bf16x8*)&sx[col * 32 + 8 * g];
      Fwd o;
      forward_block(sA, sb, xb, lane, o);
      float ti[4], wi[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        ti[k] = __shfl(t[k], col);
        wi[k] = __shfl(wl[k], col);
      }
      bf16x8 bdy;
#pragma unroll
      for (int k = 0; k < 8; ++k) bdy[k] = (__bf16)0.0f;
      if (g == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float d = o.y[i] - ti[i];
          const float dy = 2.0f * wi[i] * d * inv;
          loss += wi[i] * d * d * inv;
          gb3[i] += dy;
          bdy[i] = (__bf16)dy;
          tDY[i * 64 + col] = (__bf16)dy;
        }
      }
      bf16x8 bdz2;
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const f4v dh2 = mfma(sA[(P_W3B0 + rb) * 64 + lane], bdy, zero);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int unit = 16 * rb + 4 * g + i;
          const float dz = dh2[i] * gelu_grad(o.z2[rb][i]);
          gb2[rb][i] += dz;
          bdz2[4 * rb + i] = (__bf16)dz;
          tD2[unit * 64 + col] = (__bf16)dz;
          tH2[unit * 64 + col] = o.h2[4 * rb + i];
          tH1[unit * 64 + col] = o.h1[4 * rb + i];
        }
      }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const f4v dh1 = mfma(sA[(P_W2B0 + rb) * 64 + lane], bdz2, zero);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float dz = dh1[i] * gelu_grad(o.z1[rb][i]);
          gb1[rb][i] += dz;
          tD1[(16 * rb + 4 * g + i) * 64 + col] = (__bf16)dz;
        }
      }
    }
    wave_sync();
    // weight gradients: C[row][col] += sum_samples A[row][s] B[s][col]
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
      const int so = 32 * kc + 8 * g;
#pragma unroll
      for (int ri = 0; ri < 2; ++ri) {
        const bf16x8 ax = *(const bf16x8*)&tX[(16 * ri + r) * 64 + so];
        const bf16x8 ah1 = *(const bf16x8*)&tH1[(16 * ri + r) * 64 + so];
        const bf16x8 ah2 = *(const bf16x8*)&tH2[(16 * ri + r) * 64 + so];
#pragma unroll
        for (int co = 0; co < 2; ++co) {
          const bf16x8 bd1 = *(const bf16x8*)&tD1[(16 * co + r) * 64 + so];
          const bf16x8 bd2 = *(const bf16x8*)&tD2[(16 * co + r) * 64 + so];
          gw1[ri][co] = mfma(ax, bd1, gw1[ri][co]);
          gw2[ri][co] = mfma(ah1, bd2, gw2[ri][co]);
        }
        const bf16x8 bdy = *(const bf16x8*)&tDY[r * 64 + so];
        gw3[ri] = mfma(ah2, bdy, gw3[ri]);
      }
    }
    wave_sync();
  }
  // bias sums over the 16 columns of each lane group
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) {
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        gb1[rb][i] += __shfl_xor(gb1[rb][i], o);
        gb2[rb][i] += __shfl_xor(gb2[rb][i], o);
      }
#pragma unroll
    for (int i = 0; i < 4; ++i) gb3[i] += __shfl_xor(gb3[i], o);
    loss += __shfl_xor(loss, o);
  }
  float* out = partial + (size_t)blockIdx.x * GPART;
#pragma unroll
  for (int ri = 0; ri < 2; ++ri)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = 16 * ri + 4 * g + i;
#pragma unroll
      for (int co = 0; co < 2; ++co) {
        const int col = 16 * co + r;
        if (row < 19) out[OW1 + row * 32 + col] = gw1[ri][co][i];
        out[OW2 + row * 32 + col] = gw2[ri][co][i];
      }
      if (r < 4) out[OW3 + row * 4 + r] = gw3[ri][i];
    }
  if (r == 0) {
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        out[OB1 + 16 * rb + 4 * g + i] = gb1[rb][i];
        out[OB2 + 16 * rb + 4 * g + i] = gb2[rb][i];
      }
  }
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) out[OB3 + i] = gb3[i];
    out[CP] = loss;
  }
}

// grad[p] = sum over workgroups b (in order) of partial[b][p]
__global__ void k_closure_reduce(const float* __restrict__ partial, int blocks,
                                 float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= GPART) return;
  float s = 0.0f;
  for (int b = 0; b < blocks; ++b) s += partial[(size_t)b * GPART + p];
  grad[p] = s;
}

__global__ void k_closure_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                               const float* __restrict__ grad, float lr, float b1, float b2,
                               float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= CP) return;
  const float gi = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

template <int S> struct ClosureStepRun {
  template <int PAR, bool FORCE, bool BOX>
  static void go(const StepArgs& a, const ClosureArgs& n, hipStream_t st) {
    hipLaunchKernelGGL((k_step_closure<S, PAR, FORCE, BOX>), dim3((a.g.ncell + CBLK - 1) / CBLK),
                       dim3(CBLK), 0, st, a, n);
  }
  template <int PAR>
  static void par(const StepArgs& a, const ClosureArgs& n, hipStream_t st) {
    if (!(a.fx || a.fy || a.fz))
      go<PAR, false, false>(a, n, st);
    else if (a.box)
      go<PAR, true, true>(a, n, st);
    else
      go<PAR, true, false>(a, n, st);
  }
  static int run(const StepArgs& a, const ClosureArgs& n, int parity, hipStream_t st) {
    if (parity) par<1>(a, n, st); else par<0>(a, n, st);
    return check_launch("clbm_step_closure");
  }
};

// ---------------------------------------------------------------------------
// closure_grad_noalloc: the per-workgroup partial gradients in a file-scope
// device array instead of a caller's buffer.  Each GPU holds its own copy of
// the array (at its own address) from the load of this code object until the
// process ends.  Host code never passes the array's name to a kernel: the
// name is a host-side stand-in, so the device address is looked up once per
// device with hipGetSymbolAddress and kept.  Every workgroup of
// k_closure_grad overwrites all GPART entries of its own row before
// k_closure_reduce (same stream) reads them, so nothing needs clearing
// between calls; calls that overlap on one GPU (other streams, other
// threads) share the rows and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) float scratch[(size_t)CLOSURE_GRAD_NOALLOC_MAX_BLOCKS * GPART];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device into *out (looked up on
// the first call on that device); 0, or the error code with the message set.
static int scratch_on_current_device(float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return fail((int)e, std::string("closure_grad_noalloc: hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= MAX_DEVICES)
    return fail(-1, "closure_grad_noalloc: device ordinal beyond the library's table of 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return fail((int)e, std::string("closure_grad_noalloc: hipGetSymbolAddress(scratch): ") +
                              hipGetErrorString(e));
    if (!ptr) return fail(-1, "closure_grad_noalloc: hipGetSymbolAddress(scratch) gave NULL");
    g_scratch_ptr[dev] = (float*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace clbm

using namespace clbm;

extern "C" {

int closure_param_count(void) { return CP; }
size_t closure_packed_bytes(void) { return (size_t)P_COUNT * 64 * sizeof(bf16x8); }

int closure_pack(const float* params, void* packed, void* stream) {
  if (!params || !packed) return fail(-1, "closure_pack: NULL buffer");
  hipLaunchKernelGGL(k_closure_pack, dim3((P_COUNT * 512 + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, params, (__bf16*)packed);
  return check_launch("closure_pack");
}

int closure_infer(const void* packed, const float* params, int count, const float* X, float* Y,
                  void* stream) {
  if (!packed || !params || (count > 0 && (!X || !Y))) return fail(-1, "closure_infer: NULL buffer");
  if (count < 0) return fail(-1, "closure_infer: count < 0");
  if (count == 0) return 0;
  const ClosureArgs n{(const bf16x8*)packed, params, 0.0f, 2.0f};
  hipLaunchKernelGGL(k_closure_infer, dim3((count + CBLK - 1) / CBLK), dim3(CBLK), 0,
                     (hipStream_t)stream, n, count, X, Y);
  return check_launch("closure_infer");
}

int clbm_step_closure(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
                      const float* fx, const float* fy, const float* fz, const int32_t* force_box,
                      const void* packed, const float* params, float lo, float hi, void* stream) {
  if (!p) return fail(-1, "clbm_step_closure: params is NULL");
  if (int e = check_grid(g, p->inlet != 0)) return e;
  if (!f || !packed || !params) return fail(-1, "clbm_step_closure: NULL buffer");
  if (!(lo <= hi)) return fail(-1, "clbm_step_closure: need lo <= hi");
  Planes pl;
  if (int e = check_state(g, p, state, pl)) return e;
  if (int e = check_inlet(p, "clbm_step_closure")) return e;
  StepArgs a{};
  a.g = make_geo(*g);
  a.s_v = p->s_v;
  a.s_e = p->s_e;
  a.pl = pl;
  a.f = f;
  a.fx = fx;
  a.fy = fy;
  a.fz = fz;
  a.box = force_box;
  const ClosureArgs n{(const bf16x8*)packed, params, lo, hi};
  const int par = state & CLBM_STATE_PARITY;
  hipStream_t st = (hipStream_t)stream;
  switch (g->storage) {
    case CLBM_STORE_F32: return ClosureStepRun<CLBM_STORE_F32>::run(a, n, par, st);
    case CLBM_STORE_F16: return ClosureStepRun<CLBM_STORE_F16>::run(a, n, par, st);
    case CLBM_STORE_F16C: return ClosureStepRun<CLBM_STORE_F16C>::run(a, n, par, st);
    default: return fail(-1, "clbm_step_closure: storage must be f32, f16 or f16c");
  }
}

size_t closure_partial_bytes(int blocks) { return blocks > 0 ? (size_t)blocks * GPART * 4 : 0; }

int closure_grad(const void* packed, const float* params, int count, const float* X,
                 const float* T, const float* W, int wdim, float wsum, float* partial, int blocks,
                 float* grad, void* stream) {
  if (!packed || !params || !X || !T || !W || !partial || !grad)
    return fail(-1, "closure_grad: NULL buffer");
  if (count <= 0 || blocks <= 0) return fail(-1, "closure_grad: count and blocks must be > 0");
  if (wdim != 1 && wdim != 4) return fail(-1, "closure_grad: wdim must be 1 or 4");
  if (!(wsum > 0.0f)) return fail(-1, "closure_grad: weight sum must be > 0");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_closure_grad, dim3(blocks), dim3(64), 0, st, (const bf16x8*)packed, params,
                     count, X, T, W, wdim, wsum, partial);
  hipLaunchKernelGGL(k_closure_reduce, dim3((GPART + 255) / 256), dim3(256), 0, st, partial,
                     blocks, grad);
  return check_launch("closure_grad");
}

int closure_grad_noalloc(const void* packed, const float* params, int count, const float* X,
                         const float* T, const float* W, int wdim, float wsum, int blocks,
                         float* grad, void* stream) {
  if (!packed || !params || !X || !T || !W || !grad)
    return fail(-1, "closure_grad_noalloc: NULL buffer");
  if (count <= 0 || blocks <= 0)
    return fail(-1, "closure_grad_noalloc: count and blocks must be > 0");
  if (blocks > CLOSURE_GRAD_NOALLOC_MAX_BLOCKS)
    return fail(-1, "closure_grad_noalloc: blocks exceeds CLOSURE_GRAD_NOALLOC_MAX_BLOCKS (" +
                        std::to_string(CLOSURE_GRAD_NOALLOC_MAX_BLOCKS) +
                        "), the rows of the library's partial-gradient array");
  if (wdim != 1 && wdim != 4) return fail(-1, "closure_grad_noalloc: wdim must be 1 or 4");
  if (!(wsum > 0.0f)) return fail(-1, "closure_grad_noalloc: weight sum must be > 0");
  float* partial = nullptr;
  if (int e = scratch_on_current_device(&partial)) return e;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_closure_grad, dim3(blocks), dim3(64), 0, st, (const bf16x8*)packed, params,
                     count, X, T, W, wdim, wsum, partial);
  hipLaunchKernelGGL(k_closure_reduce, dim3((GPART + 255) / 256), dim3(256), 0, st, partial,
                     blocks, grad);
  return check_launch("closure_grad_noalloc");
}

int closure_adam(float* params, float* m, float* v, const float* grad, float lr, float beta1,
                 float beta2, float eps, int t, void* stream) {
  if (!params || !m || !v || !grad) return fail(-1, "closure_adam: NULL buffer");
  if (t < 1) return fail(-1, "closure_adam: step t must be >= 1");
  const float c1 = 1.0f / (1.0f - std::pow(beta1, (float)t));
  const float c2 = 1.0f / (1.0f - std::pow(beta2, (float)t));
  hipLaunchKernelGGL(k_closure_adam, dim3((CP + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     params, m, v, grad, lr, beta1, beta2, eps, c1, c2);
  return check_launch("closure_adam");
}

}  // extern "C"
