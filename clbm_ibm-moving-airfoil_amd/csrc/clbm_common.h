// Shared device/host definitions for the CLBM-IBM kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <string>
#include <type_traits>

#include "../../include/clbm_amd.h"

namespace clbm {

// ---------------------------------------------------------------------------
// D3Q19 lattice in the reference's direction order (clbm_moving.py:21-30).
// ---------------------------------------------------------------------------
constexpr int Q = 19;
__host__ __device__ constexpr int ex(int q) {
  constexpr int v[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  return v[q];
}
__host__ __device__ constexpr int ey(int q) {
  constexpr int v[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
  return v[q];
}
__host__ __device__ constexpr int ez(int q) {
  constexpr int v[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
  return v[q];
}
// opposite direction: e_opp(q) = -e_q
__host__ __device__ constexpr int opp(int q) {
  constexpr int v[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  return v[q];
}
template <typename R>
__host__ __device__ constexpr R wq(int q) {
  return q == 0 ? R(1) / R(3) : (q < 7 ? R(1) / R(18) : R(1) / R(36));
}

// ---------------------------------------------------------------------------
// Population storage.  Each format maps a stored element to the compute type
// R and back.
// ---------------------------------------------------------------------------
template <int S> struct Store;
// ld/st take the population's deviation base: 16-bit formats store f - base
// (base = the equilibrium at the grid's deviation velocity, the lattice
// weight when that is zero); the wide formats ignore it.
template <> struct Store<CLBM_STORE_F32> {
  using T = float;
  using R = float;
  static constexpr bool dev = false;
  __device__ static R ld(const T* p, float) { return *p; }
  __device__ static void st(T* p, R v, float) { *p = v; }
};
template <> struct Store<CLBM_STORE_F64> {
  using T = double;
  using R = double;
  static constexpr bool dev = false;
  __device__ static R ld(const T* p, float) { return *p; }
  __device__ static void st(T* p, R v, float) { *p = v; }
};
template <> struct Store<CLBM_STORE_F16> {
  using T = __half;
  using R = float;
  static constexpr bool dev = true;
  __device__ static R ld(const T* p, float b) { return b + __half2float(*p); }
  __device__ static void st(T* p, R v, float b) { *p = __float2half_rn(v - b); }
};
template <> struct Store<CLBM_STORE_BF16> {
  using T = __hip_bfloat16;
  using R = float;
  static constexpr bool dev = true;
  __device__ static R ld(const T* p, float b) { return b + __bfloat162float(*p); }
  __device__ static void st(T* p, R v, float b) { *p = __float2bfloat16(v - b); }
};

// 16-bit code for deviations |x| < 1: a sign bit and a 15-bit code u for
// the magnitude.  u is the 12-significant-bit float F = |x| + 2^-15 with
// F's exponent stored in 4 bits (F in [2^-15, 1) for exponent codes 0-14)
// and 11 mantissa bits: encode rounds fl32(|x| + 2^-15) to 12 significant
// bits (nearest, ties to even), decode returns F(u) - 2^-15 (exact in
// float32).  Steps are 2^-26 below 2^-15 and at most 2^-11 of the value
// above it (|x| >= 2^-13: 2^-12, twice IEEE half's precision).  Exponent
// code 15 is not finite: 0x7800 is infinity, above it NaN; values rounding to
// F >= 1, infinity and NaN encode there (0x7fff from F >= 2).  Decode ORs the
// code, shifted by 12, under the bias 0x78000000, so exponent code 15 lands on
// float32 exponent 0xff by itself; an ldexp by -128 undoes the bias.  Both
// directions are branch-free.
// The code in bits 31:16 of the result (bits 15:0 unspecified), for a store
// of the high half: R = fl32(|v| + 2^-15) - bits(2^-15), rounded at bit 12,
// is clamped to 0x7ffffff (the codes above 0x7fff: too large, inf, NaN) and
// shifted up by 4, so that bits 30:16 hold R >> 12 and bit 31 takes v's sign.
__device__ __forceinline__ uint32_t f16c_encode_hi(float v) {
  const uint32_t b = __float_as_uint(v);
  const uint32_t t = __float_as_uint(fabsf(v) + 3.0517578125e-05f);  // + 2^-15
  const uint32_t r = min(t + ((t >> 12) & 1u) + (0x7ffu - (112u << 23)), 0x7ffffffu);
  return (b & 0x80000000u) | (r << 4);
}
__device__ __forceinline__ uint16_t f16c_encode(float v) {
  return (uint16_t)(f16c_encode_hi(v) >> 16);
}
__device__ __forceinline__ void f16c_store(uint16_t* p, float v) { *p = f16c_encode(v); }
// The code is read sign-extended: shifted by 12, its sign copies fill bits
// 27-31, of which the bias covers 27-30, so the OR gives +-F * 2^128 with the
// code's sign.  One FMA on its magnitude scales it back and removes the offset
// 2^-15 (both exact, so the FMA's single rounding changes nothing), and the
// code's sign is put back (code 0x8000 decodes to -0, which the load's
// addition of the deviation base turns into the base itself).  The scale
// 2^-128 is a float32 denormal: the kernels run with float32 denormals
// preserved (the compiler's default for gfx950), which the GPU tests of the
// codec rely on.
__device__ __forceinline__ float f16c_decode(uint16_t hv) {
  const uint32_t hs = (uint32_t)(int32_t)(int16_t)hv;
  const float fs = __uint_as_float((hs << 12) | 0x78000000u);
  const float m = __builtin_fmaf(fabsf(fs), 0x1p-128f, -3.0517578125e-05f);
  return copysignf(m, fs);
}
template <> struct Store<CLBM_STORE_F16C> {
  using T = uint16_t;
  using R = float;
  static constexpr bool dev = true;
  __device__ static R ld(const T* p, float b) { return b + f16c_decode(*p); }
  __device__ static void st(T* p, R v, float b) { f16c_store(p, v - b); }
};

inline size_t elem_bytes(int storage) {
  switch (storage) {
    case CLBM_STORE_F32: return 4;
    case CLBM_STORE_F64: return 8;
    case CLBM_STORE_F16:
    case CLBM_STORE_BF16:
    case CLBM_STORE_F16C: return 2;
    default: return 0;
  }
}

// ---------------------------------------------------------------------------
// Unsigned division by a run-time constant (round-up multiplier method):
// q = (umulhi(n, m) + n) >> s, exact for every n < 2^31.
// ---------------------------------------------------------------------------
struct FastDiv {
  uint32_t d, m, s;
};
inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv r{d, 0, 0};
  while ((1ull << r.s) < d) r.s++;
  r.m = (uint32_t)((((1ull << 32) * ((1ull << r.s) - d)) / d) + 1);
  return r;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
  return (__umulhi(n, f.m) + n) >> f.s;
}

// Lattice geometry passed by value to every kernel.
struct Geo {
  int nx, ny, nz;
  int nxy;        // nx*ny
  int ncell;      // nx*ny*nz
  FastDiv dx;     // divide by nx
  FastDiv dy;     // divide by ny
  float dev[Q];   // deviation base of 16-bit storage per direction
};
inline Geo make_geo(const clbm_grid_t& g) {
  Geo o;
  o.nx = g.nx;
  o.ny = g.ny;
  o.nz = g.nz;
  o.nxy = g.nx * g.ny;
  o.ncell = g.nx * g.ny * g.nz;
  o.dx = make_fastdiv((uint32_t)g.nx);
  o.dy = make_fastdiv((uint32_t)g.ny);
  // second-order equilibrium at rho = 1 and the deviation velocity, in double,
  // rounded once (= the lattice weights for a zero velocity)
  const double u = g.dev_ux, v = g.dev_uy, u2 = u * u + v * v;
  for (int q = 0; q < Q; ++q) {
    const double eu = ex(q) * u + ey(q) * v;
    o.dev[q] = (float)(wq<double>(q) * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * u2));
  }
  return o;
}

// The deviation bases of a Geo as values (uniform across the launch); empty
// for the wide formats.
template <int S, bool D = Store<S>::dev> struct DevBase {
  __device__ __forceinline__ explicit DevBase(const Geo&) {}
  __device__ __forceinline__ float operator[](int) const { return 0.0f; }
};
template <int S> struct DevBase<S, true> {
  float v[Q];
  __device__ __forceinline__ explicit DevBase(const Geo& g) {
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = g.dev[q];
  }
  __device__ __forceinline__ float operator[](int q) const { return v[q]; }
};

// Linear offsets from a cell to its neighbours along +-x/y/z with periodic wrap.
struct Nbr {
  int xm, xp, ym, yp, zm, zp;
  __device__ __forceinline__ int off(int dxs, int dys, int dzs) const {
    int o = 0;
    o += dxs > 0 ? xp : (dxs < 0 ? xm : 0);
    o += dys > 0 ? yp : (dys < 0 ? ym : 0);
    o += dzs > 0 ? zp : (dzs < 0 ? zm : 0);
    return o;
  }
};
__device__ __forceinline__ Nbr make_nbr(const Geo& g, int x, int y, int z) {
  Nbr n;
  n.xm = (x == 0) ? (g.nx - 1) : -1;
  n.xp = (x == g.nx - 1) ? -(g.nx - 1) : 1;
  n.ym = (y == 0) ? (g.ny - 1) * g.nx : -g.nx;
  n.yp = (y == g.ny - 1) ? -(g.ny - 1) * g.nx : g.nx;
  n.zm = (z == 0) ? (g.nz - 1) * g.nxy : -g.nxy;
  n.zp = (z == g.nz - 1) ? -(g.nz - 1) * g.nxy : g.nxy;
  return n;
}
__device__ __forceinline__ void cell_xyz(const Geo& g, int c, int& x, int& y, int& z) {
  uint32_t row = fdiv((uint32_t)c, g.dx);     // z*ny + y
  x = c - (int)row * g.nx;
  z = (int)fdiv(row, g.dy);
  y = (int)row - z * g.ny;
}

// An inclusive box of cells [lo, hi] and the divisors that enumerate it
// (x fastest), for the two parts of a split step.
struct Region {
  int lo[3], hi[3];
  int rx, ry, vol;
  FastDiv drx, dry;
};
inline Region make_region(const int32_t* r) {
  Region o;
  for (int a = 0; a < 3; ++a) {
    o.lo[a] = r[a];
    o.hi[a] = r[3 + a];
  }
  o.rx = o.hi[0] - o.lo[0] + 1;
  o.ry = o.hi[1] - o.lo[1] + 1;
  o.vol = o.rx * o.ry * (o.hi[2] - o.lo[2] + 1);
  o.drx = make_fastdiv((uint32_t)o.rx);
  o.dry = make_fastdiv((uint32_t)o.ry);
  return o;
}
__device__ __forceinline__ bool in_region(const Region& r, int x, int y, int z) {
  return (x >= r.lo[0]) & (x <= r.hi[0]) & (y >= r.lo[1]) & (y <= r.hi[1]) & (z >= r.lo[2]) &
         (z <= r.hi[2]);
}

// Block index remap so that consecutive logical tiles run on the same XCD
// (the dispatcher hands workgroups to the 8 XCDs round robin).
__device__ __forceinline__ int xcd_block(int bid, int nblk) {
  constexpr int NX = 8;
  int per = nblk / NX, rem = nblk % NX;
  int xcd = bid % NX, loc = bid / NX;
  return (xcd < rem) ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
}

// Block order of a lattice sweep.  Steps from parity 0 sweep the cells
// forward, steps from parity 1 backward, so each step starts on the cells the
// previous one finished with.  Their populations are still in the 256 MB
// last-level cache (over half of a 6 M-cell FP32 lattice, all of a 16-bit
// one), which cuts the HBM reads at the turn.  Both sweeps keep one front for
// the whole chip.  The backward sweep hands each XCD (workgroup b runs on XCD
// b % 8) runs of 32 consecutive tiles, so the rows its x-shifted loads and
// stores share with the next tile stay in that XCD's L2.
__device__ __forceinline__ int xcd_runs(int bid, int nblk) {
  constexpr int K = 32, SG = 8 * K;
  const int full = (nblk / SG) * SG;
  if (bid >= full) return bid;
  const int sg = bid / SG, r = bid - sg * SG;
  return sg * SG + (r % 8) * K + r / 8;
}
template <int PAR>
__device__ __forceinline__ int sweep_block(int bid, int nblk) {
  return PAR == 0 ? bid : nblk - 1 - xcd_runs(bid, nblk);
}

// Element `cell` of population plane `plane` (N cells per plane).  The plane's
// base is uniform across the launch and the cell's byte offset a 32-bit value,
// so loads and stores take the scalar-base + 32-bit-offset addressing form and
// one offset register serves all planes; a plane must therefore hold less than
// 4 GiB (check_grid rejects larger grids).  The empty asm keeps each plane's
// base in a scalar register pair as a value of its own (otherwise the compiler
// chains the bases into 64-bit vector adds, one per plane), and the pointer is
// rebuilt in the global address space so that the accesses stay global, not
// flat, ones.
template <typename T>
__device__ __forceinline__ uint64_t plane_base(T* f, size_t N, int plane) {
  uint64_t base = reinterpret_cast<uint64_t>(f + (size_t)plane * N);
  asm("" : "+s"(base));
  return base;
}
template <typename T>
__device__ __forceinline__ T* base_elem(uint64_t base, int cell) {
  using B = std::conditional_t<std::is_const<T>::value, const char, char>;
  using GB = __attribute__((address_space(1))) B;
  GB* p = reinterpret_cast<GB*>(base) + (uint32_t)cell * (uint32_t)sizeof(T);
  return (T*)p;
}
template <typename T>
__device__ __forceinline__ T* plane_elem(T* f, size_t N, int plane, int cell) {
  return base_elem<T>(plane_base(f, N, plane), cell);
}

// Post-stream populations of cell c in an AA state of parity PAR: parity 0
// keeps population q at [q][c], parity 1 at [opp(q)][c - e_q].
template <int S, int PAR>
__device__ __forceinline__ void load_state(const typename Store<S>::T* __restrict__ f,
                                           const Geo& g, int c, const Nbr& nb,
                                           typename Store<S>::R* fq) {
  const DevBase<S> dv(g);
  const size_t N = (size_t)g.ncell;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if (PAR == 0)
      fq[q] = Store<S>::ld(plane_elem(f, N, q, c), dv[q]);
    else
      fq[q] = Store<S>::ld(plane_elem(f, N, opp(q), c + nb.off(-ex(q), -ey(q), -ez(q))), dv[q]);
  }
}

// ---------------------------------------------------------------------------
// Population access of the fused step (element = plane * N + cell).
// ---------------------------------------------------------------------------
template <int S> struct Pop {
  using T = typename Store<S>::T;
  using R = typename Store<S>::R;
  uint64_t base[Q];    // plane bases, computed once (uniform)
  DevBase<S> dv;
  __device__ __forceinline__ Pop(T* f, const Geo& g) : dv(g) {
#pragma unroll
    for (int q = 0; q < Q; ++q) base[q] = plane_base(f, (size_t)g.ncell, q);
  }
  __device__ __forceinline__ R ld(int plane, int cell, int q) const {
    return Store<S>::ld(base_elem<T>(base[plane], cell), dv[q]);
  }
  // AGAIN: a store of the value an earlier st of this thread stored (the
  // outlet copy).  For f16c, the encoding of the earlier store is reused, and
  // an opaque copy of it gives each store a shift of its own: the compiler
  // then stores the code straight from the high half of the register
  // (a shift shared by two stores keeps it from doing so).
  template <bool AGAIN = false>
  __device__ __forceinline__ void st(int plane, int cell, int q, R v) const {
    uint32_t off = (uint32_t)cell * (uint32_t)sizeof(T);
    asm("" : "+v"(off));
    using GT = __attribute__((address_space(1))) char;
    T* p = (T*)(reinterpret_cast<GT*>(base[plane]) + off);
    if constexpr (S == CLBM_STORE_F16C && AGAIN) {
      uint32_t h = f16c_encode_hi(v - dv[q]);
      asm("" : "+v"(h));
      *p = (uint16_t)(h >> 16);
    } else {
      Store<S>::st(p, v, dv[q]);
    }
  }
};

// Post-stream populations of cell c at parity PAR / post-collision stores of
// a step from parity PAR, through a Pop accessor.
template <int PAR, typename P>
__device__ __forceinline__ void pop_load(const P& pop, int c, const Nbr& nb, typename P::R* fq) {
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if (PAR == 0)
      fq[q] = pop.ld(q, c, q);
    else
      fq[q] = pop.ld(opp(q), c + nb.off(-ex(q), -ey(q), -ez(q)), q);
  }
}
// Stores fq as the post-stream populations of cell c (the slots pop_load reads).
template <int PAR, typename P>
__device__ __forceinline__ void pop_store_in(const P& pop, int c, const Nbr& nb,
                                             const typename P::R* fq) {
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if (PAR == 0)
      pop.st(q, c, q, fq[q]);
    else
      pop.st(opp(q), c + nb.off(-ex(q), -ey(q), -ez(q)), q, fq[q]);
  }
}
// AGAIN: the thread stored these values before (Pop::st).
template <int PAR, bool AGAIN = false, typename P>
__device__ __forceinline__ void pop_store(const P& pop, int c, const Nbr& nb, const typename P::R* fq) {
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if (PAR == 0)
      pop.template st<AGAIN>(opp(q), c, q, fq[q]);
    else
      pop.template st<AGAIN>(q, c + nb.off(ex(q), ey(q), ez(q)), q, fq[q]);
  }
}

// The reference's inlet rule at x = 0 (CLBMSolver._bc_inlet, clbm_moving.py:
// 266-294): the five east-going populations from the known ones, in place.
// rule 1 replaces it with the on-site velocity boundary of Hecht and Harting
// (J. Stat. Mech. 2010, P01018; no counterpart in the reference): the
// normal population by bounce-back of its non-equilibrium part and each
// diagonal from its opposite, the equilibrium difference rho (ux +- u_t) / 6
// and a transverse correction that makes the cell's momentum exactly
// rho_in (ux_in, uy_in, 0).  The empty asm reads the five populations the
// rule replaces, so that their loads stay with the others above the inlet
// branch: the compiler would otherwise move them into the other path, where
// they issue only after the first loads have been waited for.
template <typename R>
__device__ __forceinline__ void inlet_rule(R* a, R ux_in, R uy_in, int rule) {
  asm("" : "+v"(a[1]), "+v"(a[7]), "+v"(a[9]), "+v"(a[11]), "+v"(a[13]));
  const R s_known = a[0] + a[3] + a[4] + a[5] + a[6] + a[15] + a[16] + a[17] + a[18];
  const R s_out = a[2] + a[8] + a[10] + a[12] + a[14];
  const R rho_in = (s_known + R(2) * s_out) / (R(1) - ux_in);
  if (rule == 0) {
    const R dy = a[3] - a[4], dz = a[5] - a[6];
    const R e = (R(1) / R(6)) * rho_in * ux_in;
    a[1] = a[2] + (R(2) / R(3)) * rho_in * ux_in;
    a[7] = a[10] + e + R(0.5) * rho_in * uy_in - R(0.5) * dy;
    a[9] = a[8] + e - R(0.5) * rho_in * uy_in + R(0.5) * dy;
    a[11] = a[14] + e - R(0.5) * dz;
    a[13] = a[12] + e + R(0.5) * dz;
  } else {
    const R third = R(1) / R(3), sixth = R(1) / R(6);
    const R n_y = R(0.5) * ((a[3] - a[4]) + (a[15] - a[16]) + (a[17] - a[18])) -
                  third * rho_in * uy_in;
    const R n_z = R(0.5) * ((a[5] - a[6]) + (a[15] + a[16]) - (a[17] + a[18]));
    a[1] = a[2] + third * rho_in * ux_in;
    a[7] = a[10] + sixth * rho_in * (ux_in + uy_in) - n_y;
    a[9] = a[8] + sixth * rho_in * (ux_in - uy_in) + n_y;
    a[11] = a[14] + sixth * rho_in * ux_in - n_z;
    a[13] = a[12] + sixth * rho_in * ux_in + n_z;
  }
}

// Boundary planes of a state: `on` when the x planes still hold the raw
// streamed values, so that readers apply the inlet rule at x = 0 and take
// cell nx-2's populations for x = nx-1 (CLBMSolver._bc_outlet,
// clbm_moving.py:296-300) as they load; `rule` selects the inlet rule
// (0 the reference's, 1 Hecht-Harting).
struct Planes {
  int on;
  int rule;
  double ux_in, uy_in;
};

// Post-stream, post-boundary populations of cell (x, y, z): what the
// reference's f holds between steps.
template <int S, int PAR>
__device__ __forceinline__ void load_cell(const typename Store<S>::T* __restrict__ f, const Geo& g,
                                          int c, int x, int y, int z, const Planes& pl,
                                          typename Store<S>::R* a) {
  using R = typename Store<S>::R;
  if (pl.on && x == g.nx - 1) {
    c -= 1;
    x -= 1;
  }
  const Nbr nb = make_nbr(g, x, y, z);
  load_state<S, PAR>(f, g, c, nb, a);
  if (pl.on && x == 0) inlet_rule<R>(a, R(pl.ux_in), R(pl.uy_in), pl.rule);
}

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
void set_error(const std::string& s);
int fail(int code, const std::string& s);
int check_launch(const char* what);
int check_grid(const clbm_grid_t* g, bool need_inlet_width);
int check_inlet(const clbm_params_t* p, const char* who);
// Validates a state word and its boundary parameters; fills `pl`.
int check_state(const clbm_grid_t* g, const clbm_params_t* p, int state, Planes& pl);

}  // namespace clbm
