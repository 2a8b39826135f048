// Device side of the immersed-boundary step (DroneMovingIBM,
// airfoil_moving.py:27-255), shared by the marker kernels and the coupled
// step: the Peskin delta, marker placement, the 4x4x4 stencils, the force-box
// record and the fused marker group (K8 + K10 + K11 + K12).
#pragma once

#include "clbm_common.h"

#include <climits>
#include <cmath>

namespace clbm {

constexpr int IBLK = 256;        // threads per block of the small marker kernels
constexpr int MPB = IBLK / 64;   // markers per block of the separate stencil kernel

// Peskin 4-point regularised delta (airfoil_moving.py:27-38).
__device__ __forceinline__ float phi1d(float r) {
  const float s = fabsf(r);
  float res = 0.0f;
  if (s < 1.0f) {
    res = (3.0f - 2.0f * s + sqrtf(fmaxf(1.0f + 4.0f * s - 4.0f * s * s, 0.0f))) * 0.125f;
  } else if (s < 2.0f) {
    res = (5.0f - 2.0f * s - sqrtf(fmaxf(-7.0f + 12.0f * s - 4.0f * s * s, 0.0f))) * 0.125f;
  }
  return res;
}

__device__ __forceinline__ int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// ---------------------------------------------------------------------------
// force box: atomic maxima of (gen << 32 | key); lower corners are stored as
// INT32_MAX - coordinate so that both corners are maxima.  Each generation
// has IBM_BOX_SUBSLOTS copies, picked by block index: atomics on one address
// serialise, so spreading the workgroups' maxima over separate words keeps
// them off the critical path; readers take the maximum over the copies.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void box_grow(uint64_t* box, uint32_t gen, const int* lo, const int* hi,
                                         int sub) {
  uint64_t* slot = box + 6 * ((gen & 1u) * IBM_BOX_SUBSLOTS + (sub & (IBM_BOX_SUBSLOTS - 1)));
  const uint64_t g = (uint64_t)gen << 32;
  for (int a = 0; a < 3; ++a) {
    if (lo[a] > hi[a]) continue;
    atomicMax((unsigned long long*)&slot[a], (unsigned long long)(g | (uint32_t)(INT_MAX - lo[a])));
    atomicMax((unsigned long long*)&slot[3 + a], (unsigned long long)(g | (uint32_t)hi[a]));
  }
}
// Box of generation gen, combined by the workgroup's first wavefront (one
// lane per copy) and shared through LDS; every thread of the block must call
// it.  False when no stencil of that generation touched a cell.  COHERENT
// reads the copies with device-scope atomic loads (within the launch whose
// workgroups grew them).
template <bool COHERENT = false>
__device__ __forceinline__ bool box_get_block(const uint64_t* box, uint32_t gen, int* lo, int* hi) {
  static_assert(IBM_BOX_SUBSLOTS == 64, "one lane per box copy");
  __shared__ uint64_t sk[6];
  if (threadIdx.x < 64) {
    uint64_t k[6] = {0, 0, 0, 0, 0, 0};
    if (gen != 0) {
      const uint64_t* slot = box + 6 * ((gen & 1u) * IBM_BOX_SUBSLOTS + threadIdx.x);
      for (int a = 0; a < 6; ++a)
        k[a] = COHERENT ? __hip_atomic_load(slot + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                        : slot[a];
    }
    for (int o = 32; o > 0; o >>= 1)
      for (int a = 0; a < 6; ++a) {
        const uint64_t t = __shfl_xor(k[a], o);
        k[a] = t > k[a] ? t : k[a];
      }
    if (threadIdx.x == 0)
      for (int a = 0; a < 6; ++a) sk[a] = k[a];
  }
  __syncthreads();
  bool ok = gen != 0;
  for (int a = 0; a < 3; ++a) {
    const uint64_t kl = sk[a], kh = sk[3 + a];
    ok = ok && (uint32_t)(kl >> 32) == gen && (uint32_t)(kh >> 32) == gen;
    lo[a] = INT_MAX - (int)(uint32_t)kl;
    hi[a] = (int)(uint32_t)kh;
  }
  __syncthreads();  // sk is reused by the next call
  return ok;
}

// Inclusive index range of the 4 stencil cells along one axis.  x is
// clamped; a periodic axis whose 4 cells wrap takes the whole axis.
__device__ __forceinline__ void axis_range_clamped(int b, int n, int& lo, int& hi) {
  lo = min(max(b, 0), n - 1);
  hi = min(max(b + 3, 0), n - 1);
}
__device__ __forceinline__ void axis_range_periodic(int b, int n, int& lo, int& hi) {
  if (b >= 0 && b + 3 <= n - 1) {
    lo = b;
    hi = b + 3;
  } else {
    lo = 0;
    hi = n - 1;
  }
}
__device__ __forceinline__ void stencil_ranges(const Geo& g, int ib, int jb, int kb, int* lo, int* hi) {
  axis_range_clamped(ib, g.nx, lo[0], hi[0]);
  axis_range_periodic(jb, g.ny, lo[1], hi[1]);
  axis_range_periodic(kb, g.nz, lo[2], hi[2]);
}

// ---------------------------------------------------------------------------
// K8 arithmetic (airfoil_moving.py:114-127), float32 without contraction
// ---------------------------------------------------------------------------
struct Kin {
  float xc, yc, ct, st, vx, vy, omega;
};
__device__ __forceinline__ void place_marker(const Kin& k, float x0m, float y0m, float& xk,
                                             float& yk, float& vxk, float& vyk) {
#pragma clang fp contract(off)
  xk = (k.xc + x0m * k.ct) - y0m * k.st;
  yk = (k.yc + x0m * k.st) + y0m * k.ct;
  vxk = k.vx - k.omega * (yk - k.yc);
  vyk = k.vy + k.omega * (xk - k.xc);
}

// ---------------------------------------------------------------------------
// Stencil geometry of one lane: (di, dj, dk) = (lane & 3, lane >> 2 & 3, lane >> 4),
// so that lanes next to each other read cells next to each other in x.
// ---------------------------------------------------------------------------
struct Lane {
  int ib, jb, kb;  // floor(pos) - 1
  int iu, ju, ku;  // unwrapped stencil coordinates
  int cell;        // clamped / wrapped cell index
  int ii, jj, kk;  // clamped / wrapped cell coordinates
  float w;         // (phx * phy) * phz
};

__device__ __forceinline__ Lane lane_stencil(const Geo& g, int quirk, float xk, float yk, float zk,
                                             int lane) {
  Lane L;
  L.ib = (int)floorf(xk) - 1;
  L.jb = (int)floorf(yk) - 1;
  L.kb = (int)floorf(zk) - 1;
  L.iu = L.ib + (lane & 3);
  L.ju = L.jb + ((lane >> 2) & 3);
  L.ku = L.kb + (lane >> 4);
  L.ii = min(max(L.iu, 0), g.nx - 1);
  L.jj = pmod(L.ju, g.ny);
  L.kk = pmod(L.ku, g.nz);
  const float phx = phi1d((float)(quirk ? L.ii : L.iu) - xk);
  const float phy = phi1d((float)(quirk ? L.jj : L.ju) - yk);
  const float phz = phi1d((float)(quirk ? L.kk : L.ku) - zk);
  L.w = (phx * phy) * phz;
  L.cell = (L.kk * g.ny + L.jj) * g.nx + L.ii;
  return L;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Force-free velocity of one cell from the population state (K7's formula,
// clbm_moving.py:305-321).
template <int S, int PAR>
__device__ __forceinline__ void cell_velocity(const Geo& g, const Planes& pl,
                                              const typename Store<S>::T* f, int c, int x, int y,
                                              int z, float& uo, float& vo) {
  using R = typename Store<S>::R;
  R a[Q];
  load_cell<S, PAR>(f, g, c, x, y, z, pl, a);
  R r = R(0), mx = R(0), my = R(0);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    r += a[q];
    if (ex(q)) mx += R(ex(q)) * a[q];
    if (ey(q)) my += R(ey(q)) * a[q];
  }
  uo = (float)(mx / r);
  vo = (float)(my / r);
}

template <int S, int UMODE>
__device__ __forceinline__ void lane_velocity(const Geo& g, const Planes& pl,
                                              const typename Store<S>::T* f, const float* uxf,
                                              const float* uyf, const Lane& L, float& uc,
                                              float& vc) {
  if (UMODE == 0) {
    uc = uxf[L.cell];
    vc = uyf[L.cell];
  } else if (UMODE == 1) {
    cell_velocity<S, 0>(g, pl, f, L.cell, L.ii, L.jj, L.kk, uc, vc);
  } else {
    cell_velocity<S, 1>(g, pl, f, L.cell, L.ii, L.jj, L.kk, uc, vc);
  }
}

// One contribution into the global accumulators.
template <int SMODE>
__device__ __forceinline__ void acc_add(void* accx, void* accy, int c, float px, float py,
                                        float fscale) {
  if (SMODE == IBM_SPREAD_FIXED) {
    atomicAdd((unsigned long long*)accx + c, (unsigned long long)__float2ll_rn(px * fscale));
    atomicAdd((unsigned long long*)accy + c, (unsigned long long)__float2ll_rn(py * fscale));
  } else {
    atomicAdd((float*)((unsigned long long*)accx + c), px);
    atomicAdd((float*)((unsigned long long*)accy + c), py);
  }
}

// ---------------------------------------------------------------------------
// The fused kernel: K8 + K10 + K11 + K12 for G markers per workgroup.
// UMODE 0: velocity from ux/uy fields; 1: from the state at parity 0; 2: parity 1.
//
// The workgroup's stencils are staged in an LDS tile spanning the bounding
// box of their unwrapped coordinates: the force-free velocity of each tile
// cell is computed once (instead of once per stencil point that reaches it),
// and the spread products are summed per tile cell before one global atomic
// per touched cell.  A group whose box exceeds the tile (markers far apart)
// takes the direct path for both.
// ---------------------------------------------------------------------------
template <int G> struct Group {
  static constexpr int threads = 64 * G;
  static constexpr int tile = 128 * G;  // cells
};

// The region every stencil cell must lie in (when `on`): the cells a
// concurrent outside-part step leaves alone.
struct RegionCheck {
  int on;
  int lo[3], hi[3];
};

// Everything the marker kernel needs for one body on one lattice.
struct IbmArgs {
  Geo g;
  Planes pl;
  int quirk;
  float f_max, fscale;
  const void* f;
  const float *uxf, *uyf;
  int n;
  const int32_t* order;
  const float *X0, *Y0, *Z0;
  float *X, *Y, *Z, *VX, *VY, *U, *V, *FX, *FY;
  int32_t* base;
  void *accx, *accy;
  uint64_t* box;
  RegionCheck rc;
  const float* kin;  // batched runs: per step xc, yc, cos, sin, vx, vy, omega
  uint32_t gen0;     // batched runs: generation of step 0 of the step counter, minus one
  const int32_t* regions;  // batched runs: per step the predicted region (6 int32), or NULL
};

// Marker group `grp` (G markers, one wavefront each) of generation gen;
// `sub` picks the force-box copy this workgroup updates.
template <int S, int UMODE, int SMODE, int G>
__device__ __forceinline__ void compute_group(const IbmArgs& a, const Kin& kin, uint32_t gen,
                                              int grp, int sub) {
  const Geo& g = a.g;
  const Planes& pl = a.pl;
  const int quirk = a.quirk, n = a.n;
  const float f_max = a.f_max, fscale = a.fscale;
  const auto* __restrict__ f = (const typename Store<S>::T*)a.f;
  const float *__restrict__ uxf = a.uxf, *__restrict__ uyf = a.uyf;
  const int32_t* __restrict__ order = a.order;
  const float *__restrict__ X0 = a.X0, *__restrict__ Y0 = a.Y0, *__restrict__ Z0 = a.Z0;
  float *__restrict__ X = a.X, *__restrict__ Y = a.Y, *__restrict__ Z = a.Z;
  float *__restrict__ VX = a.VX, *__restrict__ VY = a.VY, *__restrict__ U = a.U, *__restrict__ V = a.V;
  float *__restrict__ FX = a.FX, *__restrict__ FY = a.FY;
  int32_t* __restrict__ base = a.base;
  void *accx = a.accx, *accy = a.accy;
  uint64_t* box = a.box;
  const RegionCheck& rc = a.rc;
  constexpr int NT = Group<G>::threads, TC = Group<G>::tile;
  __shared__ unsigned long long sacc[2 * TC];  // spread sums (x then y)
  __shared__ float svel[2 * TC];               // staged velocity (u then v)
  __shared__ int sbase[G][3];
  __shared__ int srng[G][6];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int slot = grp * G + wv;
  const bool live = slot < n;  // uniform per wavefront
  const int m = live ? (order ? order[slot] : slot) : 0;

  // K8: place the marker (every lane of the wave computes the same values)
  float xk = 0.f, yk = 0.f, zk = 0.f, vxk = 0.f, vyk = 0.f;
  if (live) {
    place_marker(kin, X0[m], Y0[m], xk, yk, vxk, vyk);
    zk = Z0[m];
    if (lane == 0) {
      X[m] = xk;
      Y[m] = yk;
      Z[m] = zk;
      VX[m] = vxk;
      VY[m] = vyk;
    }
  }
  const Lane L = lane_stencil(g, quirk, xk, yk, zk, lane);
  if (lane == 0) {
    sbase[wv][0] = live ? L.ib : INT_MAX;
    sbase[wv][1] = live ? L.jb : INT_MAX;
    sbase[wv][2] = live ? L.kb : INT_MAX;
    int lo[3], hi[3];
    stencil_ranges(g, L.ib, L.jb, L.kb, lo, hi);
    for (int a = 0; a < 3; ++a) {
      srng[wv][a] = live ? lo[a] : INT_MAX;
      srng[wv][3 + a] = live ? hi[a] : INT_MIN;
    }
    if (live && base) {
      base[3 * m + 0] = L.ib;
      base[3 * m + 1] = L.jb;
      base[3 * m + 2] = L.kb;
    }
  }
  __syncthreads();
  // the box of the cells this generation's stencils reach
  if (box && threadIdx.x == 0) {
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int i = 0; i < G; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = min(lo[a], srng[i][a]);
        hi[a] = max(hi[a], srng[i][3 + a]);
      }
    box_grow(box, gen, lo, hi, sub);
    if (rc.on && lo[0] != INT_MAX) {
      bool out = false;
      for (int a = 0; a < 3; ++a) out |= (lo[a] < rc.lo[a]) | (hi[a] > rc.hi[a]);
      if (out)
        atomicMax((unsigned long long*)IBM_REGION_FLAG(box), (unsigned long long)gen);
    }
  }
  // the tile: bounding box of the workgroup's unwrapped stencil coordinates
  int o[3] = {INT_MAX, INT_MAX, INT_MAX}, e[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (int i = 0; i < G; ++i) {
    if (sbase[i][0] == INT_MAX) continue;
    for (int a = 0; a < 3; ++a) {
      o[a] = min(o[a], sbase[i][a]);
      e[a] = max(e[a], sbase[i][a] + 3);
    }
  }
  const bool any = o[0] != INT_MAX;
  const long long lbx = (long long)e[0] - o[0] + 1, lby = (long long)e[1] - o[1] + 1,
                  lbz = (long long)e[2] - o[2] + 1;
  const bool use_tile = any && lbx * lby * lbz <= TC;
  const int bx = use_tile ? (int)lbx : 0, by = use_tile ? (int)lby : 0;
  const int vol = use_tile ? (int)(lbx * lby * lbz) : 0;
  // stage the velocity of every tile cell; clear the spread sums
  for (int j = threadIdx.x; j < vol; j += NT) {
    const int ix = j % bx, r = j / bx;
    const int iy = r % by, iz = r / by;
    const int ii = min(max(o[0] + ix, 0), g.nx - 1);
    const int jj = pmod(o[1] + iy, g.ny);
    const int kk = pmod(o[2] + iz, g.nz);
    Lane T;
    T.ii = ii;
    T.jj = jj;
    T.kk = kk;
    T.cell = (kk * g.ny + jj) * g.nx + ii;
    float u, v;
    lane_velocity<S, UMODE>(g, pl, f, uxf, uyf, T, u, v);
    svel[j] = u;
    svel[TC + j] = v;
    if (SMODE == IBM_SPREAD_FIXED) {
      sacc[j] = 0ull;
      sacc[TC + j] = 0ull;
    } else {
      ((float*)sacc)[j] = 0.0f;
      ((float*)sacc)[TC + j] = 0.0f;
    }
  }
  __syncthreads();

  if (live) {
    const int j = use_tile ? ((L.ku - o[2]) * by + (L.ju - o[1])) * bx + (L.iu - o[0]) : 0;
    // K10 _interpolate (airfoil_moving.py:142-170)
    float uc, vc;
    if (use_tile) {
      uc = svel[j];
      vc = svel[TC + j];
    } else {
      lane_velocity<S, UMODE>(g, pl, f, uxf, uyf, L, uc, vc);
    }
    const float su = wave_sum(uc * L.w);
    const float sv = wave_sum(vc * L.w);
    // K11 _marker_forces (airfoil_moving.py:175-183)
    const float fxk = fmaxf(-f_max, fminf(f_max, -(su - vxk)));
    const float fyk = fmaxf(-f_max, fminf(f_max, -(sv - vyk)));
    if (lane == 0) {
      U[m] = su;
      V[m] = sv;
      FX[m] = fxk;
      FY[m] = fyk;
    }
    // K12 _spread (airfoil_moving.py:188-214)
    const float px = fxk * L.w, py = fyk * L.w;
    if (use_tile) {
      if (SMODE == IBM_SPREAD_FIXED) {
        atomicAdd(&sacc[j], (unsigned long long)__float2ll_rn(px * fscale));
        atomicAdd(&sacc[TC + j], (unsigned long long)__float2ll_rn(py * fscale));
      } else {
        atomicAdd(&((float*)sacc)[j], px);
        atomicAdd(&((float*)sacc)[TC + j], py);
      }
    } else {
      acc_add<SMODE>(accx, accy, L.cell, px, py, fscale);
    }
  }
  __syncthreads();
  // one global atomic per touched cell of the tile
  for (int j = threadIdx.x; j < vol; j += NT) {
    const int ix = j % bx, r = j / bx;
    const int iy = r % by, iz = r / by;
    const int ii = min(max(o[0] + ix, 0), g.nx - 1);
    const int jj = pmod(o[1] + iy, g.ny);
    const int kk = pmod(o[2] + iz, g.nz);
    const int c = (kk * g.ny + jj) * g.nx + ii;
    if (SMODE == IBM_SPREAD_FIXED) {
      const unsigned long long a = sacc[j], b = sacc[TC + j];
      if (a) atomicAdd((unsigned long long*)accx + c, a);
      if (b) atomicAdd((unsigned long long*)accy + c, b);
    } else {
      const float a = ((const float*)sacc)[j], b = ((const float*)sacc)[TC + j];
      if (a != 0.0f) atomicAdd((float*)((unsigned long long*)accx + c), a);
      if (b != 0.0f) atomicAdd((float*)((unsigned long long*)accy + c), b);
    }
  }
}

// ---------------------------------------------------------------------------
// host-side helpers of the marker entry points
// ---------------------------------------------------------------------------
static inline int check_ibm(const ibm_params_t* ip) {
  if (!ip) return fail(-1, "ibm params are NULL");
  if (ip->spread_mode != IBM_SPREAD_FIXED && ip->spread_mode != IBM_SPREAD_ATOMIC_F32 &&
      ip->spread_mode != IBM_SPREAD_SORTED)
    return fail(-1, "unknown spread mode");
  if (!(ip->f_max > 0.0f) || !std::isfinite(ip->f_max)) return fail(-1, "f_max must be > 0");
  if (ip->fixed_shift < 0 || ip->fixed_shift > 60) return fail(-1, "fixed_shift out of [0, 60]");
  if (ip->group != 0 && ip->group != 4 && ip->group != 8 && ip->group != 16)
    return fail(-1, "group must be 0, 4, 8 or 16");
  if (ip->check_region != 0 && ip->check_region != 1) return fail(-1, "check_region must be 0 or 1");
  if (ip->check_region)
    for (int a = 0; a < 3; ++a)
      if (ip->region[a] > ip->region[3 + a]) return fail(-1, "region has lo > hi");
  return 0;
}

static inline RegionCheck region_check(const ibm_params_t* ip) {
  RegionCheck r{};
  r.on = ip->check_region;
  for (int a = 0; a < 3; ++a) {
    r.lo[a] = ip->region[a];
    r.hi[a] = ip->region[3 + a];
  }
  return r;
}

// Largest shift that keeps 2^20 clamped contributions |F w| <= f_max / 8
// (a Peskin weight is at most 1/8 per cell) below 2^62.
static inline int fixed_shift(const ibm_params_t* ip) {
  if (ip->fixed_shift > 0) return ip->fixed_shift;
  const double lim = std::ldexp(1.0, 42) / (ip->f_max * 0.125);
  int s = (int)std::floor(std::log2(lim));
  return s < 0 ? 0 : (s > 60 ? 60 : s);
}

static inline Kin make_kin(float xc, float yc, float theta, float vx, float vy, float omega) {
  // cos/sin of the float32 angle in float64, rounded once (the same for every marker)
  return Kin{xc, yc, (float)std::cos((double)theta), (float)std::sin((double)theta), vx, vy, omega};
}

}  // namespace clbm
