// Immersed-boundary kernels of the CLBM-IBM step for gfx950 (MI355X).
//
// DroneMovingIBM.compute_force (airfoil_moving.py:230-255) runs six kernels,
// three of them over the full grid.  Here every kernel works on markers or on
// the band of cells the markers' 4x4x4 Peskin stencils reach:
//   * ibm_update_markers   K8, one thread per marker; also reduces the
//                          bounding box of all stencil cells (the "force
//                          box") with a last-block reduction;
//   * ibm_band_velocity    force-free velocity (K7's formula) of the box
//                          cells only, straight from the population state;
//   * ibm_interp_spread    K10 + K11 + K12: one 64-lane wavefront per marker,
//                          one lane per stencil point (x fastest across
//                          lanes); the wave reduces U_k, V_k, forms the
//                          clamped marker force and every lane adds F_k w to
//                          its cell with the same weight;
//   * ibm_clip_forces      K9 + K13 on the union of this and the previous
//                          force box (cells outside both are already zero).
// Spreading is order-independent by default: each product F_k w is converted
// to a 64-bit fixed-point integer and summed with integer atomics, so the
// grid force has the same bit pattern on every run.  The reference's float
// atomics (airfoil_moving.py:213-214) are kept as IBM_SPREAD_ATOMIC_F32.
#include "clbm_common.h"

#include <algorithm>
#include <climits>
#include <cmath>

namespace clbm {

constexpr int IBLK = 256;       // threads per block of the marker kernels
constexpr int MPB = IBLK / 64;  // markers per block of the stencil kernel
constexpr int UBLK_MAX = IBM_BOX_PARTIALS;  // blocks of the marker-update kernel

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

// ---------------------------------------------------------------------------
// K8 _update_markers (airfoil_moving.py:100-127) + force box of the stencils
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(IBLK) void k_update_markers(
    Geo g, int n, const float* __restrict__ X0, const float* __restrict__ Y0,
    const float* __restrict__ Z0, float xc, float yc, float theta, float vx, float vy, float omega,
    float* __restrict__ X, float* __restrict__ Y, float* __restrict__ Z, float* __restrict__ VX,
    float* __restrict__ VY, int32_t* __restrict__ box) {
#pragma clang fp contract(off)
  __shared__ int sb[IBLK / 64][6];
  __shared__ int last;
  const double td = (double)theta;
  const float ct = (float)cos(td);
  const float st = (float)sin(td);
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (int m = blockIdx.x * IBLK + threadIdx.x; m < n; m += gridDim.x * IBLK) {
    const float x0m = X0[m], y0m = Y0[m];
    const float xk = (xc + x0m * ct) - y0m * st;
    const float yk = (yc + x0m * st) + y0m * ct;
    const float zk = Z0[m];
    X[m] = xk;
    Y[m] = yk;
    Z[m] = zk;
    VX[m] = vx - omega * (yk - yc);
    VY[m] = vy + omega * (xk - xc);
    if (box) {
      int l, h;
      axis_range_clamped((int)floorf(xk) - 1, g.nx, l, h);
      lo[0] = min(lo[0], l);
      hi[0] = max(hi[0], h);
      axis_range_periodic((int)floorf(yk) - 1, g.ny, l, h);
      lo[1] = min(lo[1], l);
      hi[1] = max(hi[1], h);
      axis_range_periodic((int)floorf(zk) - 1, g.nz, l, h);
      lo[2] = min(lo[2], l);
      hi[2] = max(hi[2], h);
    }
  }
  if (!box) return;
  // block reduction of the box
  for (int a = 0; a < 3; ++a) {
    for (int o = 32; o > 0; o >>= 1) {
      lo[a] = min(lo[a], __shfl_xor(lo[a], o));
      hi[a] = max(hi[a], __shfl_xor(hi[a], o));
    }
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      sb[wv][a] = lo[a];
      sb[wv][3 + a] = hi[a];
    }
  __syncthreads();
  int32_t* part = box + IBM_BOX_HEADER;
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    int v = sb[0][a];
    for (int i = 1; i < IBLK / 64; ++i) v = a < 3 ? min(v, sb[i][a]) : max(v, sb[i][a]);
    part[6 * blockIdx.x + a] = v;
  }
  // the last block to finish combines the partial boxes
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = (atomicAdd(&box[13], 1) == (int)gridDim.x - 1);
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    volatile int32_t* vp = part;
    int v = vp[a];
    for (int b = 1; b < (int)gridDim.x; ++b) v = a < 3 ? min(v, vp[6 * b + a]) : max(v, vp[6 * b + a]);
    box[6 + a] = box[a];  // previous box
    box[a] = v;
  }
  if (threadIdx.x == 0) box[13] = 0;
}

// ---------------------------------------------------------------------------
// Force-free velocity on the box cells (K7's formula, clbm_moving.py:305-321)
// ---------------------------------------------------------------------------
template <int S, int PAR>
__global__ __launch_bounds__(IBLK) void k_band_velocity(Geo g, const typename Store<S>::T* __restrict__ f,
                                                        const int32_t* __restrict__ box,
                                                        float* __restrict__ uo, float* __restrict__ vo) {
  using R = typename Store<S>::R;
  const int x0 = box[0], y0 = box[1], z0 = box[2];
  const int bx = box[3] - x0 + 1, by = box[4] - y0 + 1, bz = box[5] - z0 + 1;
  if (bx <= 0 || by <= 0 || bz <= 0) return;
  const int vol = bx * by * bz;
  for (int t = blockIdx.x * IBLK + threadIdx.x; t < vol; t += gridDim.x * IBLK) {
    const int r = t / bx;
    const int x = x0 + (t - r * bx);
    const int k = r / by;
    const int y = y0 + (r - k * by);
    const int z = z0 + k;
    const int c = (z * g.ny + y) * g.nx + x;
    const Nbr nb = make_nbr(g, x, y, z);
    R a[Q];
    load_state<S, PAR>(f, (size_t)g.ncell, c, nb, a);
    R rho = R(0), mx = R(0), my = R(0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      rho += a[q];
      if (ex(q)) mx += R(ex(q)) * a[q];
      if (ey(q)) my += R(ey(q)) * a[q];
    }
    uo[c] = (float)(mx / rho);
    vo[c] = (float)(my / rho);
  }
}

// ---------------------------------------------------------------------------
// Stencil geometry of one lane: (di, dj, dk) = (lane & 3, lane >> 2 & 3, lane >> 4),
// so that lanes next to each other read cells next to each other in x.
// ---------------------------------------------------------------------------
struct Lane {
  int ib, jb, kb;  // floor(pos) - 1
  int ii, jj, kk;  // clamped / wrapped cell indices
  float w;         // (phx * phy) * phz
  int cell;
};

__device__ __forceinline__ Lane lane_stencil(const Geo& g, int quirk, float xk, float yk, float zk,
                                             int lane) {
  Lane L;
  const int di = lane & 3, dj = (lane >> 2) & 3, dk = lane >> 4;
  L.ib = (int)floorf(xk) - 1;
  L.jb = (int)floorf(yk) - 1;
  L.kb = (int)floorf(zk) - 1;
  const int iu = L.ib + di, ju = L.jb + dj, ku = L.kb + dk;
  L.ii = min(max(iu, 0), g.nx - 1);
  L.jj = pmod(ju, g.ny);
  L.kk = pmod(ku, g.nz);
  const float phx = phi1d((float)(quirk ? L.ii : iu) - xk);
  const float phy = phi1d((float)(quirk ? L.jj : ju) - yk);
  const float phz = phi1d((float)(quirk ? L.kk : ku) - zk);
  L.w = (phx * phy) * phz;
  L.cell = (L.kk * g.ny + L.jj) * g.nx + L.ii;
  return L;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Force-free velocity of one cell from the population state.
template <int S, int PAR>
__device__ __forceinline__ void cell_velocity(const Geo& g, const typename Store<S>::T* f, int c,
                                              int x, int y, int z, float& uo, float& vo) {
  using R = typename Store<S>::R;
  const Nbr nb = make_nbr(g, x, y, z);
  R a[Q];
  load_state<S, PAR>(f, (size_t)g.ncell, c, nb, a);
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

// ---------------------------------------------------------------------------
// K10 + K11 (+ K12 when SPREAD): one wavefront per marker.
// UMODE 0: velocity from ux/uy fields; 1: from the state at parity 0; 2: parity 1.
// ---------------------------------------------------------------------------
template <int S, int UMODE, bool INTERP, bool SPREAD, int SMODE>
__global__ __launch_bounds__(IBLK) void k_marker(
    Geo g, int quirk, float f_max, float fscale, const typename Store<S>::T* __restrict__ f,
    const float* __restrict__ uxf, const float* __restrict__ uyf, int n,
    const float* __restrict__ X, const float* __restrict__ Y, const float* __restrict__ Z,
    const float* __restrict__ VX, const float* __restrict__ VY, float* __restrict__ U,
    float* __restrict__ V, float* __restrict__ FXo, float* __restrict__ FYo,
    const float* __restrict__ FXi, const float* __restrict__ FYi, int32_t* __restrict__ base,
    void* __restrict__ accx, void* __restrict__ accy) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * MPB + (threadIdx.x >> 6);
  if (m >= n) return;  // uniform per wavefront
  const float xk = X[m], yk = Y[m], zk = Z[m];
  const Lane L = lane_stencil(g, quirk, xk, yk, zk, lane);
  float fxk, fyk;
  if (INTERP) {
    float uc, vc;
    if (UMODE == 0) {
      uc = uxf[L.cell];
      vc = uyf[L.cell];
    } else if (UMODE == 1) {
      cell_velocity<S, 0>(g, f, L.cell, L.ii, L.jj, L.kk, uc, vc);
    } else {
      cell_velocity<S, 1>(g, f, L.cell, L.ii, L.jj, L.kk, uc, vc);
    }
    const float su = wave_sum(uc * L.w);
    const float sv = wave_sum(vc * L.w);
    // K11 _marker_forces (airfoil_moving.py:175-183)
    fxk = fmaxf(-f_max, fminf(f_max, -(su - VX[m])));
    fyk = fmaxf(-f_max, fminf(f_max, -(sv - VY[m])));
    if (lane == 0) {
      if (U) U[m] = su;
      if (V) V[m] = sv;
      if (FXo) FXo[m] = fxk;
      if (FYo) FYo[m] = fyk;
    }
  } else {
    fxk = FXi[m];
    fyk = FYi[m];
  }
  if (base && lane == 0) {
    base[3 * m + 0] = L.ib;
    base[3 * m + 1] = L.jb;
    base[3 * m + 2] = L.kb;
  }
  if (SPREAD) {
    const float px = fxk * L.w, py = fyk * L.w;
    if (SMODE == IBM_SPREAD_FIXED) {
      atomicAdd((unsigned long long*)accx + L.cell, (unsigned long long)__float2ll_rn(px * fscale));
      atomicAdd((unsigned long long*)accy + L.cell, (unsigned long long)__float2ll_rn(py * fscale));
    } else {
      atomicAdd((float*)((unsigned long long*)accx + L.cell), px);
      atomicAdd((float*)((unsigned long long*)accy + L.cell), py);
    }
  }
}

// ---------------------------------------------------------------------------
// K9 + K13: accumulators -> clamped float fields on the union of the boxes.
// ---------------------------------------------------------------------------
template <int SMODE>
__global__ __launch_bounds__(IBLK) void k_clip(Geo g, float f_max, float inv_scale,
                                               void* __restrict__ accx, void* __restrict__ accy,
                                               float* __restrict__ fx, float* __restrict__ fy,
                                               const int32_t* __restrict__ box) {
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = min(box[a], box[6 + a]);
    hi[a] = max(box[3 + a], box[9 + a]);
  }
  if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return;  // both boxes empty
  const int bx = hi[0] - lo[0] + 1, by = hi[1] - lo[1] + 1, bz = hi[2] - lo[2] + 1;
  const int vol = bx * by * bz;
  for (int t = blockIdx.x * IBLK + threadIdx.x; t < vol; t += gridDim.x * IBLK) {
    const int r = t / bx;
    const int i = t - r * bx;
    const int k = r / by;
    const int j = r - k * by;
    const int c = ((lo[2] + k) * g.ny + (lo[1] + j)) * g.nx + (lo[0] + i);
    float vx, vy;
    if (SMODE == IBM_SPREAD_FIXED) {
      long long* ax = (long long*)accx;
      long long* ay = (long long*)accy;
      vx = (float)((double)ax[c] * (double)inv_scale);
      vy = (float)((double)ay[c] * (double)inv_scale);
      ax[c] = 0;
      ay[c] = 0;
    } else {
      float* ax = (float*)((unsigned long long*)accx + c);
      float* ay = (float*)((unsigned long long*)accy + c);
      vx = *ax;
      vy = *ay;
      *ax = 0.0f;
      *ay = 0.0f;
    }
    if (fx) fx[c] = fmaxf(-f_max, fminf(f_max, vx));
    if (fy) fy[c] = fmaxf(-f_max, fminf(f_max, vy));
  }
}

// ---------------------------------------------------------------------------
// FP64 marker force sums, one block, fixed reduction tree (deterministic).
// ---------------------------------------------------------------------------
constexpr int SBLK = 256;
__global__ __launch_bounds__(SBLK) void k_force_sum(int n, const float* __restrict__ FX,
                                                    const float* __restrict__ FY,
                                                    const float* __restrict__ dV,
                                                    double* __restrict__ out2) {
  __shared__ double sx[SBLK], sy[SBLK];
  double ax = 0.0, ay = 0.0;
  for (int m = threadIdx.x; m < n; m += SBLK) {
    const double d = dV ? (double)dV[m] : 1.0;
    ax += (double)FX[m] * d;
    ay += (double)FY[m] * d;
  }
  sx[threadIdx.x] = ax;
  sy[threadIdx.x] = ay;
  __syncthreads();
  for (int s = SBLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sx[threadIdx.x] += sx[threadIdx.x + s];
      sy[threadIdx.x] += sy[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out2[0] = sx[0];
    out2[1] = sy[0];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int check_ibm(const ibm_params_t* ip) {
  if (!ip) return fail(-1, "ibm params are NULL");
  if (ip->spread_mode != IBM_SPREAD_FIXED && ip->spread_mode != IBM_SPREAD_ATOMIC_F32)
    return fail(-1, "unknown spread mode");
  if (!(ip->f_max > 0.0f) || !std::isfinite(ip->f_max)) return fail(-1, "f_max must be > 0");
  if (ip->fixed_shift < 0 || ip->fixed_shift > 60) return fail(-1, "fixed_shift out of [0, 60]");
  return 0;
}

// Largest shift that keeps 2^20 clamped contributions |F w| <= f_max / 8
// (a Peskin weight is at most 1/8 per cell) below 2^62.
static int fixed_shift(const ibm_params_t* ip) {
  if (ip->fixed_shift > 0) return ip->fixed_shift;
  const double lim = std::ldexp(1.0, 42) / (ip->f_max * 0.125);
  int s = (int)std::floor(std::log2(lim));
  return s < 0 ? 0 : (s > 60 ? 60 : s);
}

static inline int mblocks(int n) { return (n + MPB - 1) / MPB; }

// Blocks of the grid-stride box kernels: enough to cover a box of the whole
// lattice in a few passes, and at least one block per CU.
static inline int box_blocks(const Geo& g) {
  return std::max(1, std::min(4096, (g.ncell + IBLK - 1) / IBLK));
}

template <int S, bool INTERP, bool SPREAD, int SMODE>
static void launch_marker(const Geo& g, const ibm_params_t* ip, const void* f, int parity,
                          const float* ux, const float* uy, int n, const float* X, const float* Y,
                          const float* Z, const float* VX, const float* VY, float* U, float* V,
                          float* FXo, float* FYo, const float* FXi, const float* FYi, int32_t* base,
                          void* accx, void* accy, hipStream_t st) {
  const float scale = (float)std::ldexp(1.0, fixed_shift(ip));
  auto fp = (const typename Store<S>::T*)f;
  const dim3 gr(mblocks(n)), bl(IBLK);
  const int uq = ip->wrap_weight_quirk ? 1 : 0;
  if (!INTERP || (ux && uy))
    hipLaunchKernelGGL((k_marker<S, 0, INTERP, SPREAD, SMODE>), gr, bl, 0, st, g, uq, ip->f_max,
                       scale, fp, ux, uy, n, X, Y, Z, VX, VY, U, V, FXo, FYo, FXi, FYi, base, accx,
                       accy);
  else if (parity == 0)
    hipLaunchKernelGGL((k_marker<S, 1, INTERP, SPREAD, SMODE>), gr, bl, 0, st, g, uq, ip->f_max,
                       scale, fp, ux, uy, n, X, Y, Z, VX, VY, U, V, FXo, FYo, FXi, FYi, base, accx,
                       accy);
  else
    hipLaunchKernelGGL((k_marker<S, 2, INTERP, SPREAD, SMODE>), gr, bl, 0, st, g, uq, ip->f_max,
                       scale, fp, ux, uy, n, X, Y, Z, VX, VY, U, V, FXo, FYo, FXi, FYi, base, accx,
                       accy);
}

template <bool INTERP, bool SPREAD, int SMODE>
static int marker_dispatch(const clbm_grid_t* g, const ibm_params_t* ip, const void* f, int parity,
                           const float* ux, const float* uy, int n, const float* X, const float* Y,
                           const float* Z, const float* VX, const float* VY, float* U, float* V,
                           float* FXo, float* FYo, const float* FXi, const float* FYi,
                           int32_t* base, void* accx, void* accy, hipStream_t st) {
  const Geo geo = make_geo(*g);
#define CLBM_MK(SS)                                                                              \
  launch_marker<SS, INTERP, SPREAD, SMODE>(geo, ip, f, parity, ux, uy, n, X, Y, Z, VX, VY, U, V, \
                                           FXo, FYo, FXi, FYi, base, accx, accy, st)
  switch (g->storage) {
    case CLBM_STORE_F32: CLBM_MK(CLBM_STORE_F32); break;
    case CLBM_STORE_F16: CLBM_MK(CLBM_STORE_F16); break;
    case CLBM_STORE_BF16: CLBM_MK(CLBM_STORE_BF16); break;
    case CLBM_STORE_F64: CLBM_MK(CLBM_STORE_F64); break;
    default: return fail(-1, "unknown storage format");
  }
#undef CLBM_MK
  return check_launch("ibm marker kernel");
}

template <int S>
static void launch_band(const Geo& g, const void* f, int parity, const int32_t* box, float* uo,
                        float* vo, hipStream_t st) {
  auto fp = (const typename Store<S>::T*)f;
  if (parity == 0)
    hipLaunchKernelGGL((k_band_velocity<S, 0>), dim3(box_blocks(g)), dim3(IBLK), 0, st, g, fp, box,
                       uo, vo);
  else
    hipLaunchKernelGGL((k_band_velocity<S, 1>), dim3(box_blocks(g)), dim3(IBLK), 0, st, g, fp, box,
                       uo, vo);
}

}  // namespace clbm

using namespace clbm;

extern "C" {

int ibm_update_markers(const clbm_grid_t* g, int n, const float* X0, const float* Y0,
                       const float* Z0, float xc, float yc, float theta, float vx, float vy,
                       float omega, float* X, float* Y, float* Z, float* VX, float* VY,
                       int32_t* box, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (n < 0) return fail(-1, "ibm_update_markers: n < 0");
  if (n > 0 && (!X0 || !Y0 || !Z0 || !X || !Y || !Z || !VX || !VY))
    return fail(-1, "ibm_update_markers: NULL marker array");
  if (n == 0 && !box) return 0;
  const int nb = std::max(1, std::min(UBLK_MAX, (n + IBLK - 1) / IBLK));
  hipLaunchKernelGGL(k_update_markers, dim3(nb), dim3(IBLK), 0, (hipStream_t)stream,
                     make_geo(*g), n, X0, Y0, Z0, xc, yc, theta, vx, vy, omega, X, Y, Z, VX, VY,
                     box);
  return check_launch("ibm_update_markers");
}

int ibm_band_velocity(const clbm_grid_t* g, const void* f, int parity, const int32_t* box,
                      float* ux, float* uy, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!f || !box || !ux || !uy) return fail(-1, "ibm_band_velocity: NULL buffer");
  if (parity != 0 && parity != 1) return fail(-1, "ibm_band_velocity: parity must be 0 or 1");
  const Geo geo = make_geo(*g);
  hipStream_t st = (hipStream_t)stream;
  switch (g->storage) {
    case CLBM_STORE_F32: launch_band<CLBM_STORE_F32>(geo, f, parity, box, ux, uy, st); break;
    case CLBM_STORE_F16: launch_band<CLBM_STORE_F16>(geo, f, parity, box, ux, uy, st); break;
    case CLBM_STORE_BF16: launch_band<CLBM_STORE_BF16>(geo, f, parity, box, ux, uy, st); break;
    case CLBM_STORE_F64: launch_band<CLBM_STORE_F64>(geo, f, parity, box, ux, uy, st); break;
    default: return fail(-1, "unknown storage format");
  }
  return check_launch("ibm_band_velocity");
}

static int interp_args(const clbm_grid_t* g, const ibm_params_t* ip, const void* f, int parity,
                       const float* ux, const float* uy, int n, const float* X, const float* Y,
                       const float* Z, const float* VX, const float* VY) {
  if (int e = check_grid(g, false)) return e;
  if (int e = check_ibm(ip)) return e;
  if (n < 0) return fail(-1, "n < 0");
  if (n > 0 && (!X || !Y || !Z || !VX || !VY)) return fail(-1, "NULL marker array");
  if (!(ux && uy)) {
    if (!f) return fail(-1, "need ux and uy, or the population state f");
    if (parity != 0 && parity != 1) return fail(-1, "parity must be 0 or 1");
  }
  return 0;
}

int ibm_interp_force(const clbm_grid_t* g, const ibm_params_t* ip, const void* f, int parity,
                     const float* ux, const float* uy, int n, const float* X, const float* Y,
                     const float* Z, const float* VX, const float* VY, float* U, float* V,
                     float* FX, float* FY, int32_t* stencil_base, void* stream) {
  if (int e = interp_args(g, ip, f, parity, ux, uy, n, X, Y, Z, VX, VY)) return e;
  if (n == 0) return 0;
  return marker_dispatch<true, false, IBM_SPREAD_FIXED>(g, ip, f, parity, ux, uy, n, X, Y, Z, VX,
                                                         VY, U, V, FX, FY, nullptr, nullptr,
                                                         stencil_base, nullptr, nullptr,
                                                         (hipStream_t)stream);
}

int ibm_spread(const clbm_grid_t* g, const ibm_params_t* ip, int n, const float* X, const float* Y,
               const float* Z, const float* FX, const float* FY, void* acc_x, void* acc_y,
               void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (int e = check_ibm(ip)) return e;
  if (n < 0) return fail(-1, "ibm_spread: n < 0");
  if (n == 0) return 0;
  if (!X || !Y || !Z || !FX || !FY || !acc_x || !acc_y)
    return fail(-1, "ibm_spread: NULL buffer");
  if (ip->spread_mode == IBM_SPREAD_FIXED)
    return marker_dispatch<false, true, IBM_SPREAD_FIXED>(
        g, ip, nullptr, 0, nullptr, nullptr, n, X, Y, Z, nullptr, nullptr, nullptr, nullptr,
        nullptr, nullptr, FX, FY, nullptr, acc_x, acc_y, (hipStream_t)stream);
  return marker_dispatch<false, true, IBM_SPREAD_ATOMIC_F32>(
      g, ip, nullptr, 0, nullptr, nullptr, n, X, Y, Z, nullptr, nullptr, nullptr, nullptr, nullptr,
      nullptr, FX, FY, nullptr, acc_x, acc_y, (hipStream_t)stream);
}

int ibm_interp_spread(const clbm_grid_t* g, const ibm_params_t* ip, const void* f, int parity,
                      const float* ux, const float* uy, int n, const float* X, const float* Y,
                      const float* Z, const float* VX, const float* VY, float* U, float* V,
                      float* FX, float* FY, int32_t* stencil_base, void* acc_x, void* acc_y,
                      void* stream) {
  if (int e = interp_args(g, ip, f, parity, ux, uy, n, X, Y, Z, VX, VY)) return e;
  if (n == 0) return 0;
  if (!acc_x || !acc_y) return fail(-1, "ibm_interp_spread: NULL accumulator");
  if (ip->spread_mode == IBM_SPREAD_FIXED)
    return marker_dispatch<true, true, IBM_SPREAD_FIXED>(g, ip, f, parity, ux, uy, n, X, Y, Z, VX,
                                                          VY, U, V, FX, FY, nullptr, nullptr,
                                                          stencil_base, acc_x, acc_y,
                                                          (hipStream_t)stream);
  return marker_dispatch<true, true, IBM_SPREAD_ATOMIC_F32>(g, ip, f, parity, ux, uy, n, X, Y, Z,
                                                             VX, VY, U, V, FX, FY, nullptr, nullptr,
                                                             stencil_base, acc_x, acc_y,
                                                             (hipStream_t)stream);
}

int ibm_clip_forces(const clbm_grid_t* g, const ibm_params_t* ip, void* acc_x, void* acc_y,
                    float* fx, float* fy, const int32_t* box, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (int e = check_ibm(ip)) return e;
  if (!acc_x || !acc_y || !box) return fail(-1, "ibm_clip_forces: NULL buffer");
  const Geo geo = make_geo(*g);
  const float inv = (float)std::ldexp(1.0, -fixed_shift(ip));
  if (ip->spread_mode == IBM_SPREAD_FIXED)
    hipLaunchKernelGGL(k_clip<IBM_SPREAD_FIXED>, dim3(box_blocks(geo)), dim3(IBLK), 0,
                       (hipStream_t)stream, geo, ip->f_max, inv, acc_x, acc_y, fx, fy, box);
  else
    hipLaunchKernelGGL(k_clip<IBM_SPREAD_ATOMIC_F32>, dim3(box_blocks(geo)), dim3(IBLK), 0,
                       (hipStream_t)stream, geo, ip->f_max, inv, acc_x, acc_y, fx, fy, box);
  return check_launch("ibm_clip_forces");
}

int ibm_force_sum_f64(int n, const float* FX, const float* FY, const float* dV, double* out2,
                      void* stream) {
  if (n < 0) return fail(-1, "ibm_force_sum_f64: n < 0");
  if (!out2 || (n > 0 && (!FX || !FY))) return fail(-1, "ibm_force_sum_f64: NULL buffer");
  hipLaunchKernelGGL(k_force_sum, dim3(1), dim3(SBLK), 0, (hipStream_t)stream, n, FX, FY, dV, out2);
  return check_launch("ibm_force_sum_f64");
}

}  // extern "C"
