// Immersed-boundary kernels of the CLBM-IBM step for gfx950 (MI355X).
//
// DroneMovingIBM.compute_force (airfoil_moving.py:230-255) runs six kernels,
// three of them over the full grid.  Here the whole of it is two launches:
//   * ibm_compute_force  K8 + K10 + K11 + K12.  One 64-lane wavefront per
//                        marker, one lane per point of its 4x4x4 Peskin
//                        stencil (x fastest across lanes).  The wave places
//                        its marker, each lane reads the force-free velocity
//                        of its cell straight from the populations, the wave
//                        reduces U_k, V_k and the clamped marker force, and
//                        each lane spreads F_k w with the same weight.  A
//                        workgroup holds G = 4, 8 or 16 markers that are
//                        neighbours in the body (consecutive in the
//                        template, which runs along each section, or in a
//                        caller-supplied spatial order), pre-sums their
//                        overlapping stencils in LDS and sends one global
//                        atomic per touched cell.
//   * ibm_clip_forces    K9 + K13 on the bounding box of the touched cells.
// Spreading is order-independent by default: each product F_k w becomes a
// 64-bit fixed-point integer and all sums (LDS and global) are integer sums,
// so the grid force is the same bit pattern on every run.  The reference's
// float atomics (airfoil_moving.py:213-214) are kept as IBM_SPREAD_ATOMIC_F32,
// and IBM_SPREAD_SORTED spreads by grouping the stencil points by cell and
// summing each group once (ibm_spread_sorted, the same fixed-point sums).
#include "clbm_markers.h"
#include "clbm_step.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

namespace clbm {

// K8 on its own (airfoil_moving.py:100-127), with the box of the new stencils.
__global__ __launch_bounds__(IBLK) void k_update_markers(
    Geo g, int n, const float* __restrict__ X0, const float* __restrict__ Y0,
    const float* __restrict__ Z0, Kin kin, float* __restrict__ X, float* __restrict__ Y,
    float* __restrict__ Z, float* __restrict__ VX, float* __restrict__ VY, uint64_t* box,
    uint32_t gen) {
  __shared__ int sb[IBLK / 64][6];
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  const int m = blockIdx.x * IBLK + threadIdx.x;
  if (m < n) {
    float xk, yk, vxk, vyk;
    place_marker(kin, X0[m], Y0[m], xk, yk, vxk, vyk);
    const float zk = Z0[m];
    X[m] = xk;
    Y[m] = yk;
    Z[m] = zk;
    VX[m] = vxk;
    VY[m] = vyk;
    stencil_ranges(g, (int)floorf(xk) - 1, (int)floorf(yk) - 1, (int)floorf(zk) - 1, lo, hi);
  }
  if (!box) return;
  for (int a = 0; a < 3; ++a)
    for (int o = 32; o > 0; o >>= 1) {
      lo[a] = min(lo[a], __shfl_xor(lo[a], o));
      hi[a] = max(hi[a], __shfl_xor(hi[a], o));
    }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      sb[wv][a] = lo[a];
      sb[wv][3 + a] = hi[a];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < IBLK / 64; ++i)
      for (int a = 0; a < 3; ++a) {
        sb[0][a] = min(sb[0][a], sb[i][a]);
        sb[0][3 + a] = max(sb[0][3 + a], sb[i][3 + a]);
      }
    box_grow(box, gen, &sb[0][0], &sb[0][3], blockIdx.x);
  }
}

template <int S, int UMODE, int SMODE, int G>
__global__ __launch_bounds__(64 * G) void k_compute_force(IbmArgs a, Kin kin, uint32_t gen) {
  // consecutive groups (neighbours along the body, then the next layers) go
  // to the same XCD, so their overlapping tiles meet in that XCD's L2
  compute_group<S, UMODE, SMODE, G>(a, kin, gen, xcd_block(blockIdx.x, gridDim.x), blockIdx.x);
}

// Many bodies on many lattices in one launch (block -> instance as in
// k_step_batch); kinematics and generation come from the step counter.
template <int S, int UMODE, int SMODE, int G>
__global__ __launch_bounds__(64 * G) void k_compute_force_batch(const IbmArgs* __restrict__ args,
                                                                const uint16_t* __restrict__ inst,
                                                                const int32_t* __restrict__ first,
                                                                const int32_t* __restrict__ step,
                                                                int planes_on) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int i = inst[lb];
  IbmArgs a = args[i];
  a.pl.on &= planes_on;
  const int st = *step;
  const float* k = a.kin + 7 * (size_t)st;
  const Kin kin{k[0], k[1], k[2], k[3], k[4], k[5], k[6]};
  compute_group<S, UMODE, SMODE, G>(a, kin, a.gen0 + (uint32_t)st + 1u, lb - first[i], blockIdx.x);
}

// ---------------------------------------------------------------------------
// The separate K10/K11 and K12 kernels (one wavefront per marker, no LDS).
// ---------------------------------------------------------------------------
template <int S, int UMODE, bool INTERP, bool SPREAD, int SMODE>
__global__ __launch_bounds__(IBLK) void k_marker(
    Geo g, Planes pl, int quirk, float f_max, float fscale,
    const typename Store<S>::T* __restrict__ f, const float* __restrict__ uxf,
    const float* __restrict__ uyf, int n, const float* __restrict__ X, const float* __restrict__ Y,
    const float* __restrict__ Z, const float* __restrict__ VX, const float* __restrict__ VY,
    float* __restrict__ U, float* __restrict__ V, float* __restrict__ FXo, float* __restrict__ FYo,
    const float* __restrict__ FXi, const float* __restrict__ FYi, int32_t* __restrict__ base,
    void* __restrict__ accx, void* __restrict__ accy) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * MPB + (threadIdx.x >> 6);
  if (m >= n) return;  // uniform per wavefront
  const Lane L = lane_stencil(g, quirk, X[m], Y[m], Z[m], lane);
  float fxk, fyk;
  if (INTERP) {
    float uc, vc;
    lane_velocity<S, UMODE>(g, pl, f, uxf, uyf, L, uc, vc);
    const float su = wave_sum(uc * L.w);
    const float sv = wave_sum(vc * L.w);
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
  if (SPREAD) acc_add<SMODE>(accx, accy, L.cell, fxk * L.w, fyk * L.w, fscale);
}

// ---------------------------------------------------------------------------
// Force-free velocity on the cells of a generation's box
// ---------------------------------------------------------------------------
template <int S, int PAR>
__global__ __launch_bounds__(IBLK) void k_band_velocity(Geo g, Planes pl,
                                                        const typename Store<S>::T* __restrict__ f,
                                                        const uint64_t* __restrict__ box,
                                                        uint32_t gen, float* __restrict__ uo,
                                                        float* __restrict__ vo) {
  int lo[3], hi[3];
  if (!box_get_block(box, gen, lo, hi)) return;
  const int bx = hi[0] - lo[0] + 1, by = hi[1] - lo[1] + 1, bz = hi[2] - lo[2] + 1;
  const int vol = bx * by * bz;
  for (int t = blockIdx.x * IBLK + threadIdx.x; t < vol; t += gridDim.x * IBLK) {
    const int r = t / bx;
    const int x = lo[0] + (t - r * bx);
    const int k = r / by;
    const int y = lo[1] + (r - k * by);
    const int z = lo[2] + k;
    const int c = (z * g.ny + y) * g.nx + x;
    float u, v;
    cell_velocity<S, PAR>(g, pl, f, c, x, y, z, u, v);
    uo[c] = u;
    vo[c] = v;
  }
}

// ---------------------------------------------------------------------------
// K9 + K13: accumulators -> clamped float fields on the union of the boxes
// of generations gen and gen - 1; block 0 publishes gen's box for clbm_step.
// ---------------------------------------------------------------------------
struct ClipArgs {
  Geo g;
  float f_max, inv_scale;
  void *accx, *accy;
  float *fx, *fy;
  uint64_t* box;
  uint32_t gen0;  // batched runs, as IbmArgs::gen0
  const int32_t* regions;  // batched runs, as IbmArgs::regions
};

// Block `bid` of `nblk` working on one body.
template <int SMODE>
__device__ __forceinline__ void clip_body(const ClipArgs& a, uint32_t gen, int bid, int nblk) {
  const Geo& g = a.g;
  uint64_t* box = a.box;
  int lo[3], hi[3], plo[3], phi[3];
  const bool cur = box_get_block(box, gen, lo, hi);
  const bool prev = box_get_block(box, gen - 1, plo, phi);
  if (bid == 0 && threadIdx.x == 0) {
    int32_t* out = IBM_FORCE_BOX(box);
    for (int i = 0; i < 3; ++i) {
      out[i] = cur ? lo[i] : INT_MAX;
      out[3 + i] = cur ? hi[i] : INT_MIN;
    }
  }
  if (!cur && !prev) return;
  for (int i = 0; i < 3; ++i) {
    if (!cur) {
      lo[i] = plo[i];
      hi[i] = phi[i];
    } else if (prev) {
      lo[i] = min(lo[i], plo[i]);
      hi[i] = max(hi[i], phi[i]);
    }
  }
  const int bx = hi[0] - lo[0] + 1, by = hi[1] - lo[1] + 1, bz = hi[2] - lo[2] + 1;
  const int vol = bx * by * bz;
  float* fx = a.fx;
  float* fy = a.fy;
  const float f_max = a.f_max;
  for (int t = bid * IBLK + threadIdx.x; t < vol; t += nblk * IBLK) {
    const int r = t / bx;
    const int i = t - r * bx;
    const int k = r / by;
    const int j = r - k * by;
    const int c = ((lo[2] + k) * g.ny + (lo[1] + j)) * g.nx + (lo[0] + i);
    float vx, vy;
    if (SMODE == IBM_SPREAD_FIXED) {
      long long* ax = (long long*)a.accx;
      long long* ay = (long long*)a.accy;
      vx = (float)((double)ax[c] * (double)a.inv_scale);
      vy = (float)((double)ay[c] * (double)a.inv_scale);
      ax[c] = 0;
      ay[c] = 0;
    } else {
      float* ax = (float*)((unsigned long long*)a.accx + c);
      float* ay = (float*)((unsigned long long*)a.accy + c);
      vx = *ax;
      vy = *ay;
      *ax = 0.0f;
      *ay = 0.0f;
    }
    if (fx) fx[c] = fmaxf(-f_max, fminf(f_max, vx));
    if (fy) fy[c] = fmaxf(-f_max, fminf(f_max, vy));
  }
}

template <int SMODE>
__global__ __launch_bounds__(IBLK) void k_clip(ClipArgs a, uint32_t gen) {
  clip_body<SMODE>(a, gen, blockIdx.x, gridDim.x);
}

// Publishes generation gen for a lattice step that waits in its own launch
// for the clamped forces (clbm_step_fused): queued after the clamp on its
// stream, it runs once the clamp's kernel has ended and its writes (fx, fy,
// the decoded box) are visible to the whole device.
__global__ void k_publish(unsigned long long* ready, unsigned long long gen) {
  if (threadIdx.x == 0) __hip_atomic_store(ready, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// bpi blocks per body, body i = blockIdx.x / bpi.  Every block reads the
// batch's step counter, then counts itself in *done; the last block to
// finish (all others have read the counter by then) clears *done and
// advances the counter for the next step.  No block waits for another.
template <int SMODE>
__global__ __launch_bounds__(IBLK) void k_clip_batch(const ClipArgs* __restrict__ args, int bpi,
                                                     int32_t* step, int32_t* done) {
  const int i = blockIdx.x / bpi;
  const ClipArgs a = args[i];
  clip_body<SMODE>(a, a.gen0 + (uint32_t)*step + 1u, blockIdx.x - i * bpi, bpi);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(done, 1) == (int)gridDim.x - 1) {
      atomicExch(done, 0);
      atomicAdd(step, 1);
    }
  }
}

// ---------------------------------------------------------------------------
// A batch step without the clamp launch (ibm_batch_markers +
// clbm_batch_step_acc).  Every body carries a table of its predicted region
// per step (from the kinematics, as the split step predicts it); the lattice
// step takes the force inside that region straight from the accumulators, so
// no launch has to publish the box of the cells the stencils reached.
//
// The step counter is a pair indexed by the step's AA parity: a step of
// parity p reads cnt[p] and its marker kernel's block 0 writes cnt[1 - p] =
// cnt[p] + 1 for the next step (which has the other parity), so no workgroup
// waits for another or counts the others.
// ---------------------------------------------------------------------------
template <int S, int PAR, int G>
__global__ __launch_bounds__(64 * G) void k_compute_force_batch_reg(
    const IbmArgs* __restrict__ args, const uint16_t* __restrict__ inst,
    const int32_t* __restrict__ first, int32_t* cnt, int planes_on) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int i = inst[lb];
  IbmArgs a = args[i];
  a.pl.on &= planes_on;
  const int st = cnt[PAR];
  const float* k = a.kin + 7 * (size_t)st;
  const Kin kin{k[0], k[1], k[2], k[3], k[4], k[5], k[6]};
  // the stencils are checked against the region the lattice step will read
  const int32_t* r = a.regions + 6 * (size_t)st;
  a.rc.on = 1;
  for (int d = 0; d < 3; ++d) {
    a.rc.lo[d] = r[d];
    a.rc.hi[d] = r[3 + d];
  }
  compute_group<S, 1 + PAR, IBM_SPREAD_FIXED, G>(a, kin, a.gen0 + (uint32_t)st + 1u,
                                                 lb - first[i], blockIdx.x);
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt[1 - PAR] = st + 1;
}

// The lattice step of such a batch: as k_step_batch, with each lattice's
// force decoded and clamped straight from its body's accumulators inside the
// body's predicted region of this step (AccForce: the values ibm_clip_forces
// would write; zero where no stencil reached), which it clears for the next
// step.
template <int S, int PAR, int BK>
__global__ __launch_bounds__(BK) void k_step_batch_acc(const StepArgs* __restrict__ sargs,
                                                       const ClipArgs* __restrict__ cargs,
                                                       const uint16_t* __restrict__ inst,
                                                       const int32_t* __restrict__ first,
                                                       const int32_t* __restrict__ cnt,
                                                       int planes_on) {
  const int lb = sweep_block<PAR>(blockIdx.x, gridDim.x);
  const int i = inst[lb];
  StepArgs a = sargs[i];
  a.pl.on &= planes_on;
  const ClipArgs c = cargs[i];
  const int32_t* r = c.regions + 6 * (size_t)cnt[PAR];
  const AccForce<false> force{(long long*)c.accx, (long long*)c.accy, {r[0], r[1], r[2]},
                              {r[3], r[4], r[5]}, c.f_max, (double)c.inv_scale};
  step_cell<S, PAR, 0>(a, (lb - first[i]) * BK + threadIdx.x, force);
}

// ---------------------------------------------------------------------------
// FP64 marker force sums, one block, fixed reduction tree (deterministic).
// ---------------------------------------------------------------------------
constexpr int SBLK = 1024;
__global__ __launch_bounds__(SBLK) void k_force_sum(int n, const float* __restrict__ FX,
                                                    const float* __restrict__ FY,
                                                    const float* __restrict__ dV,
                                                    double* __restrict__ out2) {
  __shared__ double sx[SBLK / 64], sy[SBLK / 64];
  double ax = 0.0, ay = 0.0;
  for (int m = threadIdx.x; m < n; m += SBLK) {
    const double d = dV ? (double)dV[m] : 1.0;
    ax += (double)FX[m] * d;
    ay += (double)FY[m] * d;
  }
  for (int o = 32; o > 0; o >>= 1) {
    ax += __shfl_xor(ax, o);
    ay += __shfl_xor(ay, o);
  }
  if ((threadIdx.x & 63) == 0) {
    sx[threadIdx.x >> 6] = ax;
    sy[threadIdx.x >> 6] = ay;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tx = 0.0, ty = 0.0;
    for (int i = 0; i < SBLK / 64; ++i) {
      tx += sx[i];
      ty += sy[i];
    }
    out2[0] = tx;
    out2[1] = ty;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Blocks of the grid-stride box kernels.
static inline int box_blocks(const Geo& g) {
  return std::max(1, std::min(1024, (g.ncell + IBLK - 1) / IBLK));
}

template <int S, int UMODE, int SMODE>
static void launch_compute(const Geo& g, const Planes& pl, const ibm_params_t* ip, const void* f,
                           const float* ux, const float* uy, int n, const int32_t* order,
                           const float* X0, const float* Y0, const float* Z0, const Kin& kin,
                           float* X, float* Y, float* Z, float* VX, float* VY, float* U, float* V,
                           float* FX, float* FY, int32_t* base, void* ax, void* ay, uint64_t* box,
                           uint32_t gen, hipStream_t st) {
  IbmArgs a{};
  a.g = g;
  a.pl = pl;
  a.quirk = ip->wrap_weight_quirk ? 1 : 0;
  a.f_max = ip->f_max;
  a.fscale = (float)std::ldexp(1.0, fixed_shift(ip));
  a.f = f;
  a.uxf = ux;
  a.uyf = uy;
  a.n = n;
  a.order = order;
  a.X0 = X0;
  a.Y0 = Y0;
  a.Z0 = Z0;
  a.X = X;
  a.Y = Y;
  a.Z = Z;
  a.VX = VX;
  a.VY = VY;
  a.U = U;
  a.V = V;
  a.FX = FX;
  a.FY = FY;
  a.base = base;
  a.accx = ax;
  a.accy = ay;
  a.box = box;
  a.rc = region_check(ip);
#define CLBM_CG(GG)                                                                    \
  hipLaunchKernelGGL((k_compute_force<S, UMODE, SMODE, GG>), dim3((n + GG - 1) / GG), \
                     dim3(64 * GG), 0, st, a, kin, gen)
  switch (ip->group) {
    case 4: CLBM_CG(4); break;
    case 16: CLBM_CG(16); break;
    default: CLBM_CG(8); break;
  }
#undef CLBM_CG
}

template <int S, int SMODE>
static void launch_compute_u(int umode, const Geo& g, const Planes& pl, const ibm_params_t* ip,
                             const void* f, const float* ux, const float* uy, int n,
                             const int32_t* order, const float* X0, const float* Y0,
                             const float* Z0, const Kin& kin, float* X, float* Y, float* Z,
                             float* VX, float* VY, float* U, float* V, float* FX, float* FY,
                             int32_t* base, void* ax, void* ay, uint64_t* box, uint32_t gen,
                             hipStream_t st) {
#define CLBM_CF(UM)                                                                               \
  launch_compute<S, UM, SMODE>(g, pl, ip, f, ux, uy, n, order, X0, Y0, Z0, kin, X, Y, Z, VX, VY, U, \
                               V, FX, FY, base, ax, ay, box, gen, st)
  if (umode == 0)
    CLBM_CF(0);
  else if (umode == 1)
    CLBM_CF(1);
  else
    CLBM_CF(2);
#undef CLBM_CF
}

template <int S, bool INTERP, bool SPREAD, int SMODE>
static void launch_marker(const Geo& g, const ibm_params_t* ip, const Planes& pl, const void* f,
                          int umode, const float* ux, const float* uy, int n, const float* X,
                          const float* Y, const float* Z, const float* VX, const float* VY,
                          float* U, float* V, float* FXo, float* FYo, const float* FXi,
                          const float* FYi, int32_t* base, void* accx, void* accy,
                          hipStream_t st) {
  const float scale = (float)std::ldexp(1.0, fixed_shift(ip));
  auto fp = (const typename Store<S>::T*)f;
  const dim3 gr((n + MPB - 1) / MPB), bl(IBLK);
  const int uq = ip->wrap_weight_quirk ? 1 : 0;
#define CLBM_MK(UM)                                                                             \
  hipLaunchKernelGGL((k_marker<S, UM, INTERP, SPREAD, SMODE>), gr, bl, 0, st, g, pl, uq,        \
                     ip->f_max, scale, fp, ux, uy, n, X, Y, Z, VX, VY, U, V, FXo, FYo, FXi, FYi, \
                     base, accx, accy)
  if (umode == 0)
    CLBM_MK(0);
  else if (umode == 1)
    CLBM_MK(1);
  else
    CLBM_MK(2);
#undef CLBM_MK
}

// ---------------------------------------------------------------------------
// batches of bodies
// ---------------------------------------------------------------------------
constexpr int CLIP_BPI = 32;  // clamp blocks per body in a batch

static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// int32 offset of the clamp kernel's done count from the step counter (its
// own 128-byte line of the counter's 256 bytes)
constexpr int BATCH_DONE_OFFSET = 32;

struct IbmBatchLayout {
  size_t args, clip, inst, first, counter, total;
  int blocks;
};
static int ibm_batch_layout(int n, const ibm_buffers_t* bufs, int group, IbmBatchLayout& L) {
  if (n <= 0 || n > 65535) return fail(-1, "ibm batch: n must be in [1, 65535]");
  if (!bufs) return fail(-1, "ibm batch: bufs is NULL");
  if (group != 4 && group != 8) return fail(-1, "ibm batch: group must be 4 or 8");
  long long blocks = 0;
  for (int i = 0; i < n; ++i) {
    if (bufs[i].n <= 0) return fail(-1, "ibm batch: every body needs markers");
    blocks += (bufs[i].n + group - 1) / group;
  }
  if (blocks >= (1ll << 31)) return fail(-1, "ibm batch: too many blocks");
  L.blocks = (int)blocks;
  L.args = 0;
  L.clip = align256(sizeof(IbmArgs) * n);
  L.inst = L.clip + align256(sizeof(ClipArgs) * n);
  L.first = L.inst + align256(sizeof(uint16_t) * (size_t)blocks);
  L.counter = L.first + align256(sizeof(int32_t) * n);   // step counter, then the done count
  L.total = L.counter + 256;
  return 0;
}

template <int S, int SMODE, int G>
static void launch_compute_batch(const ibm_batch_t* b, int par, int planes, hipStream_t st) {
  const IbmArgs* a = (const IbmArgs*)b->args;
  if (par)
    hipLaunchKernelGGL((k_compute_force_batch<S, 2, SMODE, G>), dim3(b->blocks), dim3(64 * G), 0,
                       st, a, b->inst, b->first, b->counter, planes);
  else
    hipLaunchKernelGGL((k_compute_force_batch<S, 1, SMODE, G>), dim3(b->blocks), dim3(64 * G), 0,
                       st, a, b->inst, b->first, b->counter, planes);
}

template <int S, int SMODE>
static void launch_batch(const ibm_batch_t* b, int par, int planes, hipStream_t st) {
  if (b->group == 4)
    launch_compute_batch<S, SMODE, 4>(b, par, planes, st);
  else
    launch_compute_batch<S, SMODE, 8>(b, par, planes, st);
  hipLaunchKernelGGL(k_clip_batch<SMODE>, dim3(b->clip_blocks * b->n), dim3(IBLK), 0, st,
                     (const ClipArgs*)b->clip_args, b->clip_blocks, b->counter,
                     b->counter + BATCH_DONE_OFFSET);
}

template <int S, int G>
static void launch_markers_reg(const ibm_batch_t* b, int par, int planes, hipStream_t st) {
  const IbmArgs* a = (const IbmArgs*)b->args;
  if (par)
    hipLaunchKernelGGL((k_compute_force_batch_reg<S, 1, G>), dim3(b->blocks), dim3(64 * G), 0, st,
                       a, b->inst, b->first, b->counter, planes);
  else
    hipLaunchKernelGGL((k_compute_force_batch_reg<S, 0, G>), dim3(b->blocks), dim3(64 * G), 0, st,
                       a, b->inst, b->first, b->counter, planes);
}

template <int S>
static void launch_markers_reg_g(const ibm_batch_t* b, int par, int planes, hipStream_t st) {
  if (b->group == 4)
    launch_markers_reg<S, 4>(b, par, planes, st);
  else
    launch_markers_reg<S, 8>(b, par, planes, st);
}

template <int S, int PAR, int BK>
static void launch_step_acc(const clbm_batch_t* lb, const ibm_batch_t* ib, int planes,
                            hipStream_t st) {
  hipLaunchKernelGGL((k_step_batch_acc<S, PAR, BK>), dim3(lb->blocks), dim3(BK),
                     (unsigned)lb->lds, st, (const StepArgs*)lb->args,
                     (const ClipArgs*)ib->clip_args, lb->inst, lb->first, ib->counter, planes);
}

template <int S>
static void launch_step_acc_s(const clbm_batch_t* lb, const ibm_batch_t* ib, int par, int planes,
                              hipStream_t st) {
  if (lb->block == 128) {
    if (par) launch_step_acc<S, 1, 128>(lb, ib, planes, st);
    else launch_step_acc<S, 0, 128>(lb, ib, planes, st);
  } else {
    if (par) launch_step_acc<S, 1, 256>(lb, ib, planes, st);
    else launch_step_acc<S, 0, 256>(lb, ib, planes, st);
  }
}

#define CLBM_STORAGE_SWITCH(storage, CALL)                   \
  switch (storage) {                                         \
    case CLBM_STORE_F32: { constexpr int SS = CLBM_STORE_F32; CALL; } break;   \
    case CLBM_STORE_F16: { constexpr int SS = CLBM_STORE_F16; CALL; } break;   \
    case CLBM_STORE_BF16: { constexpr int SS = CLBM_STORE_BF16; CALL; } break; \
    case CLBM_STORE_F64: { constexpr int SS = CLBM_STORE_F64; CALL; } break;   \
    case CLBM_STORE_F16C: { constexpr int SS = CLBM_STORE_F16C; CALL; } break; \
    default: return fail(-1, "unknown storage format");      \
  }

// Velocity source: 0 = ux/uy fields, 1/2 = populations at parity 0/1.
// ---------------------------------------------------------------------------
// K12 by sorting (IBM_SPREAD_SORTED).  The 64 n (cell, F_k w) pairs of the
// markers' stencils are grouped by cell (a counting sort keyed by the cell
// index), and one thread per cell adds up its group: a segmented reduction
// instead of atomics on the accumulators.  The sums are the 64-bit
// fixed-point ones of IBM_SPREAD_FIXED (each F_k w rounded to 2^-shift, then
// added exactly), so they do not depend on the order of a group and equal
// the atomic fixed-point spread bit for bit.  Four passes:
//  * per pair: its cell, its contribution in fixed point and its rank in the
//    cell's group (the count the pair's increment finds); the pair of rank 0
//    appends the cell to the list of groups;
//  * per group: its offset in the pair list (a prefix sum of the counts
//    across the lanes of a wavefront, one atomic per wavefront on the
//    cursor); the cell's count goes back to zero for the next call;
//  * per pair: its contribution moved to its slot, the group's offset plus
//    its rank (no atomics);
//  * per group: the sum of its now contiguous contributions into the cell's
//    accumulators.
// The wavefronts append and reserve with one aggregated atomic each.
// ---------------------------------------------------------------------------
struct SortWork {
  int32_t *count, *off;               // per cell: group size, group offset
  int32_t *key, *rank;                // per pair: its cell, its rank in the cell's group
  long long *vx, *vy;                 // per pair: F_k w in fixed point, in pair order
  long long *gx, *gy;                 // the same, grouped by cell
  int32_t *ukey, *uoff, *ulen;        // per group: cell, offset, length
  int32_t* ctr;                       // [0] groups, [1] pairs reserved
};

static size_t sort_layout(long long ncell, long long pairs, unsigned char* w, SortWork* out) {
  size_t o = 0;
  auto take = [&](long long n) {
    const size_t at = o;
    o = align256(o + (size_t)n * 4);
    return w ? (int32_t*)(w + at) : nullptr;
  };
  SortWork W;
  W.count = take(ncell);
  W.off = take(ncell);
  W.key = take(pairs);
  W.rank = take(pairs);
  W.vx = (long long*)take(2 * pairs);
  W.vy = (long long*)take(2 * pairs);
  W.gx = (long long*)take(2 * pairs);
  W.gy = (long long*)take(2 * pairs);
  W.ukey = take(pairs);
  W.uoff = take(pairs);
  W.ulen = take(pairs);
  W.ctr = take(2);
  if (out) *out = W;
  return o;
}

// Exclusive prefix sum of v over the lanes of the wavefront, and its total.
__device__ __forceinline__ int wave_excl_scan(int v, int lane, int& total) {
  int incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(incl, d);
    if (lane >= d) incl += t;
  }
  total = __shfl(incl, 63);
  return incl - v;
}

__global__ __launch_bounds__(64) void k_sort_clear(int32_t* ctr) {
  if (threadIdx.x < 2) ctr[threadIdx.x] = 0;
}

__global__ __launch_bounds__(IBLK) void k_sort_pairs(Geo g, int quirk, int n,
                                                     const float* __restrict__ X,
                                                     const float* __restrict__ Y,
                                                     const float* __restrict__ Z,
                                                     const float* __restrict__ FX,
                                                     const float* __restrict__ FY, float fscale,
                                                     SortWork W) {
  const long long i = (long long)blockIdx.x * IBLK + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool first = false;
  int c = 0;
  if (i < 64ll * n) {  // every lane takes part in the wavefront's append below
    const int m = (int)(i >> 6), st = (int)(i & 63);
    const Lane L = lane_stencil(g, quirk, X[m], Y[m], Z[m], st);
    c = L.cell;
    W.key[i] = c;
    // the rounding of the fixed-point spread: (F_k w) * 2^shift to nearest
    W.vx[i] = __float2ll_rn((FX[m] * L.w) * fscale);
    W.vy[i] = __float2ll_rn((FY[m] * L.w) * fscale);
    const int r = atomicAdd(W.count + c, 1);
    W.rank[i] = r;
    first = r == 0;
  }
  const unsigned long long who = __ballot(first);
  int base = 0;
  if (lane == 0 && who) base = atomicAdd(W.ctr, __popcll(who));
  base = __shfl(base, 0);
  if (first) W.ukey[base + __popcll(who & ((1ull << lane) - 1ull))] = c;
}

__global__ __launch_bounds__(IBLK) void k_sort_claim(long long pairs, SortWork W) {
  const long long sgm = (long long)blockIdx.x * IBLK + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const long long ngroups = W.ctr[0];
  if (sgm - lane >= ngroups) return;  // the whole wavefront lies past the groups
  int c = 0, k = 0;
  if (sgm < ngroups) {
    c = W.ukey[sgm];
    k = W.count[c];
    W.count[c] = 0;  // zero again for the next call
  }
  int total;
  const int excl = wave_excl_scan(k, lane, total);
  int base = 0;
  if (lane == 0 && total > 0) base = atomicAdd(W.ctr + 1, total);
  base = __shfl(base, 0);
  if (sgm < ngroups) {
    W.off[c] = base + excl;
    W.uoff[sgm] = base + excl;
    W.ulen[sgm] = k;
  }
}

__global__ __launch_bounds__(IBLK) void k_sort_place(long long pairs, SortWork W) {
  const long long i = (long long)blockIdx.x * IBLK + threadIdx.x;
  if (i >= pairs) return;
  const int slot = W.off[W.key[i]] + W.rank[i];
  W.gx[slot] = W.vx[i];
  W.gy[slot] = W.vy[i];
}

__global__ __launch_bounds__(IBLK) void k_sort_sum(long long pairs, SortWork W, void* accx,
                                                   void* accy) {
  const long long sgm = (long long)blockIdx.x * IBLK + threadIdx.x;
  if (sgm >= pairs || sgm >= W.ctr[0]) return;
  const int c = W.ukey[sgm], o = W.uoff[sgm], k = W.ulen[sgm];
  unsigned long long sx = 0ull, sy = 0ull;
  for (int a = 0; a < k; ++a) {
    sx += (unsigned long long)W.gx[o + a];
    sy += (unsigned long long)W.gy[o + a];
  }
  // this thread is the cell's only writer
  ((unsigned long long*)accx)[c] += sx;
  ((unsigned long long*)accy)[c] += sy;
}

static int velocity_source(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state,
                           const float* ux, const float* uy, Planes& pl, int& umode) {
  pl.on = 0;
  pl.rule = 0;
  pl.ux_in = pl.uy_in = 0.0;
  if (ux && uy) {
    umode = 0;
    return 0;
  }
  if (!f) return fail(-1, "need ux and uy, or the population state f");
  if (int e = check_state(g, p, state, pl)) return e;
  umode = 1 + (state & CLBM_STATE_PARITY);
  return 0;
}

}  // namespace clbm

using namespace clbm;

extern "C" {

int ibm_compute_force(const clbm_grid_t* g, const ibm_params_t* ip, const clbm_params_t* p,
                      const void* f, int state, const float* ux, const float* uy, int n,
                      const int32_t* order, const float* X0, const float* Y0, const float* Z0,
                      float xc, float yc, float theta, float vx, float vy, float omega, float* X,
                      float* Y, float* Z, float* VX, float* VY, float* U, float* V, float* FX,
                      float* FY, int32_t* stencil_base, void* acc_x, void* acc_y, uint64_t* box,
                      uint32_t gen, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (int e = check_ibm(ip)) return e;
  if (n < 0) return fail(-1, "ibm_compute_force: n < 0");
  if (n == 0) return 0;
  if (!X0 || !Y0 || !Z0 || !X || !Y || !Z || !VX || !VY || !U || !V || !FX || !FY)
    return fail(-1, "ibm_compute_force: NULL marker array");
  if (!acc_x || !acc_y) return fail(-1, "ibm_compute_force: NULL accumulator");
  if (box && gen == 0) return fail(-1, "ibm_compute_force: generation must be >= 1");
  if (ip->check_region && !box) return fail(-1, "ibm_compute_force: check_region needs the box record");
  if (ip->spread_mode == IBM_SPREAD_SORTED)
    return fail(-1, "ibm_compute_force: the sorted spread runs as ibm_update_markers + "
                    "ibm_interp_force + ibm_spread_sorted");
  Planes pl;
  int umode;
  if (int e = velocity_source(g, p, f, state, ux, uy, pl, umode)) return e;
  const Geo geo = make_geo(*g);
  const Kin kin = make_kin(xc, yc, theta, vx, vy, omega);
  hipStream_t st = (hipStream_t)stream;
  if (ip->spread_mode == IBM_SPREAD_FIXED) {
    CLBM_STORAGE_SWITCH(g->storage, (launch_compute_u<SS, IBM_SPREAD_FIXED>(
                                        umode, geo, pl, ip, f, ux, uy, n, order, X0, Y0, Z0, kin, X,
                                        Y, Z, VX, VY, U, V, FX, FY, stencil_base, acc_x, acc_y,
                                        box, gen, st)));
  } else {
    CLBM_STORAGE_SWITCH(g->storage, (launch_compute_u<SS, IBM_SPREAD_ATOMIC_F32>(
                                        umode, geo, pl, ip, f, ux, uy, n, order, X0, Y0, Z0, kin, X,
                                        Y, Z, VX, VY, U, V, FX, FY, stencil_base, acc_x, acc_y,
                                        box, gen, st)));
  }
  return check_launch("ibm_compute_force");
}

int ibm_update_markers(const clbm_grid_t* g, int n, const float* X0, const float* Y0,
                       const float* Z0, float xc, float yc, float theta, float vx, float vy,
                       float omega, float* X, float* Y, float* Z, float* VX, float* VY,
                       uint64_t* box, uint32_t gen, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (n < 0) return fail(-1, "ibm_update_markers: n < 0");
  if (n == 0) return 0;
  if (!X0 || !Y0 || !Z0 || !X || !Y || !Z || !VX || !VY)
    return fail(-1, "ibm_update_markers: NULL marker array");
  if (box && gen == 0) return fail(-1, "ibm_update_markers: generation must be >= 1");
  hipLaunchKernelGGL(k_update_markers, dim3((n + IBLK - 1) / IBLK), dim3(IBLK), 0,
                     (hipStream_t)stream, make_geo(*g), n, X0, Y0, Z0,
                     make_kin(xc, yc, theta, vx, vy, omega), X, Y, Z, VX, VY, box, gen);
  return check_launch("ibm_update_markers");
}

int ibm_band_velocity(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state,
                      const uint64_t* box, uint32_t gen, float* ux, float* uy, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!f || !box || !ux || !uy) return fail(-1, "ibm_band_velocity: NULL buffer");
  Planes pl;
  if (int e = check_state(g, p, state, pl)) return e;
  const int par = state & CLBM_STATE_PARITY;
  const Geo geo = make_geo(*g);
  hipStream_t st = (hipStream_t)stream;
#define CLBM_BAND(PP)                                                                          \
  hipLaunchKernelGGL((k_band_velocity<SS, PP>), dim3(box_blocks(geo)), dim3(IBLK), 0, st, geo, \
                     pl, (const typename Store<SS>::T*)f, box, gen, ux, uy)
  CLBM_STORAGE_SWITCH(g->storage, if (par) { CLBM_BAND(1); } else { CLBM_BAND(0); });
#undef CLBM_BAND
  return check_launch("ibm_band_velocity");
}

int ibm_interp_force(const clbm_grid_t* g, const ibm_params_t* ip, const clbm_params_t* p,
                     const void* f, int state, const float* ux, const float* uy, int n,
                     const float* X, const float* Y, const float* Z, const float* VX,
                     const float* VY, float* U, float* V, float* FX, float* FY,
                     int32_t* stencil_base, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (int e = check_ibm(ip)) return e;
  if (n < 0) return fail(-1, "ibm_interp_force: n < 0");
  if (n == 0) return 0;
  if (!X || !Y || !Z || !VX || !VY) return fail(-1, "ibm_interp_force: NULL marker array");
  Planes pl;
  int umode;
  if (int e = velocity_source(g, p, f, state, ux, uy, pl, umode)) return e;
  const Geo geo = make_geo(*g);
  CLBM_STORAGE_SWITCH(g->storage, (launch_marker<SS, true, false, IBM_SPREAD_FIXED>(
                                      geo, ip, pl, f, umode, ux, uy, n, X, Y, Z, VX, VY, U, V, FX,
                                      FY, nullptr, nullptr, stencil_base, nullptr, nullptr,
                                      (hipStream_t)stream)));
  return check_launch("ibm_interp_force");
}

int ibm_spread(const clbm_grid_t* g, const ibm_params_t* ip, int n, const float* X, const float* Y,
               const float* Z, const float* FX, const float* FY, void* acc_x, void* acc_y,
               void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (int e = check_ibm(ip)) return e;
  if (n < 0) return fail(-1, "ibm_spread: n < 0");
  if (n == 0) return 0;
  if (!X || !Y || !Z || !FX || !FY || !acc_x || !acc_y) return fail(-1, "ibm_spread: NULL buffer");
  if (ip->spread_mode == IBM_SPREAD_SORTED)
    return fail(-1, "ibm_spread: the sorted spread needs a work buffer (ibm_spread_sorted)");
  const Geo geo = make_geo(*g);
  const Planes pl{0, 0, 0.0, 0.0};
  hipStream_t st = (hipStream_t)stream;
  if (ip->spread_mode == IBM_SPREAD_FIXED)
    launch_marker<CLBM_STORE_F32, false, true, IBM_SPREAD_FIXED>(
        geo, ip, pl, nullptr, 0, nullptr, nullptr, n, X, Y, Z, nullptr, nullptr, nullptr, nullptr,
        nullptr, nullptr, FX, FY, nullptr, acc_x, acc_y, st);
  else
    launch_marker<CLBM_STORE_F32, false, true, IBM_SPREAD_ATOMIC_F32>(
        geo, ip, pl, nullptr, 0, nullptr, nullptr, n, X, Y, Z, nullptr, nullptr, nullptr, nullptr,
        nullptr, nullptr, FX, FY, nullptr, acc_x, acc_y, st);
  return check_launch("ibm_spread");
}

size_t ibm_spread_sorted_bytes(const clbm_grid_t* g, int n) {
  if (check_grid(g, false) || n < 0) return 0;
  return sort_layout((long long)g->nx * g->ny * g->nz, 64ll * n, nullptr, nullptr);
}

int ibm_spread_sorted(const clbm_grid_t* g, const ibm_params_t* ip, int n, const float* X,
                      const float* Y, const float* Z, const float* FX, const float* FY,
                      void* acc_x, void* acc_y, void* work, size_t work_bytes, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (int e = check_ibm(ip)) return e;
  if (ip->spread_mode != IBM_SPREAD_SORTED)
    return fail(-1, "ibm_spread_sorted: ip->spread_mode must be IBM_SPREAD_SORTED");
  if (n < 0) return fail(-1, "ibm_spread_sorted: n < 0");
  if (n == 0) return 0;
  if (!X || !Y || !Z || !FX || !FY || !acc_x || !acc_y || !work)
    return fail(-1, "ibm_spread_sorted: NULL buffer");
  const long long ncell = (long long)g->nx * g->ny * g->nz, pairs = 64ll * n;
  if (pairs > INT_MAX) return fail(-1, "ibm_spread_sorted: more than 2^31 stencil points");
  SortWork W;
  if (work_bytes < sort_layout(ncell, pairs, (unsigned char*)work, &W))
    return fail(-1, "ibm_spread_sorted: work buffer smaller than ibm_spread_sorted_bytes");
  const Geo geo = make_geo(*g);
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = (unsigned)((pairs + IBLK - 1) / IBLK);
  hipLaunchKernelGGL(k_sort_clear, dim3(1), dim3(64), 0, st, W.ctr);
  hipLaunchKernelGGL(k_sort_pairs, dim3(nb), dim3(IBLK), 0, st, geo, ip->wrap_weight_quirk, n, X,
                     Y, Z, FX, FY, (float)std::ldexp(1.0, fixed_shift(ip)), W);
  hipLaunchKernelGGL(k_sort_claim, dim3(nb), dim3(IBLK), 0, st, pairs, W);
  hipLaunchKernelGGL(k_sort_place, dim3(nb), dim3(IBLK), 0, st, pairs, W);
  hipLaunchKernelGGL(k_sort_sum, dim3(nb), dim3(IBLK), 0, st, pairs, W, acc_x, acc_y);
  return check_launch("ibm_spread_sorted");
}

int ibm_clip_forces(const clbm_grid_t* g, const ibm_params_t* ip, void* acc_x, void* acc_y,
                    float* fx, float* fy, uint64_t* box, uint32_t gen, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (int e = check_ibm(ip)) return e;
  if (!acc_x || !acc_y || !box) return fail(-1, "ibm_clip_forces: NULL buffer");
  if (gen == 0) return fail(-1, "ibm_clip_forces: generation must be >= 1");
  const Geo geo = make_geo(*g);
  const ClipArgs a{geo, ip->f_max, (float)std::ldexp(1.0, -fixed_shift(ip)), acc_x, acc_y, fx, fy,
                   box, 0u};
  if (ip->spread_mode == IBM_SPREAD_FIXED || ip->spread_mode == IBM_SPREAD_SORTED)
    hipLaunchKernelGGL(k_clip<IBM_SPREAD_FIXED>, dim3(box_blocks(geo)), dim3(IBLK), 0,
                       (hipStream_t)stream, a, gen);
  else
    hipLaunchKernelGGL(k_clip<IBM_SPREAD_ATOMIC_F32>, dim3(box_blocks(geo)), dim3(IBLK), 0,
                       (hipStream_t)stream, a, gen);
  return check_launch("ibm_clip_forces");
}

size_t ibm_batch_bytes(int n, const ibm_buffers_t* bufs, int group) {
  IbmBatchLayout L;
  if (ibm_batch_layout(n, bufs, group, L)) return 0;
  return L.total;
}

int ibm_batch_build(int n, const clbm_grid_t* grids, const clbm_params_t* params, void* const* f,
                    const ibm_params_t* ip, const ibm_buffers_t* bufs, void* work,
                    size_t work_bytes, ibm_batch_t* out) {
  if (int e = check_ibm(ip)) return e;
  if (ip->check_region) return fail(-1, "ibm_batch_build: batches take no region check");
  if (ip->spread_mode == IBM_SPREAD_SORTED)
    return fail(-1, "ibm_batch_build: batches spread with fixed-point or float atomics");
  const int group = ip->group == 0 ? 8 : ip->group;
  IbmBatchLayout L;
  if (int e = ibm_batch_layout(n, bufs, group, L)) return e;
  if (!grids || !params || !f || !work || !out) return fail(-1, "ibm_batch_build: NULL argument");
  if (work_bytes < L.total) return fail(-1, "ibm_batch_build: work buffer too small");
  std::vector<unsigned char> host(L.total, 0);
  IbmArgs* args = (IbmArgs*)host.data();
  ClipArgs* clip = (ClipArgs*)(host.data() + L.clip);
  uint16_t* inst = (uint16_t*)(host.data() + L.inst);
  int32_t* first = (int32_t*)(host.data() + L.first);
  const float fscale = (float)std::ldexp(1.0, fixed_shift(ip));
  const float inv = (float)std::ldexp(1.0, -fixed_shift(ip));
  int blk = 0;
  for (int i = 0; i < n; ++i) {
    const ibm_buffers_t& m = bufs[i];
    const clbm_params_t& p = params[i];
    if (int e = check_grid(&grids[i], p.inlet != 0)) return e;
    if (int e = check_inlet(&p, "ibm_batch_build")) return e;
    if (grids[i].storage != grids[0].storage) return fail(-1, "ibm_batch_build: storage formats differ");
    if (!f[i] || !m.X0 || !m.Y0 || !m.Z0 || !m.X || !m.Y || !m.Z || !m.VX || !m.VY || !m.U || !m.V ||
        !m.FX || !m.FY || !m.acc_x || !m.acc_y || !m.box || !m.kin)
      return fail(-1, "ibm_batch_build: NULL buffer");
    IbmArgs a{};
    a.g = make_geo(grids[i]);
    a.pl.on = p.inlet ? 1 : 0;
    a.pl.rule = p.inlet == CLBM_INLET_HECHT_HARTING ? 1 : 0;
    a.pl.ux_in = p.ux_in;
    a.pl.uy_in = p.uy_in;
    a.quirk = ip->wrap_weight_quirk ? 1 : 0;
    a.f_max = ip->f_max;
    a.fscale = fscale;
    a.f = f[i];
    a.n = m.n;
    a.order = m.order;
    a.X0 = m.X0;
    a.Y0 = m.Y0;
    a.Z0 = m.Z0;
    a.X = m.X;
    a.Y = m.Y;
    a.Z = m.Z;
    a.VX = m.VX;
    a.VY = m.VY;
    a.U = m.U;
    a.V = m.V;
    a.FX = m.FX;
    a.FY = m.FY;
    a.accx = m.acc_x;
    a.accy = m.acc_y;
    a.box = m.box;
    a.kin = m.kin;
    a.gen0 = m.gen0;
    a.regions = m.regions;
    if ((m.regions != nullptr) != (bufs[0].regions != nullptr))
      return fail(-1, "ibm_batch_build: either every body or none carries a region table");
    args[i] = a;
    clip[i] = ClipArgs{a.g, ip->f_max, inv, m.acc_x, m.acc_y, m.fx, m.fy, m.box, m.gen0, m.regions};
    first[i] = blk;
    const int nb = (m.n + group - 1) / group;
    for (int j = 0; j < nb; ++j) inst[blk + j] = (uint16_t)i;
    blk += nb;
  }
  hipError_t e = hipMemcpy(work, host.data(), L.total, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail((int)e, std::string("ibm_batch_build: ") + hipGetErrorString(e));
  unsigned char* w = (unsigned char*)work;
  *out = ibm_batch_t{n, grids[0].storage, ip->spread_mode, group, L.blocks, CLIP_BPI, w + L.args,
                     w + L.clip, (uint16_t*)(w + L.inst), (int32_t*)(w + L.first),
                     (int32_t*)(w + L.counter), bufs[0].regions != nullptr ? 1 : 0};
  return 0;
}

int ibm_batch_compute(const ibm_batch_t* b, int state, void* stream) {
  if (!b || !b->args || !b->clip_args || !b->inst || !b->first || !b->counter || b->blocks <= 0)
    return fail(-1, "ibm_batch_compute: batch is not built");
  if (state & ~(CLBM_STATE_PARITY | CLBM_STATE_PLANES))
    return fail(-1, "ibm_batch_compute: state has unknown bits");
  if (b->group != 4 && b->group != 8) return fail(-1, "ibm_batch_compute: bad group");
  const int par = state & CLBM_STATE_PARITY;
  const int planes = (state & CLBM_STATE_PLANES) ? 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  if (b->spread_mode == IBM_SPREAD_FIXED) {
    CLBM_STORAGE_SWITCH(b->storage, (launch_batch<SS, IBM_SPREAD_FIXED>(b, par, planes, st)));
  } else {
    CLBM_STORAGE_SWITCH(b->storage, (launch_batch<SS, IBM_SPREAD_ATOMIC_F32>(b, par, planes, st)));
  }
  return check_launch("ibm_batch_compute");
}

int ibm_batch_markers(const ibm_batch_t* b, int state, void* stream) {
  if (!b || !b->args || !b->clip_args || !b->inst || !b->first || !b->counter || b->blocks <= 0)
    return fail(-1, "ibm_batch_markers: batch is not built");
  if (state & ~(CLBM_STATE_PARITY | CLBM_STATE_PLANES))
    return fail(-1, "ibm_batch_markers: state has unknown bits");
  if (b->group != 4 && b->group != 8) return fail(-1, "ibm_batch_markers: bad group");
  if (b->spread_mode != IBM_SPREAD_FIXED)
    return fail(-1, "ibm_batch_markers: needs the fixed-point spread mode");
  if (!b->regions) return fail(-1, "ibm_batch_markers: the bodies carry no region tables");
  const int par = state & CLBM_STATE_PARITY;
  const int planes = (state & CLBM_STATE_PLANES) ? 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  CLBM_STORAGE_SWITCH(b->storage, (launch_markers_reg_g<SS>(b, par, planes, st)));
  return check_launch("ibm_batch_markers");
}

int clbm_batch_step_acc(const clbm_batch_t* lb, const ibm_batch_t* ib, int state, void* stream) {
  if (!lb || !lb->args || !lb->inst || !lb->first || lb->blocks <= 0)
    return fail(-1, "clbm_batch_step_acc: lattice batch is not built");
  if (!ib || !ib->clip_args || !ib->counter)
    return fail(-1, "clbm_batch_step_acc: body batch is not built");
  if (lb->n != ib->n || lb->storage != ib->storage)
    return fail(-1, "clbm_batch_step_acc: the two batches differ in size or storage");
  if (ib->spread_mode != IBM_SPREAD_FIXED)
    return fail(-1, "clbm_batch_step_acc: needs the fixed-point spread mode");
  if (!ib->regions) return fail(-1, "clbm_batch_step_acc: the bodies carry no region tables");
  if (state & ~(CLBM_STATE_PARITY | CLBM_STATE_PLANES))
    return fail(-1, "clbm_batch_step_acc: state has unknown bits");
  if (lb->block != 128 && lb->block != 256) return fail(-1, "clbm_batch_step_acc: bad block size");
  if (lb->lds < 0 || lb->lds > 64 * 1024)
    return fail(-1, "clbm_batch_step_acc: lds must be in [0, 65536]");
  const int par = state & CLBM_STATE_PARITY;
  const int planes = (state & CLBM_STATE_PLANES) ? 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  CLBM_STORAGE_SWITCH(lb->storage, (launch_step_acc_s<SS>(lb, ib, par, planes, st)));
  return check_launch("clbm_batch_step_acc");
}

int ibm_step_region(const clbm_grid_t* g, const clbm_params_t* p, const float* hx,
                    const float* hy, int nh, int z0, int z1, double xc, double yc, double theta,
                    int32_t* out) {
  // float32 placement as the marker kernel does it: cos/sin of the float32
  // angle rounded once from float64, each product and sum rounded on its own
#pragma clang fp contract(off)
  if (!p || !out || (nh > 0 && (!hx || !hy))) return fail(-1, "ibm_step_region: NULL argument");
  if (int e = check_grid(g, p->inlet != 0)) return e;
  if (nh < 0 || z0 > z1) return fail(-1, "ibm_step_region: bad hull count or z range");
  if (nh == 0) return 0;
  const float t = (float)theta;
  const float c = (float)std::cos((double)t), s = (float)std::sin((double)t);
  const float fxc = (float)xc, fyc = (float)yc;
  float xmin = 0.f, xmax = 0.f, ymin = 0.f, ymax = 0.f;
  for (int i = 0; i < nh; ++i) {
    const float hxc = hx[i] * c, hys = hy[i] * s, hxs = hx[i] * s, hyc = hy[i] * c;
    const float X = (fxc + hxc) - hys;
    const float Y = (fyc + hxs) + hyc;
    if (!std::isfinite(X) || !std::isfinite(Y)) return 0;
    if (i == 0 || X < xmin) xmin = X;
    if (i == 0 || X > xmax) xmax = X;
    if (i == 0 || Y < ymin) ymin = Y;
    if (i == 0 || Y > ymax) ymax = Y;
  }
  // stencil cells floor(p) - 1 .. floor(p) + 2, plus one cell of margin
  const double x0 = std::max(std::floor((double)xmin) - 2.0, 0.0);
  const double x1 = std::min(std::floor((double)xmax) + 3.0, (double)(g->nx - 1));
  if (x0 > x1) return 0;
  double y0 = std::floor((double)ymin) - 2.0, y1 = std::floor((double)ymax) + 3.0;
  if (y0 < 0.0 || y1 > (double)(g->ny - 1)) {  // crosses the periodic boundary: whole axis
    y0 = 0.0;
    y1 = (double)(g->ny - 1);
  }
  const int32_t raw[6] = {(int32_t)x0, (int32_t)y0, z0, (int32_t)x1, (int32_t)y1, z1};
  if (int e = clbm_step_region(g, p, raw, out)) return e;
  const long long vol = (long long)(out[3] - out[0] + 1) * (out[4] - out[1] + 1) *
                        (out[5] - out[2] + 1);
  const long long ncell = (long long)g->nx * g->ny * g->nz;
  return 2 * vol > ncell ? 0 : 1;  // a region over half the lattice does not pay
}

int ibm_compute_force_split(ibm_split_plan_t* plan, int state, uint32_t gen, float xc, float yc,
                            float theta, float vx, float vy, float omega, void* main,
                            int fork_bound, int32_t* region_out) {
  if (!plan || !region_out) return fail(-1, "ibm_compute_force_split: NULL argument");
  ibm_split_plan_t& P = *plan;
  int split = ibm_step_region(&P.grid, &P.params, P.hx, P.hy, P.nh, P.z0, P.z1, (double)xc,
                              (double)yc, (double)theta, region_out);
  if (split < 0) return split;
  if (split && (!P.side || !P.fork_ev)) split = 0;
  void* st = main;
  if (split) {
    if (fork_bound) {
      // fork_ev completes with the last step's inside part (clbm_step_split's done_ev)
      const hipError_t r = hipStreamWaitEvent((hipStream_t)P.side, (hipEvent_t)P.fork_ev, 0);
      if (r != hipSuccess)
        return fail((int)r, std::string("ibm_compute_force_split: ") + hipGetErrorString(r));
    } else if (int e = clbm_stream_after(P.side, main, P.fork_ev)) {
      return e;
    }
    st = P.side;
    P.ip.check_region = 1;
    for (int i = 0; i < 6; ++i) P.ip.region[i] = region_out[i];
  } else {
    P.ip.check_region = 0;
  }
  if (int e = ibm_compute_force(&P.grid, &P.ip, &P.params, P.f, state, nullptr, nullptr, P.n,
                                P.order, P.X0, P.Y0, P.Z0, xc, yc, theta, vx, vy, omega, P.X,
                                P.Y, P.Z, P.VX, P.VY, P.U, P.V, P.FX, P.FY, P.stencil_base,
                                P.acc_x, P.acc_y, P.box, gen, st))
    return e;
  if (int e = ibm_clip_forces(&P.grid, &P.ip, P.acc_x, P.acc_y, P.fx, P.fy, P.box, gen, st))
    return e;
  if (split) {
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, (hipStream_t)st,
                       (unsigned long long*)IBM_SPLIT_READY(P.box), (unsigned long long)gen);
    if (int e = check_launch("ibm_compute_force_split")) return e;
  }
  return split;
}

int ibm_force_sum_f64(int n, const float* FX, const float* FY, const float* dV, double* out2,
                      void* stream) {
  if (n < 0) return fail(-1, "ibm_force_sum_f64: n < 0");
  if (!out2 || (n > 0 && (!FX || !FY))) return fail(-1, "ibm_force_sum_f64: NULL buffer");
  hipLaunchKernelGGL(k_force_sum, dim3(1), dim3(SBLK), 0, (hipStream_t)stream, n, FX, FY, dV, out2);
  return check_launch("ibm_force_sum_f64");
}

}  // extern "C"
