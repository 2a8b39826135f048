// Lattice-Boltzmann kernels of the CLBM-IBM step for gfx950 (MI355X).
//
// One fused pass per step replaces all of CLBMSolver.step (clbm_moving.py:
// 326-343): each thread loads the 19 post-stream populations of its cell,
// applies the reference's boundary planes on load (inlet rule at x = 0,
// outlet copy at x = Nx-1, left pending by the previous step), adds the
// force, collides and writes the post-collision values where the next step
// reads them as post-stream values (AA pattern: in place, one buffer).  The
// macroscopic fields (K2, K7) live in registers; K7 is materialised only when
// someone reads rho/u.
#include "clbm_step.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <vector>

namespace clbm {

// ---------------------------------------------------------------------------
// errors and argument checks
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;
void set_error(const std::string& s) { g_last_error = s; }
int fail(int code, const std::string& s) {
  g_last_error = s;
  return code;
}
int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}
int check_grid(const clbm_grid_t* g, bool need_inlet_width) {
  if (!g) return fail(-1, "grid is NULL");
  if (g->nx < 1 || g->ny < 1 || g->nz < 1) return fail(-1, "grid sizes must be positive");
  if (need_inlet_width && g->nx < 3) return fail(-1, "the inlet/outlet rule needs nx >= 3");
  long long nc = (long long)g->nx * g->ny * g->nz;
  if (nc >= (1ll << 31)) return fail(-1, "grid has 2^31 cells or more");
  if (elem_bytes(g->storage) == 0) return fail(-1, "unknown storage format");
  // kernels address a population plane with 32-bit byte offsets
  if ((unsigned long long)nc * elem_bytes(g->storage) > 0xffffffffull)
    return fail(-1, "grid too large for this storage: one population plane would exceed 4 GiB");
  return 0;
}
int check_inlet(const clbm_params_t* p, const char* who) {
  if (p->inlet < CLBM_INLET_NONE || p->inlet > CLBM_INLET_HECHT_HARTING)
    return fail(-1, std::string(who) + ": inlet must be 0 (periodic x), 1 (the reference's rule) or 2 (Hecht-Harting)");
  if (p->inlet && !(p->ux_in < 1.0)) return fail(-1, std::string(who) + ": inlet needs ux_in < 1");
  return 0;
}
int check_state(const clbm_grid_t* g, const clbm_params_t* p, int state, Planes& pl) {
  if (state & ~(CLBM_STATE_PARITY | CLBM_STATE_PLANES))
    return fail(-1, "state has unknown bits (expected parity | planes flag)");
  pl.on = (state & CLBM_STATE_PLANES) ? 1 : 0;
  pl.rule = 0;
  pl.ux_in = pl.uy_in = 0.0;
  if (pl.on) {
    if (!p) return fail(-1, "a state with pending boundary planes needs the params (inlet velocity)");
    if (g->nx < 3) return fail(-1, "the inlet/outlet rule needs nx >= 3");
    pl.rule = p->inlet == CLBM_INLET_HECHT_HARTING ? 1 : 0;
    pl.ux_in = p->ux_in;
    pl.uy_in = p->uy_in;
  }
  return 0;
}


// ---------------------------------------------------------------------------
// K1: equilibrium initialisation (clbm_moving.py:75-91)
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(BLK) void k_init_feq(Geo g, typename Store<S>::T* __restrict__ f,
                                                  double rho0, double ux0, double uy0, double uz0) {
  using R = typename Store<S>::R;
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  const size_t N = (size_t)g.ncell;
  const R cs2 = R(1) / R(3);
  const R u = R(ux0), v = R(uy0), w = R(uz0), r = R(rho0);
  const R u2 = u * u + v * v + w * w;
  const DevBase<S> dv(g);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const R eu = R(ex(q)) * u + R(ey(q)) * v + R(ez(q)) * w;
    const R fe = wq<R>(q) * r * (R(1) + eu / cs2 + R(0.5) * eu * eu / (cs2 * cs2) - R(0.5) * u2 / cs2);
    Store<S>::st(f + (size_t)q * N + c, fe, dv[q]);
  }
}

// ---------------------------------------------------------------------------
// Fused boundary planes + macro + collide + Guo + stream (AA pattern).
// PAR: parity of the incoming state.  FORCE: force fields present.  BOX: the
// force is read only inside box[0..5].
// ---------------------------------------------------------------------------
template <int S, int PAR, bool FORCE, bool BOX, int REG>
__global__ __launch_bounds__(BLK) void k_step(StepArgs a) {
  step_cell<S, PAR, REG>(a, sweep_block<PAR>(blockIdx.x, gridDim.x) * BLK + threadIdx.x,
                         FieldForce<FORCE, BOX>{a.fx, a.fy, a.fz, a.box});
}

// The split step in one launch (clbm_step_fused).  Blocks [0, nb_out) update
// the cells outside the region force-free, in the sweep order of part 0; the
// blocks after them update the region's cells with the force, as part 1,
// once the clamp of generation `gen` has published its forces in *ready
// (k_publish, after the clamp on the marker stream).  The region's blocks are
// dispatched after every outside block, by when the marker kernels beside
// the sweep have long finished, so the wait is normally one poll.  Thread 0
// of each region block polls (with s_sleep) for at most `ticks` of the
// constant-rate clock; when that runs out it marks the generation in
// *expired and the block steps its cells with the forces as they are, which
// the health check reports as an error.  ready and expired sit in a 128-byte
// line of their own: the loads that poll them must not bring a stale copy of
// the decoded box, written by the clamp meanwhile, into this XCD's L2.  No
// other line the region reads is touched in this launch before the poll
// succeeds, so plain loads after it read what the clamp wrote.
struct FusedWait {
  const unsigned long long* ready;
  unsigned long long* expired;
  unsigned long long gen;
  long long ticks;
};
template <int S, int PAR>
__global__ __launch_bounds__(BLK) void k_step_fused(StepArgs a, int nb_out, FusedWait w) {
  const int b = blockIdx.x;
  if (b < nb_out) {
    step_cell<S, PAR, 1>(a, sweep_block<PAR>(b, nb_out) * BLK + threadIdx.x,
                         FieldForce<false, false>{nullptr, nullptr, nullptr, nullptr});
    return;
  }
  if (threadIdx.x == 0) {
    const long long t0 = wall_clock64();
    unsigned long long v = __hip_atomic_load(w.ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (v < w.gen && wall_clock64() - t0 < w.ticks) {
      __builtin_amdgcn_s_sleep(8);
      v = __hip_atomic_load(w.ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (v < w.gen) atomicMax(w.expired, w.gen);
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  step_cell<S, PAR, 2>(a, (b - nb_out) * BLK + threadIdx.x,
                       FieldForce<true, true>{a.fx, a.fy, a.fz, a.box});
}

// The data movement of one AA step without the collision: every cell's 19
// post-stream slots loaded and stored into the slots the step writes (the
// populations permuted, their values kept), in the step's sweep order.  The
// ceiling that the fused step kernel is measured against.
template <int S, int PAR>
__global__ __launch_bounds__(BLK) void k_stream_probe(Geo g, void* f) {
  const int c = sweep_block<PAR>(blockIdx.x, gridDim.x) * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  int x, y, z;
  cell_xyz(g, c, x, y, z);
  const Pop<S> pop((typename Store<S>::T*)f, g);
  const Nbr nb = make_nbr(g, x, y, z);
  typename Store<S>::R fq[Q];
  pop_load<PAR>(pop, c, nb, fq);
  pop_store<PAR>(pop, c, nb, fq);
}

// Many lattices in one launch: block b (in sweep order) belongs to
// instance inst[b] and is that instance's block b - first[inst[b]]; every
// instance reads its force inside its force box.
template <int S, int PAR, int BK>
__global__ __launch_bounds__(BK) void k_step_batch(const StepArgs* __restrict__ args,
                                                   const uint16_t* __restrict__ inst,
                                                   const int32_t* __restrict__ first,
                                                   int planes_on) {
  const int lb = sweep_block<PAR>(blockIdx.x, gridDim.x);
  const int i = inst[lb];
  StepArgs a = args[i];
  a.pl.on &= planes_on;
  step_cell<S, PAR, 0>(a, (lb - first[i]) * BK + threadIdx.x,
                      FieldForce<true, true>{a.fx, a.fy, a.fz, a.box});
}

// ---------------------------------------------------------------------------
// K7: force-free macroscopic fields (clbm_moving.py:305-321)
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void moments01(const R* a, R& r, R& mx, R& my, R& mz) {
  r = R(0);
  mx = R(0);
  my = R(0);
  mz = R(0);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    r += a[q];
    if (ex(q)) mx += R(ex(q)) * a[q];
    if (ey(q)) my += R(ey(q)) * a[q];
    if (ez(q)) mz += R(ez(q)) * a[q];
  }
}

template <int S, int PAR>
__global__ __launch_bounds__(BLK) void k_macro(Geo g, Planes pl, const typename Store<S>::T* __restrict__ f,
                                               float* __restrict__ rho, float* __restrict__ ux,
                                               float* __restrict__ uy, float* __restrict__ uz) {
  using R = typename Store<S>::R;
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  int x, y, z;
  cell_xyz(g, c, x, y, z);
  R a[Q];
  load_cell<S, PAR>(f, g, c, x, y, z, pl, a);
  R r, mx, my, mz;
  moments01(a, r, mx, my, mz);
  if (rho) rho[c] = (float)r;
  if (ux) ux[c] = (float)(mx / r);
  if (uy) uy[c] = (float)(my / r);
  if (uz) uz[c] = (float)(mz / r);
}

template <int S, int PAR, typename OutT>
__global__ __launch_bounds__(BLK) void k_export(Geo g, Planes pl,
                                                const typename Store<S>::T* __restrict__ f,
                                                OutT* __restrict__ out) {
  using R = typename Store<S>::R;
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  int x, y, z;
  cell_xyz(g, c, x, y, z);
  R a[Q];
  load_cell<S, PAR>(f, g, c, x, y, z, pl, a);
#pragma unroll
  for (int q = 0; q < Q; ++q) out[(size_t)q * g.ncell + c] = (OutT)a[q];
}

template <int S, typename InT>
__global__ __launch_bounds__(BLK) void k_import(Geo g, typename Store<S>::T* __restrict__ f,
                                                const InT* __restrict__ in) {
  const int c = blockIdx.x * BLK + threadIdx.x;
  if (c >= g.ncell) return;
  const DevBase<S> dv(g);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const size_t i = (size_t)q * g.ncell + c;
    Store<S>::st(f + i, (typename Store<S>::R)in[i], dv[q]);
  }
}

// Order-preserving map of a float onto a signed int (for atomicMin).
__device__ __forceinline__ int fkey(float v) {
  int b = __float_as_int(v);
  return b >= 0 ? b : (b ^ 0x7fffffff);
}
__device__ __forceinline__ float fkey_inv(int k) {
  return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff));
}

__global__ void k_health_init(int* acc) {
  acc[0] = 0x7f800000;  // key of +inf
  acc[1] = 1;
}

template <int S, int PAR>
__global__ __launch_bounds__(BLK) void k_health(Geo g, Planes pl,
                                                const typename Store<S>::T* __restrict__ f,
                                                int* __restrict__ acc) {
  using R = typename Store<S>::R;
  __shared__ int smin[BLK / 64], sfin[BLK / 64];
  int key = 0x7f800000, fin = 1;
  // grid-stride: a few hundred blocks, so that the per-block atomics on the
  // two result words (which serialise) stay few
  for (int c = blockIdx.x * BLK + threadIdx.x; c < g.ncell; c += gridDim.x * BLK) {
    int x, y, z;
    cell_xyz(g, c, x, y, z);
    R a[Q];
    load_cell<S, PAR>(f, g, c, x, y, z, pl, a);
    R r = R(0);
    int ok = 1;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      r += a[q];
      ok &= (int)isfinite(a[q]);
    }
    fin &= ok;
    if (ok) key = min(key, fkey((float)r));
  }
  for (int o = 32; o > 0; o >>= 1) {
    key = min(key, __shfl_xor(key, o));
    fin = min(fin, __shfl_xor(fin, o));
  }
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  if (ln == 0) {
    smin[wv] = key;
    sfin[wv] = fin;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int k = smin[0], fl = sfin[0];
    for (int i = 1; i < BLK / 64; ++i) {
      k = min(k, smin[i]);
      fl = min(fl, sfin[i]);
    }
    atomicMin(acc, k);
    atomicMin(acc + 1, fl);
  }
}

__global__ void k_health_fin(const int* acc, float* out2) {
  out2[0] = fkey_inv(acc[0]);
  out2[1] = acc[1] ? 1.0f : 0.0f;
}

// ---------------------------------------------------------------------------
// host dispatch helpers
// ---------------------------------------------------------------------------
static inline int nblocks(long long n) { return (int)((n + BLK - 1) / BLK); }

template <template <int> class Fn, typename... A>
static int dispatch_storage(int storage, A&&... a) {
  switch (storage) {
    case CLBM_STORE_F32: return Fn<CLBM_STORE_F32>::run(a...);
    case CLBM_STORE_F16: return Fn<CLBM_STORE_F16>::run(a...);
    case CLBM_STORE_BF16: return Fn<CLBM_STORE_BF16>::run(a...);
    case CLBM_STORE_F64: return Fn<CLBM_STORE_F64>::run(a...);
    case CLBM_STORE_F16C: return Fn<CLBM_STORE_F16C>::run(a...);
  }
  return fail(-1, "unknown storage format");
}

template <int S> struct InitRun {
  static int run(const Geo& g, void* f, double r, double u, double v, double w, hipStream_t st) {
    hipLaunchKernelGGL(k_init_feq<S>, dim3(nblocks(g.ncell)), dim3(BLK), 0, st, g,
                       (typename Store<S>::T*)f, r, u, v, w);
    return check_launch("clbm_init_feq");
  }
};

// Timing events for the next launch of the outside part (clbm_step_split):
// the runtime stamps them with that kernel's own start and end, unlike events
// recorded on the stream, whose start stamp is the end of the command before.
static thread_local hipEvent_t t_outside_ev[2] = {nullptr, nullptr};
// An event for the next launch of the inside part to complete with the
// kernel itself (clbm_step_split's done_ev): no marker follows it on the stream.
static thread_local hipEvent_t t_inside_done = nullptr;

// part: -1 the whole grid, 0 the cells outside the region (force-free), 1 the
// cells inside it.
template <int S> struct StepRun {
  template <int PAR, bool FORCE, bool BOX, int REG>
  static void launch(const Geo& g, const clbm_params_t* p, const Planes& pl, void* f,
                     const float* fx, const float* fy, const float* fz, const int32_t* box,
                     const Region& rg, hipStream_t st) {
    const int n = REG == 2 ? rg.vol : g.ncell;
    const unsigned lds = REG == 1 ? (unsigned)p->outside_lds : 0u;
    const StepArgs a{g, p->s_v, p->s_e, pl, f, fx, fy, fz, box, rg};
    if (REG == 1 && (t_outside_ev[0] || t_outside_ev[1])) {
      hipExtLaunchKernelGGL((k_step<S, PAR, FORCE, BOX, REG>), dim3(nblocks(n)), dim3(BLK), lds,
                            st, t_outside_ev[0], t_outside_ev[1], 0u, a);
      t_outside_ev[0] = t_outside_ev[1] = nullptr;
      return;
    }
    if (REG == 2 && t_inside_done) {
      hipExtLaunchKernelGGL((k_step<S, PAR, FORCE, BOX, REG>), dim3(nblocks(n)), dim3(BLK), lds,
                            st, nullptr, t_inside_done, 0u, a);
      t_inside_done = nullptr;
      return;
    }
    hipLaunchKernelGGL((k_step<S, PAR, FORCE, BOX, REG>), dim3(nblocks(n)), dim3(BLK), lds, st, a);
  }
  template <int PAR, int REG>
  static void launch_force(const Geo& g, const clbm_params_t* p, const Planes& pl, void* f,
                           const float* fx, const float* fy, const float* fz, const int32_t* box,
                           const Region& rg, hipStream_t st) {
    const bool force = fx || fy || fz;
    if (!force)
      launch<PAR, false, false, REG>(g, p, pl, f, fx, fy, fz, box, rg, st);
    else if (box)
      launch<PAR, true, true, REG>(g, p, pl, f, fx, fy, fz, box, rg, st);
    else
      launch<PAR, true, false, REG>(g, p, pl, f, fx, fy, fz, box, rg, st);
  }
  template <int PAR>
  static void launch_par(const Geo& g, const clbm_params_t* p, const Planes& pl, void* f,
                         const float* fx, const float* fy, const float* fz, const int32_t* box,
                         const Region& rg, int part, hipStream_t st) {
    if (part < 0)
      launch_force<PAR, 0>(g, p, pl, f, fx, fy, fz, box, rg, st);
    else if (part == 0)
      launch<PAR, false, false, 1>(g, p, pl, f, nullptr, nullptr, nullptr, nullptr, rg, st);
    else
      launch_force<PAR, 2>(g, p, pl, f, fx, fy, fz, box, rg, st);
  }
  static int run(const Geo& g, const clbm_params_t* p, const Planes& pl, void* f, int parity,
                 const float* fx, const float* fy, const float* fz, const int32_t* box,
                 const Region& rg, int part, hipStream_t st) {
    if (parity == 0)
      launch_par<0>(g, p, pl, f, fx, fy, fz, box, rg, part, st);
    else
      launch_par<1>(g, p, pl, f, fx, fy, fz, box, rg, part, st);
    return check_launch(part < 0 ? "clbm_step" : "clbm_step_part");
  }
};

template <int S> struct FusedRun {
  static int run(const Geo& g, const clbm_params_t* p, const Planes& pl, void* f, int parity,
                 const float* fx, const float* fy, const int32_t* box, const Region& rg,
                 const FusedWait& w, hipEvent_t e0, hipEvent_t e1, hipStream_t st) {
    const StepArgs a{g, p->s_v, p->s_e, pl, f, fx, fy, nullptr, box, rg};
    const int nb_out = nblocks(g.ncell);
    const dim3 grid(nb_out + nblocks(rg.vol));
    const unsigned lds = (unsigned)p->outside_lds;
    const bool timed = e0 || e1;
    if (parity == 0) {
      if (timed)
        hipExtLaunchKernelGGL((k_step_fused<S, 0>), grid, dim3(BLK), lds, st, e0, e1, 0u, a, nb_out, w);
      else
        hipLaunchKernelGGL((k_step_fused<S, 0>), grid, dim3(BLK), lds, st, a, nb_out, w);
    } else {
      if (timed)
        hipExtLaunchKernelGGL((k_step_fused<S, 1>), grid, dim3(BLK), lds, st, e0, e1, 0u, a, nb_out, w);
      else
        hipLaunchKernelGGL((k_step_fused<S, 1>), grid, dim3(BLK), lds, st, a, nb_out, w);
    }
    return check_launch("clbm_step_fused");
  }
};

template <int S> struct ProbeRun {
  static int run(const Geo& g, void* f, int parity, hipStream_t st) {
    if (parity == 0)
      hipLaunchKernelGGL((k_stream_probe<S, 0>), dim3(nblocks(g.ncell)), dim3(BLK), 0, st, g, f);
    else
      hipLaunchKernelGGL((k_stream_probe<S, 1>), dim3(nblocks(g.ncell)), dim3(BLK), 0, st, g, f);
    return check_launch("clbm_stream_probe");
  }
};

template <int S> struct MacroRun {
  static int run(const Geo& g, const Planes& pl, const void* f, int parity, float* r, float* u,
                 float* v, float* w, hipStream_t st) {
    auto fp = (const typename Store<S>::T*)f;
    if (parity == 0)
      hipLaunchKernelGGL((k_macro<S, 0>), dim3(nblocks(g.ncell)), dim3(BLK), 0, st, g, pl, fp, r, u, v, w);
    else
      hipLaunchKernelGGL((k_macro<S, 1>), dim3(nblocks(g.ncell)), dim3(BLK), 0, st, g, pl, fp, r, u, v, w);
    return check_launch("clbm_macro");
  }
};

template <int S> struct ExportRun {
  template <typename OutT>
  static int run(const Geo& g, const Planes& pl, const void* f, int parity, OutT* out,
                 hipStream_t st) {
    auto fp = (const typename Store<S>::T*)f;
    if (parity == 0)
      hipLaunchKernelGGL((k_export<S, 0, OutT>), dim3(nblocks(g.ncell)), dim3(BLK), 0, st, g, pl, fp, out);
    else
      hipLaunchKernelGGL((k_export<S, 1, OutT>), dim3(nblocks(g.ncell)), dim3(BLK), 0, st, g, pl, fp, out);
    return check_launch("clbm_export_state");
  }
};

template <int S> struct ImportRun {
  template <typename InT>
  static int run(const Geo& g, void* f, const InT* in, hipStream_t st) {
    hipLaunchKernelGGL((k_import<S, InT>), dim3(nblocks(g.ncell)), dim3(BLK), 0, st, g,
                       (typename Store<S>::T*)f, in);
    return check_launch("clbm_import_state");
  }
};

template <int S> struct HealthRun {
  static int run(const Geo& g, const Planes& pl, const void* f, int parity, float* out2,
                 hipStream_t st) {
    // out2 is used as int scratch before the final conversion
    int* acc = (int*)out2;
    hipLaunchKernelGGL(k_health_init, dim3(1), dim3(1), 0, st, acc);
    auto fp = (const typename Store<S>::T*)f;
    if (parity == 0)
      hipLaunchKernelGGL((k_health<S, 0>), dim3(std::min(nblocks(g.ncell), 512)), dim3(BLK), 0, st, g, pl, fp, acc);
    else
      hipLaunchKernelGGL((k_health<S, 1>), dim3(std::min(nblocks(g.ncell), 512)), dim3(BLK), 0, st, g, pl, fp, acc);
    hipLaunchKernelGGL(k_health_fin, dim3(1), dim3(1), 0, st, acc, out2);
    return check_launch("clbm_health");
  }
};

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct BatchLayout {
  size_t args, inst, first, total;
  int blocks;
};
static int batch_layout(int n, const clbm_grid_t* grids, int block, BatchLayout& L) {
  if (n <= 0 || n > 65535) return fail(-1, "batch: n must be in [1, 65535]");
  if (!grids) return fail(-1, "batch: grids is NULL");
  if (block != 128 && block != 256) return fail(-1, "batch: block must be 128 or 256");
  long long blocks = 0;
  for (int i = 0; i < n; ++i) {
    if (int e = check_grid(&grids[i], false)) return e;
    if (grids[i].storage != grids[0].storage) return fail(-1, "batch: storage formats differ");
    blocks += ((long long)grids[i].nx * grids[i].ny * grids[i].nz + block - 1) / block;
  }
  if (blocks >= (1ll << 31)) return fail(-1, "batch: too many blocks");
  L.blocks = (int)blocks;
  L.args = 0;
  L.inst = align256(sizeof(StepArgs) * n);
  L.first = L.inst + align256(sizeof(uint16_t) * (size_t)blocks);
  L.total = L.first + align256(sizeof(int32_t) * n);
  return 0;
}

template <int S> struct BatchRun {
  template <int PAR, int BK>
  static void go(const clbm_batch_t* b, int planes, hipStream_t st) {
    hipLaunchKernelGGL((k_step_batch<S, PAR, BK>), dim3(b->blocks), dim3(BK), (unsigned)b->lds,
                       st, (const StepArgs*)b->args, b->inst, b->first, planes);
  }
  static int run(const clbm_batch_t* b, int state, hipStream_t st) {
    const int planes = (state & CLBM_STATE_PLANES) ? 1 : 0;
    const bool odd = state & CLBM_STATE_PARITY;
    if (b->block == 128) {
      if (odd) go<1, 128>(b, planes, st); else go<0, 128>(b, planes, st);
    } else {
      if (odd) go<1, 256>(b, planes, st); else go<0, 256>(b, planes, st);
    }
    return check_launch("clbm_batch_step");
  }
};

// ---------------------------------------------------------------------------
// the marker stream and its hardware queue
// ---------------------------------------------------------------------------
// HIP maps each stream onto one of a few hardware queues (GPU_MAX_HW_QUEUES,
// 4 by default); once they are all taken, a new stream shares the least used
// one.  A marker stream that shares the solver stream's queue still orders
// the work correctly, but its kernels wait behind the sweep instead of
// running beside it (a step measured 14-18 % slower).  The check: one wave on
// `main` waits, for at most `ticks` of the constant-rate wall clock, for a
// flag that one wave on the candidate stream sets.  Queued behind the waiting
// wave, the setter runs only after the wait has run out.
// [0] the flag, [1] what the waiting wave saw
static __device__ __align__(64) unsigned g_queue_flag[2];
__global__ void k_queue_wait(unsigned* flag, unsigned gen, long long ticks) {
  const long long t0 = wall_clock64();
  unsigned v = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (v != gen && wall_clock64() - t0 < ticks) {
    __builtin_amdgcn_s_sleep(8);
    v = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x == 0) flag[1] = (v == gen) ? gen : 0u;
}
__global__ void k_queue_signal(unsigned* flag, unsigned gen) {
  if (threadIdx.x == 0) __hip_atomic_store(flag, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct SideStream {
  hipStream_t stream = nullptr;
  int concurrent = 0;
};

}  // namespace clbm

using namespace clbm;

extern "C" {

size_t clbm_batch_bytes(int n, const clbm_grid_t* grids, int block) {
  BatchLayout L;
  if (batch_layout(n, grids, block, L)) return 0;
  return L.total;
}

int clbm_batch_build(int n, const clbm_grid_t* grids, const clbm_params_t* params, void* const* f,
                     float* const* fx, float* const* fy, uint64_t* const* boxes, int block,
                     void* work, size_t work_bytes, clbm_batch_t* out) {
  BatchLayout L;
  if (int e = batch_layout(n, grids, block, L)) return e;
  if (!params || !f || !fx || !fy || !boxes || !work || !out)
    return fail(-1, "clbm_batch_build: NULL argument");
  if (work_bytes < L.total) return fail(-1, "clbm_batch_build: work buffer too small");
  std::vector<unsigned char> host(L.total, 0);
  StepArgs* args = (StepArgs*)host.data();
  uint16_t* inst = (uint16_t*)(host.data() + L.inst);
  int32_t* first = (int32_t*)(host.data() + L.first);
  int blk = 0;
  for (int i = 0; i < n; ++i) {
    const clbm_params_t& p = params[i];
    if (int e = check_grid(&grids[i], p.inlet != 0)) return e;
    if (!f[i] || !fx[i] || !fy[i] || !boxes[i]) return fail(-1, "clbm_batch_build: NULL buffer");
    if (int e = check_inlet(&p, "clbm_batch_build")) return e;
    StepArgs a{};
    a.g = make_geo(grids[i]);
    a.s_v = p.s_v;
    a.s_e = p.s_e;
    a.pl.on = p.inlet ? 1 : 0;   // the planes of a step are those of the state AND this flag
    a.pl.rule = p.inlet == CLBM_INLET_HECHT_HARTING ? 1 : 0;
    a.pl.ux_in = p.ux_in;
    a.pl.uy_in = p.uy_in;
    a.f = f[i];
    a.fx = fx[i];
    a.fy = fy[i];
    a.fz = nullptr;
    a.box = IBM_FORCE_BOX(boxes[i]);
    args[i] = a;
    first[i] = blk;
    const int nb = (a.g.ncell + block - 1) / block;
    for (int j = 0; j < nb; ++j) inst[blk + j] = (uint16_t)i;
    blk += nb;
  }
  hipError_t e = hipMemcpy(work, host.data(), L.total, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail((int)e, std::string("clbm_batch_build: ") + hipGetErrorString(e));
  unsigned char* w = (unsigned char*)work;
  *out = clbm_batch_t{n, grids[0].storage, block, L.blocks, w, (uint16_t*)(w + L.inst),
                      (int32_t*)(w + L.first), 0};
  return 0;
}

int clbm_batch_step(const clbm_batch_t* b, int state, void* stream) {
  if (!b || !b->args || !b->inst || !b->first || b->blocks <= 0)
    return fail(-1, "clbm_batch_step: batch is not built");
  if (state & ~(CLBM_STATE_PARITY | CLBM_STATE_PLANES))
    return fail(-1, "clbm_batch_step: state has unknown bits");
  if (b->block != 128 && b->block != 256) return fail(-1, "clbm_batch_step: bad block size");
  if (b->lds < 0 || b->lds > 64 * 1024) return fail(-1, "clbm_batch_step: lds must be in [0, 65536]");
  return dispatch_storage<BatchRun>(b->storage, b, state, (hipStream_t)stream);
}

int clbm_abi_version(void) { return CLBM_ABI_VERSION; }

int clbm_event_create(void** ev) {
  if (!ev) return fail(-1, "clbm_event_create: NULL argument");
  hipEvent_t e = nullptr;
  const hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
  if (r != hipSuccess) return fail((int)r, std::string("clbm_event_create: ") + hipGetErrorString(r));
  *ev = (void*)e;
  return 0;
}

int clbm_event_destroy(void* ev) {
  if (!ev) return 0;
  const hipError_t r = hipEventDestroy((hipEvent_t)ev);
  if (r != hipSuccess) return fail((int)r, std::string("clbm_event_destroy: ") + hipGetErrorString(r));
  return 0;
}

int clbm_stream_after(void* waiter, void* signaler, void* ev) {
  if (!ev) return fail(-1, "clbm_stream_after: NULL event");
  hipError_t r = hipEventRecord((hipEvent_t)ev, (hipStream_t)signaler);
  if (r == hipSuccess) r = hipStreamWaitEvent((hipStream_t)waiter, (hipEvent_t)ev, 0);
  if (r != hipSuccess) return fail((int)r, std::string("clbm_stream_after: ") + hipGetErrorString(r));
  return 0;
}

int clbm_streams_concurrent(void* main, void* other, int* concurrent) {
  if (!concurrent) return fail(-1, "clbm_streams_concurrent: NULL argument");
  static std::mutex mu;
  static std::map<int, unsigned> gens;   // per device: the last flag value used
  std::lock_guard<std::mutex> lock(mu);
  int dev = 0, khz = 0;
  hipError_t r = hipGetDevice(&dev);
  if (r == hipSuccess) r = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  if (r == hipSuccess && khz <= 0) r = hipErrorInvalidValue;
  unsigned* flag = nullptr;
  if (r == hipSuccess) r = hipGetSymbolAddress((void**)&flag, HIP_SYMBOL(g_queue_flag));
  auto g = gens.find(dev);
  if (r == hipSuccess && g == gens.end()) {
    r = hipMemset(flag, 0, 2 * sizeof(unsigned));
    g = gens.emplace(dev, 0u).first;
  }
  if (r != hipSuccess) return fail((int)r, std::string("clbm_streams_concurrent: ") + hipGetErrorString(r));
  if (++g->second == 0) ++g->second;   // 0 is the cleared flag
  const unsigned gen = g->second;
  const long long ticks = 50ll * khz;   // 50 ms
  hipLaunchKernelGGL(k_queue_wait, dim3(1), dim3(64), 0, (hipStream_t)main, flag, gen, ticks);
  hipLaunchKernelGGL(k_queue_signal, dim3(1), dim3(64), 0, (hipStream_t)other, flag, gen);
  r = hipGetLastError();
  if (r == hipSuccess) r = hipStreamSynchronize((hipStream_t)main);
  if (r == hipSuccess) r = hipStreamSynchronize((hipStream_t)other);
  unsigned seen = 0;
  if (r == hipSuccess) r = hipMemcpy(&seen, flag + 1, sizeof(unsigned), hipMemcpyDeviceToHost);
  if (r != hipSuccess) return fail((int)r, std::string("clbm_streams_concurrent: ") + hipGetErrorString(r));
  *concurrent = seen == gen;
  return 0;
}

int clbm_side_stream(void* main, void** side, int* concurrent) {
  if (!side) return fail(-1, "clbm_side_stream: NULL argument");
  static std::mutex mu;
  static std::map<int, SideStream> made;
  std::lock_guard<std::mutex> lock(mu);
  int dev = 0;
  hipError_t r = hipGetDevice(&dev);
  if (r != hipSuccess) return fail((int)r, std::string("clbm_side_stream: ") + hipGetErrorString(r));
  auto it = made.find(dev);
  if (it == made.end()) {
    // candidates are kept until the choice is made, so that each new one is
    // placed on another queue than the ones already found sharing
    std::vector<hipStream_t> tried;
    SideStream pick;
    int e = 0;
    for (int t = 0; t < 4 && !e && !pick.concurrent; ++t) {
      hipStream_t s = nullptr;
      r = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
      if (r != hipSuccess) {
        e = fail((int)r, std::string("clbm_side_stream: ") + hipGetErrorString(r));
        break;
      }
      tried.push_back(s);
      int ok = 0;
      e = clbm_streams_concurrent(main, s, &ok);
      if (!e && ok) pick = SideStream{s, 1};
    }
    if (e) {
      for (hipStream_t s : tried) (void)hipStreamDestroy(s);
      return e;
    }
    // no candidate ran beside `main`: keep the first, which still orders right
    if (!pick.stream) pick.stream = tried.front();
    for (hipStream_t s : tried)
      if (s != pick.stream) (void)hipStreamDestroy(s);
    it = made.emplace(dev, pick).first;
  }
  *side = (void*)it->second.stream;
  if (concurrent) *concurrent = it->second.concurrent;
  return 0;
}

const char* clbm_last_error(void) { return g_last_error.c_str(); }

size_t clbm_state_bytes(const clbm_grid_t* g) {
  if (check_grid(g, false)) return 0;
  return (size_t)Q * (size_t)g->nx * g->ny * g->nz * elem_bytes(g->storage);
}

int clbm_state_after_step(int state, const clbm_params_t* p) {
  return ((state & CLBM_STATE_PARITY) ^ CLBM_STATE_PARITY) | ((p && p->inlet) ? CLBM_STATE_PLANES : 0);
}

int clbm_init_feq(const clbm_grid_t* g, void* f, double rho0, double ux0, double uy0, double uz0,
                  void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!f) return fail(-1, "clbm_init_feq: f is NULL");
  return dispatch_storage<InitRun>(g->storage, make_geo(*g), f, rho0, ux0, uy0, uz0,
                                   (hipStream_t)stream);
}

int clbm_step(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state, const float* fx,
              const float* fy, const float* fz, const int32_t* force_box, void* stream) {
  if (!p) return fail(-1, "clbm_step: params is NULL");
  if (int e = check_grid(g, p->inlet != 0)) return e;
  if (!f) return fail(-1, "clbm_step: f is NULL");
  Planes pl;
  if (int e = check_state(g, p, state, pl)) return e;
  if (int e = check_inlet(p, "clbm_step")) return e;
  const Region none{};
  return dispatch_storage<StepRun>(g->storage, make_geo(*g), p, pl, f, state & CLBM_STATE_PARITY,
                                   fx, fy, fz, force_box, none, -1, (hipStream_t)stream);
}

int clbm_step_region(const clbm_grid_t* g, const clbm_params_t* p, const int32_t* region,
                     int32_t* out) {
  if (!p || !region || !out) return fail(-1, "clbm_step_region: NULL argument");
  if (int e = check_grid(g, p->inlet != 0)) return e;
  const int n[3] = {g->nx, g->ny, g->nz};
  for (int a = 0; a < 3; ++a) {
    if (region[a] > region[3 + a]) return fail(-1, "clbm_step_region: region has lo > hi");
    out[a] = std::max(region[a], 0);
    out[3 + a] = std::min(region[3 + a], n[a] - 1);
    if (out[a] > out[3 + a]) return fail(-1, "clbm_step_region: region lies outside the grid");
  }
  // the outlet cell x = nx-1 is stepped by the thread of x = nx-2, and a reader
  // of cell nx-1 reads cell nx-2's populations: keep both in the same part
  if (p->inlet && out[3] >= g->nx - 2) {
    out[0] = std::min(out[0], g->nx - 2);
    out[3] = g->nx - 1;
  }
  return 0;
}

int clbm_step_part(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
                   const float* fx, const float* fy, const float* fz, const int32_t* force_box,
                   const int32_t* region, int part, void* stream) {
  if (!p) return fail(-1, "clbm_step_part: params is NULL");
  if (int e = check_grid(g, p->inlet != 0)) return e;
  if (!f) return fail(-1, "clbm_step_part: f is NULL");
  if (part != 0 && part != 1) return fail(-1, "clbm_step_part: part must be 0 or 1");
  if (p->outside_lds < 0 || p->outside_lds > 64 * 1024)
    return fail(-1, "clbm_step_part: outside_lds must be in [0, 65536]");
  Planes pl;
  if (int e = check_state(g, p, state, pl)) return e;
  if (int e = check_inlet(p, "clbm_step_part")) return e;
  int32_t r[6];
  if (int e = clbm_step_region(g, p, region, r)) return e;
  for (int a = 0; a < 6; ++a)
    if (r[a] != region[a])
      return fail(-1, "clbm_step_part: region is not normalised (pass it through clbm_step_region)");
  return dispatch_storage<StepRun>(g->storage, make_geo(*g), p, pl, f, state & CLBM_STATE_PARITY,
                                   fx, fy, fz, force_box, make_region(r), part,
                                   (hipStream_t)stream);
}

int clbm_step_split(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
                    const int32_t* region, const float* fx, const float* fy,
                    const int32_t* force_box, void* main, void* side, void* join_ev,
                    void* start_ev, void* end_ev, void* done_ev) {
  if (!join_ev) return fail(-1, "clbm_step_split: NULL join event");
  // the events bracket the outside part's kernel itself (timing events)
  t_outside_ev[0] = (hipEvent_t)start_ev;
  t_outside_ev[1] = (hipEvent_t)end_ev;
  const int e0 = clbm_step_part(g, p, f, state, nullptr, nullptr, nullptr, nullptr, region, 0, main);
  t_outside_ev[0] = t_outside_ev[1] = nullptr;
  if (e0) return e0;
  if (int e = clbm_stream_after(main, side, join_ev)) return e;
  t_inside_done = (hipEvent_t)done_ev;
  const int e1 = clbm_step_part(g, p, f, state, fx, fy, nullptr, force_box, region, 1, main);
  t_inside_done = nullptr;
  return e1;
}

int clbm_step_fused(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
                    const int32_t* region, const float* fx, const float* fy, uint64_t* box,
                    uint32_t gen, double wait_ms, void* main, void* start_ev, void* end_ev,
                    void* done_ev) {
  if (!p) return fail(-1, "clbm_step_fused: params is NULL");
  if (int e = check_grid(g, p->inlet != 0)) return e;
  if (!f || !region || !fx || !fy || !box) return fail(-1, "clbm_step_fused: NULL buffer");
  if (gen == 0) return fail(-1, "clbm_step_fused: gen must be >= 1");
  if (!(wait_ms > 0.0) || wait_ms > 60000.0)
    return fail(-1, "clbm_step_fused: wait_ms must be in (0, 60000]");
  if (p->outside_lds < 0 || p->outside_lds > 64 * 1024)
    return fail(-1, "clbm_step_fused: outside_lds must be in [0, 65536]");
  Planes pl;
  if (int e = check_state(g, p, state, pl)) return e;
  if (int e = check_inlet(p, "clbm_step_fused")) return e;
  int32_t r[6];
  if (int e = clbm_step_region(g, p, region, r)) return e;
  for (int a = 0; a < 6; ++a)
    if (r[a] != region[a])
      return fail(-1, "clbm_step_fused: region is not normalised (pass it through clbm_step_region)");
  int dev = 0, khz = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err == hipSuccess) err = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  if (err == hipSuccess && khz <= 0) err = hipErrorInvalidValue;
  if (err != hipSuccess) return fail((int)err, std::string("clbm_step_fused: ") + hipGetErrorString(err));
  const FusedWait w{(const unsigned long long*)IBM_SPLIT_READY(box),
                    (unsigned long long*)IBM_SPLIT_EXPIRED(box), (unsigned long long)gen,
                    (long long)(wait_ms * (double)khz)};
  // the launch carries at most one end event: the timing one when asked for,
  // then done_ev is recorded after it on the stream
  hipEvent_t e1 = (hipEvent_t)(end_ev ? end_ev : done_ev);
  if (int e = dispatch_storage<FusedRun>(g->storage, make_geo(*g), p, pl, f, state & CLBM_STATE_PARITY,
                                         fx, fy, IBM_FORCE_BOX(box), make_region(r), w,
                                         (hipEvent_t)start_ev, e1, (hipStream_t)main))
    return e;
  if (end_ev && done_ev) {
    err = hipEventRecord((hipEvent_t)done_ev, (hipStream_t)main);
    if (err != hipSuccess) return fail((int)err, std::string("clbm_step_fused: ") + hipGetErrorString(err));
  }
  return 0;
}

int clbm_macro(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state, float* rho,
               float* ux, float* uy, float* uz, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!f) return fail(-1, "clbm_macro: f is NULL");
  Planes pl;
  if (int e = check_state(g, p, state, pl)) return e;
  return dispatch_storage<MacroRun>(g->storage, make_geo(*g), pl, f, state & CLBM_STATE_PARITY, rho,
                                    ux, uy, uz, (hipStream_t)stream);
}

int clbm_export_state(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state,
                      float* out, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!f || !out) return fail(-1, "clbm_export_state: NULL buffer");
  Planes pl;
  if (int e = check_state(g, p, state, pl)) return e;
  return dispatch_storage<ExportRun>(g->storage, make_geo(*g), pl, f, state & CLBM_STATE_PARITY, out,
                                     (hipStream_t)stream);
}

int clbm_export_state_f64(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state,
                          double* out, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!f || !out) return fail(-1, "clbm_export_state_f64: NULL buffer");
  Planes pl;
  if (int e = check_state(g, p, state, pl)) return e;
  return dispatch_storage<ExportRun>(g->storage, make_geo(*g), pl, f, state & CLBM_STATE_PARITY, out,
                                     (hipStream_t)stream);
}

int clbm_import_state(const clbm_grid_t* g, void* f, const float* in, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!f || !in) return fail(-1, "clbm_import_state: NULL buffer");
  return dispatch_storage<ImportRun>(g->storage, make_geo(*g), f, in, (hipStream_t)stream);
}

int clbm_import_state_f64(const clbm_grid_t* g, void* f, const double* in, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!f || !in) return fail(-1, "clbm_import_state_f64: NULL buffer");
  return dispatch_storage<ImportRun>(g->storage, make_geo(*g), f, in, (hipStream_t)stream);
}

int clbm_stream_probe(const clbm_grid_t* g, void* f, int parity, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!f) return fail(-1, "clbm_stream_probe: f is NULL");
  if (parity != 0 && parity != 1) return fail(-1, "clbm_stream_probe: parity must be 0 or 1");
  return dispatch_storage<ProbeRun>(g->storage, make_geo(*g), f, parity, (hipStream_t)stream);
}

int clbm_health(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state,
                float* out2, void* stream) {
  if (int e = check_grid(g, false)) return e;
  if (!f || !out2) return fail(-1, "clbm_health: NULL buffer");
  Planes pl;
  if (int e = check_state(g, p, state, pl)) return e;
  return dispatch_storage<HealthRun>(g->storage, make_geo(*g), pl, f, state & CLBM_STATE_PARITY,
                                     out2, (hipStream_t)stream);
}

}  // extern "C"
