// The coupled step: DroneMovingIBM.compute_force followed by CLBMSolver.step
// (the loop body of main_moving.py:116-122) in one launch plus a short tail,
// on one stream, for gfx950 (MI355X).
//
// The immersed-boundary kernels read and spread only inside a box around the
// body (the region), and the AA update of a cell touches no population slot
// of any other cell.  So the lattice away from the body (the outside part) can
// be updated while the marker kernels run, and only the region must wait for
// their forces.  Here that concurrency lives inside one grid:
//
//  * marker workgroups: the fused marker group (K8 + K10 + K11 + K12, four
//    markers per workgroup, velocity from the populations, fixed-point
//    spreading into the accumulators, the force box of this generation);
//    they are spread over the first part of the grid, one super-block of
//    eight (one per XCD) every `interleave` super-blocks, so that their
//    latency-bound work runs beside bandwidth-bound lattice workgroups;
//  * outside workgroups: the fused step of every cell outside the region,
//    force-free, in the sweep order of the plain step (the block-to-tile map
//    keeps lattice tile b on XCD b % 8);
//  * inside workgroups (optional, after all others): the fused step of the
//    region's cells with the force read straight from the accumulators,
//    clamped as ibm_clip_forces would write it (K9 + K13 fold into the step).
//    They wait for the marker workgroups: a bounded wait on per-XCD done
//    counters.  The first inside workgroup to finish waiting decides, for all
//    of them (one generation-tagged word), whether the markers were done
//    (then every inside workgroup steps its cells) or the wait ran out (then
//    none does).
//
// The tail kernel steps the region when the inside workgroups did not
// (always, when the launch has none), and resets this generation's done
// counters.  A wait can run out only if the hardware left some marker
// workgroup undispatched while inside workgroups held the CUs; every
// marker workgroup precedes every inside workgroup in the grid, so with the
// in-order dispatch of each XCD this does not happen, and if it did the tail
// would still give the same result.  Results equal ibm_compute_force +
// ibm_clip_forces + clbm_step, bit for bit.
#include "clbm_markers.h"
#include "clbm_step.h"

#include <algorithm>
#include <cmath>

namespace clbm {

constexpr int CG = 4;  // markers per marker workgroup (4 wavefronts = BLK threads)
static_assert(64 * CG == BLK, "a marker workgroup is a lattice workgroup");
constexpr unsigned long long DECIDE_GO = 1, DECIDE_SKIP = 2;

struct CoupledLayout {
  int nmsb;        // marker super-blocks (8 marker workgroups each, one per XCD)
  int interleave;  // one marker super-block every `interleave` super-blocks, from the start
  int nl;          // outside lattice workgroups (one per 256 cells of the grid)
  int nlsb;        // their super-blocks
  int ni;          // inside workgroups (one per 256 cells of the region)
  int nisb;        // their super-blocks in the main launch (0: all in the tail kernel)
  int spread;      // one inside super-block every `spread` super-blocks, at the end
  int iwin;        // first super-block of the inside window (total - spread * nisb)
};

struct CoupledArgs {
  StepArgs st;      // the lattice (region rg: the cells the markers read)
  IbmArgs ib;       // the body
  Kin kin;
  uint32_t gen;
  CoupledLayout L;
  long long *accx, *accy;
  float f_max;
  double inv_scale;
  int wait_spins;   // polls of the done counters before giving up (< 0: give up at once)
  unsigned long long* decision;  // (gen << 2) | DECIDE_* once an inside workgroup decided
  uint32_t* done;                // 8 counters (one per XCD) of this generation's parity
};

// The inside workgroups' common decision (thread 0 of a workgroup; true:
// step the region here).
__device__ bool decide(const CoupledArgs& a) {
  const unsigned long long tag = (unsigned long long)a.gen << 2;
  unsigned long long d = __hip_atomic_load(a.decision, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if ((d >> 2) != a.gen) {
    bool done = false, decided = false;
    const uint32_t target = 8u * (uint32_t)a.L.nmsb;
    for (int it = 0; it < a.wait_spins && !done && !decided; ++it) {
      uint32_t sum = 0;
      for (int x = 0; x < 8; ++x)
        sum += __hip_atomic_load(a.done + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      done = sum >= target;
      if (!done) {
        __builtin_amdgcn_s_sleep(4);
        const unsigned long long e =
            __hip_atomic_load(a.decision, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        decided = (e >> 2) == a.gen;  // another workgroup decided meanwhile
        if (decided) d = e;
      }
    }
    if (!decided) {
      const unsigned long long mine = tag | (done ? DECIDE_GO : DECIDE_SKIP);
      const unsigned long long prev = atomicCAS(a.decision, d, mine);
      d = prev == d ? mine : prev;  // another workgroup's decision of this generation wins
    }
  }
  // the accumulators and the box record are then read with device-scope
  // atomic loads (AccForce<true>, box_get_block<true>): no acquire fence
  return (d & 3ull) == DECIDE_GO;
}

// The region's cells of inside workgroup `ib`, force from the accumulators.
template <int S, int PAR, bool COHERENT>
__device__ __forceinline__ void inside_cells(const CoupledArgs& a, int ib) {
  AccForce<COHERENT> F;
  F.accx = a.accx;
  F.accy = a.accy;
  F.f_max = a.f_max;
  F.inv_scale = a.inv_scale;
  if (!box_get_block<COHERENT>(a.ib.box, a.gen, F.lo, F.hi))
    for (int i = 0; i < 3; ++i) {
      F.lo[i] = 1;
      F.hi[i] = 0;
    }
  step_cell<S, PAR, 2>(a.st, ib * BLK + threadIdx.x, F);
}

template <int S, int PAR>
__global__ __launch_bounds__(BLK) void k_coupled(CoupledArgs a) {
  const int b = blockIdx.x;
  const int s = b >> 3, x = b & 7;
  const CoupledLayout& L = a.L;
  const int k = L.interleave, span = k * L.nmsb;
  // super-block s: a marker one at 0, k, 2k, ... below span; an inside one at
  // every spread-th position of the window from iwin on (the last of each
  // run of `spread`); a lattice one elsewhere, in order
  int sl;
  if (s < span) {
    sl = s % k == 0 ? -1 : s - s / k - 1;
  } else if (s < L.iwin) {
    sl = s - L.nmsb;
  } else {
    const int r = s - L.iwin;
    if (r % L.spread == L.spread - 1) {
      __shared__ int go;
      if (threadIdx.x == 0) go = decide(a) ? 1 : 0;
      __syncthreads();
      if (go) inside_cells<S, PAR, true>(a, (r / L.spread) * 8 + x);
      return;
    }
    sl = s - L.nmsb - r / L.spread;
  }
  if (sl < 0) {
    // marker workgroup: group x * nmsb + j runs on XCD x, so each XCD takes a
    // run of neighbouring groups (their tiles overlap in its L2)
    compute_group<S, 1 + PAR, IBM_SPREAD_FIXED, CG>(a.ib, a.kin, a.gen, x * L.nmsb + s / k, b);
    // Count this workgroup done once every wave's atomics on the accumulators
    // and the box record are acknowledged: device-scope atomics are performed
    // where every XCD sees them, and the inside workgroups read those words
    // with device-scope atomic loads, so no cache maintenance is needed (a
    // release fence here would write back this XCD's L2 under the lattice
    // sweep, once per marker workgroup).
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(a.done + x, 1u);
    return;
  }
  const int lb = sl * 8 + x;
  if (lb < L.nl)
    step_cell<S, PAR, 1>(a.st, sweep_block<PAR>(lb, L.nl) * BLK + threadIdx.x,
                         FieldForce<false, false>{nullptr, nullptr, nullptr, nullptr});
}

// The tail: resets this generation's done counters and publishes the force
// box (workgroup 0), then steps the region unless the inside workgroups did.
template <int S, int PAR>
__global__ __launch_bounds__(BLK) void k_coupled_tail(CoupledArgs a, int inside_in_main) {
  if (blockIdx.x == 0) {
    int lo[3], hi[3];
    const bool any = box_get_block<false>(a.ib.box, a.gen, lo, hi);
    if (threadIdx.x == 0) {
      int32_t* out = IBM_FORCE_BOX(a.ib.box);
      for (int i = 0; i < 3; ++i) {
        out[i] = any ? lo[i] : INT_MAX;
        out[3 + i] = any ? hi[i] : INT_MIN;
      }
    }
    if (threadIdx.x < 8) a.done[threadIdx.x] = 0u;
  }
  if (inside_in_main) {
    const unsigned long long d = *a.decision;
    if ((d >> 2) == a.gen && (d & 3ull) == DECIDE_GO) return;
  }
  inside_cells<S, PAR, false>(a, blockIdx.x);
}

template <int S> struct CoupledRun {
  template <int PAR>
  static void go(const CoupledArgs& a, int nmain, int inside_in_main, hipStream_t st) {
    hipLaunchKernelGGL((k_coupled<S, PAR>), dim3(nmain), dim3(BLK), 0, st, a);
    hipLaunchKernelGGL((k_coupled_tail<S, PAR>), dim3(a.L.ni), dim3(BLK), 0, st, a,
                       inside_in_main);
  }
  static int run(const CoupledArgs& a, int par, int nmain, int inside_in_main, hipStream_t st) {
    if (par)
      go<1>(a, nmain, inside_in_main, st);
    else
      go<0>(a, nmain, inside_in_main, st);
    return check_launch("clbm_ibm_step");
  }
};

}  // namespace clbm

using namespace clbm;

extern "C" {

int clbm_ibm_step(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
                  const ibm_params_t* ip, int n, const int32_t* order, const float* X0,
                  const float* Y0, const float* Z0, float xc, float yc, float theta, float vx,
                  float vy, float omega, float* X, float* Y, float* Z, float* VX, float* VY,
                  float* U, float* V, float* FX, float* FY, int32_t* stencil_base, void* acc_x,
                  void* acc_y, uint64_t* box, uint32_t gen, const int32_t* region,
                  int interleave, int inside_spread, int wait_spins, void* stream) {
  if (!p) return fail(-1, "clbm_ibm_step: params is NULL");
  if (int e = check_grid(g, p->inlet != 0)) return e;
  if (int e = check_inlet(p, "clbm_ibm_step")) return e;
  if (int e = check_ibm(ip)) return e;
  if (ip->spread_mode != IBM_SPREAD_FIXED)
    return fail(-1, "clbm_ibm_step: needs the fixed-point spread mode");
  if (!f) return fail(-1, "clbm_ibm_step: f is NULL");
  if (n <= 0) return fail(-1, "clbm_ibm_step: n must be > 0");
  if (!X0 || !Y0 || !Z0 || !X || !Y || !Z || !VX || !VY || !U || !V || !FX || !FY)
    return fail(-1, "clbm_ibm_step: NULL marker array");
  if (!acc_x || !acc_y || !box) return fail(-1, "clbm_ibm_step: NULL accumulator or box record");
  if (gen == 0) return fail(-1, "clbm_ibm_step: generation must be >= 1");
  if (!region) return fail(-1, "clbm_ibm_step: region is NULL");
  if (interleave < 1 || inside_spread < 1)
    return fail(-1, "clbm_ibm_step: interleave and inside_spread must be >= 1");
  Planes pl;
  if (int e = check_state(g, p, state, pl)) return e;
  int32_t r[6];
  if (int e = clbm_step_region(g, p, region, r)) return e;
  for (int i = 0; i < 6; ++i)
    if (r[i] != region[i])
      return fail(-1, "clbm_ibm_step: region is not normalised (pass it through clbm_step_region)");
  const Geo geo = make_geo(*g);
  CoupledArgs a{};
  a.st = StepArgs{geo, p->s_v, p->s_e, pl, f, nullptr, nullptr, nullptr, nullptr, make_region(r)};
  IbmArgs& b = a.ib;
  b.g = geo;
  b.pl = pl;
  b.quirk = ip->wrap_weight_quirk ? 1 : 0;
  b.f_max = ip->f_max;
  b.fscale = (float)std::ldexp(1.0, fixed_shift(ip));
  b.f = f;
  b.n = n;
  b.order = order;
  b.X0 = X0;
  b.Y0 = Y0;
  b.Z0 = Z0;
  b.X = X;
  b.Y = Y;
  b.Z = Z;
  b.VX = VX;
  b.VY = VY;
  b.U = U;
  b.V = V;
  b.FX = FX;
  b.FY = FY;
  b.base = stencil_base;
  b.accx = acc_x;
  b.accy = acc_y;
  b.box = box;
  b.rc.on = 1;  // the region is what the outside workgroups leave alone
  for (int i = 0; i < 3; ++i) {
    b.rc.lo[i] = r[i];
    b.rc.hi[i] = r[3 + i];
  }
  a.kin = make_kin(xc, yc, theta, vx, vy, omega);
  a.gen = gen;
  a.accx = (long long*)acc_x;
  a.accy = (long long*)acc_y;
  a.f_max = ip->f_max;
  a.inv_scale = std::ldexp(1.0, -fixed_shift(ip));
  a.wait_spins = wait_spins < 0 ? 0 : wait_spins;
  a.decision = (unsigned long long*)IBM_STEP_DECISION(box);
  a.done = IBM_STEP_DONE(box) + 8 * (gen & 1u);
  CoupledLayout& L = a.L;
  const long long groups = (n + CG - 1) / CG;
  L.nmsb = (int)((groups + 7) / 8);
  L.nl = (geo.ncell + BLK - 1) / BLK;
  L.nlsb = (L.nl + 7) / 8;
  L.ni = (a.st.rg.vol + BLK - 1) / BLK;
  const int inside_in_main = wait_spins != 0 ? 1 : 0;
  L.nisb = inside_in_main ? (L.ni + 7) / 8 : 0;
  // the inside window takes at most the lattice super-blocks after the markers'
  // span; both spacings shrink to fit (markers first, inside last, never mixed)
  L.interleave = std::min(interleave, 1 + L.nlsb / L.nmsb);
  const int after = L.nlsb - (L.interleave - 1) * L.nmsb;  // lattice super-blocks past the span
  L.spread = L.nisb ? std::max(1, std::min(inside_spread, 1 + after / L.nisb)) : 1;
  const int total = L.nmsb + L.nlsb + L.nisb;
  L.iwin = total - L.spread * L.nisb;
  const long long nmain = 8ll * total;
  if (nmain >= (1ll << 31)) return fail(-1, "clbm_ibm_step: too many workgroups");
  const int par = state & CLBM_STATE_PARITY;
  hipStream_t st = (hipStream_t)stream;
  switch (g->storage) {
    case CLBM_STORE_F32: return CoupledRun<CLBM_STORE_F32>::run(a, par, (int)nmain, inside_in_main, st);
    case CLBM_STORE_F16: return CoupledRun<CLBM_STORE_F16>::run(a, par, (int)nmain, inside_in_main, st);
    case CLBM_STORE_BF16: return CoupledRun<CLBM_STORE_BF16>::run(a, par, (int)nmain, inside_in_main, st);
    case CLBM_STORE_F64: return CoupledRun<CLBM_STORE_F64>::run(a, par, (int)nmain, inside_in_main, st);
    case CLBM_STORE_F16C: return CoupledRun<CLBM_STORE_F16C>::run(a, par, (int)nmain, inside_in_main, st);
  }
  return fail(-1, "unknown storage format");
}

}  // extern "C"
