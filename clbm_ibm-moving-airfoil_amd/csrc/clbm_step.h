// Pieces of the fused step shared by the plain, batched, coupled and
// learned-closure kernels.
#pragma once

#include "clbm_collide.h"

namespace clbm {

constexpr int BLK = 256;  // threads per workgroup of the lattice kernels

template <typename R, bool FORCE, bool BOX>
__device__ __forceinline__ void force_at(int c, int x, int y, int z, const float* __restrict__ fx,
                                         const float* __restrict__ fy,
                                         const float* __restrict__ fz,
                                         const int32_t* __restrict__ box, R& Fx, R& Fy, R& Fz) {
  Fx = Fy = Fz = R(0);
  if (!FORCE) return;
  if (BOX) {
    const bool in = (x >= box[0]) & (y >= box[1]) & (z >= box[2]) & (x <= box[3]) &
                    (y <= box[4]) & (z <= box[5]);
    if (!in) return;
  }
  if (fx) Fx = R(fx[c]);
  if (fy) Fy = R(fy[c]);
  if (fz) Fz = R(fz[c]);
}

// Force sources of step_cell.  FieldForce reads float32 force fields, as
// CLBMSolver.step receives them (F: any field given; B: only inside the box).
template <bool F, bool B> struct FieldForce {
  static constexpr bool present = F;
  const float *fx, *fy, *fz;
  const int32_t* box;
  template <typename R>
  __device__ __forceinline__ void operator()(int c, int x, int y, int z, R& Fx, R& Fy,
                                             R& Fz) const {
    force_at<R, F, B>(c, x, y, z, fx, fy, fz, box, Fx, Fy, Fz);
  }
};

// AccForce reads the immersed boundary's fixed-point spread accumulators
// inside the force box of one generation, converts and clamps them exactly as
// ibm_clip_forces does (K13 _clip_forces, airfoil_moving.py:219-225: the
// same float32 values ibm_clip_forces would write into fx, fy), and clears
// them for the next generation (K9 _zero_forces, airfoil_moving.py:132-137),
// so that the clamp needs no launch of its own.  A cell's force is read by
// exactly one thread of a step (its own, or the x = nx-2 thread for the
// outlet cell), so the clear races nobody.  COHERENT loads the accumulators
// with device-scope atomic loads: within the launch whose marker workgroups
// added into them (from any XCD), past the caches of this one.
template <bool COHERENT> struct AccForce {
  static constexpr bool present = true;
  long long *accx, *accy;
  int lo[3], hi[3];  // inclusive box; empty when lo > hi
  float f_max;
  double inv_scale;
  __device__ __forceinline__ long long load(const long long* p) const {
    if (COHERENT) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
  }
  template <typename R>
  __device__ __forceinline__ void operator()(int c, int x, int y, int z, R& Fx, R& Fy,
                                             R& Fz) const {
    Fx = Fy = Fz = R(0);
    const bool in = (x >= lo[0]) & (y >= lo[1]) & (z >= lo[2]) & (x <= hi[0]) & (y <= hi[1]) &
                    (z <= hi[2]);
    if (!in) return;
    const float vx = (float)((double)load(accx + c) * inv_scale);
    const float vy = (float)((double)load(accy + c) * inv_scale);
    accx[c] = 0;
    accy[c] = 0;
    Fx = R(fmaxf(-f_max, fminf(f_max, vx)));
    Fy = R(fmaxf(-f_max, fminf(f_max, vy)));
    // opaque to the compiler, like the force read from memory by FieldForce
    // (Fz = 0 and the clamp's range known at compile time would let it fold
    // the collision's force terms into another floating-point evaluation
    // order than the step with force fields)
    asm("" : "+v"(Fx), "+v"(Fy), "+v"(Fz));
  }
};

// A relaxation rate in both precisions, rounded once on the host, so that a
// float32 step reads its rate instead of converting it in every thread.
struct Rate {
  double d;
  float f;
  __host__ __device__ Rate(double v = 0.0) : d(v), f((float)v) {}
  template <typename R> __device__ __forceinline__ R as() const {
    if constexpr (sizeof(R) == sizeof(float)) return f; else return R(d);
  }
};

// Everything one launch of the fused step needs for one lattice.
struct StepArgs {
  Geo g;
  Rate s_v, s_e;
  Planes pl;
  void* f;
  const float *fx, *fy, *fz;
  const int32_t* box;
  Region rg;
};

// One thread of the fused step (boundary planes + macro + collide + Guo +
// stream, CLBMSolver.step, clbm_moving.py:326-343): t is the thread's index
// in the launch over this lattice.  REG selects the cells: 0 every cell; 1
// the cells outside region rg (one thread per cell of the grid); 2 the cells
// inside it (one thread per cell of the region).  `force` gives the body
// force of a cell (a FieldForce or an AccForce).
template <int S, int PAR, int REG, class Force>
__device__ __forceinline__ void step_cell(const StepArgs& a, int t, const Force& force) {
  using R = typename Store<S>::R;
  constexpr bool FORCE = Force::present;
  const Geo& g = a.g;
  const Region& rg = a.rg;
  const Planes& pl = a.pl;
  int c, x, y, z;
  if (REG == 2) {
    if (t >= rg.vol) return;
    const uint32_t row = fdiv((uint32_t)t, rg.drx);
    const int iz = (int)fdiv(row, rg.dry);
    x = rg.lo[0] + (t - (int)row * rg.rx);
    y = rg.lo[1] + ((int)row - iz * rg.ry);
    z = rg.lo[2] + iz;
    c = (z * g.ny + y) * g.nx + x;
  } else {
    c = t;
    if (c >= g.ncell) return;
    cell_xyz(g, c, x, y, z);
    if (REG == 1 && in_region(rg, x, y, z)) return;
  }
  const Pop<S> pop((typename Store<S>::T*)a.f, g);
  // With pending planes, cell x = nx-1 collides the outlet copy of cell nx-2's
  // post-stream populations, but those slots are overwritten by the thread of
  // x = nx-2 during this pass, so that thread handles both cells (cell nx-1's
  // own thread stands aside).  When the two cells see the same force (always,
  // unless the body reaches the outlet), their post-collision populations are
  // equal and the thread stores its result a second time, into cell nx-1's
  // slots.  Otherwise it first writes what it loaded into cell nx-1's slots
  // (the outlet copy; no other thread touches them), and after its own cell
  // reads them back and collides them with cell nx-1's force.
  if (pl.on && x == g.nx - 1) return;
  R fq[Q], Fx, Fy, Fz, Gx, Gy, Gz;
  const Nbr nb = make_nbr(g, x, y, z);
  pop_load<PAR>(pop, c, nb, fq);
  force(c, x, y, z, Fx, Fy, Fz);
  const bool outlet = pl.on && x == g.nx - 2;
  // The stores' cell index and the outlet cell's are kept opaque: a thread
  // stores into the slots it loaded from (in both AA parities), and left to
  // itself the compiler forms full 64-bit addresses once to share them
  // between the loads and the stores (and the outlet cell's stores), at a
  // vector add per population instead of the scalar-base form.
  int c1 = c + 1;
  asm("" : "+v"(c1));
  bool twin = true;
  if (outlet) {
    force(c1, x + 1, y, z, Gx, Gy, Gz);
    twin = (Gx == Fx) & (Gy == Fy) & (Gz == Fz);
    if (!twin) pop_store_in<PAR>(pop, c1, make_nbr(g, x + 1, y, z), fq);
  }
  if (pl.on && x == 0) inlet_rule<R>(fq, R(pl.ux_in), R(pl.uy_in), pl.rule);
  collide_cell<R, FORCE>(fq, Fx, Fy, Fz, a.s_v.template as<R>(), a.s_e.template as<R>());
  pop_store<PAR>(pop, c, nb, fq);
  if (outlet) {
    const Nbr nb1 = make_nbr(g, x + 1, y, z);
    if (!twin) {
      pop_load<PAR>(pop, c1, nb1, fq);
      collide_cell<R, FORCE>(fq, Gx, Gy, Gz, a.s_v.template as<R>(), a.s_e.template as<R>());
    }
    pop_store<PAR, true>(pop, c1, nb1, fq);
  }
}

}  // namespace clbm
