/*
 * clbm_amd.h -- C ABI of the MI355X CLBM-IBM time step (libclbm_amd.so).
 *
 * Plain pointers and sizes only.  Every pointer argument named as a device
 * buffer is device memory the caller owns (in the Python layer: torch-ROCm
 * tensors, passed as tensor.data_ptr()); `stream` is a hipStream_t or NULL
 * for the null stream.  The library allocates no long-lived device memory.
 *
 * Return codes: 0 = ok, < 0 = argument / configuration error, > 0 = a
 * hipError_t from a launch.  The message of the last failure on the calling
 * thread is returned by clbm_last_error().  All launches are asynchronous on
 * `stream`; no entry point synchronises.
 *
 * Memory layout (the library's own, not the reference's):
 *   populations  structure of arrays, f[q][z][y][x], x fastest, 19 planes of
 *                Nx*Ny*Nz elements of the storage type; updated in place
 *                (AA pattern).  `parity` = number of steps taken mod 2:
 *                parity 0: post-stream population q of cell c sits at f[q][c];
 *                parity 1: it sits at f[opp(q)][c - e_q] (periodic index).
 *                16-bit storage holds f_q - w_q (deviation from the lattice weight).
 *   scalar fields  [z][y][x] float32, x fastest.
 *   force box    int32[IBM_BOX_INTS]: [0..5] = inclusive box (x0,y0,z0,x1,y1,z1)
 *                of the cells the current markers' stencils reach, [6..11] = the
 *                box before it, [13] = a counter that must start at 0, the rest
 *                scratch of ibm_update_markers.  Initialise [0..11] to empty
 *                boxes (lower corner INT32_MAX, upper corner INT32_MIN).
 *
 * Reference interfaces are cited as file:line of kioxli/CLBM_IBM-moving-airfoil.
 */
#ifndef CLBM_AMD_H
#define CLBM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLBM_ABI_VERSION 1

/* population storage formats */
#define CLBM_STORE_F32 0  /* float32 populations, float32 arithmetic (the drop-in) */
#define CLBM_STORE_F16 1  /* IEEE half deviations f - w, float32 arithmetic */
#define CLBM_STORE_BF16 2 /* bfloat16 deviations f - w, float32 arithmetic */
#define CLBM_STORE_F64 3  /* float64 populations, float64 arithmetic (error reference) */

/* force spreading modes */
#define IBM_SPREAD_FIXED 0 /* order-independent 64-bit fixed-point sums (deterministic) */
#define IBM_SPREAD_ATOMIC_F32 1 /* float32 atomics, as ti.atomic_add (order-dependent) */

#define IBM_BOX_HEADER 16
#define IBM_BOX_PARTIALS 64
#define IBM_BOX_INTS (IBM_BOX_HEADER + 6 * IBM_BOX_PARTIALS)

typedef struct {
  int32_t nx, ny, nz; /* lattice size, Nx >= 3 */
  int32_t storage;    /* CLBM_STORE_* */
} clbm_grid_t;

typedef struct {
  double s_v;          /* shear rate 1/tau (config_moving.py:40) */
  double s_e;          /* bulk rate (config_moving.py:41) */
  double ux_in, uy_in; /* inlet velocity _UX_IN/_UY_IN (clbm_moving.py:33-35); rates and
                          velocities are rounded to the arithmetic type of the storage format */
  int32_t inlet;      /* 1: inlet rule at x=0 and outlet copy at x=Nx-1 after every
                         stream (clbm_moving.py:266-300); 0: x stays periodic */
  int32_t reserved;
} clbm_params_t;

typedef struct {
  float f_max;               /* marker and grid force clamp, 0.05 in the reference
                                (airfoil_moving.py:178,221) */
  int32_t wrap_weight_quirk; /* 1: delta weights use the wrapped/clamped index, as
                                airfoil_moving.py:159-165 does; 0: unwrapped offsets */
  int32_t spread_mode;       /* IBM_SPREAD_* */
  int32_t fixed_shift;       /* 0 = automatic; else scale 2^shift of the fixed-point sums */
} ibm_params_t;

/* ---- library ---------------------------------------------------------- */
int clbm_abi_version(void);
const char* clbm_last_error(void);
/* Bytes of one population buffer for `g` (19 * Nx*Ny*Nz * element size). */
size_t clbm_state_bytes(const clbm_grid_t* g);

/* ---- lattice Boltzmann (replaces CLBMSolver, clbm_moving.py:41-343) ----- */

/* Replaces CLBMSolver._init_feq (clbm_moving.py:75-91): second-order
 * equilibrium at (rho0, u0) in every cell; leaves a parity-0 state. */
int clbm_init_feq(const clbm_grid_t* g, void* f, double rho0, double ux0, double uy0, double uz0,
                  void* stream);

/* Replaces CLBMSolver.step (clbm_moving.py:326-343): K2 _macro_with_force,
 * K3 _collide, K4 _stream in one fused pass, then K5 _bc_inlet and
 * K6 _bc_outlet on the two x planes.  fx/fy/fz are [z][y][x] float32 force
 * fields (each nullable = zero).  force_box (nullable) restricts the force
 * reads to the box in its first six ints; cells outside read zero.
 * `parity` is the state's parity before the step; afterwards it is 1-parity. */
int clbm_step(const clbm_grid_t* g, const clbm_params_t* p, void* f, int parity,
              const float* fx, const float* fy, const float* fz, const int32_t* force_box,
              void* stream);

/* Replaces CLBMSolver._macro_post (clbm_moving.py:305-321): force-free rho
 * and u of the current state into [z][y][x] float32 fields (each nullable). */
int clbm_macro(const clbm_grid_t* g, const void* f, int parity, float* rho, float* ux, float* uy,
               float* uz, void* stream);

/* Replaces reads of CLBMSolver.f (clbm_moving.py:56): the post-stream state as
 * float32 planes out[q][z][y][x] in the reference's q order. */
int clbm_export_state(const clbm_grid_t* g, const void* f, int parity, float* out, void* stream);
/* The same in float64 (for the float64 storage format). */
int clbm_export_state_f64(const clbm_grid_t* g, const void* f, int parity, double* out,
                          void* stream);

/* Replaces writes of CLBMSolver.f (ti.field.from_numpy): loads float32 planes
 * in[q][z][y][x]; leaves a parity-0 state. */
int clbm_import_state(const clbm_grid_t* g, void* f, const float* in, void* stream);
/* The same from float64 planes. */
int clbm_import_state_f64(const clbm_grid_t* g, void* f, const double* in, void* stream);

/* Replaces the driver's divergence guard (main_moving.py:130-145): out2[0] =
 * min rho over all cells, out2[1] = 1.0 if every population is finite else 0.0. */
int clbm_health(const clbm_grid_t* g, const void* f, int parity, float* out2, void* stream);

/* ---- immersed boundary (replaces DroneMovingIBM, airfoil_moving.py:44-255) */

/* Replaces DroneMovingIBM._update_markers (airfoil_moving.py:100-127).  cos
 * and sin of theta are evaluated in float64 and rounded once to float32; the
 * rotation is evaluated in float32 without contraction.  box (nullable)
 * receives the force box of the new positions (box[6..11] keeps the old one). */
int ibm_update_markers(const clbm_grid_t* g, int n, const float* X0, const float* Y0,
                       const float* Z0, float xc, float yc, float theta, float vx, float vy,
                       float omega, float* X, float* Y, float* Z, float* VX, float* VY,
                       int32_t* box, void* stream);

/* Part of DroneMovingIBM._interpolate's input (the K7 velocity of
 * clbm_moving.py:305-321 that _interpolate reads): the force-free velocity
 * of the cells in box[0..5] only, computed from the population state f at
 * `parity` into [z][y][x] float32 fields ux/uy (other cells untouched). */
int ibm_band_velocity(const clbm_grid_t* g, const void* f, int parity, const int32_t* box,
                      float* ux, float* uy, void* stream);

/* Replaces DroneMovingIBM._interpolate + _marker_forces (airfoil_moving.py:
 * 142-183).  Velocity is read from ux/uy ([z][y][x] float32) when both are
 * given, else computed force-free from the population state f at `parity`.
 * stencil_base (nullable) receives floor(pos)-1 per axis, int32 n x 3. */
int ibm_interp_force(const clbm_grid_t* g, const ibm_params_t* ip, const void* f, int parity,
                     const float* ux, const float* uy, int n, const float* X, const float* Y,
                     const float* Z, const float* VX, const float* VY, float* U, float* V,
                     float* FX, float* FY, int32_t* stencil_base, void* stream);

/* Replaces DroneMovingIBM._spread (airfoil_moving.py:188-214): adds F_k * w
 * into the accumulators acc_x/acc_y (Nx*Ny*Nz 8-byte slots each, zero before
 * the first spread; IBM_SPREAD_ATOMIC_F32 uses the first 4 bytes of each slot
 * as a float). */
int ibm_spread(const clbm_grid_t* g, const ibm_params_t* ip, int n, const float* X, const float* Y,
               const float* Z, const float* FX, const float* FY, void* acc_x, void* acc_y,
               void* stream);

/* K10 + K11 + K12 in one launch: ibm_interp_force followed by ibm_spread
 * with the same stencil weights. */
int ibm_interp_spread(const clbm_grid_t* g, const ibm_params_t* ip, const void* f, int parity,
                      const float* ux, const float* uy, int n, const float* X, const float* Y,
                      const float* Z, const float* VX, const float* VY, float* U, float* V,
                      float* FX, float* FY, int32_t* stencil_base, void* acc_x, void* acc_y,
                      void* stream);

/* Replaces DroneMovingIBM._zero_forces + _clip_forces (airfoil_moving.py:
 * 132-137,219-225): over the union of box[0..5] and box[6..11], converts the
 * accumulators into the clamped float32 fields fx/fy ([z][y][x]) and clears
 * the accumulators.  fx/fy are exact everywhere: outside both boxes they are
 * already zero. */
int ibm_clip_forces(const clbm_grid_t* g, const ibm_params_t* ip, void* acc_x, void* acc_y,
                    float* fx, float* fy, const int32_t* box, void* stream);

/* No reference counterpart (the reference never reduces forces): out2[0..1]
 * = sum_k FX_k dV_k, sum_k FY_k dV_k in float64 with a fixed summation order
 * (dV nullable = 1). */
int ibm_force_sum_f64(int n, const float* FX, const float* FY, const float* dV, double* out2,
                      void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CLBM_AMD_H */
