/*
 * clbm_amd.h -- C ABI of the MI355X CLBM-IBM time step (libclbm_amd.so).
 *
 * Plain pointers and sizes only.  Every pointer argument named as a device
 * buffer is device memory the caller owns (in the Python layer: torch-ROCm
 * tensors, passed as tensor.data_ptr()); `stream` is a hipStream_t or NULL
 * for the null stream.  The library allocates no device memory; it holds two
 * fixed arrays per GPU, which come with its code object: the 7.1 MiB
 * partial-gradient rows of closure_grad_noalloc and the two-word flag of
 * clbm_side_stream's queue check.
 *
 * Return codes: 0 = ok, < 0 = argument / configuration error, > 0 = a
 * hipError_t from a launch.  The message of the last failure on the calling
 * thread is returned by clbm_last_error().  All launches are asynchronous on
 * `stream`; no entry point synchronises, except clbm_side_stream on its
 * first call per device.
 *
 * Memory layout (the library's own, not the reference's):
 *   populations  structure of arrays, f[q][z][y][x], x fastest, 19 planes of
 *                Nx*Ny*Nz elements of the storage type; updated in place
 *                (AA pattern).  A state word describes the buffer:
 *                bit 0 (CLBM_STATE_PARITY) = number of steps taken mod 2;
 *                parity 0: post-stream population q of cell c sits at f[q][c];
 *                parity 1: it sits at f[opp(q)][c - e_q] (periodic index).
 *                bit 1 (CLBM_STATE_PLANES): the x planes still hold raw streamed
 *                values; every reader applies the inlet rule at x = 0 and takes
 *                cell Nx-2's populations for x = Nx-1 as it loads (this is how
 *                K5/K6 run without a pass of their own).  Readers of such a
 *                state need the params (inlet velocity).
 *                16-bit storage holds f_q - feq_q(1, dev velocity of the grid), the
 *                deviation from the lattice weight w_q when that velocity is zero.
 *   scalar fields  [z][y][x] float32, x fastest.
 *   force box    uint64[IBM_BOX_WORDS], zero-initialised, owned by the caller.
 *                Every IBM call carries a generation number gen >= 1 (one per
 *                compute_force, increasing).  The words of parity gen&1
 *                collect the bounding box of the cells that generation's
 *                stencils reach as atomic maxima of (gen << 32 | coordinate
 *                key), spread over IBM_BOX_SUBSLOTS copies, so a slot needs no
 *                reset and a stale one is recognised by its gen.
 *                ibm_clip_forces decodes the current box into the int32[6]
 *                (x0,y0,z0,x1,y1,z1) at IBM_FORCE_BOX(box), the form clbm_step
 *                reads as its force_box.
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

#define CLBM_ABI_VERSION 15

/* population storage formats */
#define CLBM_STORE_F32 0  /* float32 populations, float32 arithmetic (the drop-in) */
#define CLBM_STORE_F16 1  /* IEEE half deviations f - w, float32 arithmetic */
#define CLBM_STORE_BF16 2 /* bfloat16 deviations f - w, float32 arithmetic */
#define CLBM_STORE_F64 3  /* float64 populations, float64 arithmetic (error reference) */
#define CLBM_STORE_F16C 4 /* 16-bit deviations in a 1-4-11 float (|x| < 1, twice IEEE half's
                             precision), float32 arithmetic */

/* force spreading modes */
#define IBM_SPREAD_FIXED 0 /* order-independent 64-bit fixed-point sums (deterministic) */
#define IBM_SPREAD_ATOMIC_F32 1 /* float32 atomics, as ti.atomic_add (order-dependent) */
#define IBM_SPREAD_SORTED 2 /* the stencil points grouped by cell, one fixed-point sum per cell
                               (the sums of IBM_SPREAD_FIXED; ibm_spread_sorted only) */

/* inlet rules at x = 0 (clbm_params_t.inlet) */
#define CLBM_INLET_NONE 0           /* x periodic, no outlet copy */
#define CLBM_INLET_REFERENCE 1      /* CLBMSolver._bc_inlet (clbm_moving.py:266-294) */
#define CLBM_INLET_HECHT_HARTING 2  /* Hecht-Harting on-site velocity inlet (opt-in) */

#define CLBM_STATE_PARITY 1
#define CLBM_STATE_PLANES 2

#define IBM_BOX_SUBSLOTS 64
#define IBM_BOX_WORDS (2 * 6 * IBM_BOX_SUBSLOTS + 32)
/* the decoded int32[6] box inside the uint64 box record */
#define IBM_FORCE_BOX(box) ((int32_t*)((box) + 2 * 6 * IBM_BOX_SUBSLOTS))
/* the word of the box record where ibm_compute_force marks (as an atomic
 * maximum of gen) a generation whose stencils left ibm_params_t.region */
#define IBM_REGION_FLAG(box) ((box) + 2 * 6 * IBM_BOX_SUBSLOTS + 3)
/* words of the box record used by clbm_ibm_step within one call: the inside
 * workgroups' generation-tagged decision, and per generation parity eight
 * uint32 counts of finished marker workgroups */
#define IBM_STEP_DECISION(box) ((box) + 2 * 6 * IBM_BOX_SUBSLOTS + 4)
#define IBM_STEP_DONE(box) ((uint32_t*)((box) + 2 * 6 * IBM_BOX_SUBSLOTS + 8))
/* words of the box record for clbm_step_fused, in a 128-byte line of their
 * own (polled while the line of the decoded box is being written): the last
 * generation whose clamped forces ibm_compute_force_split has published, and
 * (atomic maximum) the generations whose region ran out of its wait */
#define IBM_SPLIT_READY(box) ((box) + 2 * 6 * IBM_BOX_SUBSLOTS + 16)
#define IBM_SPLIT_EXPIRED(box) ((box) + 2 * 6 * IBM_BOX_SUBSLOTS + 17)

typedef struct {
  int32_t nx, ny, nz; /* lattice size, Nx >= 3; one population plane (nx*ny*nz elements of
                         the storage format) must stay under 4 GiB, or calls fail with -1 */
  int32_t storage;    /* CLBM_STORE_* */
  float dev_ux, dev_uy; /* 16-bit formats store f_q - feq_q(rho = 1, u = (dev_ux, dev_uy, 0));
                           0, 0 stores f_q - w_q.  Part of the buffer's format: every call
                           on one buffer must pass the same values.  Wide formats ignore it. */
} clbm_grid_t;

typedef struct {
  double s_v;          /* shear rate 1/tau (config_moving.py:40) */
  double s_e;          /* bulk rate (config_moving.py:41) */
  double ux_in, uy_in; /* inlet velocity _UX_IN/_UY_IN (clbm_moving.py:33-35); rates and
                          velocities are rounded to the arithmetic type of the storage format */
  int32_t inlet;      /* CLBM_INLET_REFERENCE (1): inlet rule at x=0 and outlet copy at
                         x=Nx-1 after every stream (clbm_moving.py:266-300);
                         CLBM_INLET_HECHT_HARTING (2): the same outlet copy with the
                         on-site velocity inlet of Hecht and Harting (J. Stat. Mech.
                         2010, P01018), which imposes the cell's momentum exactly (no
                         counterpart in the reference, opt-in);
                         CLBM_INLET_NONE (0): x stays periodic */
  int32_t outside_lds; /* bytes of LDS each workgroup of clbm_step_part part 0 reserves
                          (unused by it): caps that launch's workgroups per CU so that
                          kernels on another stream find room beside it; 0 = none */
} clbm_params_t;

typedef struct {
  float f_max;               /* marker and grid force clamp, 0.05 in the reference
                                (airfoil_moving.py:178,221) */
  int32_t wrap_weight_quirk; /* 1: delta weights use the wrapped/clamped index, as
                                airfoil_moving.py:159-165 does; 0: unwrapped offsets */
  int32_t spread_mode;       /* IBM_SPREAD_* */
  int32_t fixed_shift;       /* 0 = automatic; else scale 2^shift of the fixed-point sums */
  int32_t group;             /* markers per workgroup of ibm_compute_force: 4, 8 or 16
                                (0 = 8); a tuning choice, results do not depend on it */
  int32_t check_region;      /* 1: ibm_compute_force checks that every stencil cell lies in
                                region and marks IBM_REGION_FLAG otherwise */
  int32_t region[6];         /* inclusive x0,y0,z0,x1,y1,z1 (see clbm_step_part) */
} ibm_params_t;

/* ---- library ---------------------------------------------------------- */
int clbm_abi_version(void);
const char* clbm_last_error(void);
/* Bytes of one population buffer for `g` (19 * Nx*Ny*Nz * element size). */
size_t clbm_state_bytes(const clbm_grid_t* g);

/* State word after one clbm_step from `state` with params `p`. */
int clbm_state_after_step(int state, const clbm_params_t* p);

/* ---- lattice Boltzmann (replaces CLBMSolver, clbm_moving.py:41-343) ----- */

/* Replaces CLBMSolver._init_feq (clbm_moving.py:75-91): second-order
 * equilibrium at (rho0, u0) in every cell; leaves state 0. */
int clbm_init_feq(const clbm_grid_t* g, void* f, double rho0, double ux0, double uy0, double uz0,
                  void* stream);

/* Replaces CLBMSolver.step (clbm_moving.py:326-343): K2 _macro_with_force,
 * K3 _collide, K4 _stream, K5 _bc_inlet and K6 _bc_outlet in one pass (the
 * planes of this step are left pending and applied by the next reader).
 * fx/fy/fz are [z][y][x] float32 force fields (each nullable = zero).
 * force_box (nullable) restricts the force reads to the box in its first six
 * ints; cells outside read zero.  `state` describes f before the step; after
 * it, f is in state clbm_state_after_step(state, p). */
int clbm_step(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
              const float* fx, const float* fy, const float* fz, const int32_t* force_box,
              void* stream);

/* The same step in two launches, so that the marker kernels can run while the
 * lattice away from the body is updated.  part 0 steps every cell outside
 * `region` (host int32[6], inclusive x0,y0,z0,x1,y1,z1) without reading any
 * force; part 1 steps the cells inside it, with the force as clbm_step reads
 * it.  Parts 0 then 1 with the same `state` equal one clbm_step, bit for bit,
 * when no force falls outside the region; the state word advances after
 * part 1.  Each cell's populations are read and written only by that cell's
 * own update (AA pattern), so part 0 touches no population of a cell inside
 * the region: readers of those (ibm_compute_force with velocity from f, whose
 * stencils lie in the region) may run concurrently with part 0, on another
 * stream.  `region` must be normalised by clbm_step_region. */
int clbm_step_part(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
                   const float* fx, const float* fy, const float* fz, const int32_t* force_box,
                   const int32_t* region, int part, void* stream);

/* Host-only: clip `region` (int32[6]) to the grid and, when the inlet/outlet
 * planes are on and it reaches x >= Nx-2, extend it to hold both x = Nx-2 and
 * x = Nx-1 (the outlet cell is stepped with, and read through, its
 * neighbour); result in `out`. */
int clbm_step_region(const clbm_grid_t* g, const clbm_params_t* p, const int32_t* region,
                     int32_t* out);

/* Host-only: the split region of one body for one step, from its template's
 * convex-hull vertices (hx, hy: nh float32 body-frame points) placed by
 * (xc, yc, theta) with the marker kernel's float32 arithmetic, the stencil
 * reach floor(p) - 1 .. floor(p) + 2 plus one cell of margin in x and y, x
 * clipped to the grid, y the whole axis when it crosses the periodic
 * boundary, z from z0..z1, normalised as clbm_step_region does.  Returns 1
 * with the region in `out`, 0 when no split pays (non-finite kinematics, body
 * off the grid in x, or a region over half the lattice), < 0 on argument
 * errors.  No reference counterpart: the reference runs the whole step after
 * the markers (main_moving.py:116-122). */
int ibm_step_region(const clbm_grid_t* g, const clbm_params_t* p, const float* hx,
                    const float* hy, int nh, int z0, int z1, double xc, double yc, double theta,
                    int32_t* out);

/* clbm_step_part 0, then the marker stream's join, then clbm_step_part 1, as
 * one call (the split step of a coupled loop whose marker kernels ran on
 * `side`): part 0 on `main`; `join_ev` recorded on `side` and waited for by
 * `main`; part 1 on `main` with the force.  start_ev / end_ev (nullable,
 * created with timing) receive the start and end of part 0's kernel itself
 * (the launch's own time stamps, not a record on the stream, whose start
 * would be the end of the command before).  done_ev (nullable, any event)
 * completes with part 1's kernel itself, with no marker queued on `main`
 * after it: a caller whose next marker kernels read this step's populations
 * passes the plan's fork_ev here and then calls ibm_compute_force_split with
 * fork_bound = 1 (a marker between two launches on `main` holds the next
 * one back by several microseconds).  No reference counterpart (the
 * reference runs one queue). */
int clbm_step_split(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
                    const int32_t* region, const float* fx, const float* fy,
                    const int32_t* force_box, void* main, void* side, void* join_ev,
                    void* start_ev, void* end_ev, void* done_ev);

/* Host-only stream ordering for running the marker kernels beside the
 * lattice step (no reference counterpart; Taichi runs one queue):
 * clbm_event_create makes a synchronisation-only HIP event, clbm_event_destroy
 * releases it, and clbm_stream_after records `ev` on `signaler` and makes
 * `waiter` wait for it (work queued on `waiter` afterwards starts after
 * everything queued on `signaler` so far). */
int clbm_event_create(void** ev);
int clbm_event_destroy(void* ev);
int clbm_stream_after(void* waiter, void* signaler, void* ev);

/* Streams and hardware queues (no reference counterpart; Taichi runs one
 * queue).  HIP spreads streams over a few hardware queues (GPU_MAX_HW_QUEUES)
 * and two streams on one queue run their kernels one after the other, never
 * side by side.  clbm_streams_concurrent tells which: a one-wave kernel on
 * `main` waits (at most 50 ms) for a flag that a one-wave kernel queued on
 * `other` afterwards sets; *concurrent = 1 when it saw the flag.  It
 * synchronises both streams.
 * clbm_side_stream gives the library's marker stream on the current device: a
 * non-blocking stream created on the first call per device and kept until the
 * process ends, shared by every caller.  The first call checks it against
 * `main` as above and creates another candidate while the check fails (four
 * at most; that call synchronises `main`).  *concurrent (nullable) is 1 when
 * the stream passed, 0 when none did (it then still orders correctly). */
int clbm_streams_concurrent(void* main, void* other, int* concurrent);
int clbm_side_stream(void* main, void** side, int* concurrent);

/* Replaces CLBMSolver._macro_post (clbm_moving.py:305-321): force-free rho
 * and u of the current state into [z][y][x] float32 fields (each nullable). */
int clbm_macro(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state, float* rho,
               float* ux, float* uy, float* uz, void* stream);

/* Replaces reads of CLBMSolver.f (clbm_moving.py:56): the post-stream state as
 * float32 planes out[q][z][y][x] in the reference's q order. */
int clbm_export_state(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state,
                      float* out, void* stream);
/* The same in float64 (for the float64 storage format). */
int clbm_export_state_f64(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state,
                          double* out, void* stream);

/* Replaces writes of CLBMSolver.f (ti.field.from_numpy): loads float32 planes
 * in[q][z][y][x]; leaves state 0. */
int clbm_import_state(const clbm_grid_t* g, void* f, const float* in, void* stream);
/* The same from float64 planes. */
int clbm_import_state_f64(const clbm_grid_t* g, void* f, const double* in, void* stream);

/* No counterpart in the reference: a measurement probe.  Moves the data of one
 * AA step of clbm_step from parity `parity` without the collision: every
 * cell's 19 post-stream slots are loaded and stored into the slots the step
 * writes, in the step's sweep order, so that the populations end up permuted
 * (values kept).  Its duration is the ceiling clbm_step is measured against. */
int clbm_stream_probe(const clbm_grid_t* g, void* f, int parity, void* stream);

/* Replaces the driver's divergence guard (main_moving.py:130-145): out2[0] =
 * min rho over all cells, out2[1] = 1.0 if every population is finite else 0.0. */
int clbm_health(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state,
                float* out2, void* stream);

/* ---- immersed boundary (replaces DroneMovingIBM, airfoil_moving.py:44-255) */

/* Replaces DroneMovingIBM.compute_force (airfoil_moving.py:230-255) as one
 * launch: K8 _update_markers, K10 _interpolate, K11 _marker_forces and K12
 * _spread; K9/K13 are ibm_clip_forces.  Markers are processed in the order
 * `order` (nullable = 0..n-1; a spatially sorted order lets each workgroup
 * pre-sum its stencils in LDS); outputs are indexed by marker.  Velocity
 * comes from ux/uy ([z][y][x] float32) when both are given, else from the
 * population state f (state word `state`, params p).  cos/sin of theta are
 * evaluated in float64 and rounded once to float32; the rotation is float32
 * without contraction.  F_k w is added into acc_x/acc_y (Nx*Ny*Nz 8-byte
 * slots each, all zero before the first call; IBM_SPREAD_ATOMIC_F32 uses the
 * first 4 bytes of a slot as a float) and the touched box into `box` for
 * generation `gen`.  stencil_base (nullable): floor(pos)-1 per axis, n x 3. */
int ibm_compute_force(const clbm_grid_t* g, const ibm_params_t* ip, const clbm_params_t* p,
                      const void* f, int state, const float* ux, const float* uy, int n,
                      const int32_t* order, const float* X0, const float* Y0, const float* Z0,
                      float xc, float yc, float theta, float vx, float vy, float omega, float* X,
                      float* Y, float* Z, float* VX, float* VY, float* U, float* V, float* FX,
                      float* FY, int32_t* stencil_base, void* acc_x, void* acc_y, uint64_t* box,
                      uint32_t gen, void* stream);

/* The launch plan of one body on one lattice for ibm_compute_force_split:
 * what ibm_compute_force + ibm_clip_forces take that stays fixed from call to
 * call, built once by the caller (hx, hy: host memory). */
typedef struct {
  clbm_grid_t grid;
  clbm_params_t params;
  ibm_params_t ip;            /* check_region and region are set per call */
  const void* f;              /* populations (the velocity source) */
  int32_t n;
  const int32_t* order;
  const float *X0, *Y0, *Z0;
  float *X, *Y, *Z, *VX, *VY, *U, *V, *FX, *FY;
  int32_t* stencil_base;
  void *acc_x, *acc_y;
  uint64_t* box;
  float *fx, *fy;
  const float *hx, *hy;       /* template hull vertices (see ibm_step_region) */
  int32_t nh, z0, z1;
  void *side, *fork_ev;       /* the marker stream and its fork event */
} ibm_split_plan_t;

/* CLBMSolver.step (clbm_moving.py:326-343) for the split coupled loop as
 * one launch on `main` with no join of the marker stream: the cells outside
 * `region` force-free (as clbm_step_part 0), then, in the last workgroups of
 * the same launch, the region's cells with the force (as part 1) once
 * *IBM_SPLIT_READY(box) >= gen, which ibm_compute_force_split publishes on
 * the marker stream after its clamp of generation gen.  Those workgroups
 * wait at most wait_ms milliseconds; when the wait runs out they mark gen in
 * *IBM_SPLIT_EXPIRED(box) and step the region with the forces as they are
 * (the step's result is then not to be trusted: check that word, as
 * CLBMSolver.health does).  The marker stream must run beside `main`
 * (clbm_side_stream).  start_ev / end_ev (nullable, created with timing)
 * bracket the launch itself; done_ev (nullable) completes with it, as in
 * clbm_step_split.  Results equal clbm_step_split bit for bit.  No
 * reference counterpart (the reference runs one queue). */
int clbm_step_fused(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
                    const int32_t* region, const float* fx, const float* fy, uint64_t* box,
                    uint32_t gen, double wait_ms, void* main, void* start_ev, void* end_ev,
                    void* done_ev);

/* Replaces DroneMovingIBM.compute_force (airfoil_moving.py:230-255) in the
 * split coupled loop, as one call: the step region (ibm_step_region); when
 * it pays, `side` waits for `main` (fork_ev) and ibm_compute_force (region
 * checked) + ibm_clip_forces run on `side`, generation gen is published in
 * IBM_SPLIT_READY(box) for clbm_step_fused, the region goes to region_out
 * and the call returns 1; otherwise both run on `main` and it returns 0.
 * Velocity from the populations in state `state`.  fork_bound = 1: fork_ev
 * already completes with the last work on `main` that the marker kernels
 * must follow (bound to the last step's part 1 by clbm_step_split's done_ev,
 * and nothing queued on `main` since that writes the populations or touches
 * the plan's buffers), so `side` waits for it without a new record on
 * `main`; 0: fork_ev is recorded on `main` first.  < 0 on errors.  No
 * reference counterpart as one call. */
int ibm_compute_force_split(ibm_split_plan_t* plan, int state, uint32_t gen, float xc, float yc,
                            float theta, float vx, float vy, float omega, void* main,
                            int fork_bound, int32_t* region_out);

/* Replaces DroneMovingIBM._update_markers (airfoil_moving.py:100-127) on its
 * own (same arithmetic as in ibm_compute_force).  box (nullable) collects the
 * box of the new stencils for generation `gen`. */
int ibm_update_markers(const clbm_grid_t* g, int n, const float* X0, const float* Y0,
                       const float* Z0, float xc, float yc, float theta, float vx, float vy,
                       float omega, float* X, float* Y, float* Z, float* VX, float* VY,
                       uint64_t* box, uint32_t gen, void* stream);

/* Part of DroneMovingIBM._interpolate's input (the K7 velocity of
 * clbm_moving.py:305-321 that _interpolate reads): the force-free velocity
 * of the cells of generation gen's box only, computed from the population
 * state f into [z][y][x] float32 fields ux/uy (other cells untouched). */
int ibm_band_velocity(const clbm_grid_t* g, const clbm_params_t* p, const void* f, int state,
                      const uint64_t* box, uint32_t gen, float* ux, float* uy, void* stream);

/* Replaces DroneMovingIBM._interpolate + _marker_forces (airfoil_moving.py:
 * 142-183) on their own.  Velocity as in ibm_compute_force. */
int ibm_interp_force(const clbm_grid_t* g, const ibm_params_t* ip, const clbm_params_t* p,
                     const void* f, int state, const float* ux, const float* uy, int n,
                     const float* X, const float* Y, const float* Z, const float* VX,
                     const float* VY, float* U, float* V, float* FX, float* FY,
                     int32_t* stencil_base, void* stream);

/* Replaces DroneMovingIBM._spread (airfoil_moving.py:188-214) on its own:
 * adds F_k w into acc_x/acc_y as ibm_compute_force does. */
int ibm_spread(const clbm_grid_t* g, const ibm_params_t* ip, int n, const float* X, const float* Y,
               const float* Z, const float* FX, const float* FY, void* acc_x, void* acc_y,
               void* stream);

/* Replaces DroneMovingIBM._spread (airfoil_moving.py:188-214) by sorting
 * instead of atomics on the fields: the 64 n (cell, F_k w) pairs of the
 * stencils are grouped by cell (a counting sort on the cell index) and one
 * thread per cell adds its group into acc_x/acc_y, in the 64-bit fixed-point
 * format of IBM_SPREAD_FIXED (ibm_clip_forces converts them as for that
 * mode): the same sums as ibm_spread in that mode, bit for bit.
 * ip->spread_mode must be IBM_SPREAD_SORTED.  `work` (device memory of
 * ibm_spread_sorted_bytes(g, n) bytes: 8 per cell and 52 per stencil
 * point) must be zero before the first call; every call leaves the parts
 * that need it at zero.  The fused entry points (ibm_compute_force, the
 * batches, clbm_ibm_step) reject this mode: it runs as ibm_update_markers +
 * ibm_interp_force + ibm_spread_sorted + ibm_clip_forces.  Calls that overlap
 * on one work buffer race.  No reference counterpart as a mode (the
 * reference adds with ti.atomic_add). */
size_t ibm_spread_sorted_bytes(const clbm_grid_t* g, int n);
int ibm_spread_sorted(const clbm_grid_t* g, const ibm_params_t* ip, int n, const float* X,
                      const float* Y, const float* Z, const float* FX, const float* FY,
                      void* acc_x, void* acc_y, void* work, size_t work_bytes, void* stream);

/* Replaces DroneMovingIBM._zero_forces + _clip_forces (airfoil_moving.py:
 * 132-137,219-225): over the union of generation gen's box and generation
 * gen-1's, converts the accumulators into the clamped float32 fields fx/fy
 * ([z][y][x]; nullable) and clears them; fx/fy are then exact everywhere
 * (outside both boxes they were already zero).  Writes generation gen's box
 * to IBM_FORCE_BOX(box) for clbm_step. */
int ibm_clip_forces(const clbm_grid_t* g, const ibm_params_t* ip, void* acc_x, void* acc_y,
                    float* fx, float* fy, uint64_t* box, uint32_t gen, void* stream);

/* Replaces the loop body of main_moving.py:116-122 (DroneMovingIBM.
 * compute_force, airfoil_moving.py:230-255, then CLBMSolver.step,
 * clbm_moving.py:326-343) for one body on one lattice: the same results as
 * ibm_compute_force (velocity from f at `state`) + ibm_clip_forces +
 * clbm_step with the clamped fields, bit for bit, in one launch and a short
 * tail on one stream.  The marker workgroups and the update of the cells
 * outside `region` share the launch (the cells the markers read lie inside
 * it and no update touches another cell's populations); the region's cells
 * take the force straight from the accumulators, clamped as ibm_clip_forces
 * would write it, and clear them (so the clamp has no launch of its own).
 * fx/fy fields are not written (ibm_spread + ibm_clip_forces rebuild them
 * from the markers when a reader wants them); the box of generation gen goes
 * to IBM_FORCE_BOX(box).  `region`: normalised by clbm_step_region and
 * holding every stencil cell (ibm_step_region predicts it; a stencil outside
 * it marks IBM_REGION_FLAG and then results are not trustworthy).  Needs the
 * fixed-point spread mode; ip->group and ip->check_region are not used (four
 * markers per workgroup, the region always checked).  `interleave`: one
 * super-block of eight marker workgroups every `interleave` super-blocks
 * from the start of the launch (1: markers first).  `wait_spins`: 0 steps the
 * region in the tail kernel; > 0 steps it in the launch's last super-blocks
 * (one every `inside_spread`), after a wait of at most that many polls for the
 * marker workgroups, else in the tail; < 0 gives up at once (a test of the
 * tail path).  No reference counterpart as one call. */
int clbm_ibm_step(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
                  const ibm_params_t* ip, int n, const int32_t* order, const float* X0,
                  const float* Y0, const float* Z0, float xc, float yc, float theta, float vx,
                  float vy, float omega, float* X, float* Y, float* Z, float* VX, float* VY,
                  float* U, float* V, float* FX, float* FY, int32_t* stencil_base, void* acc_x,
                  void* acc_y, uint64_t* box, uint32_t gen, const int32_t* region,
                  int interleave, int inside_spread, int wait_spins, void* stream);

/* No reference counterpart (the reference never reduces forces): out2[0..1]
 * = sum_k FX_k dV_k, sum_k FY_k dV_k in float64 with a fixed summation order
 * (dV nullable = 1). */
int ibm_force_sum_f64(int n, const float* FX, const float* FY, const float* dV, double* out2,
                      void* stream);

/* ---- batched runs: many independent lattices and bodies per launch ------
 * No reference counterpart (the reference runs one case per process).  A
 * batch advances n instances, each a lattice with its own grid, rates and
 * inflow (all in one storage format) and one body, in lockstep: per step one
 * ibm_batch_compute (marker kernel + clamp of every body, which also advances
 * the batch's device step counter) and one clbm_batch_step (fused step of
 * every lattice), so that a step of the whole batch is a fixed sequence of
 * three launches that a caller can capture in a graph and replay.  Kinematics
 * come from per-body device tables indexed by the step counter, never from
 * host arguments.
 * Every instance computes what its own single-instance calls would, bit for
 * bit (fixed-point spreading).  Descriptor tables live in caller-owned device
 * work buffers of the sizes the *_bytes functions return; building copies
 * them there synchronously. */
typedef struct {
  int32_t n, storage, block, blocks;
  void* args;         /* device: per-instance step arguments */
  uint16_t* inst;     /* device: instance of each block */
  int32_t* first;     /* device: first block of each instance */
  int32_t lds;        /* bytes of LDS each workgroup of clbm_batch_step reserves, unused, to cap
                         the workgroups per CU (0 after clbm_batch_build; a caller may set it,
                         at most 65536) */
} clbm_batch_t;

/* Device bytes of the lattice tables for n grids, `block` threads per
 * workgroup (128 or 256). */
size_t clbm_batch_bytes(int n, const clbm_grid_t* grids, int block);
/* Lattice i: populations f[i], the body's clamped force fields fx[i], fy[i]
 * and its force-box record boxes[i] (read inside IBM_FORCE_BOX, as clbm_step
 * reads its force_box). */
int clbm_batch_build(int n, const clbm_grid_t* grids, const clbm_params_t* params, void* const* f,
                     float* const* fx, float* const* fy, uint64_t* const* boxes, int block,
                     void* work, size_t work_bytes, clbm_batch_t* out);
/* clbm_step of every lattice from state `state` (all lattices share it). */
int clbm_batch_step(const clbm_batch_t* b, int state, void* stream);

/* The buffers of one body (as the arguments of ibm_compute_force + ibm_clip_forces). */
typedef struct {
  int32_t n;
  const int32_t* order;
  const float *X0, *Y0, *Z0;
  float *X, *Y, *Z, *VX, *VY, *U, *V, *FX, *FY;
  void *acc_x, *acc_y;
  uint64_t* box;
  float *fx, *fy;
  const float* kin; /* device: for each step s of the counter, 7 floats
                       xc, yc, cos(theta), sin(theta), vx, vy, omega */
  uint32_t gen0;    /* the body's generation is gen0 + 1 + counter */
  const int32_t* regions; /* device, nullable: for each step of the counter, 6 int32 x0, y0, z0,
                             x1, y1, z1 of a region that holds the body's stencils (e.g. from
                             ibm_step_region); needed by ibm_batch_markers / clbm_batch_step_acc,
                             and given for every body of a batch or for none */
} ibm_buffers_t;

typedef struct {
  int32_t n, storage, spread_mode, group, blocks, clip_blocks;
  void* args;         /* device: per-body marker-kernel arguments */
  void* clip_args;    /* device: per-body clamp arguments */
  uint16_t* inst;
  int32_t* first;
  int32_t* counter;   /* device: steps taken by the batch (followed by the clamp
                         kernel's own count in the same 256 bytes); with region
                         tables, counter[p] holds it before a step of AA parity p */
  int32_t regions;    /* the bodies carry region tables */
} ibm_batch_t;

/* Device bytes of the body tables (the step counter included). */
size_t ibm_batch_bytes(int n, const ibm_buffers_t* bufs, int group);
/* Body i moves on lattice i (grids[i], params[i], populations f[i]); ip
 * (spread mode, clamp, quirk, group 4 or 8) is shared.  Sets the counter to 0. */
int ibm_batch_build(int n, const clbm_grid_t* grids, const clbm_params_t* params, void* const* f,
                    const ibm_params_t* ip, const ibm_buffers_t* bufs, void* work,
                    size_t work_bytes, ibm_batch_t* out);
/* ibm_compute_force + ibm_clip_forces of every body at the counter's step,
 * velocity from each lattice's populations in state `state`; then counter
 * += 1 (the clamp kernel's last workgroup). */
int ibm_batch_compute(const ibm_batch_t* b, int state, void* stream);
/* A batch step in two launches instead of three (no reference counterpart;
 * fixed-point spreading; bodies with region tables):
 *   ibm_batch_markers(&ib, state, stream);            markers of every body (stencils
 *                                                     checked against the step's region,
 *                                                     IBM_REGION_FLAG), next counter set
 *   clbm_batch_step_acc(&lb, &ib, state, stream);     every lattice, its force decoded
 *                                                     and clamped from the accumulators
 *                                                     inside the region
 * The populations and marker arrays equal those of ibm_batch_compute +
 * clbm_batch_step bit for bit while the stencils stay in the regions; the
 * fields fx, fy and the published force box are not written (the
 * accumulators are cleared by the lattice step).  The step counter is
 * counter[p] before a step of AA parity p (the caller sets counter[p] for
 * the first step).  lb and ib must be built over the same instances in the
 * same order. */
int ibm_batch_markers(const ibm_batch_t* b, int state, void* stream);
int clbm_batch_step_acc(const clbm_batch_t* lb, const ibm_batch_t* ib, int state, void* stream);

/* ---- learned closure (no reference counterpart: the reference fixes the
 * third- and fourth-order rates at 1, clbm_moving.py:188-196) -------------
 * A per-cell network 19 -> 32 -> 32 -> 4 (GELU in tanh form) maps the cell's
 * 19 features (central moments about the half-force velocity divided by
 * rho, minus their equilibrium values, in the raw-moment order of
 * config_moving.py:62-68) to y; the rates are s = 1 + y for (k210, k120),
 * (k201, k021, k102, k012), (k220), (k202, k022): third-order moments relax as
 * (1 - s) k, fourth-order ones as rho/9 + (1 - s)(k - rho/9).  A cell whose
 * rates leave [lo, hi] takes the plain cascade.  Parameters: one float32
 * vector of closure_param_count() = 1828 entries, W1 (19 x 32, [in][out]),
 * b1, W2 (32 x 32), b2, W3 (32 x 4), b3.  Products run on bf16 MFMA (weights
 * and activations rounded to bf16, float32 accumulation and biases); the
 * kernels read the weights through a packed copy (closure_pack). */
int closure_param_count(void);
size_t closure_packed_bytes(void);
/* params (device, float32) -> packed (device, closure_packed_bytes()). */
int closure_pack(const float* params, void* packed, void* stream);
/* Y[n][4] = network(X[n][19]) (device). */
int closure_infer(const void* packed, const float* params, int count, const float* X, float* Y,
                  void* stream);
/* clbm_step with the learned rates (storage f32, f16 or f16c). */
int clbm_step_closure(const clbm_grid_t* g, const clbm_params_t* p, void* f, int state,
                      const float* fx, const float* fy, const float* fz, const int32_t* force_box,
                      const void* packed, const float* params, float lo, float hi, void* stream);
/* Bytes of the partial-gradient buffer of `blocks` workgroups. */
size_t closure_partial_bytes(int blocks);
/* grad[0..1827] = d/dparams of L = sum_n W[n] |network(X[n]) - T[n]|^2 / wsum
 * (wdim 1: one weight per sample) or L = sum_nj W[n][j] (network(X[n])_j -
 * T[n][j])^2 / wsum (wdim 4: one weight per output), grad[1828] = L, over
 * count samples with `blocks` workgroups (fixed summation order: the same
 * inputs give the same bits). */
int closure_grad(const void* packed, const float* params, int count, const float* X,
                 const float* T, const float* W, int wdim, float wsum, float* partial, int blocks,
                 float* grad, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, training or serving loops that forbid
 * hipMalloc): closure_grad without the `partial` workspace.  The partial
 * gradients go to a device array inside the library of
 * CLOSURE_GRAD_NOALLOC_MAX_BLOCKS x (closure_param_count() + 1) floats
 * (7.1 MiB), which each GPU the library runs on holds until the process ends;
 * blocks > CLOSURE_GRAD_NOALLOC_MAX_BLOCKS is rejected with an error.  Same
 * results as closure_grad with the same blocks, bit for bit.  Allocates
 * nothing; the first call on a device looks the array's address up on that
 * (current) device.  All calls on one GPU share the array: calls that overlap
 * on one GPU (different streams or threads) race. */
#define CLOSURE_GRAD_NOALLOC_MAX_BLOCKS 1024
int closure_grad_noalloc(const void* packed, const float* params, int count, const float* X,
                         const float* T, const float* W, int wdim, float wsum, int blocks,
                         float* grad, void* stream);
/* One Adam step (bias-corrected, step number t >= 1) on params with moments m, v. */
int closure_adam(float* params, float* m, float* v, const float* grad, float lr, float beta1,
                 float beta2, float eps, int t, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CLBM_AMD_H */
