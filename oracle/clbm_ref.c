/*
 * clbm_ref.c -- reference-structured C restatement of the CLBM-IBM time step.
 *
 * TEST INFRASTRUCTURE AND CPU BASELINE ONLY; the product never links it.
 *
 * It keeps the reference's structure on purpose, so that timing it gives a
 * CPU baseline of the reference's algorithm as written:
 *   - populations f and f_post as (Nx, Ny, Nz, 19) float32, q fastest (AoS,
 *     double buffer), clbm_moving.py:56-57;
 *   - the twelve kernels in the reference's order: compute_force =
 *     K8 _update_markers, K9 _zero_forces, K10 _interpolate, K11 _marker_forces,
 *     K12 _spread (atomic adds), K13 _clip_forces (airfoil_moving.py:230-255);
 *     step = K2 _macro_with_force, K3 _collide (15 central moments, dense
 *     19x19 M_inv product, Guo), K4 _stream, K5 _bc_inlet, K6 _bc_outlet,
 *     K7 _macro_post (clbm_moving.py:326-343);
 *   - float32 arithmetic, every loop over cells or markers parallel (OpenMP),
 *     like Taichi's parallel outermost loop.
 * Build with -ffp-contract=off so each float operation rounds on its own.
 * The marker rotation uses cos/sin of the float32 angle evaluated in double
 * and rounded once, the convention the GPU path uses as well.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define Q 19
static const int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const float W[Q] = {1.0f / 3, 1.0f / 18, 1.0f / 18, 1.0f / 18, 1.0f / 18, 1.0f / 18,
                           1.0f / 18, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36,
                           1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36,
                           1.0f / 36};

typedef struct {
  int nx, ny, nz, n;
  size_t ncell;
  float s_v, s_e, ux_in, uy_in, f_max;
  int quirk;
  int inlet_rule; /* 0: the reference's rule (K5); 1: Hecht-Harting (opt-in, no counterpart) */
  float minv[Q * Q];
  float *f, *f_post, *rho, *ux, *uy, *uz;  /* lattice */
  float *fx, *fy, *fz;                     /* IBM Eulerian force */
  float *X0, *Y0, *Z0, *X, *Y, *Z, *VX, *VY, *U, *V, *FX, *FY; /* markers */
} ref_sim;

static inline size_t cidx(const ref_sim* s, int i, int j, int k) {
  return ((size_t)i * s->ny + j) * s->nz + k;
}

ref_sim* ref_create(int nx, int ny, int nz, double s_v, double s_e, double ux_in, double uy_in,
                    const float* minv, int n, const float* X0, const float* Y0, const float* Z0,
                    float f_max, int quirk) {
  ref_sim* s = (ref_sim*)calloc(1, sizeof(ref_sim));
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  s->n = n;
  s->ncell = (size_t)nx * ny * nz;
  s->s_v = (float)s_v;
  s->s_e = (float)s_e;
  s->ux_in = (float)ux_in;
  s->uy_in = (float)uy_in;
  s->f_max = f_max;
  s->quirk = quirk;
  memcpy(s->minv, minv, sizeof(s->minv));
  s->f = (float*)malloc(sizeof(float) * s->ncell * Q);
  s->f_post = (float*)malloc(sizeof(float) * s->ncell * Q);
  s->rho = (float*)malloc(sizeof(float) * s->ncell);
  s->ux = (float*)malloc(sizeof(float) * s->ncell);
  s->uy = (float*)malloc(sizeof(float) * s->ncell);
  s->uz = (float*)malloc(sizeof(float) * s->ncell);
  s->fx = (float*)calloc(s->ncell, sizeof(float));
  s->fy = (float*)calloc(s->ncell, sizeof(float));
  s->fz = (float*)calloc(s->ncell, sizeof(float));
  float** m[] = {&s->X0, &s->Y0, &s->Z0, &s->X, &s->Y, &s->Z, &s->VX, &s->VY, &s->U, &s->V, &s->FX, &s->FY};
  for (int i = 0; i < 12; ++i) *m[i] = (float*)calloc(n > 0 ? n : 1, sizeof(float));
  if (n > 0) {
    memcpy(s->X0, X0, sizeof(float) * n);
    memcpy(s->Y0, Y0, sizeof(float) * n);
    memcpy(s->Z0, Z0, sizeof(float) * n);
  }
  return s;
}

void ref_destroy(ref_sim* s) {
  if (!s) return;
  float* p[] = {s->f, s->f_post, s->rho, s->ux, s->uy, s->uz, s->fx, s->fy, s->fz, s->X0, s->Y0,
                s->Z0, s->X, s->Y, s->Z, s->VX, s->VY, s->U, s->V, s->FX, s->FY};
  for (int i = 0; i < (int)(sizeof(p) / sizeof(p[0])); ++i) free(p[i]);
  free(s);
}

void ref_set_threads(int t) { omp_set_num_threads(t > 0 ? t : 1); }

/* K1 _init_feq, clbm_moving.py:75-91 */
void ref_init_feq(ref_sim* s) {
  const float ux0 = s->ux_in, uy0 = s->uy_in, cs2 = 1.0f / 3.0f;
  const float u2 = ux0 * ux0 + uy0 * uy0;
  float fe[Q];
  for (int q = 0; q < Q; ++q) {
    const float eu = EX[q] * ux0 + EY[q] * uy0;
    fe[q] = W[q] * (1.0f + eu / cs2 + 0.5f * eu * eu / (cs2 * cs2) - 0.5f * u2 / cs2);
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)s->ncell; ++c) {
    for (int q = 0; q < Q; ++q) s->f[c * Q + q] = fe[q];
    s->rho[c] = 1.0f;
    s->ux[c] = ux0;
    s->uy[c] = uy0;
    s->uz[c] = 0.0f;
  }
}

/* K2 _macro_with_force, clbm_moving.py:96-115 */
static void k2_macro_with_force(ref_sim* s) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)s->ncell; ++c) {
    const float* fc = s->f + c * Q;
    float r = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
    for (int q = 0; q < Q; ++q) {
      r += fc[q];
      mx += EX[q] * fc[q];
      my += EY[q] * fc[q];
      mz += EZ[q] * fc[q];
    }
    s->rho[c] = r;
    s->ux[c] = mx / r + 0.5f * s->fx[c] / r;
    s->uy[c] = my / r + 0.5f * s->fy[c] / r;
    s->uz[c] = mz / r + 0.5f * s->fz[c] / r;
  }
}

/* K3 _collide, clbm_moving.py:120-249 */
static void k3_collide(ref_sim* s) {
  const float s_v = s->s_v, s_e = s->s_e, cs2 = 1.0f / 3.0f;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)s->ncell; ++c) {
    const float* fc = s->f + c * Q;
    float* fp = s->f_post + c * Q;
    const float u = s->ux[c], v = s->uy[c], w = s->uz[c], r = s->rho[c];
    const float Fx = s->fx[c], Fy = s->fy[c], Fz = s->fz[c];
    float k110 = 0, k101 = 0, k011 = 0, k200 = 0, k020 = 0, k002 = 0;
    float k210 = 0, k120 = 0, k201 = 0, k021 = 0, k102 = 0, k012 = 0;
    float k220 = 0, k202 = 0, k022 = 0;
    for (int q = 0; q < Q; ++q) {
      const float fq = fc[q];
      const float xix = EX[q] - u, xiy = EY[q] - v, xiz = EZ[q] - w;
      const float xix2 = xix * xix, xiy2 = xiy * xiy, xiz2 = xiz * xiz;
      k110 += fq * xix * xiy;
      k101 += fq * xix * xiz;
      k011 += fq * xiy * xiz;
      k200 += fq * xix2;
      k020 += fq * xiy2;
      k002 += fq * xiz2;
      k210 += fq * xix2 * xiy;
      k120 += fq * xix * xiy2;
      k201 += fq * xix2 * xiz;
      k021 += fq * xiy2 * xiz;
      k102 += fq * xix * xiz2;
      k012 += fq * xiy * xiz2;
      k220 += fq * xix2 * xiy2;
      k202 += fq * xix2 * xiz2;
      k022 += fq * xiy2 * xiz2;
    }
    (void)k210; (void)k120; (void)k201; (void)k021; (void)k102; (void)k012;
    (void)k220; (void)k202; (void)k022;
    const float k110s = k110 * (1.0f - s_v), k101s = k101 * (1.0f - s_v), k011s = k011 * (1.0f - s_v);
    const float kA = k200 + k020 + k002, kA_eq = r, kD = k200 - k020, kE = k200 - k002;
    const float kAs = kA - s_e * (kA - kA_eq), kDs = kD * (1.0f - s_v), kEs = kE * (1.0f - s_v);
    const float k200s = (kAs + kDs + kEs) / 3.0f;
    const float k020s = (kAs - 2.0f * kDs + kEs) / 3.0f;
    const float k002s = (kAs + kDs - 2.0f * kEs) / 3.0f;
    const float k210s = 0, k120s = 0, k201s = 0, k021s = 0, k102s = 0, k012s = 0;
    const float k220s = r / 9.0f, k202s = r / 9.0f, k022s = r / 9.0f;
    float T[Q];
    T[0] = r;
    T[1] = u * r;
    T[2] = v * r;
    T[3] = w * r;
    T[4] = k110s + u * v * r;
    T[5] = k101s + u * w * r;
    T[6] = k011s + v * w * r;
    T[7] = k200s + u * u * r;
    T[8] = k020s + v * v * r;
    T[9] = k002s + w * w * r;
    T[10] = k210s + 2.0f * u * k110s + v * k200s + u * u * v * r;
    T[11] = k120s + 2.0f * v * k110s + u * k020s + u * v * v * r;
    T[12] = k201s + 2.0f * u * k101s + w * k200s + u * u * w * r;
    T[13] = k021s + 2.0f * v * k011s + w * k020s + v * v * w * r;
    T[14] = k102s + 2.0f * w * k101s + u * k002s + u * w * w * r;
    T[15] = k012s + 2.0f * w * k011s + v * k002s + v * w * w * r;
    T[16] = (k220s + 2.0f * u * k120s + 2.0f * v * k210s + 4.0f * u * v * k110s + v * v * k200s +
             u * u * k020s + u * u * v * v * r);
    T[17] = (k202s + 2.0f * u * k102s + 2.0f * w * k201s + 4.0f * u * w * k101s + w * w * k200s +
             u * u * k002s + u * u * w * w * r);
    T[18] = (k022s + 2.0f * v * k012s + 2.0f * w * k021s + 4.0f * v * w * k011s + w * w * k020s +
             v * v * k002s + v * v * w * w * r);
    for (int q = 0; q < Q; ++q) {
      float val = 0.0f;
      for (int ri = 0; ri < Q; ++ri) val += s->minv[q * Q + ri] * T[ri];
      fp[q] = val;
    }
    const float coeff = 1.0f - s_v * 0.5f;
    for (int q = 0; q < Q; ++q) {
      const float eu = EX[q] * u + EY[q] * v + EZ[q] * w;
      const float cx = (EX[q] - u) / cs2 + eu * EX[q] / (cs2 * cs2);
      const float cy = (EY[q] - v) / cs2 + eu * EY[q] / (cs2 * cs2);
      const float cz = (EZ[q] - w) / cs2 + eu * EZ[q] / (cs2 * cs2);
      fp[q] += W[q] * coeff * (cx * Fx + cy * Fy + cz * Fz);
    }
  }
}

/* K4 _stream, clbm_moving.py:254-261 */
static void k4_stream(ref_sim* s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k)
        for (int q = 0; q < Q; ++q) {
          const int is = (i - EX[q] + nx) % nx, js = (j - EY[q] + ny) % ny, ks = (k - EZ[q] + nz) % nz;
          s->f[cidx(s, i, j, k) * Q + q] = s->f_post[cidx(s, is, js, ks) * Q + q];
        }
}

/* K5 _bc_inlet, clbm_moving.py:266-294 */
static void k5_bc_inlet(ref_sim* s) {
  const float ux_in = s->ux_in, uy_in = s->uy_in;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < s->ny; ++j)
    for (int k = 0; k < s->nz; ++k) {
      float* g = s->f + cidx(s, 0, j, k) * Q;
      const float s_known = g[0] + g[3] + g[4] + g[5] + g[6] + g[15] + g[16] + g[17] + g[18];
      const float s_out = g[2] + g[8] + g[10] + g[12] + g[14];
      const float rho_in = (s_known + 2.0f * s_out) / (1.0f - ux_in);
      const float diff_y = g[3] - g[4], diff_z = g[5] - g[6];
      g[1] = g[2] + (2.0f / 3.0f) * rho_in * ux_in;
      g[7] = g[10] + (1.0f / 6.0f) * rho_in * ux_in + 0.5f * rho_in * uy_in - 0.5f * diff_y;
      g[9] = g[8] + (1.0f / 6.0f) * rho_in * ux_in - 0.5f * rho_in * uy_in + 0.5f * diff_y;
      g[11] = g[14] + (1.0f / 6.0f) * rho_in * ux_in - 0.5f * diff_z;
      g[13] = g[12] + (1.0f / 6.0f) * rho_in * ux_in + 0.5f * diff_z;
    }
}

/* The opt-in Hecht-Harting velocity inlet (J. Stat. Mech. 2010, P01018; no
 * reference counterpart), as clbm_oracle.bc_inlet_hecht_harting states it. */
static void k5_bc_inlet_hecht_harting(ref_sim* s) {
  const float ux_in = s->ux_in, uy_in = s->uy_in;
  const float half = 0.5f, third = 1.0f / 3.0f, sixth = 1.0f / 6.0f;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < s->ny; ++j)
    for (int k = 0; k < s->nz; ++k) {
      float* g = s->f + cidx(s, 0, j, k) * Q;
      const float s_known = g[0] + g[3] + g[4] + g[5] + g[6] + g[15] + g[16] + g[17] + g[18];
      const float s_out = g[2] + g[8] + g[10] + g[12] + g[14];
      const float rho_in = (s_known + 2.0f * s_out) / (1.0f - ux_in);
      const float n_y = half * ((g[3] - g[4]) + (g[15] - g[16]) + (g[17] - g[18])) -
                        third * rho_in * uy_in;
      const float n_z = half * ((g[5] - g[6]) + (g[15] + g[16]) - (g[17] + g[18]));
      g[1] = g[2] + third * rho_in * ux_in;
      g[7] = g[10] + sixth * rho_in * (ux_in + uy_in) - n_y;
      g[9] = g[8] + sixth * rho_in * (ux_in - uy_in) + n_y;
      g[11] = g[14] + sixth * rho_in * ux_in - n_z;
      g[13] = g[12] + sixth * rho_in * ux_in + n_z;
    }
}

void ref_set_inlet(ref_sim* s, int rule) { s->inlet_rule = rule ? 1 : 0; }

/* K6 _bc_outlet, clbm_moving.py:296-300 */
static void k6_bc_outlet(ref_sim* s) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < s->ny; ++j)
    for (int k = 0; k < s->nz; ++k)
      memcpy(s->f + cidx(s, s->nx - 1, j, k) * Q, s->f + cidx(s, s->nx - 2, j, k) * Q,
             sizeof(float) * Q);
}

/* K7 _macro_post, clbm_moving.py:305-321 */
static void k7_macro_post(ref_sim* s) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)s->ncell; ++c) {
    const float* fc = s->f + c * Q;
    float r = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
    for (int q = 0; q < Q; ++q) {
      r += fc[q];
      mx += EX[q] * fc[q];
      my += EY[q] * fc[q];
      mz += EZ[q] * fc[q];
    }
    s->rho[c] = r;
    s->ux[c] = mx / r;
    s->uy[c] = my / r;
    s->uz[c] = mz / r;
  }
}

/* CLBMSolver.step, clbm_moving.py:326-343 */
void ref_step(ref_sim* s) {
  k2_macro_with_force(s);
  k3_collide(s);
  k4_stream(s);
  if (s->inlet_rule)
    k5_bc_inlet_hecht_harting(s);
  else
    k5_bc_inlet(s);
  k6_bc_outlet(s);
  k7_macro_post(s);
}

/* Peskin 4-point kernel, airfoil_moving.py:27-38 */
static inline float phi1d(float r) {
  const float a = fabsf(r);
  float res = 0.0f;
  if (a < 1.0f)
    res = (3.0f - 2.0f * a + sqrtf(fmaxf(1.0f + 4.0f * a - 4.0f * a * a, 0.0f))) / 8.0f;
  else if (a < 2.0f)
    res = (5.0f - 2.0f * a - sqrtf(fmaxf(-7.0f + 12.0f * a - 4.0f * a * a, 0.0f))) / 8.0f;
  return res;
}

static inline int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

/* DroneMovingIBM.compute_force, airfoil_moving.py:230-255 */
void ref_compute_force(ref_sim* s, float xc, float yc, float theta, float vx, float vy,
                       float omega) {
  const int n = s->n, nx = s->nx, ny = s->ny, nz = s->nz;
  const float ct = (float)cos((double)theta), st = (float)sin((double)theta);
  /* K8 */
#pragma omp parallel for schedule(static)
  for (int m = 0; m < n; ++m) {
    const float xk = (xc + s->X0[m] * ct) - s->Y0[m] * st;
    const float yk = (yc + s->X0[m] * st) + s->Y0[m] * ct;
    s->X[m] = xk;
    s->Y[m] = yk;
    s->Z[m] = s->Z0[m];
    s->VX[m] = vx - omega * (yk - yc);
    s->VY[m] = vy + omega * (xk - xc);
  }
  /* K9 */
#pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)s->ncell; ++c) {
    s->fx[c] = 0.0f;
    s->fy[c] = 0.0f;
    s->fz[c] = 0.0f;
  }
  /* K10 + K11 */
#pragma omp parallel for schedule(static)
  for (int m = 0; m < n; ++m) {
    const float xk = s->X[m], yk = s->Y[m], zk = s->Z[m];
    const int ib = (int)floorf(xk) - 1, jb = (int)floorf(yk) - 1, kb = (int)floorf(zk) - 1;
    float us = 0.0f, vs = 0.0f;
    for (int di = 0; di < 4; ++di) {
      const int iu = ib + di, ii = iu < 0 ? 0 : (iu > nx - 1 ? nx - 1 : iu);
      const float phx = phi1d((float)(s->quirk ? ii : iu) - xk);
      for (int dj = 0; dj < 4; ++dj) {
        const int ju = jb + dj, jj = pmod(ju, ny);
        const float phy = phi1d((float)(s->quirk ? jj : ju) - yk);
        for (int dk = 0; dk < 4; ++dk) {
          const int ku = kb + dk, kk = pmod(ku, nz);
          const float phz = phi1d((float)(s->quirk ? kk : ku) - zk);
          const float w = phx * phy * phz;
          us += s->ux[cidx(s, ii, jj, kk)] * w;
          vs += s->uy[cidx(s, ii, jj, kk)] * w;
        }
      }
    }
    s->U[m] = us;
    s->V[m] = vs;
    const float ex_ = -(us - s->VX[m]), ey_ = -(vs - s->VY[m]);
    s->FX[m] = fmaxf(-s->f_max, fminf(s->f_max, ex_));
    s->FY[m] = fmaxf(-s->f_max, fminf(s->f_max, ey_));
  }
  /* K12 */
#pragma omp parallel for schedule(static)
  for (int m = 0; m < n; ++m) {
    const float xk = s->X[m], yk = s->Y[m], zk = s->Z[m], fxk = s->FX[m], fyk = s->FY[m];
    const int ib = (int)floorf(xk) - 1, jb = (int)floorf(yk) - 1, kb = (int)floorf(zk) - 1;
    for (int di = 0; di < 4; ++di) {
      const int iu = ib + di, ii = iu < 0 ? 0 : (iu > nx - 1 ? nx - 1 : iu);
      const float phx = phi1d((float)(s->quirk ? ii : iu) - xk);
      for (int dj = 0; dj < 4; ++dj) {
        const int ju = jb + dj, jj = pmod(ju, ny);
        const float phy = phi1d((float)(s->quirk ? jj : ju) - yk);
        for (int dk = 0; dk < 4; ++dk) {
          const int ku = kb + dk, kk = pmod(ku, nz);
          const float phz = phi1d((float)(s->quirk ? kk : ku) - zk);
          const float w = phx * phy * phz;
          const size_t c = cidx(s, ii, jj, kk);
          const float px = fxk * w, py = fyk * w;
#pragma omp atomic
          s->fx[c] += px;
#pragma omp atomic
          s->fy[c] += py;
        }
      }
    }
  }
  /* K13 */
#pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)s->ncell; ++c) {
    s->fx[c] = fmaxf(-s->f_max, fminf(s->f_max, s->fx[c]));
    s->fy[c] = fmaxf(-s->f_max, fminf(s->f_max, s->fy[c]));
    s->fz[c] = 0.0f;
  }
}

/* data access for the Python wrapper */
void ref_get_f(const ref_sim* s, float* out) { memcpy(out, s->f, sizeof(float) * s->ncell * Q); }
void ref_set_f(ref_sim* s, const float* in) {
  memcpy(s->f, in, sizeof(float) * s->ncell * Q);
  k7_macro_post(s);
}
void ref_get_macro(const ref_sim* s, float* rho, float* ux, float* uy, float* uz) {
  memcpy(rho, s->rho, sizeof(float) * s->ncell);
  memcpy(ux, s->ux, sizeof(float) * s->ncell);
  memcpy(uy, s->uy, sizeof(float) * s->ncell);
  memcpy(uz, s->uz, sizeof(float) * s->ncell);
}
void ref_get_forces(const ref_sim* s, float* fx, float* fy) {
  memcpy(fx, s->fx, sizeof(float) * s->ncell);
  memcpy(fy, s->fy, sizeof(float) * s->ncell);
}
void ref_get_markers(const ref_sim* s, float* out /* 9 x n */) {
  const float* src[9] = {s->X, s->Y, s->Z, s->VX, s->VY, s->U, s->V, s->FX, s->FY};
  for (int i = 0; i < 9; ++i) memcpy(out + (size_t)i * s->n, src[i], sizeof(float) * s->n);
}
/* FP64 force sums over markers (dV nullable = 1), sequential order */
void ref_force_sum(const ref_sim* s, const float* dV, double* out2) {
  double ax = 0.0, ay = 0.0;
  for (int m = 0; m < s->n; ++m) {
    const double d = dV ? (double)dV[m] : 1.0;
    ax += (double)s->FX[m] * d;
    ay += (double)s->FY[m] * d;
  }
  out2[0] = ax;
  out2[1] = ay;
}
