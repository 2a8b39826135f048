// Cascaded (central-moment) D3Q19 collision with Guo forcing, closed form.
//
// Same mathematics as CLBMSolver._collide (clbm_moving.py:120-249), arranged
// for a bandwidth-bound kernel:
//  * Only the six second-order central moments survive the reference's
//    cascade (third order is set to 0 and fourth order to rho/9 there,
//    clbm_moving.py:188-196), so they are formed from ten raw moments of f
//    and shifted by u, instead of summing fifteen central products per q.
//  * The reference's dense 19x19 product with M_raw^-1 (clbm_moving.py:
//    233-238) is replaced by its closed form: every entry of that inverse is
//    0, +-1, +-1/2 or +-1/4 and each population depends on at most seven raw
//    moments (edge populations on four, face populations on six).
//  * Guo's term w_q (1 - s_v/2) [3 (e_q - u) + 9 (e_q.u) e_q] . F
//    (clbm_moving.py:240-249) is added per direction.
// Results agree with the reference formula to rounding; the order of the
// floating-point operations differs.  FORCE = false drops the forcing terms
// of a force-free launch (IEEE rules keep u * 0 and x + 0 in the code
// otherwise), and a cell whose force is zero skips Guo's term at run time;
// either way the result is the same except for the sign of zero sums.
#pragma once

#include "clbm_common.h"

namespace clbm {

// 1/x.  In float32: the hardware reciprocal (1 ulp) refined by one
// Newton-Raphson step, r + r (1 - x r), three instructions instead of the
// dozen of an IEEE division; within an ulp of the rounded quotient.
template <typename R>
__device__ __forceinline__ R recip(R x) {
  if constexpr (sizeof(R) == sizeof(float)) {
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
  } else {
    return R(1) / x;
  }
}

template <typename R, bool FORCE = true>
__device__ __forceinline__ void collide_cell(R* f, R Fx, R Fy, R Fz, R s_v, R s_e) {
  const R half = R(0.5), third = R(1) / R(3);
  // 1/4 is no inline constant: kept in a register of its own, so that the
  // 16-bit formats' f_q - base_q (one FMA with the scalar base) needs no copy
  // of the base into a vector register per direction
  R quarter = R(0.25);
  asm("" : "+v"(quarter));
  // ---- raw moments of the post-stream populations ----
  const R sxy = (f[7] + f[8]) + (f[9] + f[10]);
  const R sxz = (f[11] + f[12]) + (f[13] + f[14]);
  const R syz = (f[15] + f[16]) + (f[17] + f[18]);
  const R m200 = (f[1] + f[2]) + sxy + sxz;
  const R m020 = (f[3] + f[4]) + sxy + syz;
  const R m002 = (f[5] + f[6]) + sxz + syz;
  const R rho = f[0] + (f[1] + f[2]) + (f[3] + f[4]) + (f[5] + f[6]) + sxy + sxz + syz;
  const R mx = (f[1] - f[2]) + ((f[7] - f[8]) + (f[9] - f[10])) + ((f[11] - f[12]) + (f[13] - f[14]));
  const R my = (f[3] - f[4]) + ((f[7] + f[8]) - (f[9] + f[10])) + ((f[15] - f[16]) + (f[17] - f[18]));
  const R mz = (f[5] - f[6]) + ((f[11] + f[12]) - (f[13] + f[14])) + ((f[15] + f[16]) - (f[17] + f[18]));
  const R m110 = (f[7] - f[8]) - (f[9] - f[10]);
  const R m101 = (f[11] - f[12]) - (f[13] - f[14]);
  const R m011 = (f[15] - f[16]) - (f[17] - f[18]);

  // ---- half-force shifted velocity (clbm_moving.py:113-115) ----
  const R inv = recip(rho);
  const R u = (FORCE ? mx + half * Fx : mx) * inv;
  const R v = (FORCE ? my + half * Fy : my) * inv;
  const R w = (FORCE ? mz + half * Fz : mz) * inv;
  const R ru = rho * u, rv = rho * v, rw = rho * w;

  // ---- second-order central moments about u (first central moments are -F/2) ----
  const R k200 = m200 - u * (R(2) * mx - ru);
  const R k020 = m020 - v * (R(2) * my - rv);
  const R k002 = m002 - w * (R(2) * mz - rw);
  const R k110 = m110 - u * my - v * mx + ru * v;
  const R k101 = m101 - u * mz - w * mx + ru * w;
  const R k011 = m011 - v * mz - w * my + rv * w;

  // ---- cascade relaxation (clbm_moving.py:172-196) ----
  const R om = R(1) - s_v;
  const R a110 = om * k110, a101 = om * k101, a011 = om * k011;
  const R kA = k200 + k020 + k002;
  const R kAs = kA - s_e * (kA - rho);
  const R kDs = om * (k200 - k020);
  const R kEs = om * (k200 - k002);
  const R a200 = (kAs + kDs + kEs) * third;
  const R a020 = (kAs - R(2) * kDs + kEs) * third;
  const R a002 = (kAs + kDs - R(2) * kEs) * third;

  // ---- post-collision raw moments T (central -> raw shift with zero first
  //      central moments, third-order central 0, fourth-order rho/9) ----
  const R uu = u * u, vv = v * v, ww = w * w;
  const R T110 = a110 + ru * v;
  const R T101 = a101 + ru * w;
  const R T011 = a011 + rv * w;
  const R T200 = a200 + ru * u;
  const R T020 = a020 + rv * v;
  const R T002 = a002 + rw * w;
  const R T210 = R(2) * u * a110 + v * a200 + ru * u * v;
  const R T120 = R(2) * v * a110 + u * a020 + ru * v * v;
  const R T201 = R(2) * u * a101 + w * a200 + ru * u * w;
  const R T021 = R(2) * v * a011 + w * a020 + rv * v * w;
  const R T102 = R(2) * w * a101 + u * a002 + ru * w * w;
  const R T012 = R(2) * w * a011 + v * a002 + rv * w * w;
  const R r9 = rho * (R(1) / R(9));
  const R T220 = r9 + R(4) * u * v * a110 + vv * a200 + uu * a020 + ru * u * vv;
  const R T202 = r9 + R(4) * u * w * a101 + ww * a200 + uu * a002 + ru * u * ww;
  const R T022 = r9 + R(4) * v * w * a011 + ww * a020 + vv * a002 + rv * v * ww;

  // ---- closed-form M_raw^-1 T ----
  f[0] = rho - (T200 + T020 + T002) + (T220 + T202 + T022);
  {
    const R a = T200 - T220 - T202, b = ru - T120 - T102;
    f[1] = half * (a + b);
    f[2] = half * (a - b);
  }
  {
    const R a = T020 - T220 - T022, b = rv - T210 - T012;
    f[3] = half * (a + b);
    f[4] = half * (a - b);
  }
  {
    const R a = T002 - T202 - T022, b = rw - T201 - T021;
    f[5] = half * (a + b);
    f[6] = half * (a - b);
  }
  {
    const R p = T220 + T210, m = T220 - T210, s = T120 + T110, d = T120 - T110;
    f[7] = quarter * (p + s);
    f[8] = quarter * (p - s);
    f[9] = quarter * (m + d);
    f[10] = quarter * (m - d);
  }
  {
    const R p = T202 + T201, m = T202 - T201, s = T102 + T101, d = T102 - T101;
    f[11] = quarter * (p + s);
    f[12] = quarter * (p - s);
    f[13] = quarter * (m + d);
    f[14] = quarter * (m - d);
  }
  {
    const R p = T022 + T021, m = T022 - T021, s = T012 + T011, d = T012 - T011;
    f[15] = quarter * (p + s);
    f[16] = quarter * (p - s);
    f[17] = quarter * (m + d);
    f[18] = quarter * (m - d);
  }

  // ---- Guo forcing ----
  if (!FORCE || (Fx == R(0) && Fy == R(0) && Fz == R(0))) return;
  const R c = R(1) - half * s_v;
  const R uF = u * Fx + v * Fy + w * Fz;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const R eF = R(ex(q)) * Fx + R(ey(q)) * Fy + R(ez(q)) * Fz;
    const R eu = R(ex(q)) * u + R(ey(q)) * v + R(ez(q)) * w;
    f[q] += (wq<R>(q) * c) * (R(3) * (eF - uF) + R(9) * eu * eF);
  }
}

}  // namespace clbm
