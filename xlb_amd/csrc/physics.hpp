// Per-cell LBM physics on register-resident populations f[Q] (compute type C).
// Each function states the reference routine whose arithmetic it restates; the order of the
// floating-point accumulations follows the reference's loops so FP32 results agree to
// rounding.
#pragma once

#include "common.hpp"

namespace xlbh {

// rho = sum_l f_l ; u_d = (sum_l c_dl f_l) / rho.
// xlb/operator/macroscopic/zero_moment.py:22-27, first_moment.py:25-38.
// RCP: u_d = (sum_l c_dl f_l) * (1 / rho), one division instead of d.
template <int LAT, class C, bool RCP = false>
__device__ __forceinline__ void macroscopic(const C* f, C& rho, C* u) {
  constexpr LatTables T = make_tables(LAT);
  rho = C(0);
#pragma unroll
  for (int l = 0; l < T.q; ++l) rho += f[l];
#pragma unroll
  for (int d = 0; d < 3; ++d) u[d] = C(0);
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
#pragma unroll
    for (int d = 0; d < T.d; ++d) {
      if (T.c[d][l] == 1) u[d] += f[l];
      else if (T.c[d][l] == -1) u[d] -= f[l];
    }
  }
  if constexpr (RCP) {
    const C r = C(1) / rho;
#pragma unroll
    for (int d = 0; d < T.d; ++d) u[d] = u[d] * r;
  } else {
#pragma unroll
    for (int d = 0; d < T.d; ++d) u[d] = u[d] / rho;
  }
}

// 1.5 |u|^2 (the usqr term of quadratic_equilibrium.py:52-53).
template <int LAT, class C>
__device__ __forceinline__ C usqr15(const C* u) {
  constexpr LatTables T = make_tables(LAT);
  C s = C(0);
#pragma unroll
  for (int d = 0; d < T.d; ++d) s += u[d] * u[d];
  return C(1.5) * s;
}

// feq_l = rho w_l (1 + cu (1 + cu/2) - 1.5|u|^2), cu = 3 c_l.u.
// xlb/operator/equilibrium/quadratic_equilibrium.py:35-60.
template <int LAT, class C>
__device__ __forceinline__ C feq_l(int l, C rho, const C* u, C usqr) {
  constexpr LatTables T = make_tables(LAT);
  C cu = C(0);
#pragma unroll
  for (int d = 0; d < T.d; ++d) {
    if (T.c[d][l] == 1) cu += u[d];
    else if (T.c[d][l] == -1) cu -= u[d];
  }
  cu *= C(3);
  return rho * C(T.w[l]) * (C(1) + cu * (C(1) + C(0.5) * cu) - usqr);
}

// Second moment of (f - feq): Pi_t = sum_l cc_lt (f_l - feq_l).  second_moment.py:68-78.
template <int LAT, class C>
__device__ __forceinline__ void pi_neq(const C* f, C rho, const C* u, C usqr, C* pi) {
  constexpr LatTables T = make_tables(LAT);
#pragma unroll
  for (int t = 0; t < 6; ++t) pi[t] = C(0);
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    const C fneq = f[l] - feq_l<LAT>(l, rho, u, usqr);
#pragma unroll
    for (int t = 0; t < T.nt; ++t) {
      if (T.cc[l][t] == 1) pi[t] += fneq;
      else if (T.cc[l][t] == -1) pi[t] -= fneq;
    }
  }
}

// KBC shear part s_l of the non-equilibrium populations from Pi (kbc.py:192-250):
// axis directions get the normal-stress differences, in-plane diagonals +-Pi_ab/4, the rest 0.
template <int LAT, class C>
__device__ __forceinline__ C kbc_shear(int l, const C* pi) {
  constexpr LatTables T = make_tables(LAT);
  const int ax = T.c[0][l], ay = T.c[1][l], az = T.c[2][l];
  const int s = (ax != 0) + (ay != 0) + (az != 0);
  if constexpr (T.d == 3) {
    const C nxz = pi[0] - pi[5];
    const C nyz = pi[3] - pi[5];
    if (s == 1) {
      if (ax != 0) return (C(2) * nxz - nyz) / C(6);
      if (ay != 0) return (-nxz + C(2) * nyz) / C(6);
      return (-nxz - nyz) / C(6);
    }
    if (s == 2) {
      if (az == 0) return (ax * ay > 0) ? pi[1] / C(4) : -pi[1] / C(4);
      if (ay == 0) return (ax * az > 0) ? pi[2] / C(4) : -pi[2] / C(4);
      return (ay * az > 0) ? pi[4] / C(4) : -pi[4] / C(4);
    }
    return C(0);
  } else {
    const C n = pi[0] - pi[2];
    if (s == 1) return (ax != 0) ? n : -n;
    if (s == 2) return (ax * ay > 0) ? pi[1] : -pi[1];
    return C(0);
  }
}

// BGK as f_l <- (1 - omega) f_l + omega feq_l with each direction and its opposite sharing the
// even part of the equilibrium: omega feq_{l, opp l} = s_l +/- b_l, where
// s_l = omega rho w_l (1 - 1.5|u|^2 + cu_l^2 / 2) and b_l = omega rho w_l cu_l, cu_l = 3 c_l.u.
// The same update as bgk.py:30-34 in fewer FP32 operations (about half), rounded differently;
// used where populations are stored in 16 bits, whose rounding is far coarser than FP32's.
template <int LAT, class C>
__device__ __forceinline__ void bgk_paired(C* f, C rho, const C* u, C usqr, C omega) {
  constexpr LatTables T = make_tables(LAT);
  const C keep = C(1) - omega;
  const C orho = omega * rho;
  const C a = orho * (C(1) - usqr);
  C u3[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) u3[d] = C(3) * u[d];
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    const int o = T.opp[l];
    if (o < l) continue;  // updated with its partner
    const C w = C(T.w[l]);
    if (o == l) {
      f[l] = __builtin_fmaf(keep, f[l], a * w);
      continue;
    }
    C cu = C(0);
#pragma unroll
    for (int d = 0; d < T.d; ++d) {
      if (T.c[d][l] == 1) cu += u3[d];
      else if (T.c[d][l] == -1) cu -= u3[d];
    }
    const C ow = orho * w, hw = C(0.5) * ow;  // per weight class
    const C s = __builtin_fmaf(hw * cu, cu, a * w);
    f[l] = __builtin_fmaf(keep, f[l], __builtin_fmaf(ow, cu, s));
    f[o] = __builtin_fmaf(keep, f[o], __builtin_fmaf(-ow, cu, s));
  }
}

// Internal collision id, not an ABI value: BGK evaluated on deviations from the rest state
// (XLBH_COMPUTE_F32_DEVIATION).
constexpr int kBgkDeviation = 100;

// Every lattice weight is k_l / D with a small integer k_l.
constexpr int weight_den(int lat) { return lat == XLBH_D3Q27 ? 216 : 36; }

// BGK in FP32 on deviations from the rest state (bgk.py:30-34 and exact_difference_force.py:71-86,
// rearranged so that every rounding is that of a small quantity):
//   g_l = f_l - w_l (exact for w_l / 2 <= f_l <= 2 w_l), drho = sum_l g_l, j = sum_l c_l g_l
//   (accumulated in FP64, where these sums of FP32 values are exact),
//   rho = 1 + drho, u = j / rho, cu_l = 3 c_l.u,
//   geq_l = feq_l - w_l = (k_l / D) (drho + rho (cu_l (1 + cu_l / 2) - 1.5 |u|^2)),
//   f_l <- w_l + (g_l + omega (geq_l - g_l)).
// The weight k_l / D is applied as a float pair (hi + lo, exact to about 2^-48): with float weights
// the equilibrium's mass and momentum are off by a fixed relative ~1e-8 in every step, an error
// that adds up coherently over a long run.  omega multiplies only geq - g, which carries no mass
// or momentum, so its rounding shifts the viscosity by ~1e-8 and conserves both.  The force term
// feq(rho, u + F) - feq(rho, u) is written as rho w_l (cF + cF (cu + cF / 2) - 1.5 (2 u.F + |F|^2)),
// cF = 3 c_l.F.
template <int LAT>
__device__ __forceinline__ void bgk_deviation(float* f, double omega, bool has_force, const float* F) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  constexpr int D = weight_den(LAT);
  // exact moment sums: u and rho - 1 carry one rounding each
  float g[Q];
  double dr = 0.0, jd[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    g[l] = f[l] - (float)T.w[l];
    dr += (double)g[l];
#pragma unroll
    for (int d = 0; d < T.d; ++d) {
      if (T.c[d][l] == 1) jd[d] += (double)g[l];
      else if (T.c[d][l] == -1) jd[d] -= (double)g[l];
    }
  }
  const float drho = (float)dr;
  const float rho = (float)(1.0 + dr);
  float u[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int d = 0; d < T.d; ++d) u[d] = (float)(jd[d] / (1.0 + dr));
  const float us = usqr15<LAT>(u);
  const float om = (float)omega;
  float fu = 0.0f, ff = 0.0f;
  if (has_force) {
#pragma unroll
    for (int d = 0; d < T.d; ++d) {
      fu += u[d] * F[d];
      ff += F[d] * F[d];
    }
  }
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    const double wk = (double)(int)(T.w[l] * D + 0.5) / (double)D;  // w_l, exactly k_l / D
    const float hi = (float)wk, lo = (float)(wk - (double)hi);
    float cu = 0.0f, cf = 0.0f;
#pragma unroll
    for (int d = 0; d < T.d; ++d) {
      if (T.c[d][l] == 1) {
        cu += u[d];
        cf += F[d];
      } else if (T.c[d][l] == -1) {
        cu -= u[d];
        cf -= F[d];
      }
    }
    cu *= 3.0f;
    const float y = __builtin_fmaf(rho, __builtin_fmaf(0.5f * cu, cu, cu) - us, drho);
    const float geq = __builtin_fmaf(y, hi, y * lo);
    float gn = __builtin_fmaf(om, geq - g[l], g[l]);
    if (has_force) {
      cf *= 3.0f;
      const float df = cf + cf * (cu + 0.5f * cf) - 1.5f * (2.0f * fu + ff);
      gn = __builtin_fmaf(rho * (float)T.w[l], df, gn);
    }
    f[l] = (float)T.w[l] + gn;
  }
}

// In-place collision f <- collide(f) with optional exact-difference body force F (delta u).
//   BGK: f - omega (f - feq)                                      bgk.py:30-34
//   KBC: f - beta (2 ds + gamma dh)                               kbc.py:268-296
//   force: f += feq(rho, u + F) - feq(rho, u)                     exact_difference_force.py:71-86
// SHORT (FP32 arithmetic for 16-bit stores): one reciprocal of rho, and BGK in the paired form.
template <int LAT, int COLL, class C, bool SHORT = false>
__device__ __forceinline__ void collide(C* f, C omega, bool has_force, const C* F, C cparam = C(0)) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  C rho, u[3];
  macroscopic<LAT, C, SHORT>(f, rho, u);
  const C usqr = usqr15<LAT>(u);
  if constexpr (COLL == XLBH_BGK && SHORT) {
    bgk_paired<LAT>(f, rho, u, usqr, omega);
  } else if constexpr (COLL == XLBH_BGK) {
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      const C fe = feq_l<LAT>(l, rho, u, usqr);
      f[l] = f[l] - omega * (f[l] - fe);
    }
  } else if constexpr (COLL == XLBH_SMAGORINSKY_BGK) {
    // BGK with the Smagorinsky eddy relaxation time (smagorinsky_les_bgk.py:78-94): "strain"
    // sums fneq^2 over the directions whose component sum is 1 and 2 fneq^2 over those whose
    // component sum is >= 2, as the reference does; tau = tau0 + (sqrt(tau0^2 + 36 Cs^2
    // sqrt(strain)) - tau0) / 2.
    C strain = C(0);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      const int cs = T.c[0][l] + T.c[1][l] + T.c[2][l];
      const C fneq = f[l] - feq_l<LAT>(l, rho, u, usqr);
      if (cs == 1) strain += fneq * fneq;
      if (cs >= 2) strain += C(2) * fneq * fneq;
    }
    const C tau0 = C(1) / omega;
    const C tau = tau0 + C(0.5) * (sqrt(tau0 * tau0 + C(36) * (cparam * cparam) * sqrt(strain)) - tau0);
    const C inv_tau = C(1) / tau;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      const C fe = feq_l<LAT>(l, rho, u, usqr);
      f[l] = f[l] - inv_tau * (f[l] - fe);
    }
  } else {
    C pi[6];
    pi_neq<LAT>(f, rho, u, usqr, pi);
    const C beta = C(0.5) * omega;
    const C inv_beta = C(1) / beta;
    C sp1 = C(0), sp2 = C(0);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      const C fe = feq_l<LAT>(l, rho, u, usqr);
      const C fneq = f[l] - fe;
      C ds = kbc_shear<LAT>(l, pi) * rho;
      if constexpr (T.d == 2) ds = ds / C(4);
      const C dh = fneq - ds;
      const C tmp = dh / fe;
      sp1 += tmp * ds;
      sp2 += tmp * dh;
    }
    const C gamma = inv_beta - (C(2) - inv_beta) * sp1 / (C(1e-32) + sp2);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      const C fe = feq_l<LAT>(l, rho, u, usqr);
      const C fneq = f[l] - fe;
      C ds = kbc_shear<LAT>(l, pi) * rho;
      if constexpr (T.d == 2) ds = ds / C(4);
      const C dh = fneq - ds;
      f[l] = f[l] - beta * (C(2) * ds + gamma * dh);
    }
  }
  if (has_force) {
    C uf[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) uf[d] = u[d] + F[d];
    const C usqr_f = usqr15<LAT>(uf);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      const C fe = feq_l<LAT>(l, rho, u, usqr);
      const C ff = feq_l<LAT>(l, rho, uf, usqr_f);
      f[l] += ff - fe;
    }
  }
}

// ------------------------------------------------------------------------------------------
// Boundary helpers (xlb/operator/boundary_condition/helper_functions_bc.py).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool bit(uint32_t m, int l) { return (m >> l) & 1u; }

// Normal = -c of the first missing axis direction in lattice order (helper_functions_bc.py:75-86);
// zero when no axis direction is missing.
template <int LAT>
__device__ __forceinline__ void bc_normal(uint32_t mm, int* n) {
  constexpr LatTables T = make_tables(LAT);
  n[0] = n[1] = n[2] = 0;
#pragma unroll
  for (int l = T.q - 1; l >= 0; --l) {
    if (T.main[l] && bit(mm, l)) {
      n[0] = -T.c[0][l];
      n[1] = -T.c[1][l];
      n[2] = -T.c[2][l];
    }
  }
}

// 2 * sum(known) + sum(middle) (helper_functions_bc.py:62-73).
template <int LAT, class C>
__device__ __forceinline__ C bc_fsum(const C* f, uint32_t mm) {
  constexpr LatTables T = make_tables(LAT);
  C known = C(0), middle = C(0);
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    if (bit(mm, T.opp[l])) known += C(2) * f[l];
    else if (!bit(mm, l)) middle += f[l];
  }
  return known + middle;
}

// Regularisation of f (in place) around the equilibrium of its own moments:
// f_l = feq_l(rho, u) + 4.5 w_l Q_l : Pi_neq(f), with rho, u taken from f.  The projection of
// bc_grads_approximation.py:112-135 (regularize_fpop), applied after the moments are recomputed
// (:203-208).
template <int LAT, class C>
__device__ __forceinline__ void regularize_own_moments(C* f) {
  constexpr LatTables T = make_tables(LAT);
  C rho, u[3];
  macroscopic<LAT>(f, rho, u);
  const C usqr = usqr15<LAT>(u);
  C pi[6];
  pi_neq<LAT>(f, rho, u, usqr, pi);
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    C qp = C(0);
#pragma unroll
    for (int t = 0; t < T.nt; ++t) qp += C(T.qi[l][t]) * pi[t];
    f[l] = feq_l<LAT>(l, rho, u, usqr) + C(4.5) * C(T.w[l]) * qp;
  }
}

// Grad's-approximation wall (bc_grads_approximation.py:181-210, the reference's active
// functional_method1): every missing l gets the interpolated bounce-back with the fixed wall
// weight q = 1/2, ((1 - q) f_post[opp l] + q (f_pre[l] + f_pre[opp l])) / (1 + q), in place and in
// ascending l (a later missing direction sees an earlier one's new value), then the populations
// are regularised around their own moments.  `pre(l)` returns the cell's pre-stream population l.
template <int LAT, class C, class Pre>
__device__ __forceinline__ void grads_cell(C* f, uint32_t mm, Pre&& pre) {
  constexpr LatTables T = make_tables(LAT);
  const C q = C(0.5);
  const C one = C(1);
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    if (bit(mm, l)) {
      const int o = T.opp[l];
      f[l] = ((one - q) * f[o] + q * (pre(l) + pre(o))) / (one + q);
    }
  }
  regularize_own_moments<LAT>(f);
}

// Zou-He / regularized functional on post-stream populations f (in place), with the
// prescribed value `pv` (normal speed or density).  bc_zouhe.py:279-338,
// bc_regularized.py:134-196, helper_functions_bc.py:88-122.
template <int LAT, class C>
__device__ __forceinline__ void zouhe_cell(C* f, uint32_t mm, C pv, bool velocity, bool regularize) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  int ni[3];
  bc_normal<LAT>(mm, ni);
  C n[3] = {C(ni[0]), C(ni[1]), C(ni[2])};
  const C fsum = bc_fsum<LAT>(f, mm);
  C rho, u[3];
  if (velocity) {
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = -pv * n[d];
    C un = C(0);
#pragma unroll
    for (int d = 0; d < T.d; ++d) un += u[d] * n[d];
    rho = fsum / (C(1) + un);
  } else {
    rho = pv;
    const C un = -C(1) + fsum / rho;
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = un * n[d];
  }
  const C usqr = usqr15<LAT>(u);
  // Non-equilibrium bounce-back, sequential in l as in helper_functions_bc.py:88-97.
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    if (bit(mm, l)) {
      const int o = T.opp[l];
      f[l] = f[o] + feq_l<LAT>(l, rho, u, usqr) - feq_l<LAT>(o, rho, u, usqr);
    }
  }
  if (regularize) {
    C pi[6];
    pi_neq<LAT>(f, rho, u, usqr, pi);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      C qp = C(0);
#pragma unroll
      for (int t = 0; t < T.nt; ++t) qp += C(T.qi[l][t]) * pi[t];
      f[l] = feq_l<LAT>(l, rho, u, usqr) + C(4.5) * C(T.w[l]) * qp;
    }
  }
}

}  // namespace xlbh
