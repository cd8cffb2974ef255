#pragma once
// Reverse mode of the fused step (BGK; KBC and Smagorinsky BGK through collide_adjoint): given the
// state f^t that a forward step read and
// lam_out = dL/df^{t+1}, produce lam_in = dL/df^t and accumulate dL/dg for a per-cell body force.
//
// Forward per cell x (step.hpp, nse_stepper.py:344-383):  f_ps(x) = pull(f^t)  ->  h = BC_stream(f_ps)
//   -> out = BGK(h) (+ exact-difference force)  ->  BC_collision / extrapolation update  ->  f^{t+1}(x).
// Reverse per cell: lam_eff = lam_out(x) with the slots the BC overwrote after collision zeroed;
// G_h = (d out / d h)^T lam_eff (closed form below); then G_ps = (d h / d f_ps)^T G_h.
// The transpose of the pull is a push: f_ps_k(x) = f^t_k(x - c_k), so lam_in_k(x - c_k) = G_ps_k(x).
// Every slot of lam_in has exactly one pushing source, so kernel A writes lam_in as a permutation
// without atomics.  Contributions that do not follow the pull -- reads of the cell's own pre-stream
// populations (bounce-back, do-nothing, the extrapolation stash), extrapolation neighbour reads and
// the Zou-He / regularized Jacobian -- exist only at boundary cells and are added by kernel B.
//
// BGK with exact-difference force g: out_l = (1-w) h_l + a feq_l(rho, u) + [g] feq_l(rho, u+g),
// a = w - 1 with a force, w without.  With feq_l(rho, v) = rho w_l P_l(v),
// P_l(v) = 1 + 3 c_l.v + 9/2 (c_l.v)^2 - 3/2 v.v, d rho/d h_k = 1, d v/d h_k = (c_k - u)/rho:
//   sum_l lam_l d feq_l(rho, v)/d h_k = A(v) + sum_d B_d(v) (c_dk - u_d),
//   A(v) = sum_l lam_l w_l P_l(v),  B_d(v) = sum_l lam_l w_l (3 c_dl + 9 (c_l.v) c_dl - 3 v_d),
//   G_h_k = (1-w) lam_k + a [A(u) + B(u).(c_k - u)] + [g] [A(u+g) + B(u+g).(c_k - u)],
//   dL/dg_d = rho B_d(u+g).
#include "step.hpp"

namespace xlbh {

// ------------------------------------------------------------------------------------------
// Forward-mode dual numbers (value, derivative), used to differentiate the Zou-He / regularized
// functional at the few cells that carry it.
// ------------------------------------------------------------------------------------------
template <class C>
struct Dual {
  C v, d;
  __device__ Dual() : v(0), d(0) {}
  template <class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type>
  __device__ Dual(S a) : v((C)a), d(0) {}
  __device__ Dual(C a, C b) : v(a), d(b) {}
  __device__ explicit operator C() const { return v; }
};
template <class C>
__device__ __forceinline__ Dual<C> operator+(Dual<C> a, Dual<C> b) { return {a.v + b.v, a.d + b.d}; }
template <class C>
__device__ __forceinline__ Dual<C> operator-(Dual<C> a, Dual<C> b) { return {a.v - b.v, a.d - b.d}; }
template <class C>
__device__ __forceinline__ Dual<C> operator-(Dual<C> a) { return {-a.v, -a.d}; }
template <class C>
__device__ __forceinline__ Dual<C> operator*(Dual<C> a, Dual<C> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <class C>
__device__ __forceinline__ Dual<C> operator/(Dual<C> a, Dual<C> b) {
  const C inv = C(1) / b.v;
  return {a.v * inv, (a.d - a.v * inv * b.d) * inv};
}
template <class C>
__device__ __forceinline__ Dual<C>& operator+=(Dual<C>& a, Dual<C> b) { return a = a + b; }
template <class C>
__device__ __forceinline__ Dual<C>& operator-=(Dual<C>& a, Dual<C> b) { return a = a - b; }
template <class C>
__device__ __forceinline__ Dual<C>& operator*=(Dual<C>& a, Dual<C> b) { return a = a * b; }

// ------------------------------------------------------------------------------------------
// Collision adjoint (BGK + optional force), in place: lam (Q, the post-collision adjoint) ->
// G_h (Q).  Also returns dL/dg in gg (D values).
// ------------------------------------------------------------------------------------------
// The weights enter as w_l = k_l / D (small integers k_l, one division by D per moment): with
// float weights A and B would carry a fixed relative rounding (sum_l float(w_l) - 1 is ~1.5e-8
// for D3Q19), which the reverse sweep would apply coherently in every step.
template <int LAT, class C>
__device__ __forceinline__ void eq_moments(const C* lam, const C* u, const C* v, C& A, C* B) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int D = weight_den(LAT);
  A = C(0);
  B[0] = B[1] = B[2] = C(0);
  C vv = C(0);
#pragma unroll
  for (int d = 0; d < T.d; ++d) vv += v[d] * v[d];
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    C cv = C(0);
#pragma unroll
    for (int d = 0; d < T.d; ++d) cv += C(T.c[d][l]) * v[d];
    const C lw = lam[l] * C((int)(T.w[l] * D + 0.5));
    A += lw * (C(1) + C(3) * cv + C(4.5) * cv * cv - C(1.5) * vv);
#pragma unroll
    for (int d = 0; d < T.d; ++d) B[d] += lw * (C(3) * C(T.c[d][l]) + C(9) * cv * C(T.c[d][l]) - C(3) * v[d]);
  }
  A = A / C(D);
#pragma unroll
  for (int d = 0; d < T.d; ++d) B[d] = B[d] / C(D);
  (void)u;
}

// Smagorinsky BGK (physics.hpp, smagorinsky_les_bgk.py:73-89) relaxes with k = 1/tau(S) instead of
// omega, S = sum_l a_l fneq_l^2 (a_l = 1 where the signed component sum of c_l is 1, 2 where it is
// >= 2): the BGK adjoint with k for omega plus the strain term
//   -(sum_l G_l fneq_l) dk/dS dS/dh_m,  dS/dh_m = 2 a_m fneq_m - (A + B.(c_m - u))[2 a fneq],
//   dk/dS = -k^2 9 Cs^2 / (2 sqrt(tau0^2 + 36 Cs^2 sqrt S) sqrt S)  (taken as 0 at S = 0).
// KBC (physics.hpp, kbc.py:192-296): out = h - beta (2 ds + gamma dh), fneq = h - feq(rho, u),
// Pi = cc.fneq, ds = rho s(Pi) (s / 4 on D2Q9), dh = fneq - ds,
// gamma = 1/beta - (2 - 1/beta) sp1 / (eps + sp2), sp1 = sum dh ds / feq, sp2 = sum dh^2 / feq.
// Reverse in closed form: through gamma (sp1, sp2 and their 1/feq weights), the shear map's
// transpose (s is linear in Pi), rho (ds = rho s and dRho/dh = 1) and feq(rho, u) via the
// equilibrium moments A + B.(c_m - u).
template <int LAT, class C>
__device__ __forceinline__ void kbc_adjoint(const C* h, C* lam, C omega, bool has_force, const C* F, C* gg) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  C rho, u[3];
  macroscopic<LAT>(h, rho, u);
  const C usqr = usqr15<LAT>(u);
  C pi[6];
  pi_neq<LAT>(h, rho, u, usqr, pi);
  const C sc = (T.d == 2) ? C(0.25) : C(1);
  const C beta = C(0.5) * omega;
  const C inv_beta = C(1) / beta;
  C fe[Q], sh[Q], ds[Q], dh[Q];
  C sp1 = C(0), sp2 = C(0);
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    fe[l] = feq_l<LAT>(l, rho, u, usqr);
    sh[l] = kbc_shear<LAT>(l, pi);
    ds[l] = sh[l] * rho * sc;
    dh[l] = (h[l] - fe[l]) - ds[l];
    const C tmp = dh[l] / fe[l];
    sp1 += tmp * ds[l];
    sp2 += tmp * dh[l];
  }
  const C D = C(1e-32) + sp2;
  const C c2 = C(2) - inv_beta;
  const C gamma = inv_beta - c2 * sp1 / D;
  C Gg = C(0);
#pragma unroll
  for (int l = 0; l < Q; ++l) Gg += lam[l] * dh[l];
  Gg *= -beta;
  const C Gsp1 = Gg * (-c2 / D);
  const C Gsp2 = Gg * (c2 * sp1 / (D * D));
  C Gfe[Q], Gfn[Q], Gs[Q];
  C Grho = C(0);
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    const C ife = C(1) / fe[l];
    C Gds = -C(2) * beta * lam[l] + Gsp1 * dh[l] * ife;
    const C Gdh = -beta * gamma * lam[l] + (Gsp1 * ds[l] + C(2) * Gsp2 * dh[l]) * ife;
    Gfe[l] = -(Gsp1 * dh[l] * ds[l] + Gsp2 * dh[l] * dh[l]) * ife * ife;
    Gfn[l] = Gdh;
    Gds -= Gdh;
    Gs[l] = Gds * rho * sc;
    Grho += Gds * sh[l] * sc;
  }
  // s is linear in Pi: dL/dPi_t = sum_l Gs_l s_l(e_t)
  C GP[6] = {C(0), C(0), C(0), C(0), C(0), C(0)};
#pragma unroll
  for (int t = 0; t < T.nt; ++t) {
    C e[6] = {C(0), C(0), C(0), C(0), C(0), C(0)};
    e[t] = C(1);
#pragma unroll
    for (int l = 0; l < Q; ++l) GP[t] += Gs[l] * kbc_shear<LAT>(l, e);
  }
#pragma unroll
  for (int m = 0; m < Q; ++m) {
#pragma unroll
    for (int t = 0; t < T.nt; ++t) {
      if (T.cc[m][t] == 1) Gfn[m] += GP[t];
      else if (T.cc[m][t] == -1) Gfn[m] -= GP[t];
    }
    Gfe[m] -= Gfn[m];  // fneq = h - feq
  }
  C A1 = C(0), B1[3] = {C(0), C(0), C(0)};
  if (has_force) {
    C ug[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) ug[d] = u[d] + F[d];
    eq_moments<LAT>(lam, u, ug, A1, B1);
#pragma unroll
    for (int l = 0; l < Q; ++l) Gfe[l] -= lam[l];  // the - feq(rho, u) of the exact difference
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) gg[d] = has_force ? rho * B1[d] : C(0);
  C A0, B0[3];
  eq_moments<LAT>(Gfe, u, u, A0, B0);
#pragma unroll
  for (int m = 0; m < Q; ++m) {
    C s = lam[m] + Gfn[m] + Grho + A0 + A1;
#pragma unroll
    for (int d = 0; d < T.d; ++d) s += (B0[d] + B1[d]) * (C(T.c[d][m]) - u[d]);
    lam[m] = s;
  }
}

template <int LAT, int COLL, class C>
__device__ __forceinline__ void collide_adjoint(const C* h, C* lam, C omega, bool has_force, const C* F, C* gg,
                                                C smag_cs = C(0)) {
  constexpr LatTables T = make_tables(LAT);
  if constexpr (COLL == XLBH_KBC) {
    kbc_adjoint<LAT>(h, lam, omega, has_force, F, gg);
    return;
  }
  constexpr bool smag = COLL == XLBH_SMAGORINSKY_BGK;
  C rho, u[3];
  macroscopic<LAT>(h, rho, u);
  C strain[T.q];
#pragma unroll
  for (int k = 0; k < T.q; ++k) strain[k] = C(0);
  if constexpr (smag) {
    const C usqr = usqr15<LAT>(u);
    C fneq[T.q], lamS[T.q];
    C S = C(0), Gf = C(0);
#pragma unroll
    for (int l = 0; l < T.q; ++l) {
      const int csum = T.c[0][l] + T.c[1][l] + T.c[2][l];
      const C al = csum == 1 ? C(1) : (csum >= 2 ? C(2) : C(0));
      fneq[l] = h[l] - feq_l<LAT>(l, rho, u, usqr);
      S += al * fneq[l] * fneq[l];
      lamS[l] = C(2) * al * fneq[l];
      Gf += lam[l] * fneq[l];
    }
    const C tau0 = C(1) / omega;
    const C sqS = sqrt(S);
    const C root = sqrt(tau0 * tau0 + C(36) * (smag_cs * smag_cs) * sqS);
    const C tau = tau0 + C(0.5) * (root - tau0);
    const C kk = C(1) / tau;
    if (S > C(0)) {
      const C dkdS = -kk * kk * C(9) * (smag_cs * smag_cs) / (C(2) * root * sqS);
      C AS, BS[3];
      eq_moments<LAT>(lamS, u, u, AS, BS);
#pragma unroll
      for (int m = 0; m < T.q; ++m) {
        C dS = lamS[m] - AS;
#pragma unroll
        for (int d = 0; d < T.d; ++d) dS -= BS[d] * (C(T.c[d][m]) - u[d]);
        strain[m] = -Gf * dkdS * dS;
      }
    }
    omega = kk;
  }
  const C a = has_force ? omega - C(1) : omega;
  C A0, B0[3];
  eq_moments<LAT>(lam, u, u, A0, B0);
  C A1 = C(0), B1[3] = {C(0), C(0), C(0)};
  if (has_force) {
    C ug[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) ug[d] = u[d] + F[d];
    eq_moments<LAT>(lam, u, ug, A1, B1);
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) gg[d] = has_force ? rho * B1[d] : C(0);
#pragma unroll
  for (int k = 0; k < T.q; ++k) {
    C s = a * A0 + A1;
#pragma unroll
    for (int d = 0; d < T.d; ++d) s += (a * B0[d] + B1[d]) * (C(T.c[d][k]) - u[d]);
    lam[k] = (C(1) - omega) * lam[k] + s + strain[k];
  }
}

// Reverse mode of zouhe_cell (physics.hpp) in one pass: given G = dL/dh for h = zouhe_cell(f),
// returns Gf = dL/df (zero for the missing directions, which the functional overwrites).
// Forward: fsum = sum_l a_l f_l (a = 2 opposite-of-missing, 1 middle, 0 missing); velocity:
// rho = fsum / (1 + u.n) with u prescribed; pressure: u = (-1 + fsum / rho) n with rho prescribed;
// missing l (in order): f_l = f_opp(l) + feq_l - feq_opp(l); regularized: f_l = feq_l +
// 4.5 w_l sum_t qi_lt Pi_t with Pi_t = sum_m cc_mt (f_m - feq_m).  The in-place updates are undone
// in reverse order; the equilibria's dependence on (rho, u) uses the moments A, B of eq_moments.
template <int LAT, class C>
__device__ __forceinline__ void zouhe_cell_adjoint(const C* f, uint32_t mm, C pv, bool velocity, bool regularize,
                                                   const C* G, C* Gf) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  int ni[3];
  bc_normal<LAT>(mm, ni);
  const C n[3] = {C(ni[0]), C(ni[1]), C(ni[2])};
  const C fsum = bc_fsum<LAT>(f, mm);
  C rho, u[3];
  C un = C(0);
  if (velocity) {
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = -pv * n[d];
#pragma unroll
    for (int d = 0; d < T.d; ++d) un += u[d] * n[d];
    rho = fsum / (C(1) + un);
  } else {
    rho = pv;
    un = -C(1) + fsum / rho;
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = un * n[d];
  }
  C Gfeq[Q];
  if (regularize) {
    C P[6] = {C(0), C(0), C(0), C(0), C(0), C(0)};
#pragma unroll
    for (int l = 0; l < Q; ++l)
#pragma unroll
      for (int t = 0; t < T.nt; ++t) P[t] += G[l] * C(4.5) * C(T.w[l]) * C(T.qi[l][t]);
#pragma unroll
    for (int m = 0; m < Q; ++m) {
      C g = C(0);
#pragma unroll
      for (int t = 0; t < T.nt; ++t) g += C(T.cc[m][t]) * P[t];
      Gf[m] = g;
      Gfeq[m] = G[m] - g;
    }
  } else {
#pragma unroll
    for (int m = 0; m < Q; ++m) {
      Gf[m] = G[m];
      Gfeq[m] = C(0);
    }
  }
#pragma unroll
  for (int l = Q - 1; l >= 0; --l) {
    if (bit(mm, l)) {
      // zero before adding: for l = opp(l) = 0 (missing[0] at padded interior cells) the slot
      // reads itself
      const int o = T.opp[l];
      const C g = Gf[l];
      Gf[l] = C(0);
      Gf[o] += g;
      Gfeq[l] += g;
      Gfeq[o] -= g;
    }
  }
  C A, B[3];
  eq_moments<LAT>(Gfeq, u, u, A, B);
  C gfsum;
  if (velocity) {
    gfsum = A / (C(1) + un);
  } else {
    C gun = C(0);
#pragma unroll
    for (int d = 0; d < T.d; ++d) gun += rho * B[d] * n[d];
    gfsum = gun / rho;
  }
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    if (bit(mm, T.opp[l])) Gf[l] += C(2) * gfsum;
    else if (!bit(mm, l)) Gf[l] += gfsum;
  }
}

// Reverse mode of grads_cell (physics.hpp): given G = dL/dh for h = grads_cell(fps, pre), returns
// Gf = dL/dfps and Gpre = dL/dpre (the cell's own pre-stream populations).  Forward: every missing l
// (ascending, in place) takes (f_o + pre_l + pre_o) / 3 (weight 1/2), then h_l = feq_l(rho, u) +
// 4.5 w_l sum_t qi_lt Pi_t with rho, u, Pi_t = sum_m cc_mt (f_m - feq_m) of the interpolated f.
// Reverse: P_t = 4.5 sum_l G_l w_l qi_lt; dL/df_m = sum_t cc_mt P_t + A + B.(c_m - u) with A, B the
// equilibrium moments (eq_moments) of G - cc P; then the interpolations undone in descending l.
template <int LAT, class C>
__device__ __forceinline__ void grads_cell_adjoint(const C* fps, const C* pre, uint32_t mm, const C* G, C* Gf,
                                                   C* Gpre) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  const C third = C(1) / C(3);
  C f[Q];
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    f[l] = fps[l];
    Gpre[l] = C(0);
  }
  const C q = C(0.5);
  const C one = C(1);
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    if (bit(mm, l)) {
      const int o = T.opp[l];
      f[l] = ((one - q) * f[o] + q * (pre[l] + pre[o])) / (one + q);
    }
  }
  C rho, u[3];
  macroscopic<LAT>(f, rho, u);
  C P[6] = {C(0), C(0), C(0), C(0), C(0), C(0)};
#pragma unroll
  for (int l = 0; l < Q; ++l)
#pragma unroll
    for (int t = 0; t < T.nt; ++t) P[t] += G[l] * C(4.5) * C(T.w[l]) * C(T.qi[l][t]);
  C lamq[Q];
#pragma unroll
  for (int m = 0; m < Q; ++m) {
    C g = C(0);
#pragma unroll
    for (int t = 0; t < T.nt; ++t) g += C(T.cc[m][t]) * P[t];
    Gf[m] = g;
    lamq[m] = G[m] - g;
  }
  C A, B[3];
  eq_moments<LAT>(lamq, u, u, A, B);
#pragma unroll
  for (int m = 0; m < Q; ++m) {
    C s = A;
#pragma unroll
    for (int d = 0; d < T.d; ++d) s += B[d] * (C(T.c[d][m]) - u[d]);
    Gf[m] += s;
  }
#pragma unroll
  for (int l = Q - 1; l >= 0; --l) {
    if (bit(mm, l)) {
      const int o = T.opp[l];
      const C g = Gf[l] * third;
      Gf[l] = C(0);  // before adding: l = opp(l) = 0 reads itself
      Gf[o] += g;
      Gpre[l] += g;
      Gpre[o] += g;
    }
  }
}

struct AdjParams {
  const void* f_in;
  const void* lam_out;
  void* lam_in;
  const uint8_t* bc_mask;
  const uint32_t* missing;
  const void* aux;
  const xlbh_bc_desc* table;
  const void* force_field;
  double* grad_force;
  const int32_t* bc_cells;  // flat indices of the boundary cells, or nullptr (scan the grid)
  int n_bc_cells;
  // Fused reverse step (workspace given): per-cell deltas carried to the next reverse step.
  void* d_own;           // (Q, N): contributions to dL/df_in at the producing cell itself
  void* d_nb_in;         // (Q, N): neighbour contributions produced by the previous reverse step
  void* d_nb_out;        // (Q, N): neighbour contributions produced by this reverse step
  const uint8_t* nb_flag;  // (N): 1 where d_nb entries can arrive
  double fu[3];
  int force_mode;
  double omega;
  double cparam;   // Smagorinsky coefficient
  int collision;   // XLBH_BGK or XLBH_SMAGORINSKY_BGK
  Sweep sw;
  uint32_t N;
  uint64_t P;  // plane stride (elements) of f_in; the adjoint fields are dense (Q, N)
};

// The populations the forward step pulled into cell (x, y, z), f_ps (periodic wrap).
template <int LAT, int FMT, class C>
__device__ __forceinline__ void adj_pull(const AdjParams& p, int x, int y, int z, C* fps) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  using R = typename Fmt<FMT>::raw;
  const R* fin = static_cast<const R*>(p.f_in);
  const int nx = p.sw.nx, ny = p.sw.ny, nz = p.sw.nz;
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    const int sx = wrap1(x - T.c[0][l], nx), sy = wrap1(y - T.c[1][l], ny), sz = wrap1(z - T.c[2][l], nz);
    fps[l] = ld_pop<FMT, C>(fin + (size_t)l * p.P + ((uint32_t)sx * ny + sy) * nz + sz, T.w[l]);
  }
}

// The streaming-step view of one cell: post-BC populations h, the adjoint of the output with the
// slots overwritten after collision removed, and the BC kind, from the pulled populations fps
// (adj_pull) and the cell's dL/df_out lamf.  Shared by every adjoint kernel.
template <int LAT, int FMT, class C>
__device__ __forceinline__ void adj_prepare_pulled(const AdjParams& p, uint32_t cell, int kind, uint32_t mm,
                                                   const xlbh_bc_desc* d, const C* fps, C* h, C* lam,
                                                   const C* lamf) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  using R = typename Fmt<FMT>::raw;
  const R* fin = static_cast<const R*>(p.f_in);
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    lam[l] = lamf[l];
    h[l] = fps[l];
  }
  if (kind == XLBH_BC_EQUILIBRIUM) {
    const C rb = (C)d->rho;
    const C ub[3] = {(C)d->u[0], (C)d->u[1], (C)d->u[2]};
    const C us = usqr15<LAT>(ub);
#pragma unroll
    for (int l = 0; l < Q; ++l) h[l] = feq_l<LAT>(l, rb, ub, us);
  } else if (kind == XLBH_BC_HALFWAY_BB || kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
#pragma unroll
    for (int l = 0; l < Q; ++l)
      if (bit(mm, l)) h[l] = ld_pop<FMT, C>(fin + (size_t)T.opp[l] * p.P + cell, T.w[T.opp[l]]);
  } else if (kind == XLBH_BC_DO_NOTHING) {
#pragma unroll
    for (int l = 0; l < Q; ++l) h[l] = ld_pop<FMT, C>(fin + (size_t)l * p.P + cell, T.w[l]);
  } else if (kind >= XLBH_BC_ZOUHE_VELOCITY && kind <= XLBH_BC_REGULARIZED_PRESSURE) {
    const C pv = p.aux != nullptr ? static_cast<const C*>(p.aux)[cell] : C(0);
    zouhe_cell<LAT>(h, mm, pv, kind == XLBH_BC_ZOUHE_VELOCITY || kind == XLBH_BC_REGULARIZED_VELOCITY,
                    kind >= XLBH_BC_REGULARIZED_VELOCITY);
  } else if (kind == XLBH_BC_GRADS_APPROXIMATION) {
    grads_cell<LAT>(h, mm, [&](int l) { return ld_pop<FMT, C>(fin + (size_t)l * p.P + cell, T.w[l]); });
  }
  if (kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
#pragma unroll
    for (int l = 0; l < Q; ++l)
      if (bit(mm, l)) lam[T.opp[l]] = C(0);  // collision output of slot opp(l) is replaced by the stash
  }
}

// adj_prepare_pulled with the pulls and dL/df_out loaded here.
template <int LAT, int FMT, class C>
__device__ __forceinline__ void adj_prepare(const AdjParams& p, uint32_t cell, int x, int y, int z, int kind,
                                            uint32_t mm, const xlbh_bc_desc* d, C* fps, C* h, C* lam) {
  constexpr int Q = make_tables(LAT).q;
  const C* lo = static_cast<const C*>(p.lam_out);
  C lamf[Q];
  adj_pull<LAT, FMT, C>(p, x, y, z, fps);
#pragma unroll
  for (int l = 0; l < Q; ++l) lamf[l] = lo[(size_t)l * p.N + cell];
  adj_prepare_pulled<LAT, FMT, C>(p, cell, kind, mm, d, fps, h, lam, lamf);
}

template <int LAT, class C>
__device__ __forceinline__ void cell_force(const AdjParams& p, uint32_t cell, C* F) {
  constexpr LatTables T = make_tables(LAT);
  F[0] = F[1] = F[2] = C(0);
  if (p.force_mode == XLBH_FORCE_UNIFORM) {
    F[0] = (C)p.fu[0];
    F[1] = (C)p.fu[1];
    F[2] = (C)p.fu[2];
  } else if (p.force_mode == XLBH_FORCE_FIELD) {
    const C* ff = static_cast<const C*>(p.force_field);
#pragma unroll
    for (int dd = 0; dd < T.d; ++dd) F[dd] = ff[(size_t)dd * p.N + cell];
  }
}

__device__ __forceinline__ bool heavy_kind(int kind) {
  return kind >= XLBH_BC_ZOUHE_VELOCITY && kind <= XLBH_BC_REGULARIZED_PRESSURE;
}

// The cell's own pre-stream populations (Grad's approximation reads them).
template <int LAT, int FMT, class C>
__device__ __forceinline__ void own_pre(const AdjParams& p, uint32_t cell, C* pre) {
  constexpr LatTables T = make_tables(LAT);
  const typename Fmt<FMT>::raw* fin = static_cast<const typename Fmt<FMT>::raw*>(p.f_in);
#pragma unroll
  for (int l = 0; l < T.q; ++l) pre[l] = ld_pop<FMT, C>(fin + (size_t)l * p.P + cell, T.w[l]);
}

// Kernel A: every cell pushes G_ps into lam_in (a permutation) and adds its force gradient.
template <int LAT, int FMT, class C, int COLL>
__global__ __launch_bounds__(256) void k_adjoint_push(AdjParams p) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  int x, y, z;
  if (!sweep_cell(p.sw, x, y, z)) return;
  uniform_slow_axes(T.d == 3, x, y);
  const int nx = p.sw.nx, ny = p.sw.ny, nz = p.sw.nz;
  const uint32_t cell = ((uint32_t)x * ny + y) * nz + z;
  int id = p.bc_mask ? p.bc_mask[cell] : 0;
  int kind = XLBH_BC_NONE;
  uint32_t mm = 0;
  const xlbh_bc_desc* d = nullptr;
  if (id != 0) {
    d = p.table + id;
    kind = d->kind;
    mm = p.missing[cell];
  }
  C fps[Q], h[Q], lam[Q];
  adj_prepare<LAT, FMT, C>(p, cell, x, y, z, kind, mm, d, fps, h, lam);
  (void)fps;
  C G[Q];
  if (id == 255) {
    // solid cells are not updated by the forward step: nothing flows back through their pulls
#pragma unroll
    for (int k = 0; k < Q; ++k) G[k] = C(0);
  } else if (kind == XLBH_BC_FULLWAY_BB) {
#pragma unroll
    for (int k = 0; k < Q; ++k) G[k] = lam[T.opp[k]];
  } else if (heavy_kind(kind) || kind == XLBH_BC_EQUILIBRIUM || kind == XLBH_BC_DO_NOTHING ||
             kind == XLBH_BC_GRADS_APPROXIMATION) {
    // no dependence on the pulled values through this cell's pull (Zou-He and Grad's are added by
    // kernel B)
#pragma unroll
    for (int k = 0; k < Q; ++k) G[k] = C(0);
    if (kind == XLBH_BC_EQUILIBRIUM && p.force_mode != XLBH_FORCE_NONE && p.grad_force) {
      C F[3], gg[3];
      cell_force<LAT>(p, cell, F);
      collide_adjoint<LAT, COLL>(h, lam, (C)p.omega, true, F, gg, (C)p.cparam);
#pragma unroll
      for (int dd = 0; dd < T.d; ++dd) p.grad_force[(size_t)dd * p.N + cell] += (double)gg[dd];
    }
  } else {
    C F[3], gg[3];
    cell_force<LAT>(p, cell, F);
    const bool has_force = p.force_mode != XLBH_FORCE_NONE;
#pragma unroll
    for (int k = 0; k < Q; ++k) G[k] = lam[k];
    collide_adjoint<LAT, COLL>(h, G, (C)p.omega, has_force, F, gg, (C)p.cparam);
    if (has_force && p.grad_force) {
#pragma unroll
      for (int dd = 0; dd < T.d; ++dd) p.grad_force[(size_t)dd * p.N + cell] += (double)gg[dd];
    }
    if (kind == XLBH_BC_HALFWAY_BB || kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
#pragma unroll
      for (int k = 0; k < Q; ++k)
        if (bit(mm, k)) G[k] = C(0);  // replaced by the own-cell value (kernel B)
    }
  }
  C* li = static_cast<C*>(p.lam_in);
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    const int sx = wrap1(x - T.c[0][k], nx), sy = wrap1(y - T.c[1][k], ny), sz = wrap1(z - T.c[2][k], nz);
    li[(size_t)k * p.N + ((uint32_t)sx * ny + sy) * nz + sz] = G[k];
  }
}

// Boundary-cell contributions that do not follow the pull, for one boundary cell.  `part` selects
// the work: -1 everything; 0 everything except the Zou-He / regularized Jacobian; 1 + k the
// Jacobian column of input direction k only (so a cell's columns can run on separate threads).
template <int LAT, int FMT, class C, int COLL>
__device__ __forceinline__ void adjoint_bc_cell(const AdjParams& p, uint32_t cell, int x, int y, int z, int part) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  const int nx = p.sw.nx, ny = p.sw.ny, nz = p.sw.nz;
  const int id = p.bc_mask[cell];
  if (id == 0 || id == 255) return;
  const xlbh_bc_desc* d = p.table + id;
  const int kind = d->kind;
  const bool grads = kind == XLBH_BC_GRADS_APPROXIMATION;
  if (kind != XLBH_BC_HALFWAY_BB && kind != XLBH_BC_EXTRAPOLATION_OUTFLOW && kind != XLBH_BC_DO_NOTHING &&
      !heavy_kind(kind) && !grads)
    return;
  const uint32_t mm = p.missing[cell];
  if (part > 0 && !heavy_kind(kind)) return;
  C fps[Q], h[Q], lam[Q];
  adj_prepare<LAT, FMT, C>(p, cell, x, y, z, kind, mm, d, fps, h, lam);
  C* li = static_cast<C*>(p.lam_in);
  auto add = [&](int k, uint32_t c, C v) { atomicAdd(li + (size_t)k * p.N + c, v); };
  C F[3], gg[3];
  cell_force<LAT>(p, cell, F);
  const bool has_force = p.force_mode != XLBH_FORCE_NONE;
  C G[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) G[k] = lam[k];
  collide_adjoint<LAT, COLL>(h, G, (C)p.omega, has_force, F, gg, (C)p.cparam);

  if (part <= 0) {
    if (kind == XLBH_BC_HALFWAY_BB || kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
#pragma unroll
      for (int l = 0; l < Q; ++l)
        if (bit(mm, l)) add(T.opp[l], cell, G[l]);  // h_l = f_own[opp l]
    }
    if (kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
      // stash of slot opp(l): (1 - cs) f_own[opp l] + cs f^t_l(x - c_l - n)
      int n[3];
      bc_normal<LAT>(mm, n);
      const C cs = C(0.57735026918962584);
      const C* lo = static_cast<const C*>(p.lam_out);
#pragma unroll
      for (int l = 0; l < Q; ++l) {
        if (bit(mm, l)) {
          const C ls = lo[(size_t)T.opp[l] * p.N + cell];
          add(T.opp[l], cell, (C(1) - cs) * ls);
          const int sx = wrap(x - (T.c[0][l] + n[0]), nx);
          const int sy = wrap(y - (T.c[1][l] + n[1]), ny);
          const int sz = wrap(z - (T.c[2][l] + n[2]), nz);
          add(l, ((uint32_t)sx * ny + sy) * nz + sz, cs * ls);
        }
      }
    }
    if (kind == XLBH_BC_DO_NOTHING) {
#pragma unroll
      for (int k = 0; k < Q; ++k) add(k, cell, G[k]);
    }
    if (grads) {
      // the whole Grad's Jacobian in one reverse pass: pulled values and own pre-stream values
      C pre[Q], Gf[Q], Gpre[Q];
      own_pre<LAT, FMT, C>(p, cell, pre);
      grads_cell_adjoint<LAT>(fps, pre, mm, G, Gf, Gpre);
#pragma unroll
      for (int k = 0; k < Q; ++k) {
        const int sx = wrap1(x - T.c[0][k], nx), sy = wrap1(y - T.c[1][k], ny), sz = wrap1(z - T.c[2][k], nz);
        add(k, ((uint32_t)sx * ny + sy) * nz + sz, Gf[k]);
        add(k, cell, Gpre[k]);
      }
    }
    if (heavy_kind(kind) || kind == XLBH_BC_DO_NOTHING || grads) {
      if (has_force && p.grad_force) {
#pragma unroll
        for (int dd = 0; dd < T.d; ++dd) p.grad_force[(size_t)dd * p.N + cell] += (double)gg[dd];
      }
    }
  }
  if (heavy_kind(kind) && part != 0) {
    // G_ps = (d h / d f_ps)^T G_h: one dual pass per input direction k.
    const C pv = p.aux != nullptr ? static_cast<const C*>(p.aux)[cell] : C(0);
    const bool vel = (kind == XLBH_BC_ZOUHE_VELOCITY || kind == XLBH_BC_REGULARIZED_VELOCITY);
    const bool reg = kind >= XLBH_BC_REGULARIZED_VELOCITY;
    const int k0 = part > 0 ? part - 1 : 0, k1 = part > 0 ? part : Q;
    for (int k = k0; k < k1; ++k) {
      // every column, missing ones included: a missing slot's old value is still read when
      // l = opp(l) = 0 is missing (padded interior cells) or when opp(l) is missing too
      Dual<C> hd[Q];
#pragma unroll
      for (int l = 0; l < Q; ++l) hd[l] = Dual<C>(fps[l], l == k ? C(1) : C(0));
      zouhe_cell<LAT>(hd, mm, Dual<C>(pv), vel, reg);
      C s = C(0);
#pragma unroll
      for (int l = 0; l < Q; ++l) s += G[l] * hd[l].d;
      const int sx = wrap1(x - T.c[0][k], nx), sy = wrap1(y - T.c[1][k], ny), sz = wrap1(z - T.c[2][k], nz);
      add(k, ((uint32_t)sx * ny + sy) * nz + sz, s);
    }
  }
}

// Kernel B over the whole grid (no boundary-cell list given).
template <int LAT, int FMT, class C, int COLL>
__global__ __launch_bounds__(256) void k_adjoint_bc(AdjParams p) {
  int x, y, z;
  if (!sweep_cell(p.sw, x, y, z)) return;
  const uint32_t cell = ((uint32_t)x * p.sw.ny + y) * p.sw.nz + z;
  adjoint_bc_cell<LAT, FMT, C, COLL>(p, cell, x, y, z, -1);
}

// Kernel B over a list of boundary cells: Q + 1 threads per cell (part 0 and the Q Jacobian
// columns).
template <int LAT, int FMT, class C, int COLL>
__global__ __launch_bounds__(256) void k_adjoint_bc_list(AdjParams p) {
  constexpr int Q = make_tables(LAT).q;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)p.n_bc_cells * (Q + 1)) return;
  const int i = (int)(t / (Q + 1)), part = (int)(t % (Q + 1));
  const int32_t c = p.bc_cells[i];
  if (c < 0 || (uint32_t)c >= p.N) return;
  const uint32_t cell = (uint32_t)c;
  const int nz = p.sw.nz, ny = p.sw.ny;
  const int z = (int)(cell % (uint32_t)nz);
  const int y = (int)((cell / (uint32_t)nz) % (uint32_t)ny);
  const int x = (int)(cell / ((uint32_t)nz * (uint32_t)ny));
  adjoint_bc_cell<LAT, FMT, C, COLL>(p, cell, x, y, z, part);
}

// ------------------------------------------------------------------------------------------
// Fused reverse step: one kernel per step.  The contributions that kernel B adds with atomics
// land either in slots the cell pushes itself (the Zou-He / regularized Jacobian: added before
// the push), or in slots (k, x) at the producing cell x itself (bounce-back, do-nothing and the
// extrapolation stash's own part), which the next reverse step reads at cell x only: they are
// kept per cell in d_own and read back by the same thread one step later.  The extrapolation
// neighbour part lands on cell x - c_l - n; it goes to d_nb (double-buffered by step parity) and
// is read and cleared by the flagged consumer cell in the next reverse step.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool own_kind(int kind) {
  return kind == XLBH_BC_HALFWAY_BB || kind == XLBH_BC_EXTRAPOLATION_OUTFLOW || kind == XLBH_BC_DO_NOTHING ||
         kind == XLBH_BC_GRADS_APPROXIMATION;
}

template <int LAT, int FMT, class C, int COLL>
__global__ __launch_bounds__(256) void k_adjoint_fused(AdjParams p) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  int x, y, z;
  if (!sweep_cell(p.sw, x, y, z)) return;
  uniform_slow_axes(T.d == 3, x, y);
  const int nx = p.sw.nx, ny = p.sw.ny, nz = p.sw.nz;
  const uint32_t cell = ((uint32_t)x * ny + y) * nz + z;
  // Every load that needs only the cell's position goes out first, in one batch: the boundary id,
  // the neighbour-delta flag, dL/df_out and the pulled populations.  The empty asm statement keeps
  // the compiler from sinking them below the id-dependent branches (which would cost one round
  // trip per branch); lam_in, lam_out, f_in and the delta buffers are distinct allocations.
  const int id = p.bc_mask ? __builtin_nontemporal_load(p.bc_mask + cell) : 0;
  const int nbf = p.nb_flag[cell];
  const C* lo = static_cast<const C*>(p.lam_out);
  C lamf[Q], fps[Q];
#pragma unroll
  for (int l = 0; l < Q; ++l) lamf[l] = lo[(size_t)l * p.N + cell];
  adj_pull<LAT, FMT, C>(p, x, y, z, fps);
  asm volatile("" ::: "memory");
  C* li = static_cast<C*>(p.lam_in);
  if (id == 255) {
    // solid cells are not updated by the forward step: nothing flows back through their pulls.
    // (The entry points refuse masks with solid cells; this path keeps the kernel's buffers
    // consistent regardless: deltas addressed to the cell are consumed like anywhere else.)
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const int sx = wrap1(x - T.c[0][k], nx), sy = wrap1(y - T.c[1][k], ny), sz = wrap1(z - T.c[2][k], nz);
      li[(size_t)k * p.N + ((uint32_t)sx * ny + sy) * nz + sz] = C(0);
    }
    if (nbf) {
      C* dni = static_cast<C*>(p.d_nb_in);
#pragma unroll
      for (int l = 0; l < Q; ++l) dni[(size_t)l * p.N + cell] = C(0);
    }
    return;
  }
  int kind = XLBH_BC_NONE;
  uint32_t mm = 0;
  const xlbh_bc_desc* d = nullptr;
  if (id != 0) {
    d = p.table + id;
    kind = d->kind;
    mm = p.missing[cell];
  }
  // full dL/df_out at this cell: the pushes of the previous reverse step plus its deltas
  C* dob = static_cast<C*>(p.d_own);
  if (own_kind(kind)) {
#pragma unroll
    for (int l = 0; l < Q; ++l) lamf[l] += dob[(size_t)l * p.N + cell];
  }
  if (nbf) {
    C* dni = static_cast<C*>(p.d_nb_in);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      lamf[l] += dni[(size_t)l * p.N + cell];
      dni[(size_t)l * p.N + cell] = C(0);
    }
  }
  C h[Q], lam[Q];
  adj_prepare_pulled<LAT, FMT, C>(p, cell, kind, mm, d, fps, h, lam, lamf);
  C G[Q], dn[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) dn[k] = C(0);
  if (kind == XLBH_BC_FULLWAY_BB) {
#pragma unroll
    for (int k = 0; k < Q; ++k) G[k] = lam[T.opp[k]];
  } else {
    C F[3], gg[3], Gh[Q];
    cell_force<LAT>(p, cell, F);
    const bool has_force = p.force_mode != XLBH_FORCE_NONE;
#pragma unroll
    for (int k = 0; k < Q; ++k) Gh[k] = lam[k];
    collide_adjoint<LAT, COLL>(h, Gh, (C)p.omega, has_force, F, gg, (C)p.cparam);
    if (has_force && p.grad_force) {
#pragma unroll
      for (int dd = 0; dd < T.d; ++dd) p.grad_force[(size_t)dd * p.N + cell] += (double)gg[dd];
    }
    if (kind == XLBH_BC_EQUILIBRIUM) {
#pragma unroll
      for (int k = 0; k < Q; ++k) G[k] = C(0);
    } else if (kind == XLBH_BC_DO_NOTHING) {
#pragma unroll
      for (int k = 0; k < Q; ++k) {
        G[k] = C(0);
        dn[k] = Gh[k];  // h = f_own
      }
    } else if (heavy_kind(kind)) {
      // G_ps = (d h / d f_ps)^T G_h in one reverse pass
      const C pv = p.aux != nullptr ? static_cast<const C*>(p.aux)[cell] : C(0);
      zouhe_cell_adjoint<LAT>(fps, mm, pv, kind == XLBH_BC_ZOUHE_VELOCITY || kind == XLBH_BC_REGULARIZED_VELOCITY,
                              kind >= XLBH_BC_REGULARIZED_VELOCITY, Gh, G);
    } else if (kind == XLBH_BC_GRADS_APPROXIMATION) {
      // pulled values pushed as usual; own pre-stream values carried in d_own
      C pre[Q];
      own_pre<LAT, FMT, C>(p, cell, pre);
      grads_cell_adjoint<LAT>(fps, pre, mm, Gh, G, dn);
    } else {
#pragma unroll
      for (int k = 0; k < Q; ++k) G[k] = Gh[k];
      if (kind == XLBH_BC_HALFWAY_BB || kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
#pragma unroll
        for (int l = 0; l < Q; ++l)
          if (bit(mm, l)) {
            G[l] = C(0);             // not pulled
            dn[T.opp[l]] += Gh[l];   // h_l = f_own[opp l]
          }
      }
      if (kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
        // stash of slot opp(l): (1 - cs) f_own[opp l] + cs f^t_l(x - c_l - n)
        int n[3];
        bc_normal<LAT>(mm, n);
        const C cs = C(0.57735026918962584);
        C* dno = static_cast<C*>(p.d_nb_out);
#pragma unroll
        for (int l = 0; l < Q; ++l) {
          if (bit(mm, l)) {
            const C ls = lamf[T.opp[l]];
            dn[T.opp[l]] += (C(1) - cs) * ls;
            const int sx = wrap(x - (T.c[0][l] + n[0]), nx);
            const int sy = wrap(y - (T.c[1][l] + n[1]), ny);
            const int sz = wrap(z - (T.c[2][l] + n[2]), nz);
            atomicAdd(dno + (size_t)l * p.N + ((uint32_t)sx * ny + sy) * nz + sz, cs * ls);
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    const int sx = wrap1(x - T.c[0][k], nx), sy = wrap1(y - T.c[1][k], ny), sz = wrap1(z - T.c[2][k], nz);
    li[(size_t)k * p.N + ((uint32_t)sx * ny + sy) * nz + sz] = G[k];
  }
  if (own_kind(kind)) {
#pragma unroll
    for (int k = 0; k < Q; ++k) dob[(size_t)k * p.N + cell] = dn[k];
  }
}

// Marks the cells that receive extrapolation neighbour contributions (nb_flag must be zero).
template <int LAT>
__global__ __launch_bounds__(256) void k_adjoint_nb_flags(AdjParams p, uint8_t* flag) {
  constexpr LatTables T = make_tables(LAT);
  int x, y, z;
  if (!sweep_cell(p.sw, x, y, z)) return;
  const int nx = p.sw.nx, ny = p.sw.ny, nz = p.sw.nz;
  const uint32_t cell = ((uint32_t)x * ny + y) * nz + z;
  const int id = p.bc_mask[cell];
  if (id == 0 || id == 255 || p.table[id].kind != XLBH_BC_EXTRAPOLATION_OUTFLOW) return;
  const uint32_t mm = p.missing[cell];
  int n[3];
  bc_normal<LAT>(mm, n);
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    if (bit(mm, l)) {
      const int sx = wrap(x - (T.c[0][l] + n[0]), nx);
      const int sy = wrap(y - (T.c[1][l] + n[1]), ny);
      const int sz = wrap(z - (T.c[2][l] + n[2]), nz);
      flag[((uint32_t)sx * ny + sy) * nz + sz] = 1;
    }
  }
}

// dL/df of the sweep's last step: the pushed values plus the deltas it left.
template <int LAT, class C>
__global__ __launch_bounds__(256) void k_adjoint_flush(AdjParams p, C* lam) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  int x, y, z;
  if (!sweep_cell(p.sw, x, y, z)) return;
  const uint32_t cell = ((uint32_t)x * p.sw.ny + y) * p.sw.nz + z;
  const int id = p.bc_mask ? p.bc_mask[cell] : 0;
  const bool own = id != 0 && id != 255 && own_kind(p.table[id].kind);
  const bool nb = p.nb_flag[cell] != 0;
  if (!own && !nb) return;
  const C* dob = static_cast<const C*>(p.d_own);
  const C* dno = static_cast<const C*>(p.d_nb_out);
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    C v = lam[(size_t)l * p.N + cell];
    if (own) v += dob[(size_t)l * p.N + cell];
    if (nb) v += dno[(size_t)l * p.N + cell];
    lam[(size_t)l * p.N + cell] = v;
  }
}

template <int UNUSED = 0>  // a template so that every translation unit including this header may emit it
__global__ __launch_bounds__(256) void k_zero_words(uint32_t* p, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256u;
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

enum AdjMode { ADJ_SPLIT = 0, ADJ_FUSED = 1, ADJ_BEGIN = 2, ADJ_FLUSH = 3 };

template <int LAT, int FMT, class C, int COLL>
int launch_adjoint(const AdjParams& p, hipStream_t s);

// The fused path's kernels.  `ws_words` / `flag` / `lam` are used by BEGIN / FLUSH.
template <int LAT, int FMT, class C, int COLL>
int launch_adjoint_mode(const AdjParams& p, hipStream_t s, int mode, uint64_t ws_words, uint8_t* flag, void* lam) {
  if (mode == ADJ_SPLIT) return launch_adjoint<LAT, FMT, C, COLL>(p, s);
  if (mode == ADJ_FUSED) {
    hipLaunchKernelGGL((k_adjoint_fused<LAT, FMT, C, COLL>), sweep_grid(p.sw), dim3(256), 0, s, p);
    XLBH_LAUNCH_CHECK();
    return 0;
  }
  if (mode == ADJ_BEGIN) {
    const uint64_t blocks = std::min<uint64_t>((ws_words + 255) / 256, 65536);
    hipLaunchKernelGGL(k_zero_words<0>, dim3((unsigned)blocks), dim3(256), 0, s, static_cast<uint32_t*>(p.d_own),
                       ws_words);
    XLBH_LAUNCH_CHECK();
    if (p.bc_mask) {
      hipLaunchKernelGGL((k_adjoint_nb_flags<LAT>), sweep_grid(p.sw), dim3(256), 0, s, p, flag);
      XLBH_LAUNCH_CHECK();
    }
    return 0;
  }
  hipLaunchKernelGGL((k_adjoint_flush<LAT, C>), sweep_grid(p.sw), dim3(256), 0, s, p, static_cast<C*>(lam));
  XLBH_LAUNCH_CHECK();
  return 0;
}

template <int LAT, int FMT, class C, int COLL>
int launch_adjoint(const AdjParams& p, hipStream_t s) {
  hipLaunchKernelGGL((k_adjoint_push<LAT, FMT, C, COLL>), sweep_grid(p.sw), dim3(256), 0, s, p);
  XLBH_LAUNCH_CHECK();
  if (p.bc_mask && p.bc_cells) {
    if (p.n_bc_cells > 0) {
      constexpr int Q = make_tables(LAT).q;
      const int64_t threads = (int64_t)p.n_bc_cells * (Q + 1);
      hipLaunchKernelGGL((k_adjoint_bc_list<LAT, FMT, C, COLL>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, p);
      XLBH_LAUNCH_CHECK();
    }
  } else if (p.bc_mask) {
    hipLaunchKernelGGL((k_adjoint_bc<LAT, FMT, C, COLL>), sweep_grid(p.sw), dim3(256), 0, s, p);
    XLBH_LAUNCH_CHECK();
  }
  return 0;
}

// Dispatch of one lattice's reverse kernels on (store, compute, collision); each lattice is
// instantiated in its own translation unit (adjoint_d2q9.hip, ...) so they compile in parallel.
template <int LAT>
int adjoint_dispatch_lat(int store, int compute, const AdjParams& p, hipStream_t s, int mode, uint64_t ws_words,
                         uint8_t* flag, void* lam) {
  // the collision is a template parameter, so each reverse kernel carries only its own collision
  auto with_coll = [&](auto K) -> int {
    constexpr int COLL = decltype(K)::value;
    if (compute == XLBH_COMPUTE_F32) {
      switch (store) {
        case XLBH_F32: return launch_adjoint_mode<LAT, XLBH_F32, float, COLL>(p, s, mode, ws_words, flag, lam);
        case XLBH_F16: return launch_adjoint_mode<LAT, XLBH_F16, float, COLL>(p, s, mode, ws_words, flag, lam);
        case XLBH_BF16: return launch_adjoint_mode<LAT, XLBH_BF16, float, COLL>(p, s, mode, ws_words, flag, lam);
        case XLBH_F16_SHIFTED:
          return launch_adjoint_mode<LAT, XLBH_F16_SHIFTED, float, COLL>(p, s, mode, ws_words, flag, lam);
        case XLBH_BF16_SHIFTED:
          return launch_adjoint_mode<LAT, XLBH_BF16_SHIFTED, float, COLL>(p, s, mode, ws_words, flag, lam);
      }
    } else if (compute == XLBH_COMPUTE_F64) {
      switch (store) {
        case XLBH_F64: return launch_adjoint_mode<LAT, XLBH_F64, double, COLL>(p, s, mode, ws_words, flag, lam);
        case XLBH_F32: return launch_adjoint_mode<LAT, XLBH_F32, double, COLL>(p, s, mode, ws_words, flag, lam);
      }
    }
    set_error("unsupported (store, compute) combination for the adjoint step");
    return 1;
  };
  // BEGIN / FLUSH do not depend on the collision
  if (mode == ADJ_BEGIN || mode == ADJ_FLUSH || p.collision == XLBH_BGK)
    return with_coll(std::integral_constant<int, XLBH_BGK>{});
  if (p.collision == XLBH_SMAGORINSKY_BGK) return with_coll(std::integral_constant<int, XLBH_SMAGORINSKY_BGK>{});
  if constexpr (LAT != XLBH_D3Q19) {
    if (p.collision == XLBH_KBC) return with_coll(std::integral_constant<int, XLBH_KBC>{});
  }
  set_error("unsupported collision for the adjoint step");
  return 1;
}

int adjoint_dispatch_d2q9(int store, int compute, const AdjParams& p, hipStream_t s, int mode, uint64_t ws_words,
                          uint8_t* flag, void* lam);
int adjoint_dispatch_d3q19(int store, int compute, const AdjParams& p, hipStream_t s, int mode, uint64_t ws_words,
                           uint8_t* flag, void* lam);
int adjoint_dispatch_d3q27(int store, int compute, const AdjParams& p, hipStream_t s, int mode, uint64_t ws_words,
                           uint8_t* flag, void* lam);

}  // namespace xlbh
