// Standalone operators: the reference's per-operator Warp kernels (used on their own by the
// operator API and its tests; the stepper uses the fused kernel of step.hpp instead).
#include "step.hpp"

namespace xlbh {

// ------------------------------------------------------------------------------------------
// Stream: f_out[l](x) = f_in[l](x - c_l), periodic (stream.py:59-99).  A pure permutation of
// stored values, so raw elements are moved without conversion.  It is the step kernel's memory
// pattern without its arithmetic (same addressing, launch shape and occupancy, the same Q pulls
// issued back to back and non-temporal stores), so its bandwidth is the ceiling the step's
// access pattern allows.
// ------------------------------------------------------------------------------------------
template <int LAT, class R, bool WIDE>
__global__ __launch_bounds__(256) void k_stream(const R* __restrict__ fin, R* __restrict__ fout, Sweep sw,
                                                uint64_t P) {
  constexpr LatTables T = make_tables(LAT);
  using O = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
  int x, y, z;
  if (!sweep_cell(sw, x, y, z)) return;
  const int nx = sw.nx, ny = sw.ny, nz = sw.nz;
  uniform_slow_axes(T.d == 3, x, y);
  const O row = T.d == 3 ? ((O)x * ny + y) * nz : (O)x * ny;
  const O lane = T.d == 3 ? (O)z : (O)y;
  const size_t pb = (size_t)P * sizeof(R);
  const Planes<R, O> in{reinterpret_cast<const char*>(fin), pb}, out{reinterpret_cast<const char*>(fout), pb};
  const PullOffsets<O> po(x, y, z, nx, ny, nz);
  R v[T.q];
#pragma unroll
  for (int l = 0; l < T.q; ++l) v[l] = ld_stream(in.at(l, po.template uni<LAT>(l), po.template lane<LAT>(l)));
#pragma unroll
  for (int l = 0; l < T.q; ++l) st_stream(out.at_w(l, row, lane), v[l]);
}

// ------------------------------------------------------------------------------------------
// Equilibrium: f = feq(rho, u)  (quadratic_equilibrium.py:62-98).
// ------------------------------------------------------------------------------------------
template <int LAT, int FMT, class C, class FR>
__global__ __launch_bounds__(256) void k_equilibrium(const FR* __restrict__ rho, const FR* __restrict__ u,
                                                     typename Fmt<FMT>::raw* __restrict__ f, Sweep sw,
                                                     uint32_t N, uint64_t P) {
  constexpr LatTables T = make_tables(LAT);
  constexpr bool BLK = FMT == XLBH_I16_BLOCK;
  int x, y, z;
  const bool live = sweep_cell(sw, x, y, z);
  if constexpr (!BLK) {
    if (!live) return;
  } else {
    // every lane of the block stays to the segment reduction; lanes outside the grid read a cell
    // of their own row and write nothing
    x = min(x, sw.nx - 1);
    y = min(y, sw.ny - 1);
    z = min(z, sw.nz - 1);
  }
  const uint32_t cell = ((uint32_t)x * sw.ny + y) * sw.nz + z;
  C uu[3] = {C(0), C(0), C(0)};
#pragma unroll
  for (int d = 0; d < T.d; ++d) uu[d] = raw_to<FR, C>(u[(size_t)d * N + cell]);
  const C r = raw_to<FR, C>(rho[cell]);
  const C us = usqr15<LAT>(uu);
  C fe[T.q];
#pragma unroll
  for (int l = 0; l < T.q; ++l) fe[l] = feq_l<LAT>(l, r, uu, us);
  if constexpr (!BLK) {
#pragma unroll
    for (int l = 0; l < T.q; ++l) st_pop<FMT, C>(f + (size_t)l * P + cell, fe[l], T.w[l]);
  } else {
    const int k = blk_segment_k(live ? blk_cell_max<LAT>(fe) : 0.0f, sw.lbz);
    int32_t* e = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(f) + blk_exps_offset(T.q, P));
    if ((threadIdx.x & ((1u << sw.lbz) - 1)) == 0) e[blk_seg(sw, x, y, z)] = k;
    if (live) {
#pragma unroll
      for (int l = 0; l < T.q; ++l)
        f[(size_t)l * P + cell] = k == kBlkNonFinite
                                      ? (int16_t)32767
                                      : blk_round(__builtin_fmaf(fe[l], blk_escale((float)(1.0 / T.w[l]), k),
                                                                 -blk_pow2(k)));
    }
  }
}

// Population l of `cell` from field f (plane stride P) as compute type C; for the block format
// with exponents e, the cell's segment exponent is looked up.
template <int LAT, int FMT, class C>
__device__ __forceinline__ C ld_pop_at(const typename Fmt<FMT>::raw* f, uint64_t P, const int32_t* e, int k, int l,
                                       uint32_t cell) {
  constexpr LatTables T = make_tables(LAT);
  if constexpr (FMT == XLBH_I16_BLOCK)
    return (C)__builtin_fmaf((float)f[(size_t)l * P + cell], blk_dscale((float)T.w[l], k), (float)T.w[l]);
  else
    return ld_pop<FMT, C>(f + (size_t)l * P + cell, T.w[l]);
}

// ------------------------------------------------------------------------------------------
// Moments: which = 0 rho+u, 1 rho, 2 u given rho, 3 Pi (macroscopic.py:39-55,
// zero_moment.py:29-42, first_moment.py:40-56, second_moment.py:81-98).
// ------------------------------------------------------------------------------------------
template <int LAT, int FMT, class C, class FR>
__global__ __launch_bounds__(256) void k_moments(const typename Fmt<FMT>::raw* __restrict__ f, FR* rho,
                                                 FR* u, FR* pi, int which, Sweep sw, uint32_t N, uint64_t P) {
  constexpr LatTables T = make_tables(LAT);
  int x, y, z;
  if (!sweep_cell(sw, x, y, z)) return;
  const uint32_t cell = ((uint32_t)x * sw.ny + y) * sw.nz + z;
  const int32_t* e = reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(f) + blk_exps_offset(T.q, P));
  const int k = FMT == XLBH_I16_BLOCK ? e[blk_seg(sw, x, y, z)] : 0;
  C ff[T.q];
#pragma unroll
  for (int l = 0; l < T.q; ++l) ff[l] = ld_pop_at<LAT, FMT, C>(f, P, e, k, l, cell);
  if (which == 3) {
#pragma unroll
    for (int t = 0; t < T.nt; ++t) {
      C s = C(0);
#pragma unroll
      for (int l = 0; l < T.q; ++l) {
        if (T.cc[l][t] == 1) s += ff[l];
        else if (T.cc[l][t] == -1) s -= ff[l];
      }
      pi[(size_t)t * N + cell] = to_raw<FR, C>(s);
    }
    return;
  }
  C r, uu[3];
  if (which == 2) {
    r = raw_to<FR, C>(rho[cell]);
#pragma unroll
    for (int d = 0; d < 3; ++d) uu[d] = C(0);
#pragma unroll
    for (int l = 0; l < T.q; ++l)
#pragma unroll
      for (int d = 0; d < T.d; ++d) {
        if (T.c[d][l] == 1) uu[d] += ff[l];
        else if (T.c[d][l] == -1) uu[d] -= ff[l];
      }
#pragma unroll
    for (int d = 0; d < T.d; ++d) uu[d] = uu[d] / r;
  } else {
    macroscopic<LAT>(ff, r, uu);
  }
  if (which == 0 || which == 1) rho[cell] = to_raw<FR, C>(r);
  if (which == 0 || which == 2) {
#pragma unroll
    for (int d = 0; d < T.d; ++d) u[(size_t)d * N + cell] = to_raw<FR, C>(uu[d]);
  }
}

// ------------------------------------------------------------------------------------------
// Collision kernels: fout = collide(f, feq, rho, u, omega) (+ force).  BGK uses feq as given
// (bgk.py:37-62); KBC reads rho, u and feq from the fields (kbc.py:299-329).
// ------------------------------------------------------------------------------------------
template <int LAT, int COLL, int FMT, class C, class FR>
__global__ __launch_bounds__(256) void k_collide(const typename Fmt<FMT>::raw* __restrict__ f,
                                                 const typename Fmt<FMT>::raw* __restrict__ feq,
                                                 typename Fmt<FMT>::raw* __restrict__ fout,
                                                 const FR* __restrict__ rho, const FR* __restrict__ u,
                                                 C omega, C cparam, int force_mode, double fu0, double fu1,
                                                 double fu2, const C* __restrict__ force_field, Sweep sw,
                                                 uint32_t N, uint64_t P) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  int x, y, z;
  if (!sweep_cell(sw, x, y, z)) return;
  const uint32_t cell = ((uint32_t)x * sw.ny + y) * sw.nz + z;
  C ff[Q], fe[Q];
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    ff[l] = ld_pop<FMT, C>(f + (size_t)l * P + cell, T.w[l]);
    fe[l] = ld_pop<FMT, C>(feq + (size_t)l * P + cell, T.w[l]);
  }
  C out[Q];
  C r = C(0), uu[3] = {C(0), C(0), C(0)};
  const bool need_macro = (COLL == XLBH_KBC) || force_mode != XLBH_FORCE_NONE;
  if (need_macro) {
    r = raw_to<FR, C>(rho[cell]);
#pragma unroll
    for (int d = 0; d < T.d; ++d) uu[d] = raw_to<FR, C>(u[(size_t)d * N + cell]);
  }
  if constexpr (COLL == XLBH_BGK) {
#pragma unroll
    for (int l = 0; l < Q; ++l) out[l] = ff[l] - omega * (ff[l] - fe[l]);
  } else if constexpr (COLL == XLBH_SMAGORINSKY_BGK) {
    // smagorinsky_les_bgk.py:40-94 with feq as given (the fused step's branch in physics.hpp)
    C strain = C(0);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      const int cs = T.c[0][l] + T.c[1][l] + T.c[2][l];
      const C fneq = ff[l] - fe[l];
      if (cs == 1) strain += fneq * fneq;
      if (cs >= 2) strain += C(2) * fneq * fneq;
    }
    const C tau0 = C(1) / omega;
    const C tau = tau0 + C(0.5) * (sqrt(tau0 * tau0 + C(36) * (cparam * cparam) * sqrt(strain)) - tau0);
    const C inv_tau = C(1) / tau;
#pragma unroll
    for (int l = 0; l < Q; ++l) out[l] = ff[l] - inv_tau * (ff[l] - fe[l]);
  } else {
    C fneq[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) fneq[l] = ff[l] - fe[l];
    C pi[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) pi[t] = C(0);
#pragma unroll
    for (int l = 0; l < Q; ++l)
#pragma unroll
      for (int t = 0; t < T.nt; ++t) {
        if (T.cc[l][t] == 1) pi[t] += fneq[l];
        else if (T.cc[l][t] == -1) pi[t] -= fneq[l];
      }
    const C beta = C(0.5) * omega, inv_beta = C(1) / beta;
    C sp1 = C(0), sp2 = C(0);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      C ds = kbc_shear<LAT>(l, pi) * r;
      if constexpr (T.d == 2) ds = ds / C(4);
      const C dh = fneq[l] - ds;
      const C tmp = dh / fe[l];
      sp1 += tmp * ds;
      sp2 += tmp * dh;
    }
    const C gamma = inv_beta - (C(2) - inv_beta) * sp1 / (C(1e-32) + sp2);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      C ds = kbc_shear<LAT>(l, pi) * r;
      if constexpr (T.d == 2) ds = ds / C(4);
      const C dh = fneq[l] - ds;
      out[l] = ff[l] - beta * (C(2) * ds + gamma * dh);
    }
  }
  if (force_mode != XLBH_FORCE_NONE) {
    C F[3] = {C(fu0), C(fu1), C(fu2)};
    if (force_mode == XLBH_FORCE_FIELD) {
#pragma unroll
      for (int d = 0; d < T.d; ++d) F[d] = force_field[(size_t)d * N + cell];
    }
    C uf[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) uf[d] = uu[d] + F[d];
    const C usf = usqr15<LAT>(uf);
#pragma unroll
    for (int l = 0; l < Q; ++l) out[l] += feq_l<LAT>(l, r, uf, usf) - fe[l];
  }
#pragma unroll
  for (int l = 0; l < Q; ++l) st_pop<FMT, C>(fout + (size_t)l * P + cell, out[l], T.w[l]);
}

// ------------------------------------------------------------------------------------------
// Standalone boundary condition (BoundaryCondition._construct_kernel, boundary_condition.py:86-117):
// at cells with bc_mask == id, f_post = functional(f_pre, f_post); elsewhere unchanged.
// The functional of every kind is its streaming-step form (FullwayBB: flipped f_pre).
// ------------------------------------------------------------------------------------------
template <int LAT, int FMT, class C>
__global__ __launch_bounds__(256) void k_apply_bc(int kind, int id, xlbh_bc_desc desc,
                                                  const typename Fmt<FMT>::raw* fpre,
                                                  typename Fmt<FMT>::raw* fpost,
                                                  const uint8_t* __restrict__ bc_mask,
                                                  const uint32_t* __restrict__ missing, const C* aux,
                                                  Sweep sw, uint64_t P) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  int x, y, z;
  if (!sweep_cell(sw, x, y, z)) return;
  const uint32_t cell = ((uint32_t)x * sw.ny + y) * sw.nz + z;
  if (bc_mask[cell] != id) return;
  const uint32_t mm = missing ? missing[cell] : 0u;
  C f[Q];
#pragma unroll
  for (int l = 0; l < Q; ++l) f[l] = ld_pop<FMT, C>(fpost + (size_t)l * P + cell, T.w[l]);
  auto pre = [&](int l) { return ld_pop<FMT, C>(fpre + (size_t)l * P + cell, T.w[l]); };
  if (kind == XLBH_BC_EQUILIBRIUM) {
    const C rb = (C)desc.rho;
    const C ub[3] = {(C)desc.u[0], (C)desc.u[1], (C)desc.u[2]};
    const C us = usqr15<LAT>(ub);
#pragma unroll
    for (int l = 0; l < Q; ++l) f[l] = feq_l<LAT>(l, rb, ub, us);
  } else if (kind == XLBH_BC_HALFWAY_BB || kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
#pragma unroll
    for (int l = 0; l < Q; ++l)
      if (bit(mm, l)) f[l] = pre(T.opp[l]);
  } else if (kind == XLBH_BC_FULLWAY_BB) {
#pragma unroll
    for (int l = 0; l < Q; ++l) f[l] = pre(T.opp[l]);
  } else if (kind == XLBH_BC_DO_NOTHING) {
#pragma unroll
    for (int l = 0; l < Q; ++l) f[l] = pre(l);
  } else if (kind >= XLBH_BC_ZOUHE_VELOCITY && kind <= XLBH_BC_REGULARIZED_PRESSURE) {
    const C pv = aux[cell];
    zouhe_cell<LAT>(f, mm, pv, kind == XLBH_BC_ZOUHE_VELOCITY || kind == XLBH_BC_REGULARIZED_VELOCITY,
                    kind >= XLBH_BC_REGULARIZED_VELOCITY);
  } else if (kind == XLBH_BC_GRADS_APPROXIMATION) {
    grads_cell<LAT>(f, mm, pre);
  }
#pragma unroll
  for (int l = 0; l < Q; ++l) st_pop<FMT, C>(fpost + (size_t)l * P + cell, f[l], T.w[l]);
}

// ------------------------------------------------------------------------------------------
// Host dispatch helpers.
// ------------------------------------------------------------------------------------------
// (store, compute) pairs accepted by the standalone operators: all stores with FP32 compute,
// and F64/F32/F16 stores with FP64 compute.
template <class Fn>
int with_precision(int store, int compute, Fn&& fn) {
  compute = arith_compute(compute);
  if (compute == XLBH_COMPUTE_F32) {
    switch (store) {
      case XLBH_F32: return fn(std::integral_constant<int, XLBH_F32>{}, float{});
      case XLBH_F64: return fn(std::integral_constant<int, XLBH_F64>{}, float{});
      case XLBH_F16: return fn(std::integral_constant<int, XLBH_F16>{}, float{});
      case XLBH_BF16: return fn(std::integral_constant<int, XLBH_BF16>{}, float{});
      case XLBH_F16_SHIFTED: return fn(std::integral_constant<int, XLBH_F16_SHIFTED>{}, float{});
      case XLBH_BF16_SHIFTED: return fn(std::integral_constant<int, XLBH_BF16_SHIFTED>{}, float{});
    }
  } else if (compute == XLBH_COMPUTE_F64) {
    switch (store) {
      case XLBH_F64: return fn(std::integral_constant<int, XLBH_F64>{}, double{});
      case XLBH_F32: return fn(std::integral_constant<int, XLBH_F32>{}, double{});
      case XLBH_F16: return fn(std::integral_constant<int, XLBH_F16>{}, double{});
    }
  }
  set_error("unsupported (store, compute) combination");
  return 1;
}

// with_precision plus the block format (FP32 compute), for the operators that support it.
template <class Fn>
int with_precision_blk(int store, int compute, Fn&& fn) {
  if (store == XLBH_I16_BLOCK) {
    if (arith_compute(compute) == XLBH_COMPUTE_F32) return fn(std::integral_constant<int, XLBH_I16_BLOCK>{}, float{});
    set_error("the block format (XLBH_I16_BLOCK) computes in FP32 only");
    return 1;
  }
  return with_precision(store, compute, fn);
}

template <class Fn>
int with_field(int field, Fn&& fn) {
  switch (field) {
    case XLBH_FIELD_F32: return fn(float{});
    case XLBH_FIELD_F64: return fn(double{});
    case XLBH_FIELD_F16: return fn(_Float16{});
    case XLBH_FIELD_BF16: return fn(uint16_t{});
  }
  set_error("unknown field dtype");
  return 1;
}

// ------------------------------------------------------------------------------------------
// Elementwise precision cast (precision_caster.py:43-82).  Values go through float unless both
// ends are double; 16-bit ends use the same conversions as the population formats.
// ------------------------------------------------------------------------------------------
template <class I, class O>
__global__ __launch_bounds__(256) void k_cast(const I* __restrict__ in, O* __restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    if constexpr (__is_same(I, double) && __is_same(O, double))
      out[i] = in[i];
    else if constexpr (__is_same(O, double))
      out[i] = (double)raw_to<I, float>(in[i]);
    else
      out[i] = to_raw<O, float>(raw_to<I, float>(in[i]));
  }
}

}  // namespace xlbh

using namespace xlbh;

extern "C" {

int xlbh_stream(int lattice, int store, const void* f_in, void* f_out, int nx, int ny, int nz, int64_t plane,
                void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(valid_store(store), "unknown store format");
  XLBH_REQUIRE(store != XLBH_I16_BLOCK, "stream: the block format's segment exponents do not move with the "
                                        "values; stream a decoded field or use the fused step");
  XLBH_REQUIRE(f_in && f_out && f_in != f_out, "stream needs distinct non-null fields");
  const Sweep sw = make_sweep(nx, ny, nz);
  uint64_t N;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, N)) return e;
  hipStream_t s = (hipStream_t)stream;
  return with_lattice(lattice, [&](auto L) -> int {
    constexpr int LAT = decltype(L)::value;
    auto go = [&](auto r) -> int {
      using R = decltype(r);
      // the step's occupancy: 4- and 8-byte populations capped at 4 waves per SIMD
      const unsigned lds = sizeof(R) >= 4 ? kStepLdsCap : 0u;
      const bool wide = (uint64_t)Lat<LAT>::Q * N * sizeof(R) > 0xffffffffull;
      if (wide)
        hipLaunchKernelGGL((k_stream<LAT, R, true>), sweep_grid(sw), dim3(256), lds, s, (const R*)f_in, (R*)f_out,
                           sw, N);
      else
        hipLaunchKernelGGL((k_stream<LAT, R, false>), sweep_grid(sw), dim3(256), lds, s, (const R*)f_in, (R*)f_out,
                           sw, N);
      XLBH_LAUNCH_CHECK();
      return 0;
    };
    if (store == XLBH_F64) return go(double{});
    if (store == XLBH_F32) return go(float{});
    return go(uint16_t{});
  });
}

int xlbh_equilibrium(int lattice, int store, int compute, int field, const void* rho, const void* u,
                     void* f, int nx, int ny, int nz, int64_t plane, void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(rho && u && f, "equilibrium needs rho, u and f");
  const Sweep sw = make_sweep(nx, ny, nz);
  const uint32_t N = (uint32_t)((int64_t)nx * ny * nz);
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  hipStream_t s = (hipStream_t)stream;
  return with_lattice(lattice, [&](auto L) -> int {
    constexpr int LAT = decltype(L)::value;
    return with_precision_blk(store, compute, [&](auto S, auto c) -> int {
      constexpr int FMT = decltype(S)::value;
      using C = decltype(c);
      return with_field(field, [&](auto fr) -> int {
        using FR = decltype(fr);
        hipLaunchKernelGGL((k_equilibrium<LAT, FMT, C, FR>), sweep_grid(sw), dim3(256), 0, s, (const FR*)rho,
                           (const FR*)u, (typename Fmt<FMT>::raw*)f, sw, N, P);
        XLBH_LAUNCH_CHECK();
        return 0;
      });
    });
  });
}

int xlbh_moments(int lattice, int store, int compute, int field, int which, const void* f, void* rho,
                 void* u, void* pi, int nx, int ny, int nz, int64_t plane, void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(which >= 0 && which <= 3, "moments: which must be 0..3");
  XLBH_REQUIRE(f, "moments: f is null");
  XLBH_REQUIRE(which == 3 || rho, "moments: rho is null");
  XLBH_REQUIRE(!(which == 0 || which == 2) || u, "moments: u is null");
  XLBH_REQUIRE(which != 3 || pi, "moments: pi is null");
  const Sweep sw = make_sweep(nx, ny, nz);
  const uint32_t N = (uint32_t)((int64_t)nx * ny * nz);
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  hipStream_t s = (hipStream_t)stream;
  return with_lattice(lattice, [&](auto L) -> int {
    constexpr int LAT = decltype(L)::value;
    return with_precision_blk(store, compute, [&](auto S, auto c) -> int {
      constexpr int FMT = decltype(S)::value;
      using C = decltype(c);
      return with_field(field, [&](auto fr) -> int {
        using FR = decltype(fr);
        hipLaunchKernelGGL((k_moments<LAT, FMT, C, FR>), sweep_grid(sw), dim3(256), 0, s,
                           (const typename Fmt<FMT>::raw*)f, (FR*)rho, (FR*)u, (FR*)pi, which, sw, N, P);
        XLBH_LAUNCH_CHECK();
        return 0;
      });
    });
  });
}

int xlbh_collide(int lattice, int collision, int store, int compute, int field, const void* f,
                 const void* feq, void* fout, const void* rho, const void* u, double omega,
                 double collision_param, int force_mode,
                 const double* force_uniform, const void* force_field, int nx, int ny, int nz, int64_t plane,
                 void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(f && feq && fout, "collide needs f, feq and fout");
  XLBH_REQUIRE(store != XLBH_I16_BLOCK, "collide: the block format is written by the fused step and the "
                                        "equilibrium operator only");
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  XLBH_REQUIRE(collision == XLBH_BGK || collision == XLBH_KBC || collision == XLBH_SMAGORINSKY_BGK,
               "unknown collision");
  XLBH_REQUIRE(!(collision == XLBH_KBC && lattice == XLBH_D3Q19),
               "KBC collision is defined for D2Q9 and D3Q27 only (kbc.py:183-185)");
  XLBH_REQUIRE(force_mode >= XLBH_FORCE_NONE && force_mode <= XLBH_FORCE_FIELD, "unknown force mode");
  XLBH_REQUIRE(force_mode != XLBH_FORCE_UNIFORM || force_uniform, "uniform force needs 3 host doubles");
  XLBH_REQUIRE(force_mode != XLBH_FORCE_FIELD || force_field, "force field is null");
  XLBH_REQUIRE((collision != XLBH_KBC && force_mode == XLBH_FORCE_NONE) || (rho && u),
               "KBC and forced collisions need rho and u");
  const Sweep sw = make_sweep(nx, ny, nz);
  const uint32_t N = (uint32_t)((int64_t)nx * ny * nz);
  hipStream_t s = (hipStream_t)stream;
  double fu[3] = {0, 0, 0};
  if (force_mode == XLBH_FORCE_UNIFORM)
    for (int d = 0; d < 3; ++d) fu[d] = force_uniform[d];
  return with_lattice(lattice, [&](auto L) -> int {
    constexpr int LAT = decltype(L)::value;
    return with_precision(store, compute, [&](auto S, auto c) -> int {
      constexpr int FMT = decltype(S)::value;
      using C = decltype(c);
      using R = typename Fmt<FMT>::raw;
      return with_field(field, [&](auto fr) -> int {
        using FR = decltype(fr);
        const C cp = (C)collision_param;
        if (collision == XLBH_BGK)
          hipLaunchKernelGGL((k_collide<LAT, XLBH_BGK, FMT, C, FR>), sweep_grid(sw), dim3(256), 0, s,
                             (const R*)f, (const R*)feq, (R*)fout, (const FR*)rho, (const FR*)u, (C)omega, cp,
                             force_mode, fu[0], fu[1], fu[2], (const C*)force_field, sw, N, P);
        else if (collision == XLBH_SMAGORINSKY_BGK)
          hipLaunchKernelGGL((k_collide<LAT, XLBH_SMAGORINSKY_BGK, FMT, C, FR>), sweep_grid(sw), dim3(256), 0, s,
                             (const R*)f, (const R*)feq, (R*)fout, (const FR*)rho, (const FR*)u, (C)omega, cp,
                             force_mode, fu[0], fu[1], fu[2], (const C*)force_field, sw, N, P);
        else if constexpr (LAT != XLBH_D3Q19)
          hipLaunchKernelGGL((k_collide<LAT, XLBH_KBC, FMT, C, FR>), sweep_grid(sw), dim3(256), 0, s,
                             (const R*)f, (const R*)feq, (R*)fout, (const FR*)rho, (const FR*)u, (C)omega, cp,
                             force_mode, fu[0], fu[1], fu[2], (const C*)force_field, sw, N, P);
        XLBH_LAUNCH_CHECK();
        return 0;
      });
    });
  });
}

int xlbh_apply_bc(int lattice, int store, int compute, int kind, int bc_id, const xlbh_bc_desc* desc,
                  const void* f_pre, void* f_post, const uint8_t* bc_mask, const uint32_t* missing_bits,
                  const void* aux, int nx, int ny, int nz, int64_t plane, void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(f_pre && f_post && bc_mask, "apply_bc needs f_pre, f_post and bc_mask");
  XLBH_REQUIRE(store != XLBH_I16_BLOCK, "apply_bc: the block format is written by the fused step and the "
                                        "equilibrium operator only");
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  XLBH_REQUIRE(kind >= XLBH_BC_NONE && kind <= XLBH_BC_GRADS_APPROXIMATION, "unknown BC kind");
  XLBH_REQUIRE(kind != XLBH_BC_GRADS_APPROXIMATION || lattice != XLBH_D2Q9,
               "GradsApproximationBC is 3-D only (bc_grads_approximation.py:75-77)");
  XLBH_REQUIRE(bc_id >= 0 && bc_id <= 255, "bc_id out of range");
  const bool needs_missing = kind == XLBH_BC_HALFWAY_BB || kind == XLBH_BC_EXTRAPOLATION_OUTFLOW ||
                             kind == XLBH_BC_GRADS_APPROXIMATION ||
                             (kind >= XLBH_BC_ZOUHE_VELOCITY && kind <= XLBH_BC_REGULARIZED_PRESSURE);
  XLBH_REQUIRE(!needs_missing || missing_bits, "this BC kind needs the missing bits");
  XLBH_REQUIRE(!(kind >= XLBH_BC_ZOUHE_VELOCITY && kind <= XLBH_BC_REGULARIZED_PRESSURE) || aux,
               "Zou-He / regularized BCs need the aux table");
  xlbh_bc_desc d{};
  if (desc) d = *desc;
  const Sweep sw = make_sweep(nx, ny, nz);
  hipStream_t s = (hipStream_t)stream;
  return with_lattice(lattice, [&](auto L) -> int {
    constexpr int LAT = decltype(L)::value;
    return with_precision(store, compute, [&](auto S, auto c) -> int {
      constexpr int FMT = decltype(S)::value;
      using C = decltype(c);
      using R = typename Fmt<FMT>::raw;
      hipLaunchKernelGGL((k_apply_bc<LAT, FMT, C>), sweep_grid(sw), dim3(256), 0, s, kind, bc_id, d,
                         (const R*)f_pre, (R*)f_post, bc_mask, missing_bits, (const C*)aux, sw, P);
      XLBH_LAUNCH_CHECK();
      return 0;
    });
  });
}

int xlbh_cast(int field_in, int field_out, const void* in, void* out, int64_t n, void* stream) {
  clear_error();
  XLBH_REQUIRE(n >= 0, "negative element count");
  XLBH_REQUIRE(n == 0 || (in && out), "cast needs input and output");
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 20);
  return with_field(field_in, [&](auto a) -> int {
    using I = decltype(a);
    return with_field(field_out, [&](auto b) -> int {
      using O = decltype(b);
      hipLaunchKernelGGL((k_cast<I, O>), dim3(blocks), dim3(256), 0, s, (const I*)in, (O*)out, n);
      XLBH_LAUNCH_CHECK();
      return 0;
    });
  });
}

}  // extern "C"
