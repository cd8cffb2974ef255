// Checkpoint codec of the checkpointed reverse sweep: a population field in 16 bits per
// population plus the cell's moments.
//
// The reverse sweep restarts every segment from its checkpoint, so the rounding of a checkpoint
// becomes a perturbation of the recomputed trajectory.  Rounding f_l - w_l to bfloat16 perturbs
// the moments (rho, u) by about 2^-9 of |f_l - w_l|; here the moments are kept in float32 and only
// the non-equilibrium part f_l - feq_l(rho, u), which is small wherever the flow is smooth, is
// rounded to bfloat16.  The equilibrium is evaluated (in FP64) from the float32 moments on both
// sides, so decoding adds back exactly what encoding took out, up to the bfloat16 rounding, whose
// mass and momentum components the decoder removes.
//
// No counterpart in the reference: its out-of-core example keeps checkpoints as full Warp arrays
// (examples/out_of_core/autodiff_lbm.py:64-127).
#include "physics.hpp"

namespace xlbh {
namespace {

template <int LAT, int FMT>
__global__ __launch_bounds__(256) void k_ckpt_encode(const typename Fmt<FMT>::raw* __restrict__ f,
                                                     uint16_t* __restrict__ neq, float* __restrict__ mom,
                                                     uint32_t N, uint64_t P) {
  constexpr LatTables T = make_tables(LAT);
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= N) return;
  double fl[T.q];
#pragma unroll
  for (int l = 0; l < T.q; ++l) fl[l] = ld_pop<FMT, double>(f + (size_t)l * P + i, T.w[l]);
  double rho, u[3];
  macroscopic<LAT>(fl, rho, u);
  // the moments as kept, and the equilibrium evaluated from them
  const float r32 = (float)(rho - 1.0);
  float u32[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int d = 0; d < T.d; ++d) u32[d] = (float)u[d];
  mom[i] = r32;
#pragma unroll
  for (int d = 0; d < T.d; ++d) mom[(size_t)(d + 1) * N + i] = u32[d];
  const double rk = 1.0 + (double)r32;
  const double uk[3] = {(double)u32[0], (double)u32[1], (double)u32[2]};
  const double us = usqr15<LAT>(uk);
#pragma unroll
  for (int l = 0; l < T.q; ++l)
    neq[(size_t)l * N + i] = to_raw<uint16_t, float>((float)(fl[l] - feq_l<LAT>(l, rk, uk, us)));
}

template <int LAT, int FMT>
__global__ __launch_bounds__(256) void k_ckpt_decode(const uint16_t* __restrict__ neq,
                                                     const float* __restrict__ mom,
                                                     typename Fmt<FMT>::raw* __restrict__ f, uint32_t N,
                                                     uint64_t P) {
  constexpr LatTables T = make_tables(LAT);
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= N) return;
  const double rk = 1.0 + (double)mom[i];
  double uk[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int d = 0; d < T.d; ++d) uk[d] = (double)mom[(size_t)(d + 1) * N + i];
  const double us = usqr15<LAT>(uk);
  // The exact non-equilibrium part carries no mass or momentum; its rounded values do, a little.
  // Taking w_l (dr + 3 c_l.dj) off each of them removes that (sum_l w_l = 1, sum_l w_l c_l = 0,
  // sum_l w_l c_l c_l = I / 3), so the decoded field has the kept moments.
  double nq[T.q], dr = 0.0, dj[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    nq[l] = (double)bf16_bits_to_float(neq[(size_t)l * N + i]);
    dr += nq[l];
#pragma unroll
    for (int d = 0; d < T.d; ++d) dj[d] += T.c[d][l] * nq[l];
  }
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    double cj = 0.0;
#pragma unroll
    for (int d = 0; d < T.d; ++d) cj += T.c[d][l] * dj[d];
    const double v = feq_l<LAT>(l, rk, uk, us) + (nq[l] - T.w[l] * (dr + 3.0 * cj));
    st_pop<FMT, double>(f + (size_t)l * P + i, v, T.w[l]);
  }
}

// Every store format but the block format (its exponents belong to row segments, not cells).
template <class Fn>
int with_ckpt_store(int store, Fn&& fn) {
  switch (store) {
    case XLBH_F32: return fn(std::integral_constant<int, XLBH_F32>{});
    case XLBH_F64: return fn(std::integral_constant<int, XLBH_F64>{});
    case XLBH_F16: return fn(std::integral_constant<int, XLBH_F16>{});
    case XLBH_BF16: return fn(std::integral_constant<int, XLBH_BF16>{});
    case XLBH_F16_SHIFTED: return fn(std::integral_constant<int, XLBH_F16_SHIFTED>{});
    case XLBH_BF16_SHIFTED: return fn(std::integral_constant<int, XLBH_BF16_SHIFTED>{});
    default: break;
  }
  set_error("checkpoint codec: unsupported store format (every format but the block format)");
  return 1;
}

int ckpt_call(bool encode, int lattice, int store, const void* f, void* f_out, uint16_t* neq_out,
              const uint16_t* neq_in, float* mom_out, const float* mom_in, int nx, int ny, int nz, int64_t plane,
              void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  const uint32_t N = (uint32_t)((int64_t)nx * ny * nz);
  const dim3 grid((N + 255u) / 256u);
  hipStream_t s = (hipStream_t)stream;
  return with_lattice(lattice, [&](auto L) -> int {
    constexpr int LAT = decltype(L)::value;
    return with_ckpt_store(store, [&](auto S) -> int {
      constexpr int FMT = decltype(S)::value;
      using R = typename Fmt<FMT>::raw;
      if (encode)
        hipLaunchKernelGGL((k_ckpt_encode<LAT, FMT>), grid, dim3(256), 0, s, static_cast<const R*>(f), neq_out,
                           mom_out, N, P);
      else
        hipLaunchKernelGGL((k_ckpt_decode<LAT, FMT>), grid, dim3(256), 0, s, neq_in, mom_in, static_cast<R*>(f_out),
                           N, P);
      XLBH_LAUNCH_CHECK();
      return 0;
    });
  });
}

}  // namespace
}  // namespace xlbh

using namespace xlbh;

extern "C" {

int xlbh_checkpoint_encode(int lattice, int store, const void* f, uint16_t* neq, float* moments, int nx, int ny,
                           int nz, int64_t plane, void* stream) {
  if (!f || !neq || !moments) {
    set_error("checkpoint_encode: null argument");
    return 1;
  }
  return ckpt_call(true, lattice, store, f, nullptr, neq, nullptr, moments, nullptr, nx, ny, nz, plane, stream);
}

int xlbh_checkpoint_decode(int lattice, int store, const uint16_t* neq, const float* moments, void* f, int nx, int ny,
                           int nz, int64_t plane, void* stream) {
  if (!f || !neq || !moments) {
    set_error("checkpoint_decode: null argument");
    return 1;
  }
  return ckpt_call(false, lattice, store, nullptr, f, nullptr, neq, nullptr, moments, nx, ny, nz, plane, stream);
}

}  // extern "C"
