// The fused pull-scheme LBM step: one thread per cell, one launch per time step.
//
// Restates IncompressibleNavierStokesStepper._construct_warp.kernel
// (xlb/operator/stepper/nse_stepper.py:344-383) for gfx950:
//   pull stream (periodic wrap)  -> streaming-step BC -> macroscopic -> equilibrium ->
//   BGK/KBC collision (+ exact-difference force) -> collision-step BC ->
//   extrapolation-outflow post-collision update -> store.
//
// Differences in mechanism (results identical on the output field):
//  * Fluid cells (id 0) read exactly Q pulled populations and 1 byte of boundary id and write
//    Q populations: 2*Q*S + 1 bytes per update.  The reference also reads the cell's own f_0
//    and f_1 populations and Q bool mask bytes at every cell (nse_stepper.py:296-316).
//  * Boundary cells read the packed missing bits (4 bytes) and, where the BC needs them, the
//    own-cell pre-stream populations on demand.
//  * Prescribed values of Zou-He / regularized BCs come from a side table (aux) instead of
//    f_1[0] plus recovery writes into the input field (nse_stepper.py:318-342), so the input
//    field is never written (no read/write race on f_in).
#pragma once

#include "physics.hpp"

namespace xlbh {

struct StepParams {
  const void* f_in;
  void* f_out;
  const uint8_t* bc_mask;
  const uint32_t* missing;
  const void* aux;
  const xlbh_bc_desc* table;
  const void* force_field;
  double fu[3];
  int force_mode;
  double omega;
  double cparam;  // Smagorinsky constant (XLBH_SMAGORINSKY_BGK)
  Sweep sw;
  uint32_t N;
};

// Population field addressing.  Plane l of a field starts at base + l * plane_bytes (uniform,
// held in scalar registers); a cell is reached with an unsigned per-lane offset.  With
// O = uint32_t (every plane below 4 GiB) the compiler emits the scalar-base + 32-bit vector-offset
// form of global loads/stores; O = uint64_t serves larger planes.
template <class R, class O>
struct Planes {
  const char* base;
  size_t plane_bytes;
  __device__ __forceinline__ const R* at(int l, O cell) const {
    return reinterpret_cast<const R*>(base + (size_t)l * plane_bytes + cell * (O)sizeof(R));
  }
  __device__ __forceinline__ R* at_w(int l, O cell) const { return const_cast<R*>(at(l, cell)); }
};


// Pull-source offsets x - c for c in {-1, 0, +1} per axis, with periodic wrap (stream.py:64-81).
template <class O>
struct PullOffsets {
  O xo[3], yo[3], zo[3];
  __device__ __forceinline__ PullOffsets(int x, int y, int z, int nx, int ny, int nz) {
    const O plane = (O)ny * nz;
    xo[0] = (O)(x == nx - 1 ? 0 : x + 1) * plane;
    xo[1] = (O)x * plane;
    xo[2] = (O)(x == 0 ? nx - 1 : x - 1) * plane;
    yo[0] = (O)(y == ny - 1 ? 0 : y + 1) * nz;
    yo[1] = (O)y * nz;
    yo[2] = (O)(y == 0 ? ny - 1 : y - 1) * nz;
    zo[0] = (O)(z == nz - 1 ? 0 : z + 1);
    zo[1] = (O)z;
    zo[2] = (O)(z == 0 ? nz - 1 : z - 1);
  }
  template <int LAT>
  __device__ __forceinline__ O src(int l) const {
    constexpr LatTables T = make_tables(LAT);
    return xo[T.c[0][l] + 1] + yo[T.c[1][l] + 1] + zo[T.c[2][l] + 1];
  }
};

// Everything after the pull for one cell with boundary id `id` (!= 255), in place on the
// post-stream populations f: streaming-step BC, collision (+ force), collision-step BC and the
// extrapolation-outflow update (nse_stepper.py:360-378).
template <int LAT, int COLL, int FMT, class C, class O>
__device__ __forceinline__ void cell_update(C* f, int id, O cell, int x, int y, int z, const StepParams& p,
                                            const Planes<typename Fmt<FMT>::raw, O>& fin) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  const int nx = p.sw.nx, ny = p.sw.ny, nz = p.sw.nz;
  int kind = XLBH_BC_NONE;
  uint32_t mm = 0;
  if (id != 0) {
    // The whole descriptor and the missing bits are loaded together, ahead of the kind
    // branches (the empty asm statement keeps the compiler from sinking the EquilibriumBC
    // constants into their branch, which would cost those cells a second round trip).
    const xlbh_bc_desc* d = p.table + id;
    kind = d->kind;
    const double drho = d->rho, du0 = d->u[0], du1 = d->u[1], du2 = d->u[2];
    mm = p.missing[cell];
    asm volatile("" ::: "memory");
    // ---- streaming-step boundary conditions (apply_bc with is_post_streaming) ----
    if (kind == XLBH_BC_EQUILIBRIUM) {
      const C rb = (C)drho;
      const C ub[3] = {(C)du0, (C)du1, (C)du2};
      const C us = usqr15<LAT>(ub);
#pragma unroll
      for (int l = 0; l < Q; ++l) f[l] = feq_l<LAT>(l, rb, ub, us);
    } else if (kind == XLBH_BC_HALFWAY_BB || kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
#pragma unroll
      for (int l = 0; l < Q; ++l)
        if (bit(mm, l)) f[l] = ld_pop<FMT, C>(fin.at(T.opp[l], cell), T.w[T.opp[l]]);
    } else if (kind == XLBH_BC_DO_NOTHING) {
#pragma unroll
      for (int l = 0; l < Q; ++l) f[l] = ld_pop<FMT, C>(fin.at(l, cell), T.w[l]);
    } else if (kind >= XLBH_BC_ZOUHE_VELOCITY && kind <= XLBH_BC_REGULARIZED_PRESSURE) {
      const C pv = p.aux != nullptr ? static_cast<const C*>(p.aux)[cell] : C(0);
      const bool vel = (kind == XLBH_BC_ZOUHE_VELOCITY || kind == XLBH_BC_REGULARIZED_VELOCITY);
      const bool reg = (kind >= XLBH_BC_REGULARIZED_VELOCITY);
      zouhe_cell<LAT>(f, mm, pv, vel, reg);
    } else if (kind == XLBH_BC_GRADS_APPROXIMATION) {
      grads_cell<LAT>(f, mm, [&](int l) { return ld_pop<FMT, C>(fin.at(l, cell), T.w[l]); });
    }
  }

  if (kind == XLBH_BC_FULLWAY_BB) {
    // Collision-step full-way bounce-back: post-collision = flipped post-stream
    // (bc_fullway_bounce_back.py:59-72); the collision result is discarded, so skip it.
    C g[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) g[l] = f[T.opp[l]];
#pragma unroll
    for (int l = 0; l < Q; ++l) f[l] = g[l];
  } else {
    C F[3] = {C(0), C(0), C(0)};
    const bool has_force = p.force_mode != XLBH_FORCE_NONE;
    if (p.force_mode == XLBH_FORCE_UNIFORM) {
      F[0] = (C)p.fu[0];
      F[1] = (C)p.fu[1];
      F[2] = (C)p.fu[2];
    } else if (p.force_mode == XLBH_FORCE_FIELD) {
      const C* ff = static_cast<const C*>(p.force_field);
#pragma unroll
      for (int dd = 0; dd < T.d; ++dd) F[dd] = ff[(size_t)dd * p.N + cell];
    }
    collide<LAT, COLL>(f, (C)p.omega, has_force, F, (C)p.cparam);
  }

  if (kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
    // Post-collision auxiliary update (bc_extrapolation_outflow.py:163-195): the outgoing slot
    // opp(l) of every missing l receives (1 - cs) f_ps[l] + cs f_0[l](x - c_l - n).  The
    // post-stream value f_ps[l] of a missing l is the own-cell pre-stream f[opp l].
    // Out-of-domain sources wrap periodically (the reference reads them unchecked).
    int n[3];
    bc_normal<LAT>(mm, n);
    const C cs = C(0.57735026918962584);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      if (bit(mm, l)) {
        const int sx = wrap(x - (T.c[0][l] + n[0]), nx);
        const int sy = wrap(y - (T.c[1][l] + n[1]), ny);
        const int sz = wrap(z - (T.c[2][l] + n[2]), nz);
        const O src = ((O)sx * ny + sy) * nz + sz;
        const C faux = ld_pop<FMT, C>(fin.at(l, src), T.w[l]);
        const C fps = ld_pop<FMT, C>(fin.at(T.opp[l], cell), T.w[T.opp[l]]);
        f[T.opp[l]] = (C(1) - cs) * fps + cs * faux;
      }
    }
  }
}

// One thread per cell: Q pulls, the update, Q stores.
template <int LAT, int COLL, int FMT, class C, bool WIDE>
__device__ __forceinline__ void step_body(const StepParams& p) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  using R = typename Fmt<FMT>::raw;
  using O = typename std::conditional<WIDE, uint64_t, uint32_t>::type;

  int x, y, z;
  if (!sweep_cell(p.sw, x, y, z)) return;
  const int nx = p.sw.nx, ny = p.sw.ny, nz = p.sw.nz;
  const O cell = ((O)x * ny + y) * nz + z;
  const size_t plane_bytes = (size_t)p.N * sizeof(R);
  const Planes<R, O> fin{static_cast<const char*>(p.f_in), plane_bytes};
  const Planes<R, O> fout{static_cast<const char*>(p.f_out), plane_bytes};
  const PullOffsets<O> po(x, y, z, nx, ny, nz);

  // The boundary id is loaded first and the pulls right after it, with no wait in between (every
  // source address is in bounds for every cell).  The empty asm statement keeps the compiler from
  // sinking the pulls below the solid-cell exit, which would put a round trip in front of them.
  // Loads complete in issue order, so a boundary cell's table and missing-bit loads, which need
  // only the id, are issued while its pulls are still in flight.
  int id = 0;
  if (p.bc_mask != nullptr) id = __builtin_nontemporal_load(p.bc_mask + cell);
  C f[Q];
#pragma unroll
  for (int l = 0; l < Q; ++l) f[l] = ld_pop<FMT, C>(fin.at(l, po.template src<LAT>(l)), T.w[l]);
  asm volatile("" ::: "memory");
  if (id == 255) return;  // solid: no update (nse_stepper.py:356-358)
  cell_update<LAT, COLL, FMT, C, O>(f, id, cell, x, y, z, p, fin);
#pragma unroll
  for (int l = 0; l < Q; ++l) st_pop<FMT, C>(fout.at_w(l, cell), f[l], T.w[l]);
}

// Occupancy.  Measured on the 256^3 D3Q19 sweep: 4-byte populations stream best at 4 waves per
// SIMD (more waves add concurrent streams without adding bandwidth: 6 waves lost 3 %); 2-byte
// populations want more waves than that.  The 4- and 8-byte kernels are launched with an LDS
// reservation (kStepLdsCap, otherwise unused) that admits 4 blocks = 4 waves per SIMD per CU.
// The D3Q19 2-byte kernels with FP32 arithmetic are compiled for 6 waves per SIMD (80 VGPRs, no
// spills; the default allocation took 82 VGPRs = 5 waves): +2 % on the cavity, whose boundary
// lookups the extra waves hide, and -1 % on a periodic box.  D3Q27 would spill at 80 VGPRs.
constexpr unsigned kStepLdsCap = 36 * 1024;

template <int LAT, int COLL, int FMT, class C, bool WIDE>
__global__ __launch_bounds__(256) void k_step(StepParams p) {
  step_body<LAT, COLL, FMT, C, WIDE>(p);
}

template <int LAT, int COLL, int FMT, class C, bool WIDE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_step_w6(StepParams p) {
  step_body<LAT, COLL, FMT, C, WIDE>(p);
}


// Host launcher, instantiated per lattice in step_<lattice>.hip.
struct StepCall {
  int collision, store, compute;
  const void* f_in;
  void* f_out;
  const uint8_t* bc_mask;
  const uint32_t* missing;
  const void* aux;
  const xlbh_bc_desc* table;
  int force_mode;
  const double* force_uniform;
  const void* force_field;
  int nx, ny, nz;
  double omega;
  double cparam;  // Smagorinsky constant (XLBH_SMAGORINSKY_BGK)
  hipStream_t stream;
};

template <int LAT, int COLL, int FMT, class C>
int launch_step_t(const StepCall& c) {
  StepParams p;
  p.f_in = c.f_in;
  p.f_out = c.f_out;
  p.bc_mask = c.bc_mask;
  p.missing = c.missing;
  p.aux = c.aux;
  p.table = c.table;
  p.force_field = c.force_field;
  p.force_mode = c.force_mode;
  for (int d = 0; d < 3; ++d) p.fu[d] = (c.force_mode == XLBH_FORCE_UNIFORM) ? c.force_uniform[d] : 0.0;
  p.omega = c.omega;
  p.cparam = c.cparam;
  p.sw = make_sweep(c.nx, c.ny, c.nz);
  p.N = (uint32_t)((int64_t)c.nx * c.ny * c.nz);
  const bool wide = (uint64_t)p.N * sizeof(typename Fmt<FMT>::raw) > 0xffffffffull;
  const unsigned lds = sizeof(typename Fmt<FMT>::raw) >= 4 ? kStepLdsCap : 0u;
  constexpr bool w6 = LAT == XLBH_D3Q19 && sizeof(typename Fmt<FMT>::raw) == 2 && std::is_same<C, float>::value;
  if (wide)
    hipLaunchKernelGGL((k_step<LAT, COLL, FMT, C, true>), sweep_grid(p.sw), dim3(256), lds, c.stream, p);
  else if constexpr (w6)
    hipLaunchKernelGGL((k_step_w6<LAT, COLL, FMT, C, false>), sweep_grid(p.sw), dim3(256), lds, c.stream, p);
  else
    hipLaunchKernelGGL((k_step<LAT, COLL, FMT, C, false>), sweep_grid(p.sw), dim3(256), lds, c.stream, p);
  XLBH_LAUNCH_CHECK();
  return 0;
}

// Dispatch over (collision, store, compute) for one lattice.  Valid precision pairs are the
// reference's policies plus the BF16 and shifted 16-bit stores with FP32 compute.
template <int LAT>
int launch_step_lat(const StepCall& c) {
  constexpr bool has_kbc = (LAT != XLBH_D3Q19);
  if (c.collision == XLBH_KBC && !has_kbc) {
    set_error("KBC collision is defined for D2Q9 and D3Q27 only (kbc.py:183-185)");
    return 1;
  }
#define XLBH_STEP_CASES(COLL)                                                                   \
  if (c.compute == XLBH_COMPUTE_F32) {                                                          \
    switch (c.store) {                                                                          \
      case XLBH_F32: return launch_step_t<LAT, COLL, XLBH_F32, float>(c);                       \
      case XLBH_F16: return launch_step_t<LAT, COLL, XLBH_F16, float>(c);                       \
      case XLBH_BF16: return launch_step_t<LAT, COLL, XLBH_BF16, float>(c);                     \
      case XLBH_F16_SHIFTED: return launch_step_t<LAT, COLL, XLBH_F16_SHIFTED, float>(c);       \
      case XLBH_BF16_SHIFTED: return launch_step_t<LAT, COLL, XLBH_BF16_SHIFTED, float>(c);     \
      default: break;                                                                           \
    }                                                                                           \
  } else if (c.compute == XLBH_COMPUTE_F64) {                                                   \
    switch (c.store) {                                                                          \
      case XLBH_F64: return launch_step_t<LAT, COLL, XLBH_F64, double>(c);                      \
      case XLBH_F32: return launch_step_t<LAT, COLL, XLBH_F32, double>(c);                      \
      case XLBH_F16: return launch_step_t<LAT, COLL, XLBH_F16, double>(c);                      \
      default: break;                                                                           \
    }                                                                                           \
  }
  if (c.collision == XLBH_BGK) {
    XLBH_STEP_CASES(XLBH_BGK)
  } else if (c.collision == XLBH_SMAGORINSKY_BGK) {
    XLBH_STEP_CASES(XLBH_SMAGORINSKY_BGK)
  } else if constexpr (has_kbc) {
    if (c.collision == XLBH_KBC) {
      XLBH_STEP_CASES(XLBH_KBC)
    }
  }
#undef XLBH_STEP_CASES
  set_error("unsupported (collision, store, compute) combination for the fused step");
  return 1;
}

int launch_step_d2q9(const StepCall& c);
int launch_step_d3q19(const StepCall& c);
int launch_step_d3q27(const StepCall& c);

}  // namespace xlbh
