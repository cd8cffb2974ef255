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
  uint64_t plane;        // population plane stride (elements)
  const int32_t* e_in;   // XLBH_I16_BLOCK: segment exponents of f_in and f_out
  int32_t* e_out;
  int x0, x1;            // the planes x0 <= x < x1 this launch updates
};

// Reads of stored populations other than the pulls: the own cell's pre-stream values and the
// extrapolation outflow's source cell.  For the block format, `k_own` is the exponent of the
// caller's segment and other cells look theirs up.
template <int FMT, class C, class O>
struct PopIn {
  Planes<typename Fmt<FMT>::raw, O> pl;
  const int32_t* e;
  Sweep sw;
  int k_own;
  __device__ __forceinline__ C own(int l, O cell, double w) const {
    if constexpr (FMT == XLBH_I16_BLOCK)
      return __builtin_fmaf((float)*pl.at(l, cell), blk_dscale((float)w, k_own), (float)w);
    else
      return ld_pop<FMT, C>(pl.at(l, cell), w);
  }
  __device__ __forceinline__ C at(int l, O cell, int x, int y, int z, double w) const {
    if constexpr (FMT == XLBH_I16_BLOCK)
      return __builtin_fmaf((float)*pl.at(l, cell), blk_dscale((float)w, e[blk_seg(sw, x, y, z)]), (float)w);
    else
      return ld_pop<FMT, C>(pl.at(l, cell), w);
  }
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
  // src(l) split into its wave-uniform part (the slow axes) and its per-lane part (the row axis:
  // z in 3-D, y in 2-D), for Planes::at(l, uni, lane).
  template <int LAT>
  __device__ __forceinline__ O uni(int l) const {
    constexpr LatTables T = make_tables(LAT);
    if constexpr (T.d == 3) return xo[T.c[0][l] + 1] + yo[T.c[1][l] + 1];
    return xo[T.c[0][l] + 1];
  }
  template <int LAT>
  __device__ __forceinline__ O lane(int l) const {
    constexpr LatTables T = make_tables(LAT);
    if constexpr (T.d == 3) return zo[T.c[2][l] + 1];
    return yo[T.c[1][l] + 1] + zo[T.c[2][l] + 1];
  }
};

// Everything after the pull for one cell with boundary id `id` (!= 255), in place on the
// post-stream populations f: streaming-step BC, collision (+ force), collision-step BC and the
// extrapolation-outflow update (nse_stepper.py:360-378).
template <int LAT, int COLL, int FMT, class C, class O>
__device__ __forceinline__ void cell_update(C* f, int id, O cell, int x, int y, int z, const StepParams& p,
                                            const PopIn<FMT, C, O>& fin) {
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
        if (bit(mm, l)) f[l] = fin.own(T.opp[l], cell, T.w[T.opp[l]]);
    } else if (kind == XLBH_BC_DO_NOTHING) {
#pragma unroll
      for (int l = 0; l < Q; ++l) f[l] = fin.own(l, cell, T.w[l]);
    } else if (kind >= XLBH_BC_ZOUHE_VELOCITY && kind <= XLBH_BC_REGULARIZED_PRESSURE) {
      const C pv = p.aux != nullptr ? static_cast<const C*>(p.aux)[cell] : C(0);
      const bool vel = (kind == XLBH_BC_ZOUHE_VELOCITY || kind == XLBH_BC_REGULARIZED_VELOCITY);
      const bool reg = (kind >= XLBH_BC_REGULARIZED_VELOCITY);
      zouhe_cell<LAT>(f, mm, pv, vel, reg);
    } else if (kind == XLBH_BC_GRADS_APPROXIMATION) {
      grads_cell<LAT>(f, mm, [&](int l) { return fin.own(l, cell, T.w[l]); });
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
    if constexpr (COLL == kBgkDeviation) {
      bgk_deviation<LAT>(f, p.omega, has_force, F);
    } else {
      // 16-bit stores take the short FP32 form of the collision (their rounding is far coarser)
      constexpr bool SHORT = sizeof(typename Fmt<FMT>::raw) == 2 && std::is_same<C, float>::value;
      collide<LAT, COLL, C, SHORT>(f, (C)p.omega, has_force, F, (C)p.cparam);
    }
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
        const C faux = fin.at(l, src, sx, sy, sz, T.w[l]);
        const C fps = fin.own(T.opp[l], cell, T.w[T.opp[l]]);
        f[T.opp[l]] = (C(1) - cs) * fps + cs * faux;
      }
    }
  }
}

// Exponents of the Q pulls of a block-format cell: k of each pull's source segment.  (xr, yr) is
// the wave's row (3-D: x, y; 2-D: x, with yr unused), uniform across the wave; `fast` the lane's
// coordinate on the fast axis, `fb` the block's fast segment.  Pulls along the fast axis cross
// into the neighbouring segment only at segment edges, and only when the row has more than one
// segment (MSEG).
template <int LAT, bool MSEG>
__device__ __forceinline__ void blk_pull_exps(const StepParams& p, int xr, int yr, int fast, int fb,
                                              int (&kk)[Lat<LAT>::Q]) {
  constexpr LatTables T = make_tables(LAT);
  constexpr bool three = T.d == 3;
  const Sweep& sw = p.sw;
  const int S = (int)blk_segs_fast(sw);
  const int nfast = three ? sw.nz : sw.ny;
  const int fb_lo = fb == 0 ? S - 1 : fb - 1, fb_hi = fb == S - 1 ? 0 : fb + 1;
  const bool lo_cross = MSEG && ((fast == 0 ? nfast - 1 : fast - 1) >> sw.lbz) != fb;
  const bool hi_cross = MSEG && ((fast == nfast - 1 ? 0 : fast + 1) >> sw.lbz) != fb;
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    const int cf = three ? T.c[2][l] : T.c[1][l];
    uint32_t row;
    if constexpr (three) {
      const int xs = wrap1(xr - T.c[0][l], sw.nx), ys = wrap1(yr - T.c[1][l], sw.ny);
      row = ((uint32_t)xs * sw.ny + ys) * S;
    } else {
      row = (uint32_t)wrap1(xr - T.c[0][l], sw.nx) * S;
    }
    int k = p.e_in[row + fb];
    if constexpr (MSEG) {
      if (cf == 1) k = lo_cross ? p.e_in[row + fb_lo] : k;
      if (cf == -1) k = hi_cross ? p.e_in[row + fb_hi] : k;
    }
    kk[l] = k;
  }
}

// One thread per cell: Q pulls, the update, Q stores.  For the block format every lane of the
// block stays to the segment reduction: lanes outside the grid read a cell of their own row and
// write nothing, and solid cells write nothing.
template <int LAT, int COLL, int FMT, class C, bool WIDE, bool MSEG>
__device__ __forceinline__ void step_body(const StepParams& p) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  constexpr bool BLK = FMT == XLBH_I16_BLOCK;
  using R = typename Fmt<FMT>::raw;
  using O = typename std::conditional<WIDE, uint64_t, uint32_t>::type;

  int x, y, z;
  const bool live = sweep_cell_range(p.sw, p.x0, p.x1, x, y, z);
  const int nx = p.sw.nx, ny = p.sw.ny, nz = p.sw.nz;
  if constexpr (!BLK) {
    if (!live) return;
  } else {
    x = min(x, nx - 1);
    y = min(y, ny - 1);
    z = min(z, nz - 1);
  }
  uniform_slow_axes(T.d == 3, x, y);
  const O cell = ((O)x * ny + y) * nz + z;
  // cell = row + lane: the row's first cell (wave-uniform) and the offset along the row
  const O row = T.d == 3 ? ((O)x * ny + y) * nz : (O)x * ny;
  const O lane = T.d == 3 ? (O)z : (O)y;
  const size_t plane_bytes = (size_t)p.plane * sizeof(R);
  const Planes<R, O> fout{static_cast<const char*>(p.f_out), plane_bytes};
  PopIn<FMT, C, O> fin{{static_cast<const char*>(p.f_in), plane_bytes}, p.e_in, p.sw, 0};
  const PullOffsets<O> po(x, y, z, nx, ny, nz);

  // The boundary id is loaded first and the pulls right after it, with no wait in between (every
  // source address is in bounds for every cell).  The empty asm statement keeps the compiler from
  // sinking the pulls below the solid-cell exit, which would put a round trip in front of them.
  // Loads complete in issue order, so a boundary cell's table and missing-bit loads, which need
  // only the id, are issued while its pulls are still in flight.
  int id = 0;
  if (p.bc_mask != nullptr) id = __builtin_nontemporal_load(p.bc_mask + cell);
  C f[Q];
  [[maybe_unused]] int xr = 0, yr = 0, fb = 0;
  if constexpr (BLK) {
    // The exponent loads (scalar: nothing before them writes memory) go out first, the pulls
    // right behind them; the scales are formed once both are in flight.
    constexpr bool three = T.d == 3;
    xr = __builtin_amdgcn_readfirstlane(x);
    yr = three ? __builtin_amdgcn_readfirstlane(y) : 0;
    fb = blockIdx.x;
    int kk[Q];
    blk_pull_exps<LAT, MSEG>(p, xr, yr, three ? z : y, fb, kk);
    fin.k_own = kk[0];  // the rest direction pulls from the cell itself
    int m[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) m[l] = ld_stream(fin.pl.at(l, po.template uni<LAT>(l), po.template lane<LAT>(l)));
    asm volatile("" ::: "memory");
#pragma unroll
    for (int l = 0; l < Q; ++l) f[l] = __builtin_fmaf((float)m[l], blk_dscale((float)T.w[l], kk[l]), (float)T.w[l]);
  } else {
#pragma unroll
    for (int l = 0; l < Q; ++l)
      f[l] = ld_pop<FMT, C>(fin.pl.at(l, po.template uni<LAT>(l), po.template lane<LAT>(l)), T.w[l]);
  }
  asm volatile("" ::: "memory");
  if constexpr (!BLK) {
    if (id == 255) return;  // solid: no update (nse_stepper.py:356-358)
    cell_update<LAT, COLL, FMT, C, O>(f, id, cell, x, y, z, p, fin);
#pragma unroll
    for (int l = 0; l < Q; ++l) st_pop<FMT, C>(fout.at_w(l, row, lane), f[l], T.w[l]);
  } else {
    // Lanes that write nothing (outside the grid, or solid) run the fluid update on what they
    // pulled, unbranched, and stay out of the exponent and the stores.
    const bool write = live && id != 255;
    if (write) cell_update<LAT, COLL, FMT, C, O>(f, id, cell, x, y, z, p, fin);
    // x_l = f_l / w_l - 1, once: its largest magnitude sets the segment's exponent and its
    // scaled value is the mantissa.  A segment spans whole waves, so k is wave-uniform.
    float amax = 0.0f;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      f[l] = __builtin_fmaf(f[l], (float)(1.0 / T.w[l]), -1.0f);
      amax = __builtin_elementwise_maximum(amax, __builtin_fabsf(f[l]));
    }
    const int k = __builtin_amdgcn_readfirstlane(blk_segment_k(write ? amax : 0.0f, p.sw.lbz));
    if ((threadIdx.x & ((1u << p.sw.lbz) - 1)) == 0) p.e_out[blk_row_seg(p.sw, xr, yr, fb)] = k;
    if (write) {
      if (k != kBlkNonFinite) {
        const float pk = blk_pow2(k);
#pragma unroll
        for (int l = 0; l < Q; ++l) st_stream(fout.at_w(l, row, lane), blk_mantissa(f[l], pk));
      } else {
#pragma unroll
        for (int l = 0; l < Q; ++l) st_stream(fout.at_w(l, row, lane), (int16_t)32767);
      }
    }
  }
}

// Occupancy.  Measured on the 256^3 D3Q19 sweep: 4-byte populations stream best at 4 waves per
// SIMD (more waves add concurrent streams without adding bandwidth: 6 waves lost 3 %); 2-byte
// populations want more waves than that.  The 4- and 8-byte kernels are launched with an LDS
// reservation (kStepLdsCap, otherwise unused) that admits 4 blocks = 4 waves per SIMD per CU.
// The D3Q19 2-byte kernels with FP32 arithmetic are compiled for 6 waves per SIMD (80 VGPRs, no
// spills; the default allocation took 82 VGPRs = 5 waves): +2 % on the cavity, whose boundary
// lookups the extra waves hide, and -1 % on a periodic box.  D3Q27 would spill at 80 VGPRs.
constexpr unsigned kStepLdsCap = 36 * 1024;

// MSEG (block format only): rows of more than one exponent segment.  Its own kernel, so each
// form gets its own register allocation.
template <int LAT, int COLL, int FMT, class C, bool WIDE, bool MSEG = false>
__global__ __launch_bounds__(256) void k_step(StepParams p) {
  step_body<LAT, COLL, FMT, C, WIDE, MSEG>(p);
}

template <int LAT, int COLL, int FMT, class C, bool WIDE, bool MSEG = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_step_w6(StepParams p) {
  step_body<LAT, COLL, FMT, C, WIDE, MSEG>(p);
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
  uint64_t plane;  // population plane stride (elements, >= N)
  double omega;
  double cparam;  // Smagorinsky constant (XLBH_SMAGORINSKY_BGK)
  hipStream_t stream;
  // Planes x0 <= x < x1 to update (x1 < 0: all).  With 1 <= x0 and x1 <= nx - 1 no pull wraps
  // in x: planes x0 - 1 and x1 are read as neighbours and left unwritten (ghost planes).
  int x0 = 0, x1 = -1;
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
  p.x0 = c.x1 >= 0 ? c.x0 : 0;
  p.x1 = c.x1 >= 0 ? c.x1 : c.nx;
  p.N = (uint32_t)((int64_t)c.nx * c.ny * c.nz);
  p.plane = c.plane;
  p.e_in = nullptr;
  p.e_out = nullptr;
  if constexpr (FMT == XLBH_I16_BLOCK) {
    const size_t off = blk_exps_offset(Lat<LAT>::Q, c.plane);
    p.e_in = reinterpret_cast<const int32_t*>(static_cast<const char*>(c.f_in) + off);
    p.e_out = reinterpret_cast<int32_t*>(static_cast<char*>(c.f_out) + off);
  }
  // 32-bit per-lane byte offsets reach every plane of the field from its base
  const bool wide = (uint64_t)Lat<LAT>::Q * p.plane * sizeof(typename Fmt<FMT>::raw) > 0xffffffffull;
  const unsigned lds = sizeof(typename Fmt<FMT>::raw) >= 4 ? kStepLdsCap : 0u;
  constexpr bool w6 = LAT == XLBH_D3Q19 && sizeof(typename Fmt<FMT>::raw) == 2 && std::is_same<C, float>::value;
  auto launch = [&](auto mseg) {
    constexpr bool MS = decltype(mseg)::value;
    if (wide)
      hipLaunchKernelGGL((k_step<LAT, COLL, FMT, C, true, MS>), sweep_grid_range(p.sw, p.x0, p.x1), dim3(256), lds, c.stream, p);
    else if constexpr (w6)
      hipLaunchKernelGGL((k_step_w6<LAT, COLL, FMT, C, false, MS>), sweep_grid_range(p.sw, p.x0, p.x1), dim3(256), lds, c.stream, p);
    else
      hipLaunchKernelGGL((k_step<LAT, COLL, FMT, C, false, MS>), sweep_grid_range(p.sw, p.x0, p.x1), dim3(256), lds, c.stream, p);
  };
  if constexpr (FMT == XLBH_I16_BLOCK) {
    if (blk_segs_fast(p.sw) > 1)
      launch(std::true_type{});
    else
      launch(std::false_type{});
  } else {
    launch(std::false_type{});
  }
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
      case XLBH_I16_BLOCK: return launch_step_t<LAT, COLL, XLBH_I16_BLOCK, float>(c);           \
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
  if (c.compute == XLBH_COMPUTE_F32_DEVIATION) {
    if (c.collision == XLBH_BGK && c.store == XLBH_F32) return launch_step_t<LAT, kBgkDeviation, XLBH_F32, float>(c);
    set_error("XLBH_COMPUTE_F32_DEVIATION is implemented for the BGK collision with FP32 storage");
    return 1;
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
