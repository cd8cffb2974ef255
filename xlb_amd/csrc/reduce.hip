// FP64 reductions over the lattice: momentum-exchange force and global mass/momentum.
// Two passes, both in a fixed order, so results are bitwise reproducible run to run:
// pass 1 reduces each block's kRedCells cells (each thread sums kRedRun cells strided by 256,
// then a wave64 shuffle tree and the 4 wave sums in LDS) into a per-block partial; pass 2 has one
// block sum the partials in a fixed strided order + tree.  With 4096 cells per block a 256^3 grid
// leaves 4096 partials, which one block sums in a few microseconds.
#include <mutex>

#include "physics.hpp"

namespace xlbh {

constexpr int kRedWidth = 4;  // doubles per partial
constexpr int kRedRun = 16;   // cells per thread
constexpr uint32_t kRedCells = 256u * kRedRun;

inline uint32_t red_blocks(uint32_t n_cells) { return (n_cells + kRedCells - 1) / kRedCells; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block of 256 threads -> out[0..W-1] written by thread 0.
template <int W>
__device__ __forceinline__ void block_sum(double (&v)[W], double* out) {
  __shared__ double sh[4][W];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < W; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < W; ++k) sh[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < W; ++k) out[k] = (sh[0][k] + sh[1][k]) + (sh[2][k] + sh[3][k]);
}

__global__ __launch_bounds__(256) void k_final_sum(const double* __restrict__ part, int nparts, int width,
                                                   double* __restrict__ out) {
  double v[kRedWidth] = {0, 0, 0, 0};
  for (int i = threadIdx.x; i < nparts; i += 256)
    for (int k = 0; k < width; ++k) v[k] += part[(size_t)i * kRedWidth + k];
  double res[kRedWidth];
  block_sum<kRedWidth>(v, res);
  if (threadIdx.x == 0)
    for (int k = 0; k < width; ++k) out[k] = res[k];
}

// One boundary cell's momentum exchange, added to m: the post-collision populations f_pc, the
// re-pulled post-stream populations with the BC's streaming functional applied (f_ps), and the sum
// over missing l of c_{opp l} (f_pc[opp l] + f_ps[l]) (momentum_transfer.py:137-160).
// Population l of cell (x, y, z) of field fin (plane stride P); the block format looks up the
// cell's segment exponent in e.
template <int LAT, int FMT, class C>
__device__ __forceinline__ C ld_cell(const typename Fmt<FMT>::raw* __restrict__ fin, uint64_t P, const int32_t* e,
                                     const Sweep& sw, int l, int x, int y, int z) {
  constexpr LatTables T = make_tables(LAT);
  const uint32_t cell = ((uint32_t)x * sw.ny + y) * sw.nz + z;
  if constexpr (FMT == XLBH_I16_BLOCK)
    return (C)__builtin_fmaf((float)fin[(size_t)l * P + cell], blk_dscale((float)T.w[l], e[blk_seg(sw, x, y, z)]),
                             (float)T.w[l]);
  else
    return ld_pop<FMT, C>(fin + (size_t)l * P + cell, T.w[l]);
}

template <int LAT, int FMT, class C>
__device__ __forceinline__ void momentum_cell(const typename Fmt<FMT>::raw* __restrict__ fin, uint32_t mm,
                                              const C* aux, int kind, const xlbh_bc_desc& desc, const Sweep& sw,
                                              uint64_t P, const int32_t* e, uint32_t cell, double (&m)[kRedWidth]) {
  constexpr LatTables T = make_tables(LAT);
  constexpr int Q = T.q;
  const int nx = sw.nx, ny = sw.ny, nz = sw.nz;
  const int z = cell % nz;
  const int y = (cell / nz) % ny;
  const int x = cell / ((uint32_t)nz * ny);
  C fpc[Q], fps[Q];
#pragma unroll
  for (int l = 0; l < Q; ++l) {
    fpc[l] = ld_cell<LAT, FMT, C>(fin, P, e, sw, l, x, y, z);
    const int sx = wrap(x - T.c[0][l], nx), sy = wrap(y - T.c[1][l], ny), sz = wrap(z - T.c[2][l], nz);
    fps[l] = ld_cell<LAT, FMT, C>(fin, P, e, sw, l, sx, sy, sz);
  }
  if (kind == XLBH_BC_HALFWAY_BB || kind == XLBH_BC_EXTRAPOLATION_OUTFLOW) {
#pragma unroll
    for (int l = 0; l < Q; ++l)
      if (bit(mm, l)) fps[l] = fpc[T.opp[l]];
  } else if (kind == XLBH_BC_FULLWAY_BB) {
#pragma unroll
    for (int l = 0; l < Q; ++l) fps[l] = fpc[T.opp[l]];
  } else if (kind == XLBH_BC_DO_NOTHING) {
#pragma unroll
    for (int l = 0; l < Q; ++l) fps[l] = fpc[l];
  } else if (kind == XLBH_BC_EQUILIBRIUM) {
    const C rb = (C)desc.rho;
    const C ub[3] = {(C)desc.u[0], (C)desc.u[1], (C)desc.u[2]};
    const C us = usqr15<LAT>(ub);
#pragma unroll
    for (int l = 0; l < Q; ++l) fps[l] = feq_l<LAT>(l, rb, ub, us);
  } else if (kind >= XLBH_BC_ZOUHE_VELOCITY && kind <= XLBH_BC_REGULARIZED_PRESSURE) {
    zouhe_cell<LAT>(fps, mm, aux[cell], kind == XLBH_BC_ZOUHE_VELOCITY || kind == XLBH_BC_REGULARIZED_VELOCITY,
                    kind >= XLBH_BC_REGULARIZED_VELOCITY);
  } else if (kind == XLBH_BC_GRADS_APPROXIMATION) {
    grads_cell<LAT>(fps, mm, [&](int l) { return fpc[l]; });
  }
#pragma unroll
  for (int d = 0; d < T.d; ++d) {
    C md = C(0);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
      if (bit(mm, l)) {
        const C phi = fpc[T.opp[l]] + fps[l];
        if (T.c[d][T.opp[l]] == 1) md += phi;
        else if (T.c[d][T.opp[l]] == -1) md -= phi;
      }
    }
    m[d] += (double)md;
  }
}

// Momentum exchange (momentum_transfer.py:108-160): at cells with bc_mask == id whose rest
// direction is not missing, re-pull the post-stream populations, apply the BC's streaming
// functional, and sum over missing l of c_{opp l} (f_pc[opp l] + f_ps[l]).  The kRedRun ids of a
// thread are loaded together before any cell is processed.
template <int LAT, int FMT, class C>
__global__ __launch_bounds__(256) void k_momentum(const typename Fmt<FMT>::raw* __restrict__ fin,
                                                  const uint8_t* __restrict__ bc_mask,
                                                  const uint32_t* __restrict__ missing, const C* aux, int kind,
                                                  int id, xlbh_bc_desc desc, Sweep sw, uint32_t N, uint64_t P,
                                                  const int32_t* e, double* __restrict__ part) {
  const uint32_t base = blockIdx.x * kRedCells + threadIdx.x;
  uint8_t ids[kRedRun];
#pragma unroll
  for (int k = 0; k < kRedRun; ++k) {
    const uint32_t cell = base + 256u * k;
    ids[k] = cell < N ? bc_mask[cell] : 0;
  }
  double m[kRedWidth] = {0, 0, 0, 0};
  for (int k = 0; k < kRedRun; ++k) {
    const uint32_t cell = base + 256u * k;
    if (ids[k] != id) continue;
    const uint32_t mm = missing[cell];
    if (!bit(mm, 0)) momentum_cell<LAT, FMT, C>(fin, mm, aux, kind, desc, sw, P, e, cell, m);
  }
  block_sum<kRedWidth>(m, part + (size_t)blockIdx.x * kRedWidth);
}

// sum rho and sum c f over non-solid cells, accumulated in FP64 from the stored populations.
template <int LAT, int FMT>
__global__ __launch_bounds__(256) void k_mass_momentum(const typename Fmt<FMT>::raw* __restrict__ fin,
                                                       const uint8_t* __restrict__ bc_mask, Sweep sw, uint32_t N,
                                                       uint64_t P, const int32_t* e, double* __restrict__ part) {
  constexpr LatTables T = make_tables(LAT);
  double v[kRedWidth] = {0, 0, 0, 0};
  for (int k = 0; k < kRedRun; ++k) {
    const uint32_t cell = blockIdx.x * kRedCells + threadIdx.x + 256u * k;
    if (cell < N && (bc_mask == nullptr || bc_mask[cell] != 255)) {
      int kb = 0;
      if constexpr (FMT == XLBH_I16_BLOCK) {
        const int z = cell % sw.nz, y = (cell / sw.nz) % sw.ny, x = cell / ((uint32_t)sw.nz * sw.ny);
        kb = e[blk_seg(sw, x, y, z)];
      }
#pragma unroll
      for (int l = 0; l < T.q; ++l) {
        double f;
        if constexpr (FMT == XLBH_I16_BLOCK)
          f = (double)__builtin_fmaf((float)fin[(size_t)l * P + cell], blk_dscale((float)T.w[l], kb), (float)T.w[l]);
        else
          f = ld_pop<FMT, double>(fin + (size_t)l * P + cell, T.w[l]);
        v[0] += f;
#pragma unroll
        for (int d = 0; d < T.d; ++d) v[1 + d] += (double)T.c[d][l] * f;
      }
    }
  }
  block_sum<kRedWidth>(v, part + (size_t)blockIdx.x * kRedWidth);
}

}  // namespace xlbh

using namespace xlbh;

extern "C" {

int64_t xlbh_reduce_workspace_bytes(int64_t n_cells) {
  return ((n_cells + kRedCells - 1) / kRedCells) * kRedWidth * (int64_t)sizeof(double);
}

}  // extern "C"

// Momentum exchange with the per-block partials in `part` (red_blocks(N) x kRedWidth doubles).
static int momentum_transfer_into(int lattice, int store, int compute, int kind, int bc_id, const xlbh_bc_desc* desc,
                                  const void* f, const uint8_t* bc_mask, const uint32_t* missing_bits,
                                  const void* aux, double* out, double* part, int nx, int ny, int nz, int64_t plane,
                                  void* stream) {
  XLBH_REQUIRE(f && bc_mask && missing_bits && out && part, "momentum_transfer: null argument");
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  const Sweep sw = make_sweep(nx, ny, nz);
  XLBH_REQUIRE(bc_id >= 1 && bc_id <= 254, "momentum_transfer: bc_id must be 1..254");
  XLBH_REQUIRE(!(kind >= XLBH_BC_ZOUHE_VELOCITY && kind <= XLBH_BC_REGULARIZED_PRESSURE) || aux,
               "momentum_transfer: Zou-He / regularized BCs need the aux table");
  xlbh_bc_desc d{};
  if (desc) d = *desc;
  const uint32_t N = (uint32_t)((int64_t)nx * ny * nz);
  const int nb = (int)red_blocks(N);
  hipStream_t s = (hipStream_t)stream;
  return with_lattice(lattice, [&](auto L) -> int {
    constexpr int LAT = decltype(L)::value;
    auto go = [&](auto S, auto c) -> int {
      constexpr int FMT = decltype(S)::value;
      using C = decltype(c);
      const int32_t* ex = reinterpret_cast<const int32_t*>(static_cast<const char*>(f) +
                                                           blk_exps_offset(Lat<LAT>::Q, P));
      hipLaunchKernelGGL((k_momentum<LAT, FMT, C>), dim3(nb), dim3(256), 0, s,
                         (const typename Fmt<FMT>::raw*)f, bc_mask, missing_bits, (const C*)aux, kind, bc_id, d,
                         sw, N, P, ex, part);
      XLBH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, s, part, nb, Lat<LAT>::D, out);
      XLBH_LAUNCH_CHECK();
      return 0;
    };
    compute = arith_compute(compute);
    if (compute == XLBH_COMPUTE_F32) {
      switch (store) {
        case XLBH_F32: return go(std::integral_constant<int, XLBH_F32>{}, float{});
        case XLBH_F16: return go(std::integral_constant<int, XLBH_F16>{}, float{});
        case XLBH_BF16: return go(std::integral_constant<int, XLBH_BF16>{}, float{});
        case XLBH_F16_SHIFTED: return go(std::integral_constant<int, XLBH_F16_SHIFTED>{}, float{});
        case XLBH_BF16_SHIFTED: return go(std::integral_constant<int, XLBH_BF16_SHIFTED>{}, float{});
        case XLBH_I16_BLOCK: return go(std::integral_constant<int, XLBH_I16_BLOCK>{}, float{});
      }
    } else if (compute == XLBH_COMPUTE_F64) {
      switch (store) {
        case XLBH_F64: return go(std::integral_constant<int, XLBH_F64>{}, double{});
        case XLBH_F32: return go(std::integral_constant<int, XLBH_F32>{}, double{});
        case XLBH_F16: return go(std::integral_constant<int, XLBH_F16>{}, double{});
      }
    }
    set_error("unsupported (store, compute) combination");
    return 1;
  });
}

// Scratch of xlbh_momentum_transfer_noalloc: the per-block partials in a file-scope device array
// (each GPU the library runs on holds its own copy, 2 MiB, until the process ends).  Host code
// never passes the array's name to a kernel (on the host it is a stand-in symbol, not a device
// address): the entry point resolves each device's copy once with hipGetSymbolAddress, with that
// device current, and passes that pointer.
constexpr uint32_t kScratchParts = 1u << 16;  // partials: grids of up to 2^16 * 4096 = 2^28 cells
static __device__ __align__(256) double scratch[kScratchParts * kRedWidth];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_partials(double** out) {
  int dev = 0;
  XLBH_HIP_CHECK(hipGetDevice(&dev));
  XLBH_REQUIRE(dev >= 0 && dev < kMaxDevices, "momentum_transfer_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev] == nullptr) {
    void* p = nullptr;
    XLBH_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    XLBH_REQUIRE(p != nullptr, "momentum_transfer_noalloc: no device address for the scratch array");
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}
}  // namespace

extern "C" {

int xlbh_momentum_transfer(int lattice, int store, int compute, int kind, int bc_id, const xlbh_bc_desc* desc,
                           const void* f, const uint8_t* bc_mask, const uint32_t* missing_bits, const void* aux,
                           double* out, void* workspace, int nx, int ny, int nz, int64_t plane, void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  return momentum_transfer_into(lattice, store, compute, kind, bc_id, desc, f, bc_mask, missing_bits, aux, out,
                                static_cast<double*>(workspace), nx, ny, nz, plane, stream);
}

int xlbh_momentum_transfer_noalloc(int lattice, int store, int compute, int kind, int bc_id,
                                   const xlbh_bc_desc* desc, const void* f, const uint8_t* bc_mask,
                                   const uint32_t* missing_bits, const void* aux, double* out, int nx, int ny,
                                   int nz, int64_t plane, void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(red_blocks((uint32_t)((int64_t)nx * ny * nz)) <= kScratchParts,
               "momentum_transfer_noalloc: grid larger than its scratch allows (at most 2^28 cells); use "
               "xlbh_momentum_transfer with a workspace");
  double* part = nullptr;
  if (int e = scratch_partials(&part)) return e;
  return momentum_transfer_into(lattice, store, compute, kind, bc_id, desc, f, bc_mask, missing_bits, aux, out,
                                part, nx, ny, nz, plane, stream);
}

int xlbh_mass_momentum(int lattice, int store, const void* f, const uint8_t* bc_mask, double* out, void* workspace,
                       int nx, int ny, int nz, int64_t plane, void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(f && out && workspace, "mass_momentum: null argument");
  XLBH_REQUIRE(valid_store(store), "unknown store format");
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  const Sweep sw = make_sweep(nx, ny, nz);
  const uint32_t N = (uint32_t)((int64_t)nx * ny * nz);
  const int nb = (int)red_blocks(N);
  double* part = static_cast<double*>(workspace);
  hipStream_t s = (hipStream_t)stream;
  return with_lattice(lattice, [&](auto L) -> int {
    constexpr int LAT = decltype(L)::value;
    auto go = [&](auto S) -> int {
      constexpr int FMT = decltype(S)::value;
      const int32_t* ex = reinterpret_cast<const int32_t*>(static_cast<const char*>(f) +
                                                           blk_exps_offset(Lat<LAT>::Q, P));
      hipLaunchKernelGGL((k_mass_momentum<LAT, FMT>), dim3(nb), dim3(256), 0, s,
                         (const typename Fmt<FMT>::raw*)f, bc_mask, sw, N, P, ex, part);
      XLBH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, s, part, nb, 1 + Lat<LAT>::D, out);
      XLBH_LAUNCH_CHECK();
      return 0;
    };
    switch (store) {
      case XLBH_F32: return go(std::integral_constant<int, XLBH_F32>{});
      case XLBH_F64: return go(std::integral_constant<int, XLBH_F64>{});
      case XLBH_F16: return go(std::integral_constant<int, XLBH_F16>{});
      case XLBH_BF16: return go(std::integral_constant<int, XLBH_BF16>{});
      case XLBH_F16_SHIFTED: return go(std::integral_constant<int, XLBH_F16_SHIFTED>{});
      case XLBH_BF16_SHIFTED: return go(std::integral_constant<int, XLBH_BF16_SHIFTED>{});
      case XLBH_I16_BLOCK: return go(std::integral_constant<int, XLBH_I16_BLOCK>{});
    }
    set_error("unknown store format");
    return 1;
  });
}

}  // extern "C"
