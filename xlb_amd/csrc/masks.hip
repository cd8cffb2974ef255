// Boundary masks: the indices masker, bool -> bit packing, ordered cell compaction and the
// aux side-table scatter.  All integer work; results are bit-exact and deterministic.
#include "common.hpp"

namespace xlbh {

// One thread per boundary index (indices_boundary_masker.py:114-156, Warp semantics).
// Every write of one launch carries the same id / sets the same bits, so the result does not
// depend on thread order; successive launches apply successive BCs in list order.
template <int LAT>
__global__ __launch_bounds__(256) void k_mask_indices(const int32_t* __restrict__ idx,
                                                      const uint8_t* __restrict__ interior, int n, uint8_t id,
                                                      uint8_t* bc_mask, uint8_t* missing_bool,
                                                      uint32_t* missing_bits, int nx, int ny, int nz,
                                                      uint32_t N) {
  constexpr LatTables T = make_tables(LAT);
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int x = idx[i], y = idx[(size_t)n + i], z = idx[2 * (size_t)n + i];
  auto inside = [&](int a, int b, int e) { return a >= 0 && a < nx && b >= 0 && b < ny && e >= 0 && e < nz; };
  if (!inside(x, y, z)) return;
  const uint32_t p = ((uint32_t)x * ny + y) * nz + z;
  bc_mask[p] = id;
  const bool inter = interior[i] != 0;
  uint32_t own = 0;
#pragma unroll
  for (int l = 0; l < T.q; ++l) {
    const int px = x - T.c[0][l], py = y - T.c[1][l], pz = z - T.c[2][l];
    if (!inside(px, py, pz)) {
      own |= 1u << l;
      if (missing_bool) missing_bool[(size_t)l * N + p] = 1;
    } else if (inter) {
      const int qx = x + T.c[0][l], qy = y + T.c[1][l], qz = z + T.c[2][l];
      if (!inside(qx, qy, qz)) continue;  // only reachable with an inconsistent interior flag
      const uint32_t q = ((uint32_t)qx * ny + qy) * nz + qz;
      if (missing_bool) missing_bool[(size_t)l * N + q] = 1;
      if (missing_bits) atomicOr(missing_bits + q, 1u << l);
      bc_mask[q] = id;
    }
  }
  if (missing_bits && own) atomicOr(missing_bits + p, own);
}

template <int LAT>
__global__ __launch_bounds__(256) void k_pack_missing(const uint8_t* __restrict__ mb, uint32_t* __restrict__ bits,
                                                      uint32_t N) {
  constexpr LatTables T = make_tables(LAT);
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= N) return;
  uint32_t b = 0;
#pragma unroll
  for (int l = 0; l < T.q; ++l) b |= (mb[(size_t)l * N + c] != 0 ? 1u : 0u) << l;
  bits[c] = b;
}

// ---- ordered compaction: per-block counts, one-block exclusive scan, ordered writes ----------
__device__ __forceinline__ bool cell_selected(const uint8_t* m, uint32_t c, uint32_t N, int id) {
  if (c >= N) return false;
  const int v = m[c];
  return id < 0 ? (v != 0 && v != 255) : (v == id);
}

__global__ __launch_bounds__(256) void k_compact_count(const uint8_t* __restrict__ m, int id, uint32_t N,
                                                       int32_t* __restrict__ counts) {
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  const int n = __syncthreads_count(cell_selected(m, c, N, id) ? 1 : 0);
  if (threadIdx.x == 0) counts[blockIdx.x] = n;
}

__global__ __launch_bounds__(1024) void k_compact_scan(int32_t* counts, int32_t* offsets, int nblocks,
                                                       int32_t* total) {
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  const int per = (nblocks + 1023) / 1024;
  const int b0 = t * per, b1 = min(nblocks, b0 + per);
  int32_t s = 0;
  for (int b = b0; b < b1; ++b) s += counts[b];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int32_t v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int32_t run = part[t] - s;  // exclusive prefix of this thread's range
  for (int b = b0; b < b1; ++b) {
    offsets[b] = run;
    run += counts[b];
  }
  if (t == 1023) *total = part[1023];
}

__global__ __launch_bounds__(256) void k_compact_write(const uint8_t* __restrict__ m, int id, uint32_t N,
                                                       const int32_t* __restrict__ offsets,
                                                       int32_t* __restrict__ cells) {
  __shared__ int32_t wave_base[4];
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  const bool sel = cell_selected(m, c, N, id);
  const unsigned long long ball = __ballot(sel);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wave_base[wave] = __popcll(ball);
  __syncthreads();
  int base = offsets[blockIdx.x];
  for (int w = 0; w < wave; ++w) base += wave_base[w];
  const int pre = __popcll(ball & ((1ull << lane) - 1ull));
  if (sel) cells[base + pre] = (int32_t)c;
}

template <class C>
__global__ __launch_bounds__(256) void k_scatter(const int32_t* __restrict__ cells, const C* __restrict__ v, int n,
                                                 C* __restrict__ aux) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) aux[cells[i]] = v[i];
}

}  // namespace xlbh

using namespace xlbh;

extern "C" {

int xlbh_mask_indices(int lattice, const int32_t* indices, const uint8_t* interior, int n, int bc_id,
                      uint8_t* bc_mask, uint8_t* missing_bool, uint32_t* missing_bits, int nx, int ny, int nz,
                      void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(n >= 0, "negative index count");
  XLBH_REQUIRE(bc_id >= 0 && bc_id <= 255, "bc_id must fit in uint8");
  XLBH_REQUIRE(bc_mask && (missing_bool || missing_bits), "mask_indices needs bc_mask and a missing mask");
  if (n == 0) return 0;
  XLBH_REQUIRE(indices && interior, "mask_indices needs indices and interior flags");
  const uint32_t N = (uint32_t)((int64_t)nx * ny * nz);
  hipStream_t s = (hipStream_t)stream;
  const dim3 g((n + 255) / 256);
  switch (lattice) {
    case XLBH_D2Q9:
      hipLaunchKernelGGL(k_mask_indices<XLBH_D2Q9>, g, dim3(256), 0, s, indices, interior, n, (uint8_t)bc_id,
                         bc_mask, missing_bool, missing_bits, nx, ny, nz, N);
      break;
    case XLBH_D3Q19:
      hipLaunchKernelGGL(k_mask_indices<XLBH_D3Q19>, g, dim3(256), 0, s, indices, interior, n, (uint8_t)bc_id,
                         bc_mask, missing_bool, missing_bits, nx, ny, nz, N);
      break;
    default:
      hipLaunchKernelGGL(k_mask_indices<XLBH_D3Q27>, g, dim3(256), 0, s, indices, interior, n, (uint8_t)bc_id,
                         bc_mask, missing_bool, missing_bits, nx, ny, nz, N);
  }
  XLBH_LAUNCH_CHECK();
  return 0;
}

int xlbh_pack_missing(int lattice, const uint8_t* missing_bool, uint32_t* missing_bits, int nx, int ny, int nz,
                      void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(missing_bool && missing_bits, "pack_missing needs both masks");
  const uint32_t N = (uint32_t)((int64_t)nx * ny * nz);
  hipStream_t s = (hipStream_t)stream;
  const dim3 g((N + 255) / 256);
  switch (lattice) {
    case XLBH_D2Q9: hipLaunchKernelGGL(k_pack_missing<XLBH_D2Q9>, g, dim3(256), 0, s, missing_bool, missing_bits, N); break;
    case XLBH_D3Q19: hipLaunchKernelGGL(k_pack_missing<XLBH_D3Q19>, g, dim3(256), 0, s, missing_bool, missing_bits, N); break;
    default: hipLaunchKernelGGL(k_pack_missing<XLBH_D3Q27>, g, dim3(256), 0, s, missing_bool, missing_bits, N);
  }
  XLBH_LAUNCH_CHECK();
  return 0;
}

int64_t xlbh_compact_workspace_bytes(int64_t n_cells) {
  const int64_t nb = (n_cells + 255) / 256;
  return 2 * nb * (int64_t)sizeof(int32_t);
}

int xlbh_compact_cells(const uint8_t* bc_mask, int bc_id, int32_t* cells, int32_t* count, void* workspace, int nx,
                       int ny, int nz, void* stream) {
  clear_error();
  XLBH_REQUIRE(nx > 0 && ny > 0 && nz > 0, "grid extents must be positive");
  XLBH_REQUIRE(bc_mask && cells && count && workspace, "compact_cells needs mask, outputs and workspace");
  XLBH_REQUIRE(bc_id >= -1 && bc_id <= 255, "bc_id out of range");
  const int64_t N64 = (int64_t)nx * ny * nz;
  XLBH_REQUIRE(N64 < (int64_t)1 << 31, "compact_cells supports fewer than 2^31 cells");
  const uint32_t N = (uint32_t)N64;
  const int nb = (int)((N + 255) / 256);
  int32_t* counts = static_cast<int32_t*>(workspace);
  int32_t* offsets = counts + nb;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(256), 0, s, bc_mask, bc_id, N, counts);
  XLBH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(1024), 0, s, counts, offsets, nb, count);
  XLBH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_compact_write, dim3(nb), dim3(256), 0, s, bc_mask, bc_id, N, offsets, cells);
  XLBH_LAUNCH_CHECK();
  return 0;
}

int xlbh_scatter_aux(int compute, const int32_t* cells, const void* values, int n, void* aux, void* stream) {
  clear_error();
  XLBH_REQUIRE(n >= 0, "negative count");
  if (n == 0) return 0;
  XLBH_REQUIRE(cells && values && aux, "scatter_aux needs cells, values and aux");
  hipStream_t s = (hipStream_t)stream;
  const dim3 g((n + 255) / 256);
  compute = arith_compute(compute);
  if (compute == XLBH_COMPUTE_F64)
    hipLaunchKernelGGL(k_scatter<double>, g, dim3(256), 0, s, cells, (const double*)values, n, (double*)aux);
  else if (compute == XLBH_COMPUTE_F32)
    hipLaunchKernelGGL(k_scatter<float>, g, dim3(256), 0, s, cells, (const float*)values, n, (float*)aux);
  else {
    set_error("unknown compute type");
    return 1;
  }
  XLBH_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
