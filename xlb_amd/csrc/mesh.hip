// Surface voxelisation of a triangle mesh into the boundary masks (the reference's
// MeshBoundaryMasker, xlb/operator/boundary_masker/mesh_boundary_masker.py:49-236).
//
// A voxel (i, j, k) is the unit box [i, i+1] x [j, j+1] x [k, k+1].  A voxel the mesh touches
// becomes solid (bc_mask = 255); a voxel that is not solid but has a solid neighbour in direction
// c_l receives the boundary id and missing[opp l] (mesh_boundary_masker.py:168-185).  The
// reference evaluates, for every voxel, a BVH query plus a triangle test for the voxel and each of
// its Q-1 neighbours.  Here the solid set is computed once, triangle-parallel: one wave per
// triangle tests the voxels of the triangle's bounding box (plus one voxel of margin) and marks a
// solid grid that carries one halo voxel per side, so neighbours outside the domain are covered.
// A voxel-parallel pass then sets the masks from that grid.
//
// Triangle / box overlap: Schwarz & Seidel, "Fast parallel surface and solid voxelization on
// GPUs" (SIGGRAPH Asia 2010), sec. 3.1, the test the reference cites: bounding-box overlap (the
// reference's BVH query, mesh_boundary_masker.py:139), plane overlap through the two critical box
// corners, then the three axis-aligned projections, each with the three edge functions.
// Evaluated in FP32 without contraction, in a fixed operation order.
#include "common.hpp"

namespace xlbh {

struct TriPre {
  float lo[3], hi[3];  // bounding box
  float n[3];
  float d1, d2;
  float ne0[3][3], ne1[3][3], de[3][3];  // [projection p][edge i]
  bool valid;
};

__device__ __forceinline__ float dot3(const float* a, const float* b) {
#pragma clang fp contract(off)
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

__device__ __forceinline__ void tri_precompute(const float* v /* 9: v0, v1, v2 */, TriPre& t) {
#pragma clang fp contract(off)
  const float* v0 = v;
  const float* v1 = v + 3;
  const float* v2 = v + 6;
  float e[3][3];
  for (int a = 0; a < 3; ++a) {
    e[0][a] = v1[a] - v0[a];
    e[1][a] = v2[a] - v1[a];
    e[2][a] = v0[a] - v2[a];
  }
  float w[3];
  for (int a = 0; a < 3; ++a) {
    w[a] = v2[a] - v0[a];
    t.lo[a] = fminf(fminf(v0[a], v1[a]), v2[a]);
    t.hi[a] = fmaxf(fmaxf(v0[a], v1[a]), v2[a]);
  }
  float cr[3];
  cr[0] = e[0][1] * w[2] - e[0][2] * w[1];
  cr[1] = e[0][2] * w[0] - e[0][0] * w[2];
  cr[2] = e[0][0] * w[1] - e[0][1] * w[0];
  const float len = sqrtf(dot3(cr, cr));
  t.valid = len > 0.0f;
  for (int a = 0; a < 3; ++a) t.n[a] = t.valid ? cr[a] / len : 0.0f;
  float c[3], cm[3];
  for (int a = 0; a < 3; ++a) {
    c[a] = t.n[a] > 0.0f ? 1.0f : 0.0f;
    cm[a] = (1.0f - c[a]) - v0[a];
    c[a] = c[a] - v0[a];
  }
  t.d1 = dot3(t.n, c);
  t.d2 = dot3(t.n, cm);
  const float* vv[3] = {v0, v1, v2};
  for (int p = 0; p < 3; ++p) {
    const int a0 = p, a1 = (p + 1) % 3, a2 = (p + 2) % 3;
    const float s = t.n[a2] < 0.0f ? -1.0f : 1.0f;
    for (int i = 0; i < 3; ++i) {
      const float n0 = -s * e[i][a1];
      const float n1 = s * e[i][a0];
      t.ne0[p][i] = n0;
      t.ne1[p][i] = n1;
      t.de[p][i] = (-(n0 * vv[i][a0] + n1 * vv[i][a1]) + fmaxf(0.0f, n0)) + fmaxf(0.0f, n1);
    }
  }
}

// Does the triangle touch the unit box with minimum corner p?
__device__ __forceinline__ bool tri_box(const TriPre& t, const float* p) {
#pragma clang fp contract(off)
  if (!t.valid) return false;
  // bounding boxes must overlap (closed); the plane and projection tests below assume it
  for (int a = 0; a < 3; ++a)
    if (t.lo[a] > p[a] + 1.0f || t.hi[a] < p[a]) return false;
  const float np = dot3(t.n, p);
  if ((np + t.d1) * (np + t.d2) > 0.0f) return false;
  for (int q = 0; q < 3; ++q) {
    const int a0 = q, a1 = (q + 1) % 3;
    for (int i = 0; i < 3; ++i)
      if ((t.ne0[q][i] * p[a0] + t.ne1[q][i] * p[a1]) + t.de[q][i] < 0.0f) return false;
  }
  return true;
}

struct MeshGrid {
  int nx, ny, nz;
  __device__ __forceinline__ uint64_t halo(int x, int y, int z) const {
    return ((uint64_t)(x + 1) * (uint64_t)(ny + 2) + (uint64_t)(y + 1)) * (uint64_t)(nz + 2) + (uint64_t)(z + 1);
  }
};

__global__ __launch_bounds__(256) void k_mesh_zero(uint8_t* p, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * 256u;
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// One wave per triangle: every voxel of its bounding box (+1 voxel margin, clipped to the haloed
// grid [-1, n]) is tested; voxels the triangle touches are marked solid.
__global__ __launch_bounds__(256) void k_mesh_surface(const float* __restrict__ tri, int64_t n_tri, MeshGrid g,
                                                      uint8_t* __restrict__ solid) {
  const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (t >= n_tri) return;
  float v[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) v[k] = tri[t * 9 + k];
  TriPre pre;
  tri_precompute(v, pre);
  if (!pre.valid) return;
  const int n[3] = {g.nx, g.ny, g.nz};
  int lo[3], ext[3];
  for (int a = 0; a < 3; ++a) {
    const float mn = fminf(fminf(v[a], v[3 + a]), v[6 + a]);
    const float mx = fmaxf(fmaxf(v[a], v[3 + a]), v[6 + a]);
    // clamp in float before converting so that out-of-range coordinates cannot overflow int
    const float flo = fminf(fmaxf(floorf(mn) - 1.0f, -1.0f), (float)n[a]);
    const float fhi = fminf(fmaxf(floorf(mx) + 1.0f, -1.0f), (float)n[a]);
    lo[a] = (int)flo;
    ext[a] = (int)fhi - lo[a] + 1;
  }
  const int64_t count = (int64_t)ext[0] * ext[1] * ext[2];
  for (int64_t i = lane; i < count; i += 64) {
    const int z = lo[2] + (int)(i % ext[2]);
    const int y = lo[1] + (int)((i / ext[2]) % ext[1]);
    const int x = lo[0] + (int)(i / ((int64_t)ext[2] * ext[1]));
    const float p[3] = {(float)x, (float)y, (float)z};
    if (tri_box(pre, p)) solid[g.halo(x, y, z)] = 1;
  }
}

// Voxel-parallel marking from the solid grid (mesh_boundary_masker.py:168-185).
template <int LAT>
__global__ __launch_bounds__(256) void k_mesh_mark(const uint8_t* __restrict__ solid, MeshGrid g, Sweep sw,
                                                   int bc_id, uint8_t* __restrict__ bc_mask,
                                                   uint8_t* __restrict__ missing) {
  constexpr LatTables T = make_tables(LAT);
  int x, y, z;
  if (!sweep_cell(sw, x, y, z)) return;
  const uint64_t N = (uint64_t)g.nx * g.ny * g.nz;
  const uint64_t cell = ((uint64_t)x * g.ny + y) * g.nz + z;
  if (solid[g.halo(x, y, z)]) {
    bc_mask[cell] = 255;
    return;
  }
  bool touched = false;
#pragma unroll
  for (int l = 1; l < T.q; ++l) {
    if (solid[g.halo(x + T.c[0][l], y + T.c[1][l], z + T.c[2][l])]) {
      missing[(uint64_t)T.opp[l] * N + cell] = 1;
      touched = true;
    }
  }
  if (touched) bc_mask[cell] = (uint8_t)bc_id;
}

}  // namespace xlbh

using namespace xlbh;

extern "C" int64_t xlbh_mesh_workspace_bytes(int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return -1;
  return (((int64_t)nx + 2) * ((int64_t)ny + 2) * ((int64_t)nz + 2) + 255) / 256 * 256;
}

extern "C" int xlbh_mesh_mask(int lattice, const float* triangles, int64_t n_triangles, int bc_id, uint8_t* bc_mask,
                              uint8_t* missing_bool, void* workspace, int nx, int ny, int nz, void* stream) {
  clear_error();
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(lattice != XLBH_D2Q9, "the mesh masker is 3-D only (mesh_boundary_masker.py:27-28)");
  XLBH_REQUIRE(n_triangles >= 0, "negative triangle count");
  XLBH_REQUIRE(n_triangles == 0 || triangles, "triangles pointer is null");
  XLBH_REQUIRE(bc_mask && missing_bool && workspace, "mesh_mask needs bc_mask, missing_bool and workspace");
  XLBH_REQUIRE(bc_id >= 1 && bc_id <= 254, "boundary id must be 1..254");
  hipStream_t s = (hipStream_t)stream;
  uint8_t* solid = static_cast<uint8_t*>(workspace);
  const MeshGrid g{nx, ny, nz};
  const uint64_t halo_cells = ((uint64_t)nx + 2) * ((uint64_t)ny + 2) * ((uint64_t)nz + 2);
  hipLaunchKernelGGL(k_mesh_zero, dim3((unsigned)std::min<uint64_t>((halo_cells + 255) / 256, 65536)), dim3(256), 0,
                     s, solid, halo_cells);
  XLBH_LAUNCH_CHECK();
  if (n_triangles > 0) {
    const int64_t blocks = (n_triangles + 3) / 4;
    XLBH_REQUIRE(blocks < ((int64_t)1 << 31), "too many triangles");
    hipLaunchKernelGGL(k_mesh_surface, dim3((unsigned)blocks), dim3(256), 0, s, triangles, n_triangles, g, solid);
    XLBH_LAUNCH_CHECK();
  }
  const Sweep sw = make_sweep(nx, ny, nz);
  switch (lattice) {
    case XLBH_D3Q19:
      hipLaunchKernelGGL((k_mesh_mark<XLBH_D3Q19>), sweep_grid(sw), dim3(256), 0, s, solid, g, sw, bc_id, bc_mask,
                         missing_bool);
      break;
    case XLBH_D3Q27:
      hipLaunchKernelGGL((k_mesh_mark<XLBH_D3Q27>), sweep_grid(sw), dim3(256), 0, s, solid, g, sw, bc_id, bc_mask,
                         missing_bool);
      break;
  }
  XLBH_LAUNCH_CHECK();
  return 0;
}
