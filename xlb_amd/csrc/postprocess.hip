// Post-processing stencils on the velocity field: vorticity and Q-criterion, replacing the
// reference's Vorticity (xlb/operator/postprocess/vorticity.py:30-80) and QCriterion
// (q_criterion.py:35-132) kernels.
//
// Both run over the interior cells 1 .. n-2 of each axis, skip a cell when any of its six face
// neighbours has a non-zero boundary id (the cell's own id is not examined), and leave skipped
// and edge cells untouched.  Derivatives are central differences (a - b) / 2; the arithmetic
// follows the reference's expressions term by term in the field's precision.
#include "common.hpp"

namespace xlbh {

template <class T>
struct Grad {
  T xx, xy, xz, yx, yy, yz, zx, zy, zz;  // d u_a / d b as u_a_db
};

// Returns false when the cell is on the domain edge or, with MASK, next to a boundary cell.  x and
// y are wave-uniform (uniform_slow_axes), so every neighbour's address is a scalar row base plus
// the lane's z offset (Planes::at(l, uni, lane)); O = uint32_t while the three velocity planes stay
// below 4 GiB.  At an interior cell every neighbour is in the grid, so the velocity loads are
// issued together with the six mask loads instead of after their test: the mask values pass
// through an empty asm placed after the velocity loads, which is where the test waits for them.
// g is left unused where the test fails.
template <class T, class O, bool MASK>
__device__ __forceinline__ bool central_gradient(const T* u, const uint8_t* bc_mask, int x, int y, int z, int nx,
                                                 int ny, int nz, uint64_t N, Grad<T>& g) {
  if (x < 1 || y < 1 || z < 1 || x > nx - 2 || y > ny - 2 || z > nz - 2) return false;
  const O row = ((O)x * ny + y) * nz, dx = (O)ny * nz, dy = (O)nz;
  const O zc = (O)z;
  uint32_t m[6] = {0, 0, 0, 0, 0, 0};
  if constexpr (MASK) {
    const Planes<uint8_t, O> mk{reinterpret_cast<const char*>(bc_mask), 0};
    m[0] = *mk.at(0, row + dx, zc);
    m[1] = *mk.at(0, row, zc + 1);
    m[2] = *mk.at(0, row + dy, zc);
    m[3] = *mk.at(0, row - dx, zc);
    m[4] = *mk.at(0, row - dy, zc);
    m[5] = *mk.at(0, row, zc - 1);
  }
  const Planes<T, O> v{reinterpret_cast<const char*>(u), (size_t)N * sizeof(T)};
  auto d = [&](int comp, O plus_u, O plus_l, O minus_u, O minus_l) {
    return (*v.at(comp, plus_u, plus_l) - *v.at(comp, minus_u, minus_l)) / T(2);
  };
  g.xx = d(0, row + dx, zc, row - dx, zc);
  g.xy = d(0, row + dy, zc, row - dy, zc);
  g.xz = d(0, row, zc + 1, row, zc - 1);
  g.yx = d(1, row + dx, zc, row - dx, zc);
  g.yy = d(1, row + dy, zc, row - dy, zc);
  g.yz = d(1, row, zc + 1, row, zc - 1);
  g.zx = d(2, row + dx, zc, row - dx, zc);
  g.zy = d(2, row + dy, zc, row - dy, zc);
  g.zz = d(2, row, zc + 1, row, zc - 1);
  if constexpr (MASK) {
    asm volatile("" : "+v"(m[0]), "+v"(m[1]), "+v"(m[2]), "+v"(m[3]), "+v"(m[4]), "+v"(m[5])::"memory");
    return (m[0] | m[1] | m[2] | m[3] | m[4] | m[5]) == 0;
  }
  return true;
}

template <class T, class O, bool MASK>
__global__ __launch_bounds__(256) void k_vorticity(const T* __restrict__ u, const uint8_t* __restrict__ bc_mask,
                                                   T* __restrict__ vort, T* __restrict__ mag, Sweep sw) {
  int x, y, z;
  if (!sweep_cell(sw, x, y, z)) return;
  uniform_slow_axes(true, x, y);
  const uint64_t N = (uint64_t)sw.nx * sw.ny * sw.nz;
  Grad<T> g;
  if (!central_gradient<T, O, MASK>(u, bc_mask, x, y, z, sw.nx, sw.ny, sw.nz, N, g)) return;
  const O row = ((O)x * sw.ny + y) * sw.nz;
  const T vx = g.zy - g.yz;
  const T vy = g.xz - g.zx;
  const T vz = g.yx - g.xy;
  const Planes<T, O> vo{reinterpret_cast<const char*>(vort), (size_t)N * sizeof(T)};
  const Planes<T, O> mo{reinterpret_cast<const char*>(mag), 0};
  *vo.at_w(0, row, (O)z) = vx;
  *vo.at_w(1, row, (O)z) = vy;
  *vo.at_w(2, row, (O)z) = vz;
  *mo.at_w(0, row, (O)z) = sqrt(vx * vx + vy * vy + vz * vz);
}

template <class T, class O, bool MASK>
__global__ __launch_bounds__(256) void k_q_criterion(const T* __restrict__ u, const uint8_t* __restrict__ bc_mask,
                                                     T* __restrict__ norm_mu, T* __restrict__ q, Sweep sw) {
  int x, y, z;
  if (!sweep_cell(sw, x, y, z)) return;
  uniform_slow_axes(true, x, y);
  const uint64_t N = (uint64_t)sw.nx * sw.ny * sw.nz;
  Grad<T> g;
  if (!central_gradient<T, O, MASK>(u, bc_mask, x, y, z, sw.nx, sw.ny, sw.nz, N, g)) return;
  const O row = ((O)x * sw.ny + y) * sw.nz;
  const T mx = g.zy - g.yz, my = g.xz - g.zx, mz = g.yx - g.xy;
  const T mu = sqrt(mx * mx + my * my + mz * mz);
  const T half = T(0.5);
  const T s00 = g.xx, s01 = half * (g.xy + g.yx), s02 = half * (g.xz + g.zx);
  const T s11 = g.yy, s12 = half * (g.yz + g.zy), s22 = g.zz;
  const T ss = s00 * s00 + s01 * s01 + s02 * s02 + s01 * s01 + s11 * s11 + s12 * s12 + s02 * s02 + s12 * s12 + s22 * s22;
  const T o01 = half * (g.xy - g.yx), o02 = half * (g.xz - g.zx), o12 = half * (g.yz - g.zy);
  const T zero = T(0);
  const T oo = zero * zero + o01 * o01 + o02 * o02 + (-o01) * (-o01) + zero * zero + o12 * o12 + (-o02) * (-o02) +
               (-o12) * (-o12) + zero * zero;
  const Planes<T, O> no{reinterpret_cast<const char*>(norm_mu), 0}, qo{reinterpret_cast<const char*>(q), 0};
  *no.at_w(0, row, (O)z) = mu;
  *qo.at_w(0, row, (O)z) = half * (oo - ss);
}

// Calls fn(T{}, O{}, integral_constant<bool, MASK>{}) for the field's element type, the offset
// type its three velocity planes need and whether a boundary mask is given; false for an unknown
// field type.
template <class Fn>
bool with_stencil_types(int field, uint64_t N, bool mask, Fn&& fn) {
  auto pick = [&](auto t) {
    using T = decltype(t);
    const bool narrow = 3 * N * sizeof(T) <= 0xffffffffull;
    if (narrow && mask) fn(t, uint32_t{}, std::true_type{});
    else if (narrow) fn(t, uint32_t{}, std::false_type{});
    else if (mask) fn(t, uint64_t{}, std::true_type{});
    else fn(t, uint64_t{}, std::false_type{});
    return true;
  };
  if (field == XLBH_FIELD_F32) return pick(float{});
  if (field == XLBH_FIELD_F64) return pick(double{});
  return false;
}

}  // namespace xlbh

using namespace xlbh;

extern "C" {

int xlbh_vorticity(int field, const void* u, const uint8_t* bc_mask, void* vorticity, void* magnitude, int nx, int ny,
                   int nz, void* stream) {
  clear_error();
  XLBH_REQUIRE(nx > 2 && ny > 2 && nz > 2, "vorticity needs a 3-D grid of at least 3 cells per axis");
  // the cell-count and launch-grid bounds of the 3-D sweep
  if (int e = check_grid(XLBH_D3Q19, nx, ny, nz)) return e;
  XLBH_REQUIRE(u && vorticity && magnitude, "vorticity needs u, vorticity and magnitude");
  const Sweep sw = make_sweep(nx, ny, nz);
  hipStream_t s = (hipStream_t)stream;
  const uint64_t N = (uint64_t)nx * ny * nz;
  const bool launched = with_stencil_types(field, N, bc_mask != nullptr, [&](auto t, auto o, auto m) {
    using T = decltype(t);
    hipLaunchKernelGGL((k_vorticity<T, decltype(o), decltype(m)::value>), sweep_grid(sw), dim3(256), 0, s,
                       (const T*)u, bc_mask, (T*)vorticity, (T*)magnitude, sw);
  });
  if (!launched) {
    set_error("vorticity supports float32 and float64 fields");
    return 1;
  }
  XLBH_LAUNCH_CHECK();
  return 0;
}

int xlbh_q_criterion(int field, const void* u, const uint8_t* bc_mask, void* norm_mu, void* q, int nx, int ny, int nz,
                     void* stream) {
  clear_error();
  XLBH_REQUIRE(nx > 2 && ny > 2 && nz > 2, "q_criterion needs a 3-D grid of at least 3 cells per axis");
  // the cell-count and launch-grid bounds of the 3-D sweep
  if (int e = check_grid(XLBH_D3Q19, nx, ny, nz)) return e;
  XLBH_REQUIRE(u && norm_mu && q, "q_criterion needs u, norm_mu and q");
  const Sweep sw = make_sweep(nx, ny, nz);
  hipStream_t s = (hipStream_t)stream;
  const uint64_t N = (uint64_t)nx * ny * nz;
  const bool launched = with_stencil_types(field, N, bc_mask != nullptr, [&](auto t, auto o, auto m) {
    using T = decltype(t);
    hipLaunchKernelGGL((k_q_criterion<T, decltype(o), decltype(m)::value>), sweep_grid(sw), dim3(256), 0, s,
                       (const T*)u, bc_mask, (T*)norm_mu, (T*)q, sw);
  });
  if (!launched) {
    set_error("q_criterion supports float32 and float64 fields");
    return 1;
  }
  XLBH_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
