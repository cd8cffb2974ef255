// Shared definitions of the gfx950 LBM backend: lattice tables, population storage formats,
// launch geometry and error reporting.  Everything here is header-only and used by every
// translation unit of libxlbhip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/xlb_hip.h"

namespace xlbh {

// ------------------------------------------------------------------------------------------
// Error reporting: every C entry point returns 0 or a non-zero code and leaves a message in a
// thread-local string read by xlbh_last_error().
// ------------------------------------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

#define XLBH_REQUIRE(cond, msg)        \
  do {                                 \
    if (!(cond)) {                     \
      ::xlbh::set_error(msg);          \
      return 1;                        \
    }                                  \
  } while (0)

#define XLBH_HIP_CHECK(expr)                                                              \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      ::xlbh::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));               \
      return 2;                                                                           \
    }                                                                                     \
  } while (0)

#define XLBH_LAUNCH_CHECK() XLBH_HIP_CHECK(hipGetLastError())

// ------------------------------------------------------------------------------------------
// Lattice tables, built at compile time from the construction rules of
// xlb/velocity_set/d2q9.py:16-24 (explicit list), d3q19.py:17-30 and d3q27.py:17-32
// (itertools.product([0, -1, 1], repeat=3), D3Q19 keeping |c|_1 <= 2), with the opposite
// index, the symmetric second-moment basis cc (order xx, xy, xz, yy, yz, zz / xx, xy, yy) and
// qi = cc - cs^2 delta with doubled off-diagonals (velocity_set.py:124-164).
// ------------------------------------------------------------------------------------------
struct LatTables {
  int d = 0, q = 0, nt = 0;
  int c[3][27] = {};
  double w[27] = {};
  int opp[27] = {};
  int cc[27][6] = {};
  double qi[27][6] = {};
  bool main[27] = {};  // |c|_1 == 1
};

constexpr LatTables make_tables(int lat) {
  LatTables t{};
  if (lat == XLBH_D2Q9) {
    const int cx[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
    const int cy[9] = {0, 1, -1, 0, 1, -1, 0, 1, -1};
    const double w[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 36,
                         1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 36};
    t.d = 2;
    t.q = 9;
    for (int i = 0; i < 9; ++i) {
      t.c[0][i] = cx[i];
      t.c[1][i] = cy[i];
      t.c[2][i] = 0;
      t.w[i] = w[i];
    }
  } else {
    const int v[3] = {0, -1, 1};
    int n = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int e = 0; e < 3; ++e) {
          const int s = (v[a] != 0) + (v[b] != 0) + (v[e] != 0);
          if (lat == XLBH_D3Q19 && s > 2) continue;
          t.c[0][n] = v[a];
          t.c[1][n] = v[b];
          t.c[2][n] = v[e];
          if (lat == XLBH_D3Q19)
            t.w[n] = s == 0 ? 1.0 / 3.0 : (s == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
          else
            t.w[n] = s == 0 ? 8.0 / 27.0 : (s == 1 ? 2.0 / 27.0 : (s == 2 ? 1.0 / 54.0 : 1.0 / 216.0));
          ++n;
        }
    t.d = 3;
    t.q = n;
  }
  for (int i = 0; i < t.q; ++i) {
    for (int j = 0; j < t.q; ++j)
      if (t.c[0][j] == -t.c[0][i] && t.c[1][j] == -t.c[1][i] && t.c[2][j] == -t.c[2][i]) t.opp[i] = j;
    const int s = (t.c[0][i] != 0) + (t.c[1][i] != 0) + (t.c[2][i] != 0);
    t.main[i] = (s == 1);
  }
  t.nt = t.d * (t.d + 1) / 2;
  for (int i = 0; i < t.q; ++i) {
    int k = 0;
    for (int a = 0; a < t.d; ++a)
      for (int b = a; b < t.d; ++b) {
        t.cc[i][k] = t.c[a][i] * t.c[b][i];
        const bool diag = (a == b);
        t.qi[i][k] = diag ? (double)t.cc[i][k] - 1.0 / 3.0 : 2.0 * (double)t.cc[i][k];
        ++k;
      }
  }
  return t;
}

template <int LAT>
struct Lat {
  static constexpr LatTables T = make_tables(LAT);
  static constexpr int Q = T.q;
  static constexpr int D = T.d;
  static constexpr int NT = T.nt;
};

// ------------------------------------------------------------------------------------------
// Population storage formats.  Loads convert to the compute type C (adding w_l back for the
// shifted formats), stores round from C.  bf16 is kept as its raw 16 bits.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float bf16_bits_to_float(uint16_t b) {
  return __uint_as_float(((uint32_t)b) << 16);
}


template <int FMT>
struct Fmt;
template <>
struct Fmt<XLBH_F32> {
  using raw = float;
  static constexpr bool shifted = false;
};
template <>
struct Fmt<XLBH_F64> {
  using raw = double;
  static constexpr bool shifted = false;
};
template <>
struct Fmt<XLBH_F16> {
  using raw = _Float16;
  static constexpr bool shifted = false;
};
template <>
struct Fmt<XLBH_BF16> {
  using raw = uint16_t;
  static constexpr bool shifted = false;
};
template <>
struct Fmt<XLBH_F16_SHIFTED> {
  using raw = _Float16;
  static constexpr bool shifted = true;
};
template <>
struct Fmt<XLBH_BF16_SHIFTED> {
  using raw = uint16_t;
  static constexpr bool shifted = true;
};
// 16-bit block floating point: int16 mantissas m of x = f_l / w_l - 1 with one exponent k per row
// segment (blk_* below), f_l = w_l + m * w_l * 2^-k.
template <>
struct Fmt<XLBH_I16_BLOCK> {
  using raw = int16_t;
  static constexpr bool shifted = false;
};

template <class R, class C>
__device__ __forceinline__ C raw_to(R v) {
  if constexpr (sizeof(R) == 2 && !__is_same(R, _Float16))
    return (C)bf16_bits_to_float(v);
  else if constexpr (__is_same(R, _Float16))
    return (C)(float)v;
  else
    return (C)v;
}

// float -> bf16 uses gfx950's conversion instruction (v_cvt_pk_bf16_f32): round to nearest even,
// the same bits as the integer form (u + 0x7fff + ((u >> 16) & 1)) >> 16 for every non-NaN float
// (compared on the GPU over 2^30 bit patterns and every tie).  One instruction instead of five; the
// shifted-BF16 step ran 7 % faster with it.
template <class R, class C>
__device__ __forceinline__ R to_raw(C v) {
  if constexpr (sizeof(R) == 2 && !__is_same(R, _Float16))
    return __builtin_bit_cast(uint16_t, (__bf16)(float)v);
  else
    return (R)v;
}

// Cache policy of the population sweeps.  Every population is read once and written once per
// step and the fields are far larger than L2 + MALL, so the streaming (non-temporal) forms keep
// the sweep from evicting the lines that neighbouring lanes are still reusing.
#ifndef XLBH_NT_LOADS
#define XLBH_NT_LOADS 0
#endif
#ifndef XLBH_NT_STORES
#define XLBH_NT_STORES 1
#endif

template <class R>
__device__ __forceinline__ R ld_stream(const R* p) {
#if XLBH_NT_LOADS
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

template <class R>
__device__ __forceinline__ void st_stream(R* p, R v) {
#if XLBH_NT_STORES
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// Load population l (weight w) of format FMT as compute type C.
template <int FMT, class C>
__device__ __forceinline__ C ld_pop(const typename Fmt<FMT>::raw* p, double w) {
  C v = raw_to<typename Fmt<FMT>::raw, C>(ld_stream(p));
  if constexpr (Fmt<FMT>::shifted) v += (C)w;
  return v;
}

template <int FMT, class C>
__device__ __forceinline__ void st_pop(typename Fmt<FMT>::raw* p, C v, double w) {
  if constexpr (Fmt<FMT>::shifted) v -= (C)w;
  st_stream(p, to_raw<typename Fmt<FMT>::raw, C>(v));
}

// Macroscopic-field element types (rho, u, Pi outputs of the standalone operators).
template <int FT>
struct FieldT;
template <>
struct FieldT<XLBH_FIELD_F32> {
  using raw = float;
};
template <>
struct FieldT<XLBH_FIELD_F64> {
  using raw = double;
};
template <>
struct FieldT<XLBH_FIELD_F16> {
  using raw = _Float16;
};
template <>
struct FieldT<XLBH_FIELD_BF16> {
  using raw = uint16_t;
};

// ------------------------------------------------------------------------------------------
// Launch geometry of the cell sweeps.  A 256-thread block covers a (by x bz) tile of the two
// fastest axes; bz (a power of two, 64..256) runs along the unit-stride axis so a block reads
// whole cache lines and keeps the +-1 shifted pulls of neighbouring lanes inside the block.
// 3-D: fast = z, mid = y, slow = x.  2-D (nz == 1): fast = y, mid = x.
// ------------------------------------------------------------------------------------------
struct Sweep {
  int nx, ny, nz;
  int lbz;  // log2(bz)
};

inline Sweep make_sweep(int nx, int ny, int nz) {
  const int fast = (nz > 1) ? nz : ny;
  int lbz = 6;
  while (lbz < 8 && (1 << lbz) < fast) ++lbz;
  return Sweep{nx, ny, nz, lbz};
}

// The sweep's launch grid over planes x0 <= x < x1 of the grid (sweep_grid: all planes).
inline dim3 sweep_grid_range(const Sweep& s, int x0, int x1) {
  const int bz = 1 << s.lbz, by = 256 / bz;
  if (s.nz > 1) return dim3((s.nz + bz - 1) / bz, (s.ny + by - 1) / by, x1 - x0);
  return dim3((s.ny + bz - 1) / bz, (x1 - x0 + by - 1) / by, 1);
}

inline dim3 sweep_grid(const Sweep& s) { return sweep_grid_range(s, 0, s.nx); }

// Thread -> cell coordinates of tile (bx, by, bz); false when the thread lies outside the grid.
__device__ __forceinline__ bool sweep_cell_of(const Sweep& s, int bx, int by, int bzi, int& x, int& y, int& z) {
  const int bz = 1 << s.lbz;
  const int t = threadIdx.x;
  const int fast = bx * bz + (t & (bz - 1));
  const int mid = by * (256 >> s.lbz) + (t >> s.lbz);
  if (s.nz > 1) {
    z = fast;
    y = mid;
    x = bzi;
    return z < s.nz && y < s.ny;
  }
  z = 0;
  y = fast;
  x = mid;
  return y < s.ny && x < s.nx;
}
__device__ __forceinline__ bool sweep_cell(const Sweep& s, int& x, int& y, int& z) {
  return sweep_cell_of(s, blockIdx.x, blockIdx.y, blockIdx.z, x, y, z);
}
// sweep_cell for a launch of sweep_grid_range(s, x0, x1).
__device__ __forceinline__ bool sweep_cell_range(const Sweep& s, int x0, int x1, int& x, int& y, int& z) {
  const bool in = sweep_cell_of(s, blockIdx.x, blockIdx.y, blockIdx.z, x, y, z);
  x += x0;
  return in && x < x1;
}

__device__ __forceinline__ int wrap(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}
// wrap() for v in [-1, n]
__device__ __forceinline__ int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

inline bool valid_lattice(int lat) { return lat == XLBH_D2Q9 || lat == XLBH_D3Q19 || lat == XLBH_D3Q27; }
inline bool valid_store(int s) { return s >= XLBH_F32 && s <= XLBH_I16_BLOCK; }
inline int lattice_q(int lat) { return lat == XLBH_D2Q9 ? 9 : (lat == XLBH_D3Q19 ? 19 : 27); }
inline size_t store_bytes(int s) { return s == XLBH_F64 ? 8 : (s == XLBH_F32 ? 4 : 2); }

// ------------------------------------------------------------------------------------------
// Block floating point (XLBH_I16_BLOCK).  A row segment is the run of bz cells along the fast
// axis that one sweep tile covers: 3-D (x, y, z >> lbz), 2-D (x, y >> lbz).  Each segment holds
// one int32 exponent k, shared by every direction; the exponents of a field follow its last
// plane, at byte offset blk_exps_offset(q, plane).  k is chosen from the largest
// |x| = |f_l / w_l - 1| over the segment's written cells and directions, M: with
// M' = M * (1 + 2^-15) rounded to float and E its unbiased exponent, k = 14 - E clamped to
// [kBlkKMin, kBlkKMax], so every |m| <= 32767.  m = rint_even(t) with t = fma(f, 2^k / w_l, -2^k)
// rounded to float (1 / w_l is exact in float for every lattice weight).
// ------------------------------------------------------------------------------------------
constexpr int kBlkKMin = -64;
constexpr int kBlkKMax = 60;
// Exponent of a segment whose largest |x| is not finite (int16 mantissas cannot hold NaN or
// infinity): its cells are stored with m = 32767, which this exponent decodes to +inf, so a
// diverging run stays visible.
constexpr int kBlkNonFinite = -127;

__host__ __device__ inline uint32_t blk_segs_fast(const Sweep& s) {
  const uint32_t fast = (s.nz > 1) ? (uint32_t)s.nz : (uint32_t)s.ny;
  return (fast + (1u << s.lbz) - 1) >> s.lbz;
}
__host__ __device__ inline uint64_t blk_seg_count(const Sweep& s) {
  const uint64_t rows = (s.nz > 1) ? (uint64_t)s.nx * s.ny : (uint64_t)s.nx;
  return rows * blk_segs_fast(s);
}
// Segment of cell (x, y, z).
__host__ __device__ inline uint32_t blk_seg(const Sweep& s, int x, int y, int z) {
  if (s.nz > 1) return ((uint32_t)x * s.ny + y) * blk_segs_fast(s) + ((uint32_t)z >> s.lbz);
  return (uint32_t)x * blk_segs_fast(s) + ((uint32_t)y >> s.lbz);
}
// Segment fb of row (xr, yr) (3-D) or row xr (2-D).
__host__ __device__ inline uint32_t blk_row_seg(const Sweep& s, int xr, int yr, int fb) {
  if (s.nz > 1) return ((uint32_t)xr * s.ny + yr) * blk_segs_fast(s) + fb;
  return (uint32_t)xr * blk_segs_fast(s) + fb;
}
__host__ __device__ inline size_t blk_exps_offset(int q, size_t plane) { return ((size_t)q * plane * 2 + 255) / 256 * 256; }

__host__ __device__ inline uint32_t f2u(float f) {
  uint32_t b;
  __builtin_memcpy(&b, &f, 4);
  return b;
}
__host__ __device__ inline float u2f(uint32_t b) {
  float f;
  __builtin_memcpy(&f, &b, 4);
  return f;
}

// Exponent from the segment's largest |x|, m >= 0 or NaN.
__host__ __device__ inline int blk_k_from_max(float m) {
  if (!(m <= 3.0e38f)) return kBlkNonFinite;  // NaN or infinite
  const float mp = m * 1.000030517578125f;    // (1 + 2^-15), one rounding
  const int e = (int)((f2u(mp) >> 23) & 0xffu) - 127;
  const int k = 14 - e;
  return k < kBlkKMin ? kBlkKMin : (k > kBlkKMax ? kBlkKMax : k);
}
// 2^k (exact, normal for k in [kBlkKMin, kBlkKMax]).
__host__ __device__ inline float blk_pow2(int k) { return u2f((uint32_t)(k + 127) << 23); }
// Decode scale w * 2^-k (exact): f = fma(m, scale, w).  For kBlkNonFinite the scale exceeds
// 2^119, so m = 32767 decodes to +inf.
__host__ __device__ inline float blk_dscale(float w, int k) { return u2f(f2u(w) - ((uint32_t)k << 23)); }
// Encode scale 2^k / w (exact; 1/w is a float for every lattice weight): t = fma(f, scale, -2^k),
// m = rint_even(t).
__host__ __device__ inline float blk_escale(float inv_w, int k) { return u2f(f2u(inv_w) + ((uint32_t)k << 23)); }
// rint_even(t) as int16 for |t| < 2^15: t + 1.5 * 2^23 lies in [2^23, 2^24), where the float
// spacing is 1, so the addition rounds t to an integer held in the low mantissa bits.
__host__ __device__ inline int16_t blk_round(float t) { return (int16_t)(f2u(t + 12582912.0f) & 0xffffu); }
// The mantissa of x = fma(f, 1/w, -1) at exponent k, from x itself: x * 2^k is exact, so
// rint_even(fma(x, 2^k, 1.5 * 2^23)) equals blk_round(fma(f, blk_escale(1/w, k), -2^k)) bit for bit.
__device__ __forceinline__ int16_t blk_mantissa(float x, float pow2k) {
  return (int16_t)(f2u(__builtin_fmaf(x, pow2k, 12582912.0f)) & 0xffffu);
}

// Largest of v over the wave's 64 lanes (all lanes active), as a uniform value.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xf, false);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max(v, dpp_u32<0xb1, 0xf>(v));   // quad_perm [1,0,3,2]
  v = max(v, dpp_u32<0x4e, 0xf>(v));   // quad_perm [2,3,0,1]
  v = max(v, dpp_u32<0x141, 0xf>(v));  // row_half_mirror
  v = max(v, dpp_u32<0x140, 0xf>(v));  // row_mirror: each 16-lane row holds its maximum
  v = max(v, dpp_u32<0x142, 0xa>(v));  // row_bcast:15 into rows 1 and 3
  v = max(v, dpp_u32<0x143, 0xc>(v));  // row_bcast:31 into rows 2 and 3: lane 63 holds the maximum
  return __builtin_amdgcn_readlane(v, 63);
}

// Largest |f_l / w_l - 1| of one cell (NaN if any is NaN).
template <int LAT>
__device__ __forceinline__ float blk_cell_max(const float* f) {
  constexpr LatTables T = make_tables(LAT);
  float a = 0.0f;
#pragma unroll
  for (int l = 0; l < T.q; ++l)
    a = __builtin_elementwise_maximum(a, __builtin_fabsf(__builtin_fmaf(f[l], (float)(1.0 / T.w[l]), -1.0f)));
  return a;
}

// Exponent of the caller's segment from each lane's largest |x| (0 for lanes that write
// nothing).  Every thread of the block must call it: for segments of 128 or 256 cells it
// synchronises the block.
__device__ __forceinline__ int blk_segment_k(float amax, int lbz) {
  __shared__ uint32_t red[4];
  // |x| >= 0, so the float order is the order of the bit patterns; NaN sorts above infinity.
  uint32_t b = wave_max_u32(f2u(amax));
  if (lbz > 6) {
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wave] = b;
    __syncthreads();
    const int first = wave & ~((1 << (lbz - 6)) - 1);
    b = max(red[first], red[first + 1]);
    if (lbz > 7) b = max(b, max(red[first + 2], red[first + 3]));
  }
  return blk_k_from_max(u2f(b));
}

// Population field addressing.  Plane l of a field starts at base + l * plane_bytes (uniform,
// held in scalar registers); a cell is reached with an unsigned per-lane offset.  With
// O = uint32_t (every plane below 4 GiB) the compiler emits the scalar-base + 32-bit vector-offset
// form of global loads/stores; O = uint64_t serves larger planes.
#ifndef XLBH_ONE_BASE
#define XLBH_ONE_BASE 1
#endif
template <class R, class O>
struct Planes {
  const char* base;
  size_t plane_bytes;
  __device__ __forceinline__ const R* at(int l, O cell) const {
#if XLBH_ONE_BASE
    if constexpr (sizeof(O) == 4)
      return reinterpret_cast<const R*>(base + (uint32_t)((uint32_t)l * (uint32_t)plane_bytes + cell * (O)sizeof(R)));
#endif
    return reinterpret_cast<const R*>(base + (size_t)l * plane_bytes + cell * (O)sizeof(R));
  }
  __device__ __forceinline__ R* at_w(int l, O cell) const { return const_cast<R*>(at(l, cell)); }
  // The same address from a wave-uniform cell offset `uni` and a per-lane one `lane` (their sum is
  // the cell): plane base plus uniform part form one scalar pointer, and the lane part is the
  // instruction's 32-bit vector offset, so the address costs no vector arithmetic.
  __device__ __forceinline__ const R* at(int l, O uni, O lane) const {
    if constexpr (sizeof(O) == 4) {
      using gchar = const __attribute__((address_space(1))) char;
      gchar* b = (gchar*)(base + (uint32_t)((uint32_t)l * (uint32_t)plane_bytes + uni * (O)sizeof(R)));
      // an opaque scalar (global) pointer: keeps the compiler from re-associating the sum into a
      // 64-bit vector add per access (the empty statement emits nothing)
      asm("" : "+s"(b));
      return reinterpret_cast<const R*>((const char*)(b + (uint32_t)(lane * (O)sizeof(R))));
    }
    return at(l, uni + lane);
  }
  // Stores: the lane offset is restated next to each store, so that the store's own block sees a
  // 32-bit offset and selects the scalar-base form too.
  __device__ __forceinline__ R* at_w(int l, O uni, O lane) const {
    if constexpr (sizeof(O) == 4) {
      uint32_t off = (uint32_t)(lane * (O)sizeof(R));
      asm("" : "+v"(off));
      return reinterpret_cast<R*>(const_cast<char*>(reinterpret_cast<const char*>(at(l, uni, (O)0))) + off);
    }
    return const_cast<R*>(at(l, uni, lane));
  }
};

// A wave covers 64 consecutive cells of one row (row segments are at least 64 cells), so the
// slow coordinates are wave-uniform: x and y in 3-D, x in 2-D.  Stating it puts everything
// derived from them in scalar registers.
__device__ __forceinline__ void uniform_slow_axes(bool three, int& x, int& y) {
  x = __builtin_amdgcn_readfirstlane(x);
  if (three) y = __builtin_amdgcn_readfirstlane(y);
}

int check_grid(int lattice, int nx, int ny, int nz);

// The arithmetic type of a compute id: XLBH_COMPUTE_F32_DEVIATION computes in FP32 (only the fused
// step's collision distinguishes it).
inline int arith_compute(int compute) {
  return compute == XLBH_COMPUTE_F32_DEVIATION ? XLBH_COMPUTE_F32 : compute;
}

// Host dispatch over the lattice id: fn(std::integral_constant<int, LAT>{}).
template <class Fn>
int with_lattice(int lat, Fn&& fn) {
  switch (lat) {
    case XLBH_D2Q9: return fn(std::integral_constant<int, XLBH_D2Q9>{});
    case XLBH_D3Q19: return fn(std::integral_constant<int, XLBH_D3Q19>{});
    case XLBH_D3Q27: return fn(std::integral_constant<int, XLBH_D3Q27>{});
  }
  set_error("unknown lattice id");
  return 1;
}
// Plane stride (elements) of a population field: 0 selects the dense stride N; otherwise
// plane >= N.  Returns 0 and sets `out`, or non-zero with the error set.
int resolve_plane(int lattice, int store, int nx, int ny, int nz, int64_t plane, uint64_t& out);

}  // namespace xlbh
