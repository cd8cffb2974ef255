// xlb_oracle.cpp -- CPU restatement of XLB's Warp-backend algorithm for the fused LBM step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity reference of the tests (and the CPU baseline
// timed beside the GPU): only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
// product (xlb_amd, libxlbhip.so) never calls it.
//
// It follows the reference's per-cell functionals line by line, including the reference's own
// mechanism for boundary auxiliary data (prescribed values stashed in f_1[0] and recovered into
// f_0 every step, nse_stepper.py:318-342, boundary_condition.py:130-161), so agreement with the
// HIP path (which keeps those values in a side table) also pins that design change.  Where the
// reference races (aux recovery writes into f_0 while other threads pull from it), the oracle
// takes the "all reads before writes" order.  Citations are paths relative to the reference
// repository root.
//
// Every function is pinned by tests/test_oracle_cpu.py, tests/test_mesh_cpu.py and the golden
// fixtures (known-answer tests of the reference's own test suite, lattice identities,
// conservation, closed-form flows and special cases).  Where a check was written to catch a
// particular transcription slip, the slip can be switched on here (orc_plant, off by default) so
// the test shows that its check fails on it.  One aspect is parity unpinned: bit-level
// agreement of the triangle / box test (orc_tri_box) with the reference's Warp formulation of it,
// whose matrix-from-vectors convention cannot be run here.  The test itself is pinned against an
// independent float64 separating-axis test and a closed-form box voxelisation.
//
// Build: see oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#include <omp.h>

namespace {

// Planted transcription slips, off (0) except while a pinning test shows that its check fails on
// one of them (orc_plant):
//   1  KBC: sign of the (2 - 1/beta) term of gamma flipped
//   2  KBC: two shear-basis assignments exchanged (s[12] <-> s[15] in D3Q27, s[8] <-> s[4] in D2Q9)
//   3  momentum exchange: f_pc[l] taken for f_pc[opp l]
//   4  momentum exchange: the post-streaming term f_ps[l] dropped
//   5  extrapolation outflow: weights cs and 1 - cs exchanged
//   6  extrapolation outflow: source x - c_l instead of x - c_l - n
int g_plant = 0;

// ------------------------------------------------------------------------------------------
// Lattices, written out from the reference's constructors (d2q9.py:16-24; d3q19.py:17-30 =
// itertools.product([0,-1,1], repeat=3) filtered to |c|_1 <= 2; d3q27.py:17-32 = the full
// product).  Opposite indices, cc and qi are derived as in velocity_set.py:124-195.
// ------------------------------------------------------------------------------------------
struct Lattice {
  int d, q, nt;
  int c[3][27];
  double w[27];
  int opp[27];
  double cc[27][6];
  double qi[27][6];
};

const int D3Q19_C[19][3] = {{0, 0, 0},  {0, 0, -1},  {0, 0, 1},  {0, -1, 0}, {0, -1, -1}, {0, -1, 1}, {0, 1, 0},
                            {0, 1, -1}, {0, 1, 1},   {-1, 0, 0}, {-1, 0, -1}, {-1, 0, 1}, {-1, -1, 0},
                            {-1, 1, 0}, {1, 0, 0},   {1, 0, -1}, {1, 0, 1},   {1, -1, 0}, {1, 1, 0}};

Lattice build_lattice(int lat) {
  Lattice L{};
  if (lat == 0) {
    const int cx[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
    const int cy[9] = {0, 1, -1, 0, 1, -1, 0, 1, -1};
    const double w[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 36};
    L.d = 2;
    L.q = 9;
    for (int i = 0; i < 9; ++i) {
      L.c[0][i] = cx[i];
      L.c[1][i] = cy[i];
      L.c[2][i] = 0;
      L.w[i] = w[i];
    }
  } else if (lat == 1) {
    L.d = 3;
    L.q = 19;
    for (int i = 0; i < 19; ++i) {
      for (int a = 0; a < 3; ++a) L.c[a][i] = D3Q19_C[i][a];
      const int s = std::abs(D3Q19_C[i][0]) + std::abs(D3Q19_C[i][1]) + std::abs(D3Q19_C[i][2]);
      L.w[i] = s == 0 ? 1.0 / 3.0 : (s == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
    }
  } else {
    L.d = 3;
    L.q = 27;
    const int v[3] = {0, -1, 1};
    for (int i = 0; i < 27; ++i) {
      L.c[0][i] = v[i / 9];
      L.c[1][i] = v[(i / 3) % 3];
      L.c[2][i] = v[i % 3];
      const int s = std::abs(L.c[0][i]) + std::abs(L.c[1][i]) + std::abs(L.c[2][i]);
      const double ws[4] = {8.0 / 27.0, 2.0 / 27.0, 1.0 / 54.0, 1.0 / 216.0};
      L.w[i] = ws[s];
    }
  }
  for (int i = 0; i < L.q; ++i)
    for (int j = 0; j < L.q; ++j)
      if (L.c[0][j] == -L.c[0][i] && L.c[1][j] == -L.c[1][i] && L.c[2][j] == -L.c[2][i]) L.opp[i] = j;
  L.nt = L.d * (L.d + 1) / 2;
  for (int i = 0; i < L.q; ++i) {
    int k = 0;
    for (int a = 0; a < L.d; ++a)
      for (int b = a; b < L.d; ++b) {
        L.cc[i][k] = (double)(L.c[a][i] * L.c[b][i]);
        L.qi[i][k] = L.cc[i][k];
        ++k;
      }
    // Qi = cc - cs^2 I with doubled off-diagonals (velocity_set.py:124-138)
    if (L.d == 3) {
      const int diag[3] = {0, 3, 5}, off[3] = {1, 2, 4};
      for (int t : diag) L.qi[i][t] += -1.0 / 3.0;
      for (int t : off) L.qi[i][t] *= 2.0;
    } else {
      const int diag[2] = {0, 2};
      for (int t : diag) L.qi[i][t] += -1.0 / 3.0;
      L.qi[i][1] *= 2.0;
    }
  }
  return L;
}

// ------------------------------------------------------------------------------------------
// Storage formats: 0 f32, 1 f64, 2 f16, 3 bf16, 4 f16 shifted (f - w), 5 bf16 shifted.
// 16-bit conversions are written out (round to nearest even).
// ------------------------------------------------------------------------------------------
inline uint32_t fbits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline float bitsf(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

uint16_t f32_to_f16(float f) {
  uint32_t x = fbits(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  x &= 0x7fffffffu;
  if (x >= 0x7f800000u) return (uint16_t)(sign | (x > 0x7f800000u ? 0x7e00u : 0x7c00u));
  const int e = (int)(x >> 23);
  if (e >= 113) {  // normal half range (or overflow)
    uint32_t r = (x >> 13) - (112u << 10);
    const uint32_t rem = x & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
    if (r >= 0x7c00u) r = 0x7c00u;
    return (uint16_t)(sign | r);
  }
  const int s = 126 - e;  // subnormal: value = m * 2^(e-150), half ulp = 2^-24
  if (s >= 25) return (uint16_t)sign;
  const uint32_t m = (x & 0x7fffffu) | 0x800000u;
  uint32_t r = m >> s;
  const uint32_t rem = m & ((1u << s) - 1u), half = 1u << (s - 1);
  if (rem > half || (rem == half && (r & 1u))) ++r;
  return (uint16_t)(sign | r);
}

float f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  if (e == 0) {
    const float v = (float)m * 5.9604644775390625e-08f;  // m * 2^-24
    return sign ? -v : v;
  }
  if (e == 31) return bitsf(sign | 0x7f800000u | (m << 13));
  return bitsf(sign | ((e + 112u) << 23) | (m << 13));
}

uint16_t f32_to_bf16(float f) {
  uint32_t u = fbits(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf16_to_f32(uint16_t b) { return bitsf((uint32_t)b << 16); }

template <int FMT>
struct Store;
template <>
struct Store<0> {
  using T = float;
  static double get(T v) { return v; }
  static T put(double v) { return (float)v; }
};
template <>
struct Store<1> {
  using T = double;
  static double get(T v) { return v; }
  static T put(double v) { return v; }
};
template <>
struct Store<2> {
  using T = uint16_t;
  static double get(T v) { return f16_to_f32(v); }
  static T put(double v) { return f32_to_f16((float)v); }
};
template <>
struct Store<3> {
  using T = uint16_t;
  static double get(T v) { return bf16_to_f32(v); }
  static T put(double v) { return f32_to_bf16((float)v); }
};
template <>
struct Store<4> : Store<2> {};
template <>
struct Store<5> : Store<3> {};
// Format 6, 16-bit block floating point: int16 mantissas whose values need the segment exponent
// (Ctx::rd / the step's segment quantisation below); get/put are never used for it.
template <>
struct Store<6> {
  using T = int16_t;
  static double get(T) { return 0.0; }
  static T put(double) { return 0; }
};

constexpr bool shifted(int fmt) { return fmt == 4 || fmt == 5; }
constexpr bool block(int fmt) { return fmt == 6; }

// ------------------------------------------------------------------------------------------
// Block floating point (format 6; no counterpart in the reference).  Element l of a cell holds
// m = rint_even(t), t = float(2^k (f_l / w_l - 1)) formed as one fused multiply-add, and every row
// segment (the bz cells along the fast axis that one 256-thread tile of the GPU sweep covers)
// shares one exponent k, chosen from the largest |x| = |fma(f_l, 1/w_l, -1)| the step writes
// there: M' = float(M * (1 + 2^-15)), k = 14 - (unbiased exponent of M'), clamped to [-64, 60];
// a segment whose M is not finite gets k = -127 and m = 32767 everywhere (decodes to +inf).
// f = fma(m, w 2^-k, w).  w is the float-rounded lattice weight; 1/w is exact in float.  The
// exponents follow the Q planes, at byte offset ceil(Q N 2 / 256) 256.
// ------------------------------------------------------------------------------------------
constexpr int kBlkKMin = -64, kBlkKMax = 60, kBlkNonFinite = -127;

struct BlkGeom {
  int nx = 1, ny = 1, nz = 1, lbz = 6, S = 1;
  size_t seg(int x, int y, int z) const {
    if (nz > 1) return ((size_t)x * ny + y) * S + (size_t)(z >> lbz);
    return (size_t)x * S + (size_t)(y >> lbz);
  }
  size_t count() const { return (nz > 1 ? (size_t)nx * ny : (size_t)nx) * S; }
};

BlkGeom blk_geom(int nx, int ny, int nz) {
  BlkGeom g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  const int fast = nz > 1 ? nz : ny;
  g.lbz = 6;
  while (g.lbz < 8 && (1 << g.lbz) < fast) ++g.lbz;
  g.S = (fast + (1 << g.lbz) - 1) >> g.lbz;
  return g;
}

size_t blk_offset(int q, size_t N) { return ((size_t)q * N * 2 + 255) / 256 * 256; }

int blk_k_from_max(float m) {
  if (!(m <= 3.0e38f)) return kBlkNonFinite;
  const float mp = m * 1.000030517578125f;
  const int e = (int)((fbits(mp) >> 23) & 0xffu) - 127;
  const int k = 14 - e;
  return k < kBlkKMin ? kBlkKMin : (k > kBlkKMax ? kBlkKMax : k);
}

float blk_decode(int16_t m, float w, int k) {
  return std::fmaf((float)m, bitsf(fbits(w) - ((uint32_t)k << 23)), w);
}

int16_t blk_encode(float v, float inv_w, int k) {
  if (k == kBlkNonFinite) return 32767;
  const float t = std::fmaf(v, bitsf(fbits(inv_w) + ((uint32_t)k << 23)), -bitsf((uint32_t)(k + 127) << 23));
  const float y = t + 12582912.0f;  // 1.5 * 2^23: the float sum rounds t to an integer (ties to even)
  return (int16_t)(fbits(y) & 0xffffu);
}

// Quantise the float values v (q, N) of the written cells into the block field f (mantissas, then
// the exponents at blk_offset), segment by segment.
void blk_store(const double* w, int q, const BlkGeom& g, const float* v, const uint8_t* written, int16_t* f) {
  const size_t N = (size_t)g.nx * g.ny * g.nz;
  int32_t* e = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(f) + blk_offset(q, N));
  const int bz = 1 << g.lbz, fast = g.nz > 1 ? g.nz : g.ny;
  const int rows = g.nz > 1 ? g.nx * g.ny : g.nx;
#pragma omp parallel for schedule(static)
  for (int r = 0; r < rows; ++r)
    for (int sb = 0; sb < g.S; ++sb) {
      const int f0 = sb * bz, f1 = std::min(fast, f0 + bz);
      // cells of the segment: consecutive along the fast axis (z in 3-D, y in 2-D: also unit stride)
      const size_t base = (size_t)r * fast;
      float M = 0.0f;
      bool nan = false;
      for (int i = f0; i < f1; ++i) {
        if (!written[base + i]) continue;
        for (int l = 0; l < q; ++l) {
          const float x = std::fabs(std::fmaf(v[(size_t)l * N + base + i], (float)(1.0 / w[l]), -1.0f));
          if (x != x) nan = true;
          else if (x > M) M = x;
        }
      }
      const int k = nan ? kBlkNonFinite : blk_k_from_max(M);
      e[(size_t)r * g.S + sb] = k;
      for (int i = f0; i < f1; ++i) {
        if (!written[base + i]) continue;
        for (int l = 0; l < q; ++l)
          f[(size_t)l * N + base + i] = blk_encode(v[(size_t)l * N + base + i], (float)(1.0 / w[l]), k);
      }
    }
}

// ------------------------------------------------------------------------------------------
// Boundary descriptions passed from Python (one per BC instance, in the stepper's list order).
// step: 0 = STREAMING, 1 = COLLISION (boundary_condition.py:22-25).
// ------------------------------------------------------------------------------------------
struct OrcBc {
  int32_t id;
  int32_t kind;  // 1 equilibrium, 2 halfway, 3 fullway, 4 do-nothing, 5/6 zou-he v/p, 7/8 regularized v/p,
                 // 9 extrapolation, 10 Grad's approximation
  double rho;
  double u[3];
};

inline int bc_step(int kind) { return kind == 3 ? 1 : 0; }
inline bool needs_aux_recovery(int kind) { return kind >= 5 && kind <= 8; }

template <class C>
using Vec = C[27];

template <int FMT, class C>
struct Ctx {
  const Lattice* L;
  using S = typename Store<FMT>::T;
  S* f0;
  S* f1;
  const uint8_t* bc_mask;
  const uint8_t* missing;  // (q, N) bool
  const OrcBc* bcs;
  int nbc;
  int nx, ny, nz;
  size_t N;
  C omega;
  C cparam = 0;  // Smagorinsky coefficient
  BlkGeom g;     // block format: segment geometry
  int force_mode;
  C force_u[3];
  const double* force_field;  // (d, N)
  const double* aux_side = nullptr;  // block format: prescribed values per cell (no f_1[0] stash)

  size_t cell(int x, int y, int z) const { return ((size_t)x * ny + y) * nz + z; }
  // compute-type read of a population (shifted formats add w_l back; the block format decodes
  // with the cell's segment exponent, stored after the field's planes)
  C rd(const S* f, int l, size_t c) const {
    if constexpr (block(FMT)) {
      const int32_t* e = reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(f) + blk_offset(L->q, N));
      const int z = (int)(c % nz), y = (int)((c / nz) % ny), x = (int)(c / ((size_t)nz * ny));
      return (C)blk_decode(f[(size_t)l * N + c], (float)L->w[l], e[g.seg(x, y, z)]);
    }
    C v = (C)Store<FMT>::get(f[(size_t)l * N + c]);
    if (shifted(FMT)) v += (C)L->w[l];
    return v;
  }
  void wr(S* f, int l, size_t c, C v) const {
    if (shifted(FMT)) v -= (C)L->w[l];
    f[(size_t)l * N + c] = Store<FMT>::put((double)v);
  }
  // raw (unshifted) read/write of a stored value, used for the aux stash in f_1[0]
  C rd_raw(const S* f, int l, size_t c) const { return (C)Store<FMT>::get(f[(size_t)l * N + c]); }
};

// ---- operators --------------------------------------------------------------------------------

// Stream functional (stream.py:59-82): pull with periodic wrap.
template <int FMT, class C>
void stream_functional(const Ctx<FMT, C>& X, const typename Ctx<FMT, C>::S* f, int x, int y, int z, C* out) {
  const Lattice& L = *X.L;
  const int shape[3] = {X.nx, X.ny, X.nz};
  for (int l = 0; l < L.q; ++l) {
    int idx[3] = {x, y, z};
    int pull[3] = {0, 0, 0};
    for (int d = 0; d < L.d; ++d) {
      pull[d] = idx[d] - L.c[d][l];
      if (pull[d] < 0) pull[d] = shape[d] - 1;
      else if (pull[d] >= shape[d]) pull[d] = 0;
    }
    out[l] = X.rd(f, l, X.cell(pull[0], pull[1], pull[2]));
  }
}

// Macroscopic (zero_moment.py:22-27, first_moment.py:25-38).
template <class C>
void macroscopic(const Lattice& L, const C* f, C& rho, C* u) {
  rho = 0;
  for (int l = 0; l < L.q; ++l) rho += f[l];
  u[0] = u[1] = u[2] = 0;
  for (int l = 0; l < L.q; ++l)
    for (int d = 0; d < L.d; ++d) {
      if (L.c[d][l] == 1) u[d] += f[l];
      else if (L.c[d][l] == -1) u[d] -= f[l];
    }
  for (int d = 0; d < L.d; ++d) u[d] /= rho;
}

// Quadratic equilibrium (quadratic_equilibrium.py:35-60).
template <class C>
void equilibrium(const Lattice& L, C rho, const C* u, C* feq) {
  for (int l = 0; l < L.q; ++l) {
    C cu = 0;
    for (int d = 0; d < L.d; ++d) {
      if (L.c[d][l] == 1) cu += u[d];
      else if (L.c[d][l] == -1) cu -= u[d];
    }
    cu *= (C)3.0;
    C uu = 0;
    for (int d = 0; d < L.d; ++d) uu += u[d] * u[d];
    const C usqr = (C)1.5 * uu;
    feq[l] = rho * (C)L.w[l] * ((C)1.0 + cu * ((C)1.0 + (C)0.5 * cu) - usqr);
  }
}

// Second moment (second_moment.py:68-78).
template <class C>
void second_moment(const Lattice& L, const C* fneq, C* pi) {
  for (int t = 0; t < L.nt; ++t) {
    pi[t] = 0;
    for (int l = 0; l < L.q; ++l) pi[t] += (C)L.cc[l][t] * fneq[l];
  }
}

// KBC shear decomposition with the reference's literal index assignments
// (kbc.py:192-250; D3Q27 indices refer to the product([0,-1,1]) order).
template <class C>
void kbc_shear(const Lattice& L, const C* fneq, C* s) {
  C pi[6];
  second_moment(L, fneq, pi);
  for (int l = 0; l < L.q; ++l) s[l] = 0;
  if (L.d == 3) {
    const C nxz = pi[0] - pi[5], nyz = pi[3] - pi[5];
    const C two = 2, four = 4, six = 6;
    s[9] = (two * nxz - nyz) / six;
    s[18] = (two * nxz - nyz) / six;
    s[3] = (-nxz + two * nyz) / six;
    s[6] = (-nxz + two * nyz) / six;
    s[1] = (-nxz - nyz) / six;
    s[2] = (-nxz - nyz) / six;
    s[12] = pi[1] / four;
    s[24] = pi[1] / four;
    s[21] = -pi[1] / four;
    s[15] = -pi[1] / four;
    if (g_plant == 2) std::swap(s[12], s[15]);
    s[10] = pi[2] / four;
    s[20] = pi[2] / four;
    s[19] = -pi[2] / four;
    s[11] = -pi[2] / four;
    s[8] = pi[4] / four;
    s[4] = pi[4] / four;
    s[7] = -pi[4] / four;
    s[5] = -pi[4] / four;
  } else {
    const C n = pi[0] - pi[2];
    s[3] = n;
    s[6] = n;
    s[2] = -n;
    s[1] = -n;
    s[8] = pi[1];
    s[4] = -pi[1];
    s[5] = -pi[1];
    s[7] = pi[1];
    if (g_plant == 2) std::swap(s[8], s[4]);
  }
}

// Collision: 0 BGK (bgk.py:30-34), 1 KBC (kbc.py:268-296), 2 BGK with the Smagorinsky eddy
// relaxation time (smagorinsky_les_bgk.py:73-89, coefficient cs); then exact-difference forcing
// (forced_collision.py:41-47, exact_difference_force.py:71-78).
template <class C>
void collision(const Lattice& L, int coll, const C* f, const C* feq, C rho, const C* u, C omega, C* out,
               C cs = (C)0) {
  if (coll == 0) {
    for (int l = 0; l < L.q; ++l) {
      const C fneq = f[l] - feq[l];
      out[l] = f[l] - omega * fneq;
    }
    return;
  }
  if (coll == 2) {
    // "strain": fneq^2 over directions whose signed component sum is 1, 2 fneq^2 where it is >= 2
    C strain = 0;
    for (int l = 0; l < L.q; ++l) {
      const int csum = L.c[0][l] + L.c[1][l] + (L.d == 3 ? L.c[2][l] : 0);
      const C fneq = f[l] - feq[l];
      if (csum == 1) strain += fneq * fneq;
      if (csum >= 2) strain += (C)2.0 * fneq * fneq;
    }
    const C tau0 = (C)1.0 / omega;
    const C tau = tau0 + (C)0.5 * (std::sqrt(tau0 * tau0 + (C)36.0 * (cs * cs) * std::sqrt(strain)) - tau0);
    for (int l = 0; l < L.q; ++l) out[l] = f[l] - ((C)1.0 / tau) * (f[l] - feq[l]);
    return;
  }
  C fneq[27], shear[27], ds[27], dh[27];
  for (int l = 0; l < L.q; ++l) fneq[l] = f[l] - feq[l];
  kbc_shear(L, fneq, shear);
  for (int l = 0; l < L.q; ++l) ds[l] = (L.d == 3) ? shear[l] * rho : shear[l] * rho / (C)4.0;
  const C beta = (C)0.5 * omega;
  const C inv_beta = (C)1.0 / beta;
  for (int l = 0; l < L.q; ++l) dh[l] = fneq[l] - ds[l];
  C sp1 = 0, sp2 = 0;
  for (int l = 0; l < L.q; ++l) {
    const C temp = dh[l] / feq[l];
    sp1 += temp * ds[l];
    sp2 += temp * dh[l];
  }
  const C eps = (C)1e-32;
  const C gamma = g_plant == 1 ? inv_beta + ((C)2.0 - inv_beta) * sp1 / (eps + sp2)
                               : inv_beta - ((C)2.0 - inv_beta) * sp1 / (eps + sp2);
  for (int l = 0; l < L.q; ++l) out[l] = f[l] - beta * ((C)2.0 * ds[l] + gamma * dh[l]);
}

template <class C>
void exact_difference(const Lattice& L, C* fpc, const C* feq, C rho, const C* u, const C* F) {
  C uf[3] = {0, 0, 0};
  for (int d = 0; d < L.d; ++d) uf[d] = u[d] + F[d];
  C feqF[27];
  equilibrium(L, rho, uf, feqF);
  for (int l = 0; l < L.q; ++l) fpc[l] += feqF[l] - feq[l];
}

// ---- boundary helpers (helper_functions_bc.py) ---------------------------------------------
template <class C>
C get_bc_fsum(const Lattice& L, const C* f, const uint8_t* mm) {
  C known = 0, middle = 0;
  for (int l = 0; l < L.q; ++l) {
    if (mm[L.opp[l]] == 1) known += (C)2.0 * f[l];
    else if (mm[l] != 1) middle += f[l];
  }
  return known + middle;
}

template <class C>
void get_normal(const Lattice& L, const uint8_t* mm, C* n) {
  n[0] = n[1] = n[2] = 0;
  for (int l = 0; l < L.q; ++l) {
    int s = 0;
    for (int d = 0; d < L.d; ++d) s += std::abs(L.c[d][l]);
    if (mm[l] == 1 && s == 1) {
      for (int d = 0; d < L.d; ++d) n[d] = -(C)L.c[d][l];
      return;
    }
  }
}

template <class C>
void bounceback_nonequilibrium(const Lattice& L, C* f, const C* feq, const uint8_t* mm) {
  for (int l = 0; l < L.q; ++l)
    if (mm[l] == 1) f[l] = f[L.opp[l]] + feq[l] - feq[L.opp[l]];
}

template <class C>
void regularize_fpop(const Lattice& L, C* f, const C* feq) {
  C fneq[27], pi[6];
  for (int l = 0; l < L.q; ++l) fneq[l] = f[l] - feq[l];
  second_moment(L, fneq, pi);
  for (int l = 0; l < L.q; ++l) {
    C qp = 0;
    for (int t = 0; t < L.nt; ++t) qp += (C)L.qi[l][t] * pi[t];
    const C fpop1 = (C)4.5 * (C)L.w[l] * qp;
    f[l] = feq[l] + fpop1;
  }
}

// One BC functional (signature of the reference's warp functionals:
// index, timestep, missing_mask, f_0, f_1, f_pre, f_post).  Writes the result to out.
template <int FMT, class C>
void bc_functional(const Ctx<FMT, C>& X, const OrcBc& bc, int x, int y, int z, const uint8_t* mm,
                   const typename Ctx<FMT, C>::S* f1, const C* f_pre, const C* f_post, C* out) {
  const Lattice& L = *X.L;
  const size_t c = X.cell(x, y, z);
  for (int l = 0; l < L.q; ++l) out[l] = f_post[l];
  switch (bc.kind) {
    case 1: {  // EquilibriumBC (bc_equilibrium.py:75-86)
      C u[3] = {(C)bc.u[0], (C)bc.u[1], (C)bc.u[2]};
      equilibrium(L, (C)bc.rho, u, out);
      break;
    }
    case 2:    // HalfwayBounceBackBC (bc_halfway_bounce_back.py:67-85)
    case 9:    // ExtrapolationOutflowBC streaming part (bc_extrapolation_outflow.py:152-161)
      for (int l = 0; l < L.q; ++l)
        if (mm[l] == 1) out[l] = f_pre[L.opp[l]];
      break;
    case 3:  // FullwayBounceBackBC (bc_fullway_bounce_back.py:59-72)
      for (int l = 0; l < L.q; ++l) out[l] = f_pre[L.opp[l]];
      break;
    case 4:  // DoNothingBC (bc_do_nothing.py:52-62)
      for (int l = 0; l < L.q; ++l) out[l] = f_pre[l];
      break;
    case 5:
    case 7: {  // velocity: ZouHe (bc_zouhe.py:279-312), Regularized (bc_regularized.py:134-168)
      C n[3];
      get_normal(L, mm, n);
      const C fsum = get_bc_fsum(L, out, mm);
      const C pv = block(FMT) ? (C)X.aux_side[c] : X.rd_raw(f1, 0, c);
      C u[3] = {0, 0, 0};
      for (int d = 0; d < L.d; ++d) u[d] = -pv * n[d];
      C un = 0;
      for (int d = 0; d < L.d; ++d) un += u[d] * n[d];
      const C rho = fsum / ((C)1.0 + un);
      C feq[27];
      equilibrium(L, rho, u, feq);
      bounceback_nonequilibrium(L, out, feq, mm);
      if (bc.kind == 7) regularize_fpop(L, out, feq);
      break;
    }
    case 6:
    case 8: {  // pressure: ZouHe (bc_zouhe.py:314-338), Regularized (bc_regularized.py:170-196)
      C n[3];
      get_normal(L, mm, n);
      const C rho = block(FMT) ? (C)X.aux_side[c] : X.rd_raw(f1, 0, c);
      const C fsum = get_bc_fsum(L, out, mm);
      const C un = -(C)1.0 + fsum / rho;
      C u[3] = {0, 0, 0};
      for (int d = 0; d < L.d; ++d) u[d] = un * n[d];
      C feq[27];
      equilibrium(L, rho, u, feq);
      bounceback_nonequilibrium(L, out, feq, mm);
      if (bc.kind == 8) regularize_fpop(L, out, feq);
      break;
    }
    case 10: {  // GradsApproximationBC, functional_method1 (bc_grads_approximation.py:181-210)
      const C one = (C)1.0;
      const C weight = (C)0.5;
      for (int l = 0; l < L.q; ++l)
        if (mm[l] == 1)
          out[l] = ((one - weight) * out[L.opp[l]] + weight * (f_pre[l] + f_pre[L.opp[l]])) / (one + weight);
      C rho, u[3], feq[27];
      macroscopic(L, out, rho, u);
      equilibrium(L, rho, u, feq);
      regularize_fpop(L, out, feq);  // regularize_fpop of bc_grads_approximation.py:112-135
      break;
    }
    default:
      break;
  }
}

// ExtrapolationOutflowBC.update_bc_auxilary_data, Warp form (bc_extrapolation_outflow.py:163-195).
// Source indices outside the domain wrap periodically (the reference reads them unchecked).
template <int FMT, class C>
void extrapolation_aux(const Ctx<FMT, C>& X, int x, int y, int z, const uint8_t* mm, const C* f_pre,
                       const C* f_post, C* out) {
  const Lattice& L = *X.L;
  for (int l = 0; l < L.q; ++l) out[l] = f_post[l];
  const C cs = (C)(1.0 / std::sqrt(3.0));
  int nv[3] = {0, 0, 0};
  for (int l = 0; l < L.q; ++l) {
    int s = 0;
    for (int d = 0; d < L.d; ++d) s += std::abs(L.c[d][l]);
    if (mm[l] == 1 && s == 1) {
      for (int d = 0; d < L.d; ++d) nv[d] = -L.c[d][l];
      break;
    }
  }
  const int shape[3] = {X.nx, X.ny, X.nz};
  for (int l = 0; l < L.q; ++l) {
    if (mm[l] == 1) {
      int idx[3] = {x, y, z}, p[3] = {x, y, z};
      for (int d = 0; d < L.d; ++d) {
        p[d] = idx[d] - (L.c[d][l] + (g_plant == 6 ? 0 : nv[d]));
        p[d] = ((p[d] % shape[d]) + shape[d]) % shape[d];
      }
      const C faux = X.rd(X.f0, l, X.cell(p[0], p[1], p[2]));
      out[L.opp[l]] = g_plant == 5 ? cs * f_pre[l] + ((C)1.0 - cs) * faux : ((C)1.0 - cs) * f_pre[l] + cs * faux;
    }
  }
}

// apply_bc of the stepper (nse_stepper.py:265-294).
template <int FMT, class C>
void apply_bc(const Ctx<FMT, C>& X, int x, int y, int z, int id, const uint8_t* mm, const typename Ctx<FMT, C>::S* f1,
              const C* f_pre, const C* f_post, bool post_streaming, C* result) {
  const Lattice& L = *X.L;
  for (int l = 0; l < L.q; ++l) result[l] = f_post[l];
  for (int i = 0; i < X.nbc; ++i) {
    const OrcBc& bc = X.bcs[i];
    if (post_streaming) {
      if (bc_step(bc.kind) == 0 && id == bc.id) bc_functional(X, bc, x, y, z, mm, f1, f_pre, f_post, result);
    } else {
      if (bc_step(bc.kind) == 1 && id == bc.id) bc_functional(X, bc, x, y, z, mm, f1, f_pre, f_post, result);
      if (bc.kind == 9 && id == bc.id) extrapolation_aux(X, x, y, z, mm, f_pre, f_post, result);
    }
  }
}

// One fused step (nse_stepper.py:344-383) over the whole grid.
template <int FMT, class C>
void step_grid(Ctx<FMT, C>& X, int coll) {
  const Lattice& L = *X.L;
  const int Q = L.q;
  using S = typename Ctx<FMT, C>::S;
  // Snapshot f_1 at aux-recovery cells before it is overwritten (nse_stepper.py:310, 318-342).
  std::vector<size_t> aux_cells;
  for (size_t c = 0; c < (block(FMT) ? 0 : X.N); ++c) {
    const int id = X.bc_mask[c];
    for (int i = 0; i < X.nbc; ++i)
      if (needs_aux_recovery(X.bcs[i].kind) && X.bcs[i].id == id) {
        aux_cells.push_back(c);
        break;
      }
  }
  std::vector<S> f1_snap(aux_cells.size() * Q);
  for (size_t k = 0; k < aux_cells.size(); ++k)
    for (int l = 0; l < Q; ++l) f1_snap[k * Q + l] = X.f1[(size_t)l * X.N + aux_cells[k]];
  // block format: the step's float outputs, quantised segment by segment once all are known
  std::vector<float> blk_out;
  std::vector<uint8_t> blk_written;
  if constexpr (block(FMT)) {
    blk_out.assign((size_t)Q * X.N, 0.0f);
    blk_written.assign(X.N, 0);
  }

#pragma omp parallel for schedule(static)
  for (int x = 0; x < X.nx; ++x) {
    for (int y = 0; y < X.ny; ++y)
      for (int z = 0; z < X.nz; ++z) {
        const size_t c = X.cell(x, y, z);
        const int id = X.bc_mask ? X.bc_mask[c] : 0;
        if (id == 255) continue;
        C fps[27], f0t[27], fpc[27], tmp[27];
        uint8_t mm[27] = {};
        stream_functional(X, X.f0, x, y, z, fps);
        for (int l = 0; l < Q; ++l) {
          f0t[l] = X.rd(X.f0, l, c);
          mm[l] = X.missing ? (X.missing[(size_t)l * X.N + c] ? 1 : 0) : 0;
        }
        for (int l = 0; l < Q; ++l) fpc[l] = f0t[l];
        apply_bc(X, x, y, z, id, mm, X.f1, fpc, fps, true, tmp);
        for (int l = 0; l < Q; ++l) fps[l] = tmp[l];
        C rho, u[3], feq[27];
        macroscopic(L, fps, rho, u);
        equilibrium(L, rho, u, feq);
        collision(L, coll, fps, feq, rho, u, X.omega, fpc, X.cparam);
        if (X.force_mode != 0) {
          C F[3] = {X.force_u[0], X.force_u[1], X.force_u[2]};
          if (X.force_mode == 2)
            for (int d = 0; d < L.d; ++d) F[d] = (C)X.force_field[(size_t)d * X.N + c];
          exact_difference(L, fpc, feq, rho, u, F);
        }
        apply_bc(X, x, y, z, id, mm, X.f1, fps, fpc, false, tmp);
        if constexpr (block(FMT)) {
          for (int l = 0; l < Q; ++l) blk_out[(size_t)l * X.N + c] = (float)tmp[l];
          blk_written[c] = 1;
        } else {
          for (int l = 0; l < Q; ++l) X.wr(X.f1, l, c, tmp[l]);
        }
      }
  }
  if constexpr (block(FMT)) blk_store(L.w, Q, X.g, blk_out.data(), blk_written.data(), X.f1);
  // Aux recovery into f_0 (after every read of f_0 by this step).
  for (size_t k = 0; k < aux_cells.size(); ++k) {
    const size_t c = aux_cells[k];
    X.f0[c] = f1_snap[k * Q + 0];
    for (int l = 1; l < Q; ++l)
      if (X.missing[(size_t)l * X.N + c]) {
        const int o = L.opp[l];
        X.f0[(size_t)o * X.N + c] = f1_snap[k * Q + o];
      }
  }
}

template <class Fn>
int dispatch(int store, int compute, Fn&& fn) {
  if (compute == 0) {
    switch (store) {
      case 0: return fn(std::integral_constant<int, 0>{}, float{});
      case 1: return fn(std::integral_constant<int, 1>{}, float{});
      case 2: return fn(std::integral_constant<int, 2>{}, float{});
      case 3: return fn(std::integral_constant<int, 3>{}, float{});
      case 4: return fn(std::integral_constant<int, 4>{}, float{});
      case 5: return fn(std::integral_constant<int, 5>{}, float{});
      case 6: return fn(std::integral_constant<int, 6>{}, float{});
    }
  } else {
    switch (store) {
      case 0: return fn(std::integral_constant<int, 0>{}, double{});
      case 1: return fn(std::integral_constant<int, 1>{}, double{});
      case 2: return fn(std::integral_constant<int, 2>{}, double{});
      case 3: return fn(std::integral_constant<int, 3>{}, double{});
      case 4: return fn(std::integral_constant<int, 4>{}, double{});
      case 5: return fn(std::integral_constant<int, 5>{}, double{});
    }
  }
  return 1;
}

Lattice g_lat[3] = {build_lattice(0), build_lattice(1), build_lattice(2)};

}  // namespace

extern "C" {

int orc_lattice(int lat, int* c /*3*27*/, double* w, int* opp, double* qi /*27*6*/) {
  if (lat < 0 || lat > 2) return 1;
  const Lattice& L = g_lat[lat];
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 27; ++i) c[a * 27 + i] = L.c[a][i];
  for (int i = 0; i < 27; ++i) {
    w[i] = L.w[i];
    opp[i] = L.opp[i];
    for (int t = 0; t < 6; ++t) qi[i * 6 + t] = L.qi[i][t];
  }
  return L.q;
}

// Test hook: switch on planted slip `which` (0 = none); returns the previous setting.
int orc_plant(int which) {
  const int old = g_plant;
  g_plant = which;
  return old;
}

int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

// One fused step.  f0 is read (and receives aux recovery writes), f1 is written.
// aux: the block format's prescribed values per cell (N doubles), in place of the f_1[0] stash an
// int16 mantissa cannot hold; NULL for every other format.
int orc_step(int lat, int coll, int store, int compute, void* f0, void* f1, const uint8_t* bc_mask,
             const uint8_t* missing, const OrcBc* bcs, int nbc, int force_mode, const double* force_u,
             const double* force_field, int nx, int ny, int nz, double omega, double cparam, const double* aux) {
  return dispatch(store, compute, [&](auto S, auto c) -> int {
    constexpr int FMT = decltype(S)::value;
    using C = decltype(c);
    Ctx<FMT, C> X;
    X.L = &g_lat[lat];
    X.f0 = static_cast<typename Ctx<FMT, C>::S*>(f0);
    X.f1 = static_cast<typename Ctx<FMT, C>::S*>(f1);
    X.bc_mask = bc_mask;
    X.missing = missing;
    X.bcs = bcs;
    X.nbc = nbc;
    X.nx = nx;
    X.ny = ny;
    X.nz = nz;
    X.N = (size_t)nx * ny * nz;
    X.omega = (C)omega;
    X.cparam = (C)cparam;
    X.force_mode = force_mode;
    for (int d = 0; d < 3; ++d) X.force_u[d] = force_u ? (C)force_u[d] : (C)0;
    X.force_field = force_field;
    X.g = blk_geom(nx, ny, nz);
    if constexpr (block(FMT)) {
      X.aux_side = aux;
      for (int i = 0; i < nbc; ++i)
        if (needs_aux_recovery(bcs[i].kind) && !aux) return 3;  // prescribed values need the side table
    }
    step_grid(X, coll);
    return 0;
  });
}

// BoundaryCondition.aux_data_init (boundary_condition.py:130-161) with num_of_aux_data = 1:
// f_1[0] = store(value) at cells with bc_mask == id.  values: (N) doubles.
int orc_aux_init(int store, void* f1, const uint8_t* bc_mask, int id, const double* values, int nx, int ny, int nz) {
  const size_t N = (size_t)nx * ny * nz;
  return dispatch(store, 0, [&](auto S, auto) -> int {
    constexpr int FMT = decltype(S)::value;
    if constexpr (block(FMT)) return 3;
    auto* f = static_cast<typename Store<FMT>::T*>(f1);
    for (size_t c = 0; c < N; ++c)
      if (bc_mask[c] == id) f[c] = Store<FMT>::put(values[c]);
    return 0;
  });
}

// f = feq(rho, u) with rho (N), u (d, N) doubles (quadratic_equilibrium.py:62-98).
int orc_equilibrium(int lat, int store, int compute, const double* rho, const double* u, void* f, int nx, int ny,
                    int nz) {
  return dispatch(store, compute, [&](auto S, auto c) -> int {
    constexpr int FMT = decltype(S)::value;
    using C = decltype(c);
    Ctx<FMT, C> X;
    X.L = &g_lat[lat];
    X.N = (size_t)nx * ny * nz;
    const Lattice& L = *X.L;
    auto* ff = static_cast<typename Ctx<FMT, C>::S*>(f);
    std::vector<float> out(block(FMT) ? (size_t)L.q * X.N : 0);
    for (size_t cc = 0; cc < X.N; ++cc) {
      C u3[3] = {0, 0, 0}, feq[27];
      for (int d = 0; d < L.d; ++d) u3[d] = (C)u[(size_t)d * X.N + cc];
      equilibrium(L, (C)rho[cc], u3, feq);
      for (int l = 0; l < L.q; ++l) {
        if constexpr (block(FMT))
          out[(size_t)l * X.N + cc] = (float)feq[l];
        else
          X.wr(ff, l, cc, feq[l]);
      }
    }
    if constexpr (block(FMT)) {
      const std::vector<uint8_t> all(X.N, 1);
      blk_store(L.w, L.q, blk_geom(nx, ny, nz), out.data(), all.data(), ff);
    }
    return 0;
  });
}

// rho (N), u (d, N) as doubles computed in the compute type (macroscopic.py:39-55).
int orc_macroscopic(int lat, int store, int compute, const void* f, double* rho, double* u, int nx, int ny, int nz) {
  return dispatch(store, compute, [&](auto S, auto c) -> int {
    constexpr int FMT = decltype(S)::value;
    using C = decltype(c);
    Ctx<FMT, C> X;
    X.L = &g_lat[lat];
    X.N = (size_t)nx * ny * nz;
    const Lattice& L = *X.L;
    X.nx = nx;
    X.ny = ny;
    X.nz = nz;
    X.g = blk_geom(nx, ny, nz);
    auto* ff = static_cast<const typename Ctx<FMT, C>::S*>(f);
    for (size_t cc = 0; cc < X.N; ++cc) {
      C fl[27], r, uu[3];
      for (int l = 0; l < L.q; ++l) fl[l] = X.rd(ff, l, cc);
      macroscopic(L, fl, r, uu);
      rho[cc] = (double)r;
      for (int d = 0; d < L.d; ++d) u[(size_t)d * X.N + cc] = (double)uu[d];
    }
    return 0;
  });
}

// Standalone pull streaming of raw stored values (stream.py:85-99).
int orc_stream(int lat, int store, const void* fin, void* fout, int nx, int ny, int nz) {
  return dispatch(store, 0, [&](auto S, auto) -> int {
    constexpr int FMT = decltype(S)::value;
    if constexpr (block(FMT)) return 3;  // the block format is written by the step and equilibrium only
    using T = typename Store<FMT>::T;
    const Lattice& L = g_lat[lat];
    const T* a = static_cast<const T*>(fin);
    T* b = static_cast<T*>(fout);
    const size_t N = (size_t)nx * ny * nz;
    const int shape[3] = {nx, ny, nz};
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z) {
          const int idx[3] = {x, y, z};
          for (int l = 0; l < L.q; ++l) {
            int p[3] = {x, y, z};
            for (int d = 0; d < L.d; ++d) {
              p[d] = idx[d] - L.c[d][l];
              if (p[d] < 0) p[d] = shape[d] - 1;
              else if (p[d] >= shape[d]) p[d] = 0;
            }
            b[(size_t)l * N + ((size_t)x * ny + y) * nz + z] = a[(size_t)l * N + ((size_t)p[0] * ny + p[1]) * nz + p[2]];
          }
        }
    return 0;
  });
}

// Standalone collision: fout = collision(f, feq, rho, u) (+ force) with rho/u read from fields
// (doubles); BGK ignores them unless forced.
int orc_collide(int lat, int coll, int store, int compute, const void* f, const void* feq, void* fout,
                const double* rho, const double* u, double omega, double cparam, int force_mode,
                const double* force_u, const double* force_field, int nx, int ny, int nz) {
  return dispatch(store, compute, [&](auto S, auto c) -> int {
    constexpr int FMT = decltype(S)::value;
    if constexpr (block(FMT)) return 3;  // the block format is written by the step and equilibrium only
    using C = decltype(c);
    Ctx<FMT, C> X;
    X.L = &g_lat[lat];
    X.N = (size_t)nx * ny * nz;
    const Lattice& L = *X.L;
    using T = typename Ctx<FMT, C>::S;
    for (size_t cc = 0; cc < X.N; ++cc) {
      C fl[27], fe[27], out[27], uu[3] = {0, 0, 0};
      for (int l = 0; l < L.q; ++l) {
        fl[l] = X.rd(static_cast<const T*>(f), l, cc);
        fe[l] = X.rd(static_cast<const T*>(feq), l, cc);
      }
      C r = rho ? (C)rho[cc] : (C)0;
      if (u)
        for (int d = 0; d < L.d; ++d) uu[d] = (C)u[(size_t)d * X.N + cc];
      collision(L, coll, fl, fe, r, uu, (C)omega, out, (C)cparam);
      if (force_mode != 0) {
        C F[3] = {0, 0, 0};
        for (int d = 0; d < L.d; ++d)
          F[d] = (force_mode == 2) ? (C)force_field[(size_t)d * X.N + cc] : (C)force_u[d];
        exact_difference(L, out, fe, r, uu, F);
      }
      for (int l = 0; l < L.q; ++l) X.wr(static_cast<T*>(fout), l, cc, out[l]);
    }
    return 0;
  });
}

// Standalone BC (boundary_condition.py:92-117): aux value read from `aux` (N doubles) instead of
// f_post[0] so it can be compared with the product's side table.
int orc_apply_bc(int lat, int store, int compute, const OrcBc* bc, const void* fpre, void* fpost,
                 const uint8_t* bc_mask, const uint8_t* missing, const double* aux, int nx, int ny, int nz) {
  return dispatch(store, compute, [&](auto S, auto c) -> int {
    constexpr int FMT = decltype(S)::value;
    if constexpr (block(FMT)) return 3;  // the block format is written by the step and equilibrium only
    using C = decltype(c);
    using T = typename Ctx<FMT, C>::S;
    Ctx<FMT, C> X;
    X.L = &g_lat[lat];
    X.nx = nx;
    X.ny = ny;
    X.nz = nz;
    X.N = (size_t)nx * ny * nz;
    const Lattice& L = *X.L;
    std::vector<T> aux_field(X.N);  // a private f_1 carrying the aux value in slot 0
    for (size_t cc = 0; cc < X.N; ++cc) aux_field[cc] = Store<FMT>::put(aux ? aux[cc] : 0.0);
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z) {
          const size_t cc = X.cell(x, y, z);
          if (bc_mask[cc] != bc->id) continue;
          C pre[27], post[27], out[27];
          uint8_t mm[27] = {};
          for (int l = 0; l < L.q; ++l) {
            pre[l] = X.rd(static_cast<const T*>(fpre), l, cc);
            post[l] = X.rd(static_cast<T*>(fpost), l, cc);
            mm[l] = missing[(size_t)l * X.N + cc] ? 1 : 0;
          }
          bc_functional(X, *bc, x, y, z, mm, aux_field.data(), pre, post, out);
          for (int l = 0; l < L.q; ++l) X.wr(static_cast<T*>(fpost), l, cc, out[l]);
        }
    return 0;
  });
}

// IndicesBoundaryMasker, Warp semantics, BCs applied in list order (indices_boundary_masker.py:103-224).
// indices: (3, n) int32 (z row zeros in 2-D); interior: (n) flags.
int orc_mask_indices(int lat, const int32_t* indices, const uint8_t* interior, int n, int id, uint8_t* bc_mask,
                     uint8_t* missing, int nx, int ny, int nz) {
  const Lattice& L = g_lat[lat];
  const size_t N = (size_t)nx * ny * nz;
  auto inb = [&](int a, int b, int e) { return a >= 0 && a < nx && b >= 0 && b < ny && e >= 0 && e < nz; };
  for (int i = 0; i < n; ++i) {
    const int idx[3] = {indices[i], indices[n + i], indices[2 * n + i]};
    if (!inb(idx[0], idx[1], idx[2])) continue;
    for (int l = 0; l < L.q; ++l) {
      int pull[3] = {0, 0, 0}, push[3] = {0, 0, 0};
      for (int d = 0; d < L.d; ++d) {
        pull[d] = idx[d] - L.c[d][l];
        push[d] = idx[d] + L.c[d][l];
      }
      bc_mask[((size_t)idx[0] * ny + idx[1]) * nz + idx[2]] = (uint8_t)id;
      if (!inb(pull[0], pull[1], pull[2])) {
        missing[(size_t)l * N + ((size_t)idx[0] * ny + idx[1]) * nz + idx[2]] = 1;
      } else if (interior[i]) {
        missing[(size_t)l * N + ((size_t)push[0] * ny + push[1]) * nz + push[2]] = 1;
        bc_mask[((size_t)push[0] * ny + push[1]) * nz + push[2]] = (uint8_t)id;
      }
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------------------------
// MeshBoundaryMasker (mesh_boundary_masker.py:49-236), voxel-parallel as the reference: a voxel
// (unit box with minimum corner at its index) touched by any triangle is solid (255); otherwise
// every direction l whose neighbour box (index + c_l) is touched sets missing[opp l] and the id.
// Triangle / box test: bounding-box overlap (the reference's BVH query, :139), then Schwarz &
// Seidel 2010, sec. 3.1 (plane overlap through the critical corners, then the three axis
// projections with three edge functions each), FP32, fixed order.
// ---------------------------------------------------------------------------------------------
struct OrcTri {
  float lo[3], hi[3], n[3], d1, d2, ne0[3][3], ne1[3][3], de[3][3];
  bool valid;
};

static float orc_dot3(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

static OrcTri orc_tri_prepare(const float* v) {
  OrcTri t{};
  const float* vv[3] = {v, v + 3, v + 6};
  float e[3][3], w[3], cr[3];
  for (int a = 0; a < 3; ++a) {
    e[0][a] = vv[1][a] - vv[0][a];
    e[1][a] = vv[2][a] - vv[1][a];
    e[2][a] = vv[0][a] - vv[2][a];
    w[a] = vv[2][a] - vv[0][a];
    t.lo[a] = std::min(std::min(vv[0][a], vv[1][a]), vv[2][a]);
    t.hi[a] = std::max(std::max(vv[0][a], vv[1][a]), vv[2][a]);
  }
  cr[0] = e[0][1] * w[2] - e[0][2] * w[1];
  cr[1] = e[0][2] * w[0] - e[0][0] * w[2];
  cr[2] = e[0][0] * w[1] - e[0][1] * w[0];
  const float len = std::sqrt(orc_dot3(cr, cr));
  t.valid = len > 0.0f;
  for (int a = 0; a < 3; ++a) t.n[a] = t.valid ? cr[a] / len : 0.0f;
  float c[3], cm[3];
  for (int a = 0; a < 3; ++a) {
    const float k = t.n[a] > 0.0f ? 1.0f : 0.0f;
    c[a] = k - vv[0][a];
    cm[a] = (1.0f - k) - vv[0][a];
  }
  t.d1 = orc_dot3(t.n, c);
  t.d2 = orc_dot3(t.n, cm);
  for (int p = 0; p < 3; ++p) {
    const int a0 = p, a1 = (p + 1) % 3, a2 = (p + 2) % 3;
    const float sg = t.n[a2] < 0.0f ? -1.0f : 1.0f;
    for (int i = 0; i < 3; ++i) {
      const float n0 = -sg * e[i][a1];
      const float n1 = sg * e[i][a0];
      t.ne0[p][i] = n0;
      t.ne1[p][i] = n1;
      t.de[p][i] = (-(n0 * vv[i][a0] + n1 * vv[i][a1]) + std::max(0.0f, n0)) + std::max(0.0f, n1);
    }
  }
  return t;
}

static bool orc_tri_box_pre(const OrcTri& t, const float* p) {
  if (!t.valid) return false;
  for (int a = 0; a < 3; ++a)  // bounding boxes overlap (the reference's BVH query, :139)
    if (t.lo[a] > p[a] + 1.0f || t.hi[a] < p[a]) return false;
  const float np = orc_dot3(t.n, p);
  if ((np + t.d1) * (np + t.d2) > 0.0f) return false;
  for (int q = 0; q < 3; ++q) {
    const int a0 = q, a1 = (q + 1) % 3;
    for (int i = 0; i < 3; ++i)
      if ((t.ne0[q][i] * p[a0] + t.ne1[q][i] * p[a1]) + t.de[q][i] < 0.0f) return false;
  }
  return true;
}

// Single test, exported for checking against an independent overlap test.
int orc_tri_box(const float* tri9, const float* p) { return orc_tri_box_pre(orc_tri_prepare(tri9), p) ? 1 : 0; }

int orc_mesh_mask(int lat, const float* tri, int64_t n_tri, int id, uint8_t* bc_mask, uint8_t* missing, int nx, int ny,
                  int nz) {
  const Lattice& L = g_lat[lat];
  const size_t N = (size_t)nx * ny * nz;
  std::vector<OrcTri> pre((size_t)n_tri);
  for (int64_t t = 0; t < n_tri; ++t) pre[(size_t)t] = orc_tri_prepare(tri + 9 * t);
  const int hx = nx + 2, hy = ny + 2, hz = nz + 2;
  std::vector<uint8_t> solid((size_t)hx * hy * hz, 0);
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int x = -1; x <= nx; ++x)
    for (int y = -1; y <= ny; ++y)
      for (int z = -1; z <= nz; ++z) {
        const float p[3] = {(float)x, (float)y, (float)z};
        for (int64_t t = 0; t < n_tri; ++t)
          if (orc_tri_box_pre(pre[(size_t)t], p)) {
            solid[((size_t)(x + 1) * hy + (y + 1)) * hz + (z + 1)] = 1;
            break;
          }
      }
  auto sol = [&](int x, int y, int z) { return solid[((size_t)(x + 1) * hy + (y + 1)) * hz + (z + 1)] != 0; };
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const size_t cell = ((size_t)x * ny + y) * nz + z;
        if (sol(x, y, z)) {
          bc_mask[cell] = 255;
          continue;
        }
        for (int l = 1; l < L.q; ++l) {
          if (sol(x + L.c[0][l], y + L.c[1][l], z + L.c[2][l])) {
            bc_mask[cell] = (uint8_t)id;
            missing[(size_t)L.opp[l] * N + cell] = 1;
          }
        }
      }
  return 0;
}

// MomentumTransfer (momentum_transfer.py:108-176), FP64 accumulation over cells.
int orc_momentum_transfer(int lat, int store, int compute, const OrcBc* bc, const void* f0, const uint8_t* bc_mask,
                          const uint8_t* missing, const double* aux, int nx, int ny, int nz, double* out) {
  return dispatch(store, compute, [&](auto S, auto c) -> int {
    constexpr int FMT = decltype(S)::value;
    using C = decltype(c);
    using T = typename Ctx<FMT, C>::S;
    Ctx<FMT, C> X;
    X.L = &g_lat[lat];
    X.nx = nx;
    X.ny = ny;
    X.nz = nz;
    X.N = (size_t)nx * ny * nz;
    X.g = blk_geom(nx, ny, nz);
    const Lattice& L = *X.L;
    std::vector<T> aux_field(X.N);
    for (size_t cc = 0; cc < X.N; ++cc) aux_field[cc] = Store<FMT>::put(aux ? aux[cc] : 0.0);
    double acc[3] = {0, 0, 0};
    const T* f = static_cast<const T*>(f0);
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z) {
          const size_t cc = X.cell(x, y, z);
          if (bc_mask[cc] != bc->id) continue;
          uint8_t mm[27] = {};
          for (int l = 0; l < L.q; ++l) mm[l] = missing[(size_t)l * X.N + cc] ? 1 : 0;
          if (mm[0] != 0) continue;
          C fpc[27], fps[27], tmp[27];
          for (int l = 0; l < L.q; ++l) fpc[l] = X.rd(f, l, cc);
          stream_functional(X, f, x, y, z, fps);
          bc_functional(X, *bc, x, y, z, mm, aux_field.data(), fpc, fps, tmp);
          for (int d = 0; d < L.d; ++d) {
            C m = 0;
            for (int l = 0; l < L.q; ++l)
              if (mm[l] == 1) {
                const C phi = g_plant == 3 ? fpc[l] + tmp[l] : (g_plant == 4 ? fpc[L.opp[l]] : fpc[L.opp[l]] + tmp[l]);
                if (L.c[d][L.opp[l]] == 1) m += phi;
                else if (L.c[d][L.opp[l]] == -1) m -= phi;
              }
            acc[d] += (double)m;
          }
        }
    for (int d = 0; d < 3; ++d) out[d] = acc[d];
    return 0;
  });
}

// Round trips of the 16-bit conversions (for pinning them against numpy).
void orc_f32_to_f16(const float* in, uint16_t* out, int n) {
  for (int i = 0; i < n; ++i) out[i] = f32_to_f16(in[i]);
}
void orc_f32_to_bf16(const float* in, uint16_t* out, int n) {
  for (int i = 0; i < n; ++i) out[i] = f32_to_bf16(in[i]);
}

}  // extern "C"
