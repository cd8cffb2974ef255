// Reverse-mode entry points (xlbh_adjoint_step, xlbh_adjoint_run, xlbh_adjoint_workspace_bytes):
// argument checks, the checkpoint-segment driver and the per-lattice dispatch of the kernels in
// adjoint.hpp.
#include <mutex>

#include "adjoint.hpp"

using namespace xlbh;

static int adjoint_dispatch(int lattice, int store, int compute, const AdjParams& p, hipStream_t s,
                            int mode = ADJ_SPLIT, uint64_t ws_words = 0, uint8_t* flag = nullptr,
                            void* lam = nullptr) {
  compute = arith_compute(compute);  // the reverse kernels of the deviation form are the FP32 ones
  switch (lattice) {
    case XLBH_D2Q9: return adjoint_dispatch_d2q9(store, compute, p, s, mode, ws_words, flag, lam);
    case XLBH_D3Q19: return adjoint_dispatch_d3q19(store, compute, p, s, mode, ws_words, flag, lam);
    case XLBH_D3Q27: return adjoint_dispatch_d3q27(store, compute, p, s, mode, ws_words, flag, lam);
  }
  set_error("unknown lattice id");
  return 1;
}

static int adjoint_validate(int lattice, int collision, int force_mode, const double* force_uniform,
                            const void* force_field, const uint8_t* bc_mask, const uint32_t* missing_bits,
                            const xlbh_bc_desc* bc_table, int nx, int ny, int nz) {
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(collision == XLBH_BGK || collision == XLBH_KBC || collision == XLBH_SMAGORINSKY_BGK,
               "unknown collision");
  XLBH_REQUIRE(!(collision == XLBH_KBC && lattice == XLBH_D3Q19),
               "KBC collision is defined for D2Q9 and D3Q27 only (kbc.py:183-185)");
  XLBH_REQUIRE(!bc_mask || (missing_bits && bc_table), "bc_mask given without missing bits or BC table");
  XLBH_REQUIRE(force_mode >= XLBH_FORCE_NONE && force_mode <= XLBH_FORCE_FIELD, "unknown force mode");
  XLBH_REQUIRE(force_mode != XLBH_FORCE_UNIFORM || force_uniform, "uniform force needs 3 host doubles");
  XLBH_REQUIRE(force_mode != XLBH_FORCE_FIELD || force_field, "force field pointer is null");
  XLBH_REQUIRE((int64_t)nx * ny * nz < ((int64_t)1 << 31), "the adjoint supports fewer than 2^31 cells");
  return 0;
}

// Solid cells (id 255) are refused by the reverse-mode entry points: the forward step leaves them
// unwritten, so in a rollout their populations are whatever the output buffer held two steps
// earlier -- a dependence on the buffer schedule that no per-step adjoint can express, and which
// xlbh_record (whose state buffers are never written there) cannot reproduce.  The scan below reads
// the mask once and reports whether any byte is 255; the flag lives in a per-device __device__
// word, cleared by a one-thread kernel before each scan.
static __device__ int g_solid_found;

__global__ void k_clear_solid_flag(int* flag) { *flag = 0; }

__global__ __launch_bounds__(256) void k_find_solid(const uint8_t* mask, uint64_t n, int* flag) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n16 = ((uintptr_t)mask & 15) == 0 ? n / 16 : 0;
  const uint4* v = reinterpret_cast<const uint4*>(mask);
  bool found = false;
  for (uint64_t i = tid; i < n16; i += stride) {
    const uint4 q = v[i];
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t inv = ~w[k];  // a byte of w is 0xff <=> that byte of inv is zero
      found |= ((inv - 0x01010101u) & ~inv & 0x80808080u) != 0;
    }
  }
  for (uint64_t i = n16 * 16 + tid; i < n; i += stride) found |= mask[i] == 255;
  if (found) *flag = 1;
}

static int refuse_solid_cells(const uint8_t* bc_mask, int nx, int ny, int nz, hipStream_t s) {
  if (!bc_mask) return 0;
  static std::mutex serial;  // one scan at a time: the flag word is shared by every caller on a device
  std::lock_guard<std::mutex> hold(serial);
  int dev = 0;
  XLBH_HIP_CHECK(hipGetDevice(&dev));
  static int* flag_of[64] = {};
  XLBH_REQUIRE(dev >= 0 && dev < 64, "device index out of range");
  if (!flag_of[dev]) XLBH_HIP_CHECK(hipGetSymbolAddress(reinterpret_cast<void**>(&flag_of[dev]), HIP_SYMBOL(g_solid_found)));
  int* flag = flag_of[dev];
  const uint64_t n = (uint64_t)nx * ny * nz;
  hipLaunchKernelGGL(k_clear_solid_flag, dim3(1), dim3(1), 0, s, flag);
  XLBH_LAUNCH_CHECK();
  const uint64_t blocks = std::min<uint64_t>((n / 16 + 255) / 256 + 1, 4096);
  hipLaunchKernelGGL(k_find_solid, dim3((unsigned)blocks), dim3(256), 0, s, bc_mask, n, flag);
  XLBH_LAUNCH_CHECK();
  int found = 0;
  XLBH_HIP_CHECK(hipMemcpyAsync(&found, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  XLBH_HIP_CHECK(hipStreamSynchronize(s));
  XLBH_REQUIRE(!found, "the adjoint does not support solid (id 255) cells");
  return 0;
}

static AdjParams adjoint_params(int collision, double cparam, const void* f_in, const void* lam_out, void* lam_in,
                                const uint8_t* bc_mask, const uint32_t* missing_bits, const void* aux,
                                const xlbh_bc_desc* bc_table, const int32_t* bc_cells, int n_bc_cells, int force_mode,
                                const double* force_uniform, const void* force_field, double* grad_force, int nx,
                                int ny, int nz, uint64_t plane, double omega) {
  AdjParams p;
  p.collision = collision;
  p.cparam = cparam;
  p.f_in = f_in;
  p.lam_out = lam_out;
  p.lam_in = lam_in;
  p.bc_mask = bc_mask;
  p.missing = missing_bits;
  p.aux = aux;
  p.table = bc_table;
  p.bc_cells = bc_cells;
  p.n_bc_cells = n_bc_cells;
  p.d_own = nullptr;
  p.d_nb_in = nullptr;
  p.d_nb_out = nullptr;
  p.nb_flag = nullptr;
  p.force_field = force_field;
  p.grad_force = grad_force;
  p.force_mode = force_mode;
  for (int d = 0; d < 3; ++d) p.fu[d] = force_mode == XLBH_FORCE_UNIFORM ? force_uniform[d] : 0.0;
  p.omega = omega;
  p.sw = make_sweep(nx, ny, nz);
  p.N = (uint32_t)((int64_t)nx * ny * nz);
  p.P = plane;
  return p;
}

extern "C" int xlbh_adjoint_run(int lattice, int collision, int store, int compute, const void* f_ckpt, void* states,
                                void* lam_a, void* lam_b, const uint8_t* bc_mask, const uint32_t* missing_bits,
                                const void* aux, const xlbh_bc_desc* bc_table, const int32_t* bc_cells,
                                int n_bc_cells, int force_mode, const double* force_uniform,
                                const void* force_field, double* grad_force, int nx, int ny, int nz, int64_t plane,
                                double omega, double collision_param, int timestep0, int n_steps, void* workspace,
                                int flags, void* stream) {
  clear_error();
  if (int e = adjoint_validate(lattice, collision, force_mode, force_uniform, force_field, bc_mask, missing_bits,
                               bc_table, nx, ny, nz))
    return e;
  XLBH_REQUIRE(store != XLBH_I16_BLOCK, "the adjoint does not read the block format");
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  XLBH_REQUIRE(n_steps >= 1, "n_steps must be >= 1");
  XLBH_REQUIRE(n_bc_cells >= 0 && (n_bc_cells == 0 || bc_cells), "bc_cells is null with n_bc_cells > 0");
  XLBH_REQUIRE(f_ckpt && lam_a && lam_b, "adjoint_run needs f_ckpt, lam_a and lam_b");
  XLBH_REQUIRE(lam_a != lam_b, "lam_a must not alias lam_b");
  XLBH_REQUIRE(n_steps == 1 || states, "adjoint_run needs a state workspace for n_steps > 1");
  XLBH_REQUIRE(valid_store(store), "unknown store format");
  // the mask scan synchronises the stream: once per sweep (XLBH_ADJ_BEGIN) with a workspace,
  // every call without one
  if (!workspace || (flags & XLBH_ADJ_BEGIN))
    if (int e = refuse_solid_cells(bc_mask, nx, ny, nz, (hipStream_t)stream)) return e;
  const size_t field_bytes = (size_t)lattice_q(lattice) * P * store_bytes(store);
  auto state = [&](int j) -> const void* {
    return j == 0 ? f_ckpt : static_cast<const void*>(static_cast<const char*>(states) + (size_t)(j - 1) * field_bytes);
  };
  // recompute f^{t0+1} .. f^{t0+n-1} (unless the caller recorded them, e.g. with xlbh_record)
  const bool recompute = (flags & XLBH_ADJ_RECOMPUTE) != 0;
  for (int j = 1; recompute && j < n_steps; ++j) {
    if (int e = xlbh_step(lattice, collision, store, compute, state(j - 1), const_cast<void*>(state(j)), bc_mask,
                          missing_bits, aux, bc_table, force_mode, force_uniform, force_field, nx, ny, nz,
                          (int64_t)P, omega, collision_param, timestep0 + j - 1, stream))
      return e;
  }
  void* cur = lam_a;
  void* nxt = lam_b;
  const hipStream_t s = (hipStream_t)stream;
  // fused path: workspace = d_own | d_nb[0] | d_nb[1] (each Q N compute values) | nb_flag (N bytes)
  const size_t qn = (size_t)lattice_q(lattice) * (size_t)nx * ny * nz;
  const size_t cb = compute == XLBH_COMPUTE_F64 ? 8 : 4;
  char* ws = static_cast<char*>(workspace);
  uint8_t* flag = ws ? reinterpret_cast<uint8_t*>(ws + 3 * qn * cb) : nullptr;
  auto fused_params = [&](const void* f, const void* lo, void* li, int t) {
    AdjParams p = adjoint_params(collision, collision_param, f, lo, li, bc_mask, missing_bits, aux, bc_table, bc_cells,
                                 n_bc_cells, force_mode, force_uniform, force_field, grad_force, nx, ny, nz, P, omega);
    p.d_own = ws;
    p.d_nb_out = ws + (1 + (t & 1)) * qn * cb;
    p.d_nb_in = ws + (1 + ((t + 1) & 1)) * qn * cb;
    p.nb_flag = flag;
    return p;
  };
  if (ws && (flags & XLBH_ADJ_BEGIN)) {
    const AdjParams p = fused_params(f_ckpt, lam_a, lam_b, timestep0);
    const uint64_t words = (3 * qn * cb + (size_t)nx * ny * nz + 3) / 4;
    if (int e = adjoint_dispatch(lattice, store, compute, p, s, ADJ_BEGIN, words, flag)) return e;
  }
  for (int j = n_steps - 1; j >= 0; --j) {
    if (ws) {
      const AdjParams p = fused_params(state(j), cur, nxt, timestep0 + j);
      if (int e = adjoint_dispatch(lattice, store, compute, p, s, ADJ_FUSED)) return e;
    } else {
      const AdjParams p = adjoint_params(collision, collision_param, state(j), cur, nxt, bc_mask, missing_bits, aux,
                                         bc_table, bc_cells, n_bc_cells, force_mode, force_uniform, force_field,
                                         grad_force, nx, ny, nz, P, omega);
      if (int e = adjoint_dispatch(lattice, store, compute, p, s)) return e;
    }
    void* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (ws && (flags & XLBH_ADJ_END)) {
    const AdjParams p = fused_params(f_ckpt, nullptr, nullptr, timestep0);
    if (int e = adjoint_dispatch(lattice, store, compute, p, s, ADJ_FLUSH, 0, nullptr, cur)) return e;
  }
  return 0;
}

extern "C" int64_t xlbh_adjoint_workspace_bytes(int lattice, int compute, int nx, int ny, int nz) {
  if (!valid_lattice(lattice) || nx <= 0 || ny <= 0 || nz <= 0) return -1;
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t cb = compute == XLBH_COMPUTE_F64 ? 8 : 4;
  return ((3 * (int64_t)lattice_q(lattice) * n * cb + n + 255) / 256) * 256;
}

extern "C" int xlbh_adjoint_step(int lattice, int collision, int store, int compute, const void* f_in,
                                 const void* lam_out, void* lam_in, const uint8_t* bc_mask,
                                 const uint32_t* missing_bits, const void* aux, const xlbh_bc_desc* bc_table,
                                 const int32_t* bc_cells, int n_bc_cells, int force_mode,
                                 const double* force_uniform, const void* force_field, double* grad_force, int nx,
                                 int ny, int nz, int64_t plane, double omega, double collision_param, void* stream) {
  clear_error();
  if (int e = adjoint_validate(lattice, collision, force_mode, force_uniform, force_field, bc_mask, missing_bits,
                               bc_table, nx, ny, nz))
    return e;
  XLBH_REQUIRE(store != XLBH_I16_BLOCK, "the adjoint does not read the block format");
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  XLBH_REQUIRE(f_in && lam_out && lam_in, "adjoint_step needs f_in, lam_out and lam_in");
  XLBH_REQUIRE(lam_in != lam_out, "lam_in must not alias lam_out");
  XLBH_REQUIRE(n_bc_cells >= 0 && (n_bc_cells == 0 || bc_cells), "bc_cells is null with n_bc_cells > 0");
  if (int e = refuse_solid_cells(bc_mask, nx, ny, nz, (hipStream_t)stream)) return e;
  const AdjParams p = adjoint_params(collision, collision_param, f_in, lam_out, lam_in, bc_mask, missing_bits, aux,
                                     bc_table, bc_cells, n_bc_cells, force_mode, force_uniform, force_field,
                                     grad_force, nx, ny, nz, P, omega);
  return adjoint_dispatch(lattice, store, compute, p, (hipStream_t)stream);
}
