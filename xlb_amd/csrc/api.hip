// C entry points of the fused step, the error channel and grid validation.
#include "step.hpp"

namespace xlbh {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int check_grid(int lattice, int nx, int ny, int nz) {
  XLBH_REQUIRE(valid_lattice(lattice), "unknown lattice id");
  XLBH_REQUIRE(nx > 0 && ny > 0 && nz > 0, "grid extents must be positive");
  XLBH_REQUIRE(lattice != XLBH_D2Q9 || nz == 1, "2-D lattices use fields of shape (q, nx, ny, 1)");
  XLBH_REQUIRE(lattice == XLBH_D2Q9 || nz > 1, "3-D lattices need nz > 1");
  // 1-D launches round the cell count up to whole 256-thread blocks in 32-bit arithmetic.
  XLBH_REQUIRE((int64_t)nx * ny * nz <= ((int64_t)1 << 32) - 256, "more than 2^32 - 256 cells per field");
  XLBH_REQUIRE(nx <= 65535, "nx must be <= 65535");
  // the sweep's grid: (z tiles, y tiles, x) in 3-D, (y tiles, x tiles, 1) in 2-D
  const dim3 g = sweep_grid(make_sweep(nx, ny, nz));
  XLBH_REQUIRE(g.y <= 65535, "ny too large for the sweep's launch grid (more than 65535 tiles of y rows)");
  return 0;
}

int resolve_plane(int lattice, int store, int nx, int ny, int nz, int64_t plane, uint64_t& out) {
  const uint64_t N = (uint64_t)nx * ny * nz;
  XLBH_REQUIRE(plane == 0 || (plane > 0 && (uint64_t)plane >= N), "plane stride must be 0 (dense) or >= nx*ny*nz");
  out = plane == 0 ? N : (uint64_t)plane;
  // every plane offset l * plane * S and the block exponents stay below 2^63
  XLBH_REQUIRE(out <= ((uint64_t)1 << 40), "plane stride too large");
  (void)lattice;
  (void)store;
  return 0;
}

// Pad after each plane of the recommended plane stride, in bytes.  Measured on the 256^3 D3Q19
// cavity (MI355X, best of 3 x 100 steps): with 2-byte populations a 256-byte pad shifts each of
// the Q plane starts off the power-of-two alignment every plane otherwise shares, +4.8 % for
// shifted FP16 (71.9 -> 75.4 k MLUPS) and +2.4 % for the block format; pads of 1 KB to 1 MB gave
// +2 to +4.7 %.  With 4-byte populations no pad from 256 B to 1 MB helped (0 to -4 %).
static uint64_t plane_pad_bytes(int store) { return store_bytes(store) == 2 ? 256 : 0; }

// Planes a little longer than a power of two stream badly: an x-slab of a 256^3 share with its
// two ghost planes has FP32 planes of 64 MiB + 512 KiB, and its D3Q19 step took 516-916 us
// against 430 us for 256^3 (shifted FP16: 32 MiB + 256 KiB, 262-273 us against 225 us; FP64 at
// 128^2: 16 MiB + 256 KiB, 120 us against 113 us).  Padded to 2^k + 1 MiB they step at the
// whole-domain rate (435, 223 and 113 us).  Sweeping the pad after a 64 MiB plane put the slow
// band at about 320-640 KiB.  So a plane of more than 16 MiB that exceeds a power of two 2^k by
// less than 1 MiB is padded to 2^k + 1 MiB; other sizes keep the 2-byte pad above.
static int64_t recommended_plane(int store, int nx, int ny, int nz) {
  const uint64_t N = (uint64_t)nx * ny * nz;
  const uint64_t S = store_bytes(store), dense = N * S, MiB = 1ull << 20;
  if (dense > 16 * MiB) {
    const uint64_t p2 = 1ull << (63 - __builtin_clzll(dense));
    if (dense > p2 && dense - p2 < MiB) return (int64_t)((p2 + MiB) / S);
  }
  return (int64_t)(N + plane_pad_bytes(store) / S);
}

static int dispatch_step(int lattice, const StepCall& c) {
  switch (lattice) {
    case XLBH_D2Q9: return launch_step_d2q9(c);
    case XLBH_D3Q19: return launch_step_d3q19(c);
    case XLBH_D3Q27: return launch_step_d3q27(c);
  }
  set_error("unknown lattice id");
  return 1;
}

static int validate_step(int lattice, int collision, int store, int compute, const void* f_a, const void* f_b,
                         const uint8_t* bc_mask, const uint32_t* missing_bits, const xlbh_bc_desc* bc_table,
                         int force_mode, const double* force_uniform, const void* force_field, int nx, int ny,
                         int nz) {
  if (int e = check_grid(lattice, nx, ny, nz)) return e;
  XLBH_REQUIRE(collision == XLBH_BGK || collision == XLBH_KBC || collision == XLBH_SMAGORINSKY_BGK,
               "unknown collision id");
  XLBH_REQUIRE(valid_store(store), "unknown store format");
  XLBH_REQUIRE(compute == XLBH_COMPUTE_F32 || compute == XLBH_COMPUTE_F64 || compute == XLBH_COMPUTE_F32_DEVIATION,
               "unknown compute type");
  XLBH_REQUIRE(f_a && f_b, "population fields must be non-null");
  XLBH_REQUIRE(f_a != f_b, "the step cannot run in place: f_out must not alias f_in");
  XLBH_REQUIRE(!bc_mask || (missing_bits && bc_table), "bc_mask given without missing bits or BC table");
  XLBH_REQUIRE(force_mode >= XLBH_FORCE_NONE && force_mode <= XLBH_FORCE_FIELD, "unknown force mode");
  XLBH_REQUIRE(force_mode != XLBH_FORCE_UNIFORM || force_uniform, "uniform force needs 3 host doubles");
  XLBH_REQUIRE(force_mode != XLBH_FORCE_FIELD || force_field, "force field pointer is null");
  return 0;
}

}  // namespace xlbh

using namespace xlbh;

extern "C" {

const char* xlbh_last_error(void) { return g_last_error.c_str(); }

int xlbh_abi_version(void) { return 4; }

int64_t xlbh_plane_stride(int lattice, int store, int nx, int ny, int nz) {
  clear_error();
  if (check_grid(lattice, nx, ny, nz) || !valid_store(store)) return -1;
  return recommended_plane(store, nx, ny, nz);
}

int64_t xlbh_population_bytes(int lattice, int store, int nx, int ny, int nz, int64_t plane) {
  clear_error();
  if (check_grid(lattice, nx, ny, nz)) return -1;
  if (!valid_store(store)) {
    set_error("unknown store format");
    return -1;
  }
  uint64_t P;
  if (resolve_plane(lattice, store, nx, ny, nz, plane, P)) return -1;
  const int q = lattice_q(lattice);
  if (store == XLBH_I16_BLOCK)
    return (int64_t)(blk_exps_offset(q, P) + blk_seg_count(make_sweep(nx, ny, nz)) * sizeof(int32_t));
  return (int64_t)((uint64_t)q * P * store_bytes(store));
}

int xlbh_step(int lattice, int collision, int store, int compute, const void* f_in, void* f_out,
              const uint8_t* bc_mask, const uint32_t* missing_bits, const void* aux, const xlbh_bc_desc* bc_table,
              int force_mode, const double* force_uniform, const void* force_field, int nx, int ny, int nz,
              int64_t plane, double omega, double collision_param, int timestep, void* stream) {
  (void)timestep;
  clear_error();
  if (int e = validate_step(lattice, collision, store, compute, f_in, f_out, bc_mask, missing_bits, bc_table,
                            force_mode, force_uniform, force_field, nx, ny, nz))
    return e;
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  StepCall c{collision, store, compute, f_in, f_out, bc_mask, missing_bits, aux, bc_table,
             force_mode, force_uniform, force_field, nx, ny, nz, P, omega, collision_param, (hipStream_t)stream};
  return dispatch_step(lattice, c);
}

int xlbh_step_slab(int lattice, int collision, int store, int compute, const void* f_in, void* f_out,
                   const uint8_t* bc_mask, const uint32_t* missing_bits, const void* aux,
                   const xlbh_bc_desc* bc_table, int force_mode, const double* force_uniform,
                   const void* force_field, int nx, int ny, int nz, int64_t plane, double omega,
                   double collision_param, int timestep, int x_begin, int x_end, void* stream) {
  (void)timestep;
  clear_error();
  if (int e = validate_step(lattice, collision, store, compute, f_in, f_out, bc_mask, missing_bits, bc_table,
                            force_mode, force_uniform, force_field, nx, ny, nz))
    return e;
  XLBH_REQUIRE(x_begin >= 1 && x_begin < x_end && x_end <= nx - 1,
               "step_slab: planes x_begin..x_end-1 need 1 <= x_begin < x_end <= nx - 1 (a ghost plane on each side)");
  XLBH_REQUIRE(store != XLBH_I16_BLOCK,
               "step_slab: the block format is not supported (its row exponents would need an exchange of their own)");
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  StepCall c{collision, store, compute, f_in, f_out, bc_mask, missing_bits, aux, bc_table,
             force_mode, force_uniform, force_field, nx, ny, nz, P, omega, collision_param, (hipStream_t)stream};
  c.x0 = x_begin;
  c.x1 = x_end;
  return dispatch_step(lattice, c);
}

int xlbh_run(int lattice, int collision, int store, int compute, void* f_a, void* f_b, const uint8_t* bc_mask,
             const uint32_t* missing_bits, const void* aux, const xlbh_bc_desc* bc_table, int force_mode,
             const double* force_uniform, const void* force_field, int nx, int ny, int nz, int64_t plane,
             double omega, double collision_param, int timestep0, int n_steps, void* stream) {
  (void)timestep0;
  clear_error();
  if (int e = validate_step(lattice, collision, store, compute, f_a, f_b, bc_mask, missing_bits, bc_table,
                            force_mode, force_uniform, force_field, nx, ny, nz))
    return e;
  XLBH_REQUIRE(n_steps >= 0, "n_steps must be >= 0");
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  StepCall c{collision, store, compute, f_a, f_b, bc_mask, missing_bits, aux, bc_table,
             force_mode, force_uniform, force_field, nx, ny, nz, P, omega, collision_param, (hipStream_t)stream};
  for (int t = 0; t < n_steps; ++t) {
    c.f_in = (t % 2 == 0) ? f_a : f_b;
    c.f_out = (t % 2 == 0) ? f_b : f_a;
    if (int e = dispatch_step(lattice, c)) return e;
  }
  return 0;
}

int xlbh_record(int lattice, int collision, int store, int compute, const void* f_start, void* states,
                const uint8_t* bc_mask, const uint32_t* missing_bits, const void* aux, const xlbh_bc_desc* bc_table,
                int force_mode, const double* force_uniform, const void* force_field, int nx, int ny, int nz,
                int64_t plane, double omega, double collision_param, int timestep0, int n_states, void* stream) {
  clear_error();
  XLBH_REQUIRE(n_states >= 0, "n_states must be >= 0");
  if (n_states == 0) return 0;
  XLBH_REQUIRE(f_start && states, "record needs f_start and states");
  XLBH_REQUIRE(valid_store(store), "unknown store format");
  if (int e = validate_step(lattice, collision, store, compute, f_start, states, bc_mask, missing_bits, bc_table,
                            force_mode, force_uniform, force_field, nx, ny, nz))
    return e;
  uint64_t P;
  if (int e = resolve_plane(lattice, store, nx, ny, nz, plane, P)) return e;
  const size_t field_bytes = (size_t)xlbh_population_bytes(lattice, store, nx, ny, nz, (int64_t)P);
  StepCall c{collision, store, compute, f_start, states, bc_mask, missing_bits, aux, bc_table,
             force_mode, force_uniform, force_field, nx, ny, nz, P, omega, collision_param, (hipStream_t)stream};
  for (int j = 0; j < n_states; ++j) {
    c.f_in = j == 0 ? f_start : static_cast<const char*>(states) + (size_t)(j - 1) * field_bytes;
    c.f_out = static_cast<char*>(states) + (size_t)j * field_bytes;
    if (int e = dispatch_step(lattice, c)) return e;
  }
  return 0;
}

}  // extern "C"
