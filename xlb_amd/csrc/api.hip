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
  XLBH_REQUIRE(compute == XLBH_COMPUTE_F32 || compute == XLBH_COMPUTE_F64, "unknown compute type");
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

int xlbh_abi_version(void) { return 3; }

int xlbh_step(int lattice, int collision, int store, int compute, const void* f_in, void* f_out,
              const uint8_t* bc_mask, const uint32_t* missing_bits, const void* aux, const xlbh_bc_desc* bc_table,
              int force_mode, const double* force_uniform, const void* force_field, int nx, int ny, int nz,
              double omega, double collision_param, int timestep, void* stream) {
  (void)timestep;
  clear_error();
  if (int e = validate_step(lattice, collision, store, compute, f_in, f_out, bc_mask, missing_bits, bc_table,
                            force_mode, force_uniform, force_field, nx, ny, nz))
    return e;
  StepCall c{collision, store, compute, f_in, f_out, bc_mask, missing_bits, aux, bc_table,
             force_mode, force_uniform, force_field, nx, ny, nz, omega, collision_param, (hipStream_t)stream};
  return dispatch_step(lattice, c);
}

int xlbh_run(int lattice, int collision, int store, int compute, void* f_a, void* f_b, const uint8_t* bc_mask,
             const uint32_t* missing_bits, const void* aux, const xlbh_bc_desc* bc_table, int force_mode,
             const double* force_uniform, const void* force_field, int nx, int ny, int nz, double omega,
             double collision_param, int timestep0, int n_steps, void* stream) {
  (void)timestep0;
  clear_error();
  if (int e = validate_step(lattice, collision, store, compute, f_a, f_b, bc_mask, missing_bits, bc_table,
                            force_mode, force_uniform, force_field, nx, ny, nz))
    return e;
  XLBH_REQUIRE(n_steps >= 0, "n_steps must be >= 0");
  StepCall c{collision, store, compute, f_a, f_b, bc_mask, missing_bits, aux, bc_table,
             force_mode, force_uniform, force_field, nx, ny, nz, omega, collision_param, (hipStream_t)stream};
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
                double omega, double collision_param, int timestep0, int n_states, void* stream) {
  clear_error();
  XLBH_REQUIRE(n_states >= 0, "n_states must be >= 0");
  if (n_states == 0) return 0;
  XLBH_REQUIRE(f_start && states, "record needs f_start and states");
  XLBH_REQUIRE(valid_store(store), "unknown store format");
  if (int e = validate_step(lattice, collision, store, compute, f_start, states, bc_mask, missing_bits, bc_table,
                            force_mode, force_uniform, force_field, nx, ny, nz))
    return e;
  const size_t field_bytes = (size_t)lattice_q(lattice) * (size_t)nx * ny * nz * store_bytes(store);
  StepCall c{collision, store, compute, f_start, states, bc_mask, missing_bits, aux, bc_table,
             force_mode, force_uniform, force_field, nx, ny, nz, omega, collision_param, (hipStream_t)stream};
  for (int j = 0; j < n_states; ++j) {
    c.f_in = j == 0 ? f_start : static_cast<const char*>(states) + (size_t)(j - 1) * field_bytes;
    c.f_out = static_cast<char*>(states) + (size_t)j * field_bytes;
    if (int e = dispatch_step(lattice, c)) return e;
  }
  return 0;
}

}  // extern "C"
