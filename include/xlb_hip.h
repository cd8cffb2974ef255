/*
 * xlb_hip.h -- C ABI of the MI355X (gfx950) lattice-Boltzmann backend, libxlbhip.so.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and returns an int
 * status: 0 = success, non-zero = failure with a message available from xlbh_last_error()
 * (thread-local).  All field pointers are device pointers (HBM) unless stated otherwise.
 * `stream` is a hipStream_t passed as void* (NULL = the legacy default stream); every
 * launch is asynchronous on that stream, exactly as the reference's `wp.launch` is
 * asynchronous on Warp's current stream.
 *
 * Field layout (identical to the reference's WarpGrid fields, xlb/grid/warp_grid.py:17-32):
 * structure of arrays, C order, shape (card, nx, ny, nz) with z unit-stride; 2-D lattices use
 * (card, nx, ny, 1).  `N = nx*ny*nz` is the number of cells.  Population l of cell
 * (x,y,z) lives at element  l*plane + (x*ny + y)*nz + z, where `plane` >= N is the field's plane
 * stride in elements (0 = N, the dense layout).  Entry points that take population fields take
 * one `plane` for all of them; other fields (rho, u, Pi, adjoint fields, masks) are dense.  A
 * padded stride (xlbh_plane_stride) staggers the planes across HBM channels; the torch view
 * (Q, nx, ny, nz) with strides (plane, ny*nz, nz, 1) keeps the reference's indexing.
 *
 * Grid limits: 1 <= nx <= 65535, N <= 2^32 - 256, and at most 65535 launch tiles of y rows
 * (ny / 4, ny / 2 or ny rows per tile for nz <= 64, <= 128 or more); entry points that build cell lists
 * (xlbh_compact_cells and the adjoint's boundary pass) need N < 2^31.  Planes above 4 GiB are
 * addressed with 64-bit offsets.
 *
 * Lattice direction order, weights and opposite indices follow
 * xlb/velocity_set/d2q9.py:16-24, d3q19.py:17-30, d3q27.py:17-32.
 */
#ifndef XLB_HIP_H
#define XLB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Lattices (xlb/velocity_set/d2q9.py, d3q19.py, d3q27.py). */
enum { XLBH_D2Q9 = 0, XLBH_D3Q19 = 1, XLBH_D3Q27 = 2 };

/* Collision operators (xlb/operator/collision/bgk.py, kbc.py). KBC: D2Q9 and D3Q27 only,
 * as in kbc.py:183-185. */
enum { XLBH_BGK = 0, XLBH_KBC = 1, XLBH_SMAGORINSKY_BGK = 2 };

/* Storage formats of population fields.  The *_SHIFTED formats store the deviation
 * f_l - w_l (no counterpart in the reference; its FP16 store is a plain cast,
 * xlb/precision_policy.py:46-89). */
enum {
  XLBH_F32 = 0,
  XLBH_F64 = 1,
  XLBH_F16 = 2,
  XLBH_BF16 = 3,
  XLBH_F16_SHIFTED = 4,
  XLBH_BF16_SHIFTED = 5,
  XLBH_I16_BLOCK = 6
};
/* XLBH_I16_BLOCK (no counterpart in the reference): 16-bit block floating point.  Element l of a
 * cell holds an int16 mantissa m of x = f_l / w_l - 1; one int32 exponent k per row segment gives
 * f_l = w_l + m * (w_l * 2^-k) (rounded once to float).  A row segment is the run of bz cells
 * along the fast axis (z in 3-D, y in 2-D) that one 256-thread sweep tile covers, bz = the
 * smallest of 64, 128, 256 that is >= the fast extent (else 256).  The segments are numbered
 * ((x * ny + y) * S + z / bz) in 3-D and (x * S + y / bz) in 2-D, S = ceil(fast extent / bz).
 * The exponents (int32, one per segment) follow the last plane of the field, at byte offset
 * ceil(Q * plane * 2 / 256) * 256 from its base; xlbh_population_bytes() gives the whole size.
 * Each writing step chooses k per segment from the largest |x| it writes there (every
 * direction, cells with id 255 excluded), so |m| <= 32767: the relative precision is 2^-15 of
 * the segment's largest deviation instead of 2^-11 of each value (FP16). */

/* Arithmetic precision of the per-cell update (PrecisionPolicy.compute_precision).
 * XLBH_COMPUTE_F32_DEVIATION: the fused step's BGK collision in FP32 on deviations from the rest
 * state (g_l = f_l - w_l, rho - 1), with the cell's moment sums accumulated exactly (in FP64:
 * each g_l is an FP32 value, so the 19- or 27-term sums are exact), the weights applied as exact
 * k_l / D float pairs and omega multiplying only geq - g; FP32 storage, BGK only.  No rounding
 * of it is systematic: on a 1000-step 64^3 cavity the velocity error against FP64 drops from
 * 1.1e-4 to 9e-7 (relative L2).  Every other entry point treats it as XLBH_COMPUTE_F32; the
 * reverse kernels differentiate the same update in FP32. */
enum { XLBH_COMPUTE_F32 = 0, XLBH_COMPUTE_F64 = 1, XLBH_COMPUTE_F32_DEVIATION = 2 };

/* Element types of macroscopic fields (rho, u, Pi) passed to the standalone operators. */
enum { XLBH_FIELD_F32 = 0, XLBH_FIELD_F64 = 1, XLBH_FIELD_F16 = 2, XLBH_FIELD_BF16 = 3 };

/* Boundary-condition kinds; one kind per boundary id (xlb/operator/boundary_condition/). */
enum {
  XLBH_BC_NONE = 0,
  XLBH_BC_EQUILIBRIUM = 1,            /* bc_equilibrium.py:75-86            (streaming step) */
  XLBH_BC_HALFWAY_BB = 2,             /* bc_halfway_bounce_back.py:67-85    (streaming step) */
  XLBH_BC_FULLWAY_BB = 3,             /* bc_fullway_bounce_back.py:59-72    (collision step) */
  XLBH_BC_DO_NOTHING = 4,             /* bc_do_nothing.py:52-62             (streaming step) */
  XLBH_BC_ZOUHE_VELOCITY = 5,         /* bc_zouhe.py:279-312                (streaming step) */
  XLBH_BC_ZOUHE_PRESSURE = 6,         /* bc_zouhe.py:314-338                (streaming step) */
  XLBH_BC_REGULARIZED_VELOCITY = 7,   /* bc_regularized.py:134-168          (streaming step) */
  XLBH_BC_REGULARIZED_PRESSURE = 8,   /* bc_regularized.py:170-196          (streaming step) */
  XLBH_BC_EXTRAPOLATION_OUTFLOW = 9,  /* bc_extrapolation_outflow.py:152-195 (streaming + post-collision) */
  XLBH_BC_GRADS_APPROXIMATION = 10    /* bc_grads_approximation.py:181-210  (streaming step, 3-D) */
};

/* Body-force modes of the exact-difference forcing (xlb/operator/force/exact_difference_force.py). */
enum { XLBH_FORCE_NONE = 0, XLBH_FORCE_UNIFORM = 1, XLBH_FORCE_FIELD = 2 };

/* One entry of the boundary table: a device array of 256 entries indexed by boundary id
 * (bc_mask value).  Id 0 = fluid, 255 = solid (cell skipped, nse_stepper.py:356-358).
 * `rho`, `u` are the EquilibriumBC constants; other kinds ignore them. */
typedef struct xlbh_bc_desc {
  int32_t kind;
  int32_t reserved;
  double rho;
  double u[3];
} xlbh_bc_desc;

/* ---------------------------------------------------------------------------------------
 * Library
 * ------------------------------------------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char* xlbh_last_error(void);

/* ABI version of this header (incremented on incompatible change). */
int xlbh_abi_version(void);

/* Plane stride (elements) this library recommends for population fields of the given grid and
 * store format (-1 for invalid arguments): N plus a pad that staggers the Q plane starts across
 * HBM channels (256 bytes for 2-byte populations), except that a plane of more than 16 MiB
 * which exceeds a power of two 2^k by less than 1 MiB (an x-slab of a power-of-two share with
 * its ghost planes) is padded to 2^k + 1 MiB.  No counterpart in the reference (WarpGrid fields
 * are dense). */
int64_t xlbh_plane_stride(int lattice, int store, int nx, int ny, int nz);

/* Bytes of one population field with plane stride `plane` (0 = N): Q * plane * S, and for
 * XLBH_I16_BLOCK the segment exponents after it (-1 for invalid arguments). */
int64_t xlbh_population_bytes(int lattice, int store, int nx, int ny, int nz, int64_t plane);

/* ---------------------------------------------------------------------------------------
 * The fused step.  Replaces IncompressibleNavierStokesStepper.warp_implementation and its
 * kernel (xlb/operator/stepper/nse_stepper.py:245-392): pull streaming with periodic wrap,
 * streaming-step BCs, macroscopic, quadratic equilibrium, BGK/KBC collision (+ exact-
 * difference force), collision-step BCs, extrapolation-outflow auxiliary update, store.
 *
 *   f_in, f_out   : (Q, N) population fields in format `store`; f_out must not alias f_in.
 *   bc_mask       : (N) uint8 boundary ids, or NULL when every cell is fluid.
 *   missing_bits  : (N) uint32, bit l set <=> direction l is missing at the cell
 *                   (packed form of the reference's (Q, N) bool missing_mask).  Read only
 *                   at cells with a non-zero id.  May be NULL when bc_mask is NULL.
 *   aux           : (N) prescribed values (velocity magnitude or density) for Zou-He and
 *                   regularized cells, in the compute type (float or double); NULL if unused.
 *                   Replaces the reference's stash in f_1[0] (boundary_condition.py:151-159,
 *                   nse_stepper.py:318-342): the input field is never written.
 *   bc_table      : device array of 256 xlbh_bc_desc, or NULL when bc_mask is NULL.
 *   force_mode    : XLBH_FORCE_*; `force_uniform` (host, 3 doubles) for UNIFORM, `force_field`
 *                   (device, (D, N) in the compute type) for FIELD.
 *   omega         : relaxation rate.  timestep is accepted for interface parity (unused by
 *                   the in-scope boundary conditions, as in the reference).
 *   collision_param: the Smagorinsky constant for XLBH_SMAGORINSKY_BGK
 *                   (smagorinsky_les_bgk.py:19-26, default 0.17 there); ignored otherwise.
 * ------------------------------------------------------------------------------------- */
int xlbh_step(int lattice, int collision, int store, int compute,
              const void* f_in, void* f_out,
              const uint8_t* bc_mask, const uint32_t* missing_bits,
              const void* aux, const xlbh_bc_desc* bc_table,
              int force_mode, const double* force_uniform, const void* force_field,
              int nx, int ny, int nz, int64_t plane, double omega, double collision_param, int timestep, void* stream);

/* One fused step of an x-slab: xlbh_step restricted to planes x_begin <= x < x_end, with
 * 1 <= x_begin < x_end <= nx - 1, so that no pull wraps in x.  Planes x_begin - 1 and x_end are
 * read as the x-neighbours (ghost planes, which the caller fills with the adjacent slab's
 * populations before the call) and every plane outside the range is left unwritten.  All other
 * arguments, and the per-cell arrays (bc_mask, missing_bits, aux, force_field), are those of the
 * slab's own (nx, ny, nz) field, ghost planes included.  The per-shard operator of the
 * reference's x-sharded run (distribute.py:18-47, where each shard steps its block and the
 * populations that crossed a shard face are exchanged afterwards with lax.ppermute); the pull
 * scheme needs the crossing populations before the step, hence the ghost planes.  Calling it
 * for the two edge planes and then for the interior lets the exchange of the edge planes overlap
 * the interior's update.  The block format is refused. */
int xlbh_step_slab(int lattice, int collision, int store, int compute,
                   const void* f_in, void* f_out,
                   const uint8_t* bc_mask, const uint32_t* missing_bits,
                   const void* aux, const xlbh_bc_desc* bc_table,
                   int force_mode, const double* force_uniform, const void* force_field,
                   int nx, int ny, int nz, int64_t plane, double omega, double collision_param, int timestep,
                   int x_begin, int x_end, void* stream);

/* n_steps fused steps with the buffer swap done in C (the caller's loop of
 * mlups_3d.py:78-80 moved below the ABI).  Step t reads f_a when t is even and f_b when t is
 * odd; after the call the newest populations are in f_b if n_steps is odd, else in f_a. */
int xlbh_run(int lattice, int collision, int store, int compute,
             void* f_a, void* f_b,
             const uint8_t* bc_mask, const uint32_t* missing_bits,
             const void* aux, const xlbh_bc_desc* bc_table,
             int force_mode, const double* force_uniform, const void* force_field,
             int nx, int ny, int nz, int64_t plane, double omega, double collision_param, int timestep0, int n_steps,
             void* stream);

/* n_states steps from f_start, keeping every state: state j (the output of step t0 + j) is
 * written to states + j * xlbh_population_bytes(..., plane) bytes, for j = 0 .. n_states - 1.  The forward pass of a
 * checkpointed reverse sweep (xlbh_adjoint_run without XLBH_ADJ_RECOMPUTE); no counterpart in the
 * reference, whose tape keeps the segment's states itself. */
int xlbh_record(int lattice, int collision, int store, int compute, const void* f_start, void* states,
                const uint8_t* bc_mask, const uint32_t* missing_bits, const void* aux,
                const xlbh_bc_desc* bc_table, int force_mode, const double* force_uniform,
                const void* force_field, int nx, int ny, int nz, int64_t plane, double omega, double collision_param,
                int timestep0,
                int n_states, void* stream);

/* ---------------------------------------------------------------------------------------
 * Standalone operators (the reference's per-operator Warp kernels).
 * ------------------------------------------------------------------------------------- */

/* Pull streaming with periodic wrap: f_out[l](x) = f_in[l](x - c_l).  stream.py:59-99. */
int xlbh_stream(int lattice, int store, const void* f_in, void* f_out,
                int nx, int ny, int nz, int64_t plane, void* stream);

/* f = feq(rho, u).  rho: (1, N), u: (D, N) in field type `field`.  quadratic_equilibrium.py:35-98. */
int xlbh_equilibrium(int lattice, int store, int compute, int field,
                     const void* rho, const void* u, void* f,
                     int nx, int ny, int nz, int64_t plane, void* stream);

/* Moments of f.  which: 0 = rho and u (macroscopic.py:28-66), 1 = rho only
 * (zero_moment.py:22-47), 2 = u given rho (first_moment.py:25-66, rho is an input),
 * 3 = second moment Pi (second_moment.py:68-101, written to `pi` with D(D+1)/2 rows). */
int xlbh_moments(int lattice, int store, int compute, int field, int which,
                 const void* f, void* rho, void* u, void* pi,
                 int nx, int ny, int nz, int64_t plane, void* stream);

/* fout = collision(f, feq, rho, u, omega) (+ exact-difference force).  bgk.py:30-81,
 * kbc.py:268-347, forced_collision.py:41-104.  rho, u: (1,N), (D,N) in type `field`. */
int xlbh_collide(int lattice, int collision, int store, int compute, int field,
                 const void* f, const void* feq, void* fout,
                 const void* rho, const void* u, double omega, double collision_param,
                 int force_mode, const double* force_uniform, const void* force_field,
                 int nx, int ny, int nz, int64_t plane, void* stream);

/* Standalone boundary-condition application (BoundaryCondition._construct_kernel,
 * boundary_condition.py:86-117): at cells with bc_mask == bc_id, f_post = functional(
 * f_pre, f_post) of the streaming-step form of `kind`; elsewhere f_post is unchanged.
 * `desc` is a HOST pointer to the BC constants. */
int xlbh_apply_bc(int lattice, int store, int compute, int kind, int bc_id,
                  const xlbh_bc_desc* desc,
                  const void* f_pre, void* f_post,
                  const uint8_t* bc_mask, const uint32_t* missing_bits, const void* aux,
                  int nx, int ny, int nz, int64_t plane, void* stream);

/* ---------------------------------------------------------------------------------------
 * Boundary masks (integer; bit-exact).
 * ------------------------------------------------------------------------------------- */

/* One boundary condition's indices -> bc_mask / missing mask, Warp semantics of
 * IndicesBoundaryMasker (indices_boundary_masker.py:103-224): for each in-domain index p,
 * bc_mask[p] = bc_id; for each l with p - c_l outside the domain, missing[l,p] = 1; if
 * interior[i] (needs padding and p strictly inside the domain) then for each l with
 * p - c_l inside, missing[l, p + c_l] = 1 and bc_mask[p + c_l] = bc_id.
 *   indices : device int32 (3, n) (rows x, y, z; z row zeros for 2-D lattices)
 *   interior: device uint8 (n)
 *   missing_bool (Q, N) uint8 and/or missing_bits (N) uint32 are updated (either may be NULL).
 * Calls for successive BCs on one stream apply in order, so a later BC's ids win. */
int xlbh_mask_indices(int lattice, const int32_t* indices, const uint8_t* interior, int n,
                      int bc_id, uint8_t* bc_mask, uint8_t* missing_bool, uint32_t* missing_bits,
                      int nx, int ny, int nz, void* stream);

/* (Q, N) bool missing mask -> (N) uint32 bits. */
int xlbh_pack_missing(int lattice, const uint8_t* missing_bool, uint32_t* missing_bits,
                      int nx, int ny, int nz, void* stream);

/* Deterministic, ordered list of the cells whose bc_mask equals `bc_id` (bc_id < 0: every
 * cell with 0 < id < 255).  `cells` must hold the worst case N entries; `count` is a device
 * int32.  `workspace` must hold xlbh_compact_workspace_bytes(N) bytes. */
int64_t xlbh_compact_workspace_bytes(int64_t n_cells);
int xlbh_compact_cells(const uint8_t* bc_mask, int bc_id, int32_t* cells, int32_t* count,
                       void* workspace, int nx, int ny, int nz, void* stream);

/* aux[cells[i]] = values[i] for i < n (values in the compute type; `compute` selects it). */
int xlbh_scatter_aux(int compute, const int32_t* cells, const void* values, int n, void* aux,
                     void* stream);

/* ---------------------------------------------------------------------------------------
 * Reductions (FP64 accumulation, deterministic: fixed-order two-pass).
 * ------------------------------------------------------------------------------------- */

/* Momentum-exchange force on boundary id `bc_id` whose streaming-step functional is `kind`
 * (momentum_transfer.py:92-176).  f is the post-collision field (the stepper's newest
 * output).  Result: out[0..D-1] (device doubles).  `workspace` must hold
 * xlbh_reduce_workspace_bytes(N) bytes. */
int64_t xlbh_reduce_workspace_bytes(int64_t n_cells);
int xlbh_momentum_transfer(int lattice, int store, int compute, int kind, int bc_id,
                           const xlbh_bc_desc* desc,
                           const void* f, const uint8_t* bc_mask, const uint32_t* missing_bits,
                           const void* aux, double* out, void* workspace,
                           int nx, int ny, int nz, int64_t plane, void* stream);

/* xlbh_momentum_transfer without a workspace argument, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It calls no allocating routine: the per-block partials go to a
 * 2 MiB array the library holds per device (for the process lifetime), so grids of at most 2^28
 * cells are accepted (larger ones fail with a message).  Calls on one GPU share that array: calls
 * that overlap on one GPU (different streams, or host threads) race and give wrong results; calls
 * on one stream are ordered and safe.  Results are bit-identical to xlbh_momentum_transfer. */
int xlbh_momentum_transfer_noalloc(int lattice, int store, int compute, int kind, int bc_id,
                                   const xlbh_bc_desc* desc,
                                   const void* f, const uint8_t* bc_mask, const uint32_t* missing_bits,
                                   const void* aux, double* out,
                                   int nx, int ny, int nz, int64_t plane, void* stream);

/* Global mass and momentum: out[0] = sum rho, out[1..3] = sum rho*u (= sum c f) over every
 * cell whose bc_mask is not 255 (bc_mask may be NULL). */
int xlbh_mass_momentum(int lattice, int store, const void* f, const uint8_t* bc_mask,
                       double* out, void* workspace, int nx, int ny, int nz, int64_t plane, void* stream);

/* Surface voxelisation of a triangle mesh into the masks, replacing the reference's
 * MeshBoundaryMasker (xlb/operator/boundary_masker/mesh_boundary_masker.py:49-236).
 *   triangles: (T, 3, 3) float32 device array, the vertices of each triangle in lattice units
 *              (voxel (i,j,k) is the unit box [i,i+1] x [j,j+1] x [k,k+1]).
 * A voxel some triangle touches (Schwarz-Seidel triangle/box test) gets bc_mask = 255; any other
 * voxel whose neighbour x + c_l is touched gets bc_mask = bc_id and missing[opp l] = 1.  Other
 * cells are left as they are.  3-D lattices only.  `workspace` holds
 * xlbh_mesh_workspace_bytes(nx, ny, nz) bytes (a solid grid with one halo voxel per side). */
int64_t xlbh_mesh_workspace_bytes(int nx, int ny, int nz);
int xlbh_mesh_mask(int lattice, const float* triangles, int64_t n_triangles, int bc_id, uint8_t* bc_mask,
                   uint8_t* missing_bool, void* workspace, int nx, int ny, int nz, void* stream);

/* Elementwise cast of n values from field type `field_in` to `field_out` (XLBH_FIELD_*), the
 * reference's PrecisionCaster (xlb/operator/precision_caster/precision_caster.py:43-82).  Each
 * value is converted with one rounding, except FP64 -> FP16 / BF16, which rounds to FP32 first. */
int xlbh_cast(int field_in, int field_out, const void* in, void* out, int64_t n, void* stream);

/* Post-processing of a velocity field u (3, nx, ny, nz) of type `field` (F32 or F64): over the
 * interior cells, skipping every cell with a non-zero boundary id among its six face neighbours
 * (bc_mask may be NULL), central differences give
 *   xlbh_vorticity:   vorticity (3, N) = curl u and magnitude (1, N)
 *                     (xlb/operator/postprocess/vorticity.py:30-80);
 *   xlbh_q_criterion: norm_mu (1, N) = |curl u| and q (1, N) = (|Omega|^2 - |S|^2) / 2
 *                     (q_criterion.py:35-132).
 * Cells not computed keep their previous output values. */
int xlbh_vorticity(int field, const void* u, const uint8_t* bc_mask, void* vorticity, void* magnitude,
                   int nx, int ny, int nz, void* stream);
int xlbh_q_criterion(int field, const void* u, const uint8_t* bc_mask, void* norm_mu, void* q,
                     int nx, int ny, int nz, void* stream);

/* ---------------------------------------------------------------------------------------
 * Reverse mode of the fused step (every collision and boundary kind; `collision` and
 * `collision_param` as for xlbh_step, KBC on D3Q19 rejected).  No counterpart function in the
 * reference, whose gradients come from Warp's tape replaying generated adjoint kernels
 * (examples/out_of_core/subroutine/autodiff_stepper_subroutine.py:459-466).
 *
 *   f_in      : the (Q, N) state the forward step read (store format `store`).
 *   lam_out   : dL/df_out, (Q, N) in the compute type.
 *   lam_in    : written with dL/df_in, (Q, N) in the compute type (every element is written).
 *   grad_force: (D, N) FP64; dL/dg of the step is ADDED (FIELD and UNIFORM force modes; for the
 *               uniform vector sum the field).  May be NULL.
 *   bc_cells  : flat indices of every cell with a non-zero id, each once (xlbh_compact_cells with
 *               bc_id = -1), so the boundary pass runs over them only; NULL (n_bc_cells = 0)
 *               makes it scan the grid.  A list that misses boundary cells gives wrong gradients.
 * Other arguments as for xlbh_step.  Masks with solid cells (id 255) are refused: the forward
 * step leaves those cells unwritten, so their populations depend on the caller's buffer schedule,
 * which a per-step adjoint cannot differentiate.  The check reads bc_mask once and synchronises
 * the stream: on every xlbh_adjoint_step call, and on xlbh_adjoint_run calls with XLBH_ADJ_BEGIN or
 * without a workspace.
 * ------------------------------------------------------------------------------------- */
int xlbh_adjoint_step(int lattice, int collision, int store, int compute,
                      const void* f_in, const void* lam_out, void* lam_in,
                      const uint8_t* bc_mask, const uint32_t* missing_bits,
                      const void* aux, const xlbh_bc_desc* bc_table,
                      const int32_t* bc_cells, int n_bc_cells,
                      int force_mode, const double* force_uniform, const void* force_field,
                      double* grad_force, int nx, int ny, int nz, int64_t plane, double omega, double collision_param,
                      void* stream);

/* Reverse sweep over one checkpoint segment of n_steps steps, below the ABI (the counterpart
 * of xlbh_run for the backward pass; the reference replays a Warp tape per segment,
 * autodiff_stepper_subroutine.py:459-466).
 *   f_ckpt : the state f^{t0} at the segment start (store format), read only.
 *   states : workspace of (n_steps - 1) population fields of xlbh_population_bytes(..., plane)
 *            bytes each, in the store format; receives the
 *            recomputed f^{t0+1} .. f^{t0+n_steps-1}.  May be NULL when n_steps == 1.
 *   lam_a  : dL/df^{t0+n_steps} on entry.  lam_b: a second (Q, N) compute-type buffer.
 * The result dL/df^{t0} is left in lam_a when n_steps is even and in lam_b when it is odd; the
 * other buffer is overwritten.  grad_force and bc_cells as in xlbh_adjoint_step.
 * flags: XLBH_ADJ_RECOMPUTE recomputes the states first; without it they are taken as already
 *   recorded (xlbh_record), so the recording of one segment can run on another stream while the
 *   reverse sweep of the next runs on this one.
 * workspace: NULL, or xlbh_adjoint_workspace_bytes() bytes that select the fused reverse step
 *   (one kernel per step; boundary contributions are carried to the next reverse step in the
 *   workspace instead of being added by a second kernel).  A sweep over several segments passes
 *   the same workspace to every call, with XLBH_ADJ_BEGIN on its first call (clears it) and
 *   XLBH_ADJ_END on its last (adds the carried contributions into the result).  Calls that share
 *   one workspace must not run concurrently. */
int xlbh_adjoint_run(int lattice, int collision, int store, int compute,
                     const void* f_ckpt, void* states, void* lam_a, void* lam_b,
                     const uint8_t* bc_mask, const uint32_t* missing_bits,
                     const void* aux, const xlbh_bc_desc* bc_table,
                     const int32_t* bc_cells, int n_bc_cells,
                     int force_mode, const double* force_uniform, const void* force_field,
                     double* grad_force, int nx, int ny, int nz, int64_t plane, double omega, double collision_param,
                     int timestep0, int n_steps, void* workspace, int flags, void* stream);

#define XLBH_ADJ_RECOMPUTE 1
#define XLBH_ADJ_BEGIN 2
#define XLBH_ADJ_END 4

/* Bytes of the fused reverse sweep's workspace: three (Q, N) fields of the compute type and N
 * flag bytes, rounded up to 256 (-1 for invalid arguments). */
int64_t xlbh_adjoint_workspace_bytes(int lattice, int compute, int nx, int ny, int nz);

/* Checkpoint codec of checkpointed reverse sweeps (16 bits per population).  Replaces keeping
 * each segment's start state as a full field (the reference keeps Warp arrays,
 * examples/out_of_core/autodiff_lbm.py:64-127).  encode: from f (any store format but the block
 * format; plane stride as elsewhere) write
 *   moments (float32, (d + 1) planes of N cells): rho - 1, u_x, u_y (, u_z);
 *   neq (bfloat16 bits, Q planes of N cells): f_l - feq_l(rho, u), with feq evaluated in FP64
 *     from the float32 moments.
 * decode: f_l = feq_l(rho, u) + neq_l - w_l (dr + 3 c_l.dj), rounded to the store format, where
 *   dr = sum_l neq_l and dj = sum_l c_l neq_l (the mass and momentum of the rounded
 *   non-equilibrium part, which the exact one does not carry).  The decoded field has the kept
 *   moments and its non-equilibrium part to bfloat16 precision. */
int xlbh_checkpoint_encode(int lattice, int store, const void* f, uint16_t* neq, float* moments,
                           int nx, int ny, int nz, int64_t plane, void* stream);
int xlbh_checkpoint_decode(int lattice, int store, const uint16_t* neq, const float* moments, void* f,
                           int nx, int ny, int nz, int64_t plane, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* XLB_HIP_H */
