/*
 * lbm.h - C ABI of the MI355X lattice Boltzmann wind-tunnel solver.
 *
 * One opaque handle per tunnel. Every call returns an int status (LBM_OK = 0,
 * negative on failure); lbm_last_error() gives the message of the last failure
 * on the calling thread. Plain pointers and sizes only: no torch, no HIP types.
 *
 * The handle replaces the reference's WebGPU-backed classes:
 *   D2Q9  Solver   (src/solver.ts:50-340)
 *   D3Q19 Solver3D (src/solver3d.ts:42-319)
 * Each entry point names the method it stands in for. GPU work is enqueued on
 * one HIP stream per handle and runs in order, as the reference's single
 * WebGPU queue does; setters are stream-ordered (they affect the steps
 * enqueued after them), like queue.writeBuffer.
 */
#ifndef LBM_H
#define LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 2 /* 2: lbm_config3d gained velocity_set */

/* Status codes. */
#define LBM_OK 0
#define LBM_ERR_INVALID (-1)     /* bad argument or configuration */
#define LBM_ERR_ALLOC (-2)       /* device allocation failed ("Solver3D allocation failed: ...") */
#define LBM_ERR_DEVICE (-3)      /* HIP runtime error */
#define LBM_ERR_UNSUPPORTED (-4) /* feature not available (e.g. precision) */

/* Boundary modes (src/lattice.ts:17-28). */
#define LBM_X_PERIODIC 0
#define LBM_X_WALL 1
#define LBM_X_OPEN 2 /* Zou-He inlet on the left, Zou-He pressure outlet on the right */
#define LBM_Y_PERIODIC 0
#define LBM_Y_NOSLIP 1
#define LBM_Y_SLIP 2
#define LBM_Y_MOVING 3

/* Open x planes of a 3D tunnel (lbm_set_open_faces). */
#define LBM_FACES_EQUILIBRIUM 0 /* the reference's: inlet stores feq(1, uIn), outlet feq(1, u) */
#define LBM_FACES_ZOU_HE 1      /* Hecht-Harting Zou-He velocity inlet and rho = 1 pressure outlet */

/* Storage precision of the distribution buffers (src/shaders/common3d.ts:4, plus bf16). */
#define LBM_FP32 0
#define LBM_FP16 1 /* f16 of (f - w) * 2^15, clamped to +-1.99 before scaling */
#define LBM_BF16 2 /* bf16 of (f - w), round to nearest even, no clamp */
/* The LBM_FP16 format, stochastically rounded in the time step: a value in the f16 normal range is
 * rounded up with probability equal to the fraction f16 drops, using a random word that is a hash of
 * (cell, population, step), so runs repeat exactly; initial and refilled values round to nearest. */
#define LBM_FP16_SR 3

/* Length of the on-device force history ring (src/shaders/aux.ts:4, aux3d.ts:5). */
#define LBM_HISTORY_LEN 8192

/* SolverConfig (src/solver.ts:15-40) with its defaults (src/solver.ts:98-112). */
typedef struct lbm_config2d {
  int32_t width, height;
  int32_t left, right, bottom, top; /* LBM_X_* / LBM_Y_* */
  int32_t parabolic;                /* bool: parabolic inlet profile */
  int32_t workgroup_size;           /* threads per block; 0 = autotuned */
  int32_t force_every;              /* sample the obstacle force every n steps; 0 = off */
  int32_t _pad0;
  double tau;
  double lambda;      /* TRT magic parameter, default 3/16 */
  double smagorinsky; /* C_s; 0 disables the subgrid model */
  double u_in;
  double u_lid;
  double sponge_fraction;
  double sponge_tau;
  double body_force[2];
  double absorb;
  double inlet_layer_fraction;
} lbm_config2d;

/* Solver3DConfig (src/solver3d.ts:11-31) with its defaults (src/solver3d.ts:110-119). */
typedef struct lbm_config3d {
  int32_t width, height, depth;
  int32_t precision;      /* LBM_FP32 / LBM_FP16 / LBM_BF16 */
  int32_t workgroup_size; /* 0 = autotuned */
  int32_t force_every;
  double tau;
  double lambda; /* NaN = follow tau through min(3/16, 50 (tau - 1/2)^2) (src/lattice3d.ts:29-31) */
  double smagorinsky;
  double u_in;
  double sponge_fraction;
  double sponge_tau;
  double absorb;
  int32_t velocity_set; /* 19 (D3Q19, the reference's; 0 means 19) or 27 (D3Q27, its roadmap switch) */
  int32_t _pad1;
} lbm_config3d;

typedef struct lbm_solver lbm_solver;

/* Library-wide. */
int lbm_abi_version(void);
const char* lbm_last_error(void);
/* The adapter line of the reference's benchmark page (src/bench.ts:100-102): vendor, architecture and
 * product name of a GPU ordinal, NUL-terminated into out[cap]. */
int lbm_device_name(int device, char* out, size_t cap);
/* Fills the reference defaults; the caller then sets width/height/(depth)/tau (and modes in 2D). */
void lbm_config2d_defaults(lbm_config2d* cfg);
void lbm_config3d_defaults(lbm_config3d* cfg);

/* Solver.create (src/solver.ts:87-92): allocates, builds flags for an empty SDF (all FAR). */
int lbm_create_2d(const lbm_config2d* cfg, int device, lbm_solver** out);
/* Solver3D.create (src/solver3d.ts:79-104). On failure nothing stays allocated. */
int lbm_create_3d(const lbm_config3d* cfg, int device, lbm_solver** out);
/* Solver.destroy / Solver3D.destroy (src/solver.ts:337-339, src/solver3d.ts:317-319). NULL is a no-op. */
void lbm_destroy(lbm_solver* s);

/* Solver.update (src/solver.ts:175-197): replaces the whole config (width/height must not change). */
int lbm_update_2d(lbm_solver* s, const lbm_config2d* cfg);
/* Solver3D.setTau / setSmagorinsky / setInlet (src/solver3d.ts:182-208): replaces the whole config. */
int lbm_update_3d(lbm_solver* s, const lbm_config3d* cfg);

/* No counterpart in the reference, whose drivers ramp the inlet per chunk through update / setInlet
 * (src/cases.ts:52-54, src/cases3d.ts:46): evaluate the inlet speed per step as
 * u_target * r(min(1, step / steps)), r = t (kind 1) or t^2 (3 - 2t) (kind 2); kind 0 turns it off and
 * restores the configured u_in. The step index is that of the step being taken (0 for the first). */
int lbm_set_inlet_ramp(lbm_solver* s, int kind, double u_target, double steps);

/* 3D only. The reference's planes (LBM_FACES_EQUILIBRIUM, the default, src/shaders/step3d.ts:190-209)
 * or the Zou-He faces its 3D spec lists as the follow-up for validation runs (LBM_FACES_ZOU_HE; spec
 * docs/dietpowers/2026-09-27-3d-tunnel-spec.md:287, Hecht and Harting, arXiv:0811.4593): the
 * populations entering from outside are rebuilt so the inlet node carries exactly (uIn, 0, 0) and
 * the outlet node rho = 1 with no transverse velocity, and both nodes then collide. Takes effect
 * from the next enqueued step; rebuilds the boundary tables. The adjoint accepts only the default. */
int lbm_set_open_faces(lbm_solver* s, int kind);
/* If a geometry rebuild (lbm_set_sdf, lbm_rebuild_flags, lbm_brush*, lbm_set_open_faces, a 2D mode
 * change in lbm_update_2d) fails after the flag kernel has rewritten the device flags, the handle
 * keeps its previous boundary tables but refuses to step (LBM_ERR_INVALID) until a later rebuild
 * succeeds; lbm_set_open_faces rebuilds with the previous kind by itself. */

/* No counterpart in the reference: a test hook. The next n boundary-table rebuilds of this handle
 * fail where a device allocation of the new tables would (LBM_ERR_ALLOC): after the flag kernel ran,
 * before the handle's tables change. */
int lbm_debug_fail_rebuilds(lbm_solver* s, int n);

/* setSdf (src/solver.ts:200-203, src/solver3d.ts:211-214): n = N floats, negative inside solids.
 * Rebuilds flags, the boundary-cell list (in cell order) and refills cells the body uncovered. */
int lbm_set_sdf(lbm_solver* s, const float* sdf, size_t n);

/* rebuildFlags (src/solver.ts:205-217): flags, boundary tables and refill from the current SDF. */
int lbm_rebuild_flags(lbm_solver* s);

/* Solver.brush (src/solver.ts:249-268, src/shaders/aux.ts:176-196), 2D only: unions (erase = 0) or
 * subtracts (erase = 1) the disc of radius r at (cx, cy) over its box grown by 2 cells, clipped to the
 * grid; then rebuilds as lbm_rebuild_flags when rebuild != 0 and the box is not empty. */
int lbm_brush(lbm_solver* s, double cx, double cy, double r, int erase, int rebuild);
/* Solver.brushStroke (src/solver.ts:243-247), 2D only: the discs at the n_points (x, y) pairs of xy,
 * in order, then one rebuild. */
int lbm_brush_stroke(lbm_solver* s, const double* xy, size_t n_points, double r, int erase);

/* initField (src/solver.ts:220-234, src/solver3d.ts:231-242). 2D: macro = N*4 floats (rho, ux, uy, -)
 * or NULL for rest; 3D: macro must be NULL (inflow equilibrium). Resets step, parity and forces. */
int lbm_init_field(lbm_solver* s, const float* macro, size_t n);

/* resetForces (src/solver.ts:236-241, src/solver3d.ts:244-249). */
int lbm_reset_forces(lbm_solver* s);

/* encodeSteps (src/solver.ts:271-286): enqueue n steps (+ force samples every force_every). */
int lbm_enqueue_steps(lbm_solver* s, int64_t n);
/* No counterpart in the reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation): lbm_enqueue_steps, with the same results
 * bit for bit, whose force reduction keeps its scratch (FP64 per-wave partials and a completion
 * counter) in one static device array per GPU instead of the handle's buffers, and which allocates
 * nothing, host or device. The array holds 32768 partials (768 KiB): a tunnel whose reduction needs
 * more (one partial per wave of body-link nodes; D3Q27: per wave of the grid, above 2 M cells) is
 * refused with LBM_ERR_UNSUPPORTED. All calls on one GPU share the array and the counter is cleared
 * at the start of every call, so calls that overlap on one GPU (two handles on different streams)
 * race; the call itself is stream-ordered like lbm_enqueue_steps. */
int lbm_enqueue_steps_noalloc(lbm_solver* s, int64_t n);
/* encodeMacro (src/solver.ts:288-292): enqueue the macro-field kernel for the current buffer. */
int lbm_enqueue_macro(lbm_solver* s);
/* run (src/solver.ts:295-307): n steps in chunks, macro after each chunk, then wait for the GPU. */
int lbm_run(lbm_solver* s, int64_t n, int64_t chunk);
/* Waits for all enqueued work (queue.onSubmittedWorkDone). */
int lbm_sync(lbm_solver* s);

/* readMacro (src/solver.ts:309-311, src/solver3d.ts:290-292): N*4 floats, 2D (rho, ux, uy, solid),
 * 3D (ux, uy, uz, rho). Reads the macro buffer as of the last enqueued macro kernel. */
int lbm_read_macro(lbm_solver* s, float* out, size_t n);

/* readForces (src/solver.ts:315-335, src/solver3d.ts:295-315), split so that the copy can be waited
 * for off the caller's thread while the generation check stays on it:
 *   begin  enqueues a copy of the ring and its counter and returns a ticket (main thread);
 *   wait   blocks until that copy landed (any thread);
 *   end    returns the samples produced since the last read, oldest first, as step numbers and
 *          dim = 2 (2D) or 3 (3D) doubles per sample; a read overtaken by lbm_reset_forces returns 0.
 * cap is the capacity of steps[] in samples; at most LBM_HISTORY_LEN samples are ever returned. */
int lbm_read_forces_begin(lbm_solver* s, int64_t* ticket);
int lbm_read_forces_wait(lbm_solver* s, int64_t ticket);
int lbm_read_forces_end(lbm_solver* s, int64_t ticket, int64_t* steps, double* values, int64_t cap,
                        int64_t* count);
/* begin + wait + end. */
int lbm_read_forces(lbm_solver* s, int64_t* steps, double* values, int64_t cap, int64_t* count);

/* hasDiverged (src/tunnel3d.ts:97-99), kept on the device: *first_step receives the number of the
 * first step whose force sample had a non-finite component since the last lbm_init_field /
 * lbm_reset_forces, or 0 while every sample was finite. Samples are checked as the step kernels
 * append them, so a flow that blew up between two reads is caught even if the read is skipped.
 * Waits for the work enqueued before it. */
int lbm_diverged(lbm_solver* s, int64_t* first_step);

/* Readonly fields W, H, D, N, step (src/solver.ts:51-57, src/solver3d.ts:43-50). */
int lbm_info(const lbm_solver* s, int32_t* dims /* W, H, D */, int64_t* n_cells, int64_t* step,
             int32_t* lattice_q, int32_t* precision);

/* No counterpart in the reference (its buffers are read only by the renderer): copies the current
 * distribution buffer, Q*N values in storage format (float / f16 bits / bf16 bits), for parity tests. */
int lbm_read_populations(lbm_solver* s, void* out, size_t bytes);
/* No counterpart in the reference: copies the per-cell flag words (N uint32, same bit layout). */
int lbm_read_flags(lbm_solver* s, uint32_t* out, size_t n);

/* Timing, replacing the timestamp queries around one compute pass of src/bench.ts:35-58.
 * lbm_time_steps enqueues n steps between two HIP events on the handle's stream and returns the
 * elapsed milliseconds. lbm_profile_steps does the same with an event after every step (one kernel
 * launch per step) and returns the total, the summed per-step times and the slowest step; it
 * perturbs the schedule slightly and is for per-kernel roofline figures. */
int lbm_time_steps(lbm_solver* s, int64_t n, double* elapsed_ms);
int lbm_profile_steps(lbm_solver* s, int64_t n, double* total_ms, double* step_ms, double* max_step_ms);
/* No counterpart in the reference (its bench sweeps workgroup sizes 64/128/256 by hand,
 * src/bench.ts:110-118): times every bulk-kernel block width usable for this grid (in 3D, each with
 * the three row orders, in 3D also the two store variants and the two pull orders of lbm_launch_info;
 * D3Q27: its one block width in the three row orders) on a copy of the solver's
 * state, with force samples as configured, keeps the fastest (cached for the process per lattice,
 * velocity set, precision, grid size, open-face kind, subgrid model on/off and force cadence), and restores
 * populations, step and force ring (also when it fails part way). For 3D grids of at least 65,536
 * cells it then times the three fastest shapes in each population layout (strides N, and N rounded
 * up to 1024 plus 4096 or plus 69,632; for D3Q19 rows of a multiple of 64 cells also tiles of 64
 * cells; see lbm_set_population_stride and lbm_set_population_layout) and keeps the fastest pair;
 * not while an adjoint state exists. chosen (may be NULL) receives the width. Block width, row order
 * and population stride never change a result; the store variants, and the two pull orders of the
 * shared-store variant, agree to floating-point rounding (their compiled code may fuse different
 * multiply-adds). */
int lbm_autotune(lbm_solver* s, int32_t trial_steps, int32_t* chosen);
/* No counterpart in the reference: the launch shape in use, threads per bulk block and the
 * variant: bits 0-1 the row order, 0 always bottom-up, 1 (2) top-down on the steps that read the
 * population buffer at the higher (lower) address, bottom-up on the others; in 3D also bit 2, set
 * when the inlet and outlet nodes share their wave's stores, and bit 3, set when a node's 19 loads
 * are issued before its kind (flag word) is known (no effect with the subgrid model or on recording
 * steps). D3Q27 has one block width (whole rows of at most 256 threads) and the row order as its
 * variant; D2Q9 the row order. */
int lbm_launch_info(lbm_solver* s, int32_t* block, int32_t* sweep);
/* No counterpart in the reference: sets the launch shape by hand instead of lbm_autotune (block as
 * workgroup_size, variant as lbm_launch_info reports it; a row order 0-2 in 2D; D3Q27: its one block
 * width and a row order 0-2), e.g. to pin a run's shape. */
int lbm_set_launch(lbm_solver* s, int32_t block, int32_t sweep);
/* No counterpart in the reference (its distribution buffers are N-strided, src/solver3d.ts:131-132): the
 * device buffers' population stride in elements, population i of cell idx at element i * stride +
 * idx. N unless lbm_autotune (3D) or lbm_set_population_stride (any lattice, also tunnels of a batch)
 * chose a padded stride, which changes which memory channels and cache sets a node's streams share. Every host-visible layout stays Q x N (lbm_read_populations copies the padded buffers row by
 * row) and every result the same bits. Setting 0 or N restores the unpadded layout; a padded stride
 * is a multiple of 64 elements, at least N, with Q x stride below 2^32. The populations are moved to
 * the new stride (a device copy into two new buffers) and the boundary tables' indices re-strided; on
 * failure the old buffers, stride and tables stay. Refused while an adjoint state exists
 * (lbm_adjoint_release first). */
int lbm_population_stride(lbm_solver* s, int64_t* stride);
int lbm_set_population_stride(lbm_solver* s, int64_t stride);
/* No counterpart in the reference: the device population layout. tiled = 0: one array per
 * population, stride elements apart (lbm_population_stride). tiled = 1 (a row width that is a
 * multiple of 64 and Q x N x element bytes below 2^32): tiles of 64 cells, each
 * holding its cells' Q populations one after another, population i of cell c at element
 * c + (Q - 1) (c & ~63) + 64 i (stride reads 64); a tile is one wave of the step's bulk rows, and the
 * Q streams of a wave fall in one block of memory instead of Q arrays. lbm_autotune times it next to
 * the strides for D3Q19 and D3Q27; set by hand here (stride 0 or 64 with tiled = 1, as lbm_set_population_stride
 * otherwise; batched tunnels too). Same guarantees as the stride: host-visible layouts and results unchanged, refused
 * while an adjoint state exists. */
int lbm_population_layout(lbm_solver* s, int32_t* tiled, int64_t* stride);
int lbm_set_population_layout(lbm_solver* s, int32_t tiled, int64_t stride);
/* Copy roofline (src/bench.ts:61-95): iters launches (after 50 untimed ones) of a kernel that reads q
 * structure-of-arrays populations of elem_bytes each plus a u32 flag word per cell and writes the q
 * populations, with no arithmetic, one thread per cell in blocks of workgroup (64, 128 or 256; 0 for
 * 256); returns their total milliseconds. The memory-only ceiling of a step. */
int lbm_copy_roofline(int device, int64_t n_cells, int32_t q, int32_t elem_bytes, int32_t workgroup, int32_t iters,
                      double* ms_total);
/* No counterpart in the reference: the copy roofline with a pull step's access pattern. A D3Q19 grid of
 * width x height x depth (width <= 1024), elem_bytes 4 or 2; one launch reads, for every node of the
 * rows with a neighbour row on both sides in y and z, population i from the upstream node x - c_i and
 * the node's flag word, and writes the 19 populations at the node: the bytes of a step, no arithmetic,
 * the shifted streams of the pull. fill 0 copies zeroed buffers (as the reference page's copy does);
 * fill 1 first fills them with the kind of values a step moves (f - w of a uniform stream plus a
 * small hashed perturbation), which the memory system moves measurably more slowly than zeros: the
 * ceiling for a step. iters launches after 50 untimed ones; *cells receives the nodes one launch
 * updates. */
int lbm_pull_copy_roofline(int device, int32_t width, int32_t height, int32_t depth, int32_t elem_bytes,
                           int32_t fill, int32_t iters, double* ms_total, int64_t* cells);
/* No counterpart in the reference: lbm_copy_roofline with the fill choice of lbm_pull_copy_roofline
 * (fill 0 is lbm_copy_roofline). */
int lbm_copy_roofline_filled(int device, int64_t n_cells, int32_t q, int32_t elem_bytes, int32_t workgroup,
                             int32_t fill, int32_t iters, double* ms_total);
/* Counts of the node classes the step kernels see: plain (bulk), boundary list, solid. */
int lbm_node_counts(lbm_solver* s, int64_t* bulk, int64_t* list, int64_t* solid);

/* ---- Discrete adjoint (shape gradients) ----
 * No counterpart in the reference. A forward/backward pair with autograd-Function semantics:
 * lbm_adjoint_forward runs the tunnel from its initial field (lbm_init_field: rest in 2D, the
 * equilibrium at the first step's inlet speed in 3D) for `steps` steps with the given inlet ramp and a force sample every step, and
 * returns J = scale * mean over the samples of steps window_start+1 .. steps of direction . F (the
 * sample of step k is taken from the populations entering it, as lbm_read_forces numbers them). It
 * records the last segments directly while record_budget_bytes allows (the moments of the collision
 * input, 16 B per cell per step with FP32 storage, 8 B with 16-bit storage, plus 4 B per Bouzidi link
 * per step), so that the backward pass does not recompute them, keeps a checkpoint of the
 * populations at the start of every other segment of checkpoint_every steps, and records the
 * segment before the recorded ones into the one segment of scratch the backward pass recomputes
 * into (so that, if no later backward pass overwrote it, that segment is not recomputed either).
 * lbm_adjoint_backward may be called any number of times after one forward run.
 * lbm_adjoint_backward(seed) then returns seed * dJ/dsdf at every node (N doubles) of the discrete
 * step: it recomputes each remaining segment from its checkpoint, runs the reverse-time kernels
 * (FP32 adjoint populations, FP64 accumulation of dJ/dq per Bouzidi link) and chains dq/dsdf through
 * q = clamp(dF / max(dF - dS, 1e-6), 1e-3, 1) (src/shaders/step.ts:140-142); the flag topology is
 * that of the current SDF. Afterwards the solver holds the state after the forward run.
 * Not supported: Smagorinsky (C_s > 0) and the 2D body force. */
typedef struct lbm_adjoint_config {
  int64_t steps;
  int64_t window_start;
  double direction[3]; /* 2D uses the first two */
  double scale;
  int32_t checkpoint_every;
  int32_t ramp_kind; /* as lbm_set_inlet_ramp; 0 keeps u_in constant */
  double ramp_u, ramp_steps;
  int64_t record_budget_bytes; /* device memory for segments recorded by the forward pass (0: none) */
} lbm_adjoint_config;
int lbm_adjoint_forward(lbm_solver* s, const lbm_adjoint_config* cfg, double* J);
int lbm_adjoint_backward(lbm_solver* s, double seed, double* dJ_dsdf, size_t n);
/* Node counts of the reverse-time kernel (streamed and listed nodes, list entries) and the number of
 * recorded and total segments of the last forward run. */
int lbm_adjoint_info(lbm_solver* s, int64_t* regular_nodes, int64_t* list_nodes, int64_t* entries,
                     int64_t* stored_segments, int64_t* segments);
/* Frees the checkpoints, records and tables (also freed by lbm_destroy). */
int lbm_adjoint_release(lbm_solver* s);

/* ---- Batched D2Q9 tunnels ----
 * No counterpart in the reference (one tunnel per page); the nearest analogue is encoding many steps
 * into one pass (src/solver.ts:271-286). n tunnels of any sizes and parameters advance together, one
 * kernel launch per time step for all of them, replayed from hipGraphs. Each tunnel is an ordinary
 * 2D solver handle (lbm_batch_tunnel) sharing the batch's stream: use lbm_set_sdf, lbm_init_field,
 * lbm_update_2d, lbm_enqueue_macro, lbm_read_macro and lbm_read_forces on it as on a lone solver, but
 * step it only through the batch. After lbm_init_field on the tunnels call lbm_batch_reset_clock. */
typedef struct lbm_batch lbm_batch;
int lbm_batch_create_2d(const lbm_config2d* cfgs, int32_t n, int device, lbm_batch** out);
void lbm_batch_destroy(lbm_batch* b);
int lbm_batch_size(const lbm_batch* b, int32_t* n_tunnels, int64_t* total_cells);
/* Borrowed handle of tunnel i (owned by the batch: never lbm_destroy it). */
lbm_solver* lbm_batch_tunnel(lbm_batch* b, int32_t i);
/* Per-step inlet ramp for every tunnel toward its own configured u_in (kinds as lbm_set_inlet_ramp). */
int lbm_batch_set_inlet_ramp(lbm_batch* b, int kind, double steps);
int lbm_batch_reset_clock(lbm_batch* b);
int lbm_batch_enqueue_steps(lbm_batch* b, int64_t n);
/* lbm_diverged for every tunnel of the batch (first_steps holds n >= the tunnel count values). */
int lbm_batch_diverged(lbm_batch* b, int64_t* first_steps, size_t n);
int lbm_batch_sync(lbm_batch* b);
int lbm_batch_time_steps(lbm_batch* b, int64_t n, double* elapsed_ms);
/* No counterpart in the reference (its page sweeps workgroup sizes by hand, src/bench.ts:110-118):
 * times the tile widths 64, 128, 256 and 512 (cells or list nodes per block) with the tunnels in
 * creation order and largest first, each with the bulk tiles walked in one direction or in
 * alternating directions, trial_steps launches each on the tunnels' own state, keeps the
 * fastest and restores every tunnel's populations and force ring; tile and order (may be NULL)
 * receive the choice. Neither changes a result. */
int lbm_batch_autotune(lbm_batch* b, int32_t trial_steps, int32_t* tile, int32_t* order);
/* The launch shape in use, and setting it by hand (tile 64/128/256/512; order bit 0: 0 creation
 * order, 1 largest tunnel first; bit 1: bulk tiles walked from the end of the table on every other
 * step). */
int lbm_batch_launch_info(lbm_batch* b, int32_t* tile, int32_t* order);
int lbm_batch_set_launch(lbm_batch* b, int32_t tile, int32_t order);

#ifdef __cplusplus
}
#endif
#endif /* LBM_H */
