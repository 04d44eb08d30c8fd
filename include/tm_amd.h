/*
 * tm_amd.h - C ABI of the MI355X (gfx950) MD hot path.
 *
 * Each entry point replaces one method of the reference's pybind11 module
 * (timemachine/cpp/src/wrap_kernels.cpp) or of the C++ class behind it; the
 * citation above each declaration names it. Conventions:
 *   - every function returns TM_OK (0) or an error code; tm_last_error()
 *     returns the message (thread-local), which for validation errors is the
 *     reference's own std::runtime_error text;
 *   - handles are opaque and reference counted internally: a Context keeps
 *     its integrator, bound potentials and barostat alive;
 *   - arrays are C-contiguous; x is [N, 3] f64, box [3, 3] f64, params as the
 *     potential defines them; uint64 outputs are 64-bit fixed point
 *     (2^36 for du_dx, u, du_dl; nonbonded du_dp uses 2^36 / 2^37 / 2^38 for
 *     charge / sigma / epsilon columns) and accumulate into caller-zeroed buffers;
 *   - a NULL output pointer means "not requested".
 * The library owns no global state besides the HIP device context.
 */
#ifndef TM_AMD_H
#define TM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TM_OK 0
#define TM_ERR_INVALID 1 /* validation error (reference message in tm_last_error) */
#define TM_ERR_HIP 2     /* HIP runtime error / device fault */
#define TM_ERR_INTERNAL 3

typedef struct tm_potential tm_potential;
typedef struct tm_bound_potential tm_bound_potential;
typedef struct tm_neighborlist tm_neighborlist;
typedef struct tm_integrator tm_integrator;
typedef struct tm_barostat tm_barostat;
typedef struct tm_context tm_context;
typedef struct tm_batch tm_batch;

/* Last error message of this thread. */
const char *tm_last_error(void);
/* Tile width of the neighborlist / pair kernel (64: one wave64 per tile;
 * the reference uses 32, kernels/k_neighborlist.cuh:7). */
int tm_tile_size(void);
/* Number of visible HIP devices (0 when no GPU). */
int tm_device_count(int *n);

/* ---- potentials ------------------------------------------------------- */

/* NonbondedAllPairs_f32(kernel_dir, lambda_plane_idxs, lambda_offset_idxs, beta, cutoff, atom_idxs)
 * wrap_kernels.cpp:835-895, nonbonded_all_pairs.cu:16-141. n_atom_idxs < 0: all atoms. */
int tm_nonbonded_all_pairs_create(int N, const int32_t *lambda_plane_idxs, const int32_t *lambda_offset_idxs,
                                  double beta, double cutoff, const int32_t *atom_idxs, int n_atom_idxs,
                                  tm_potential **out);
/* NonbondedAllPairs_f32_interpolated: parameters [2N, 3] (initial, final) used as
 * p = (1 - lambda) p0 + lambda p1 with dp/dlambda entering du/dlambda; du/dp returned for both
 * halves (nonbonded_all_pairs.cu:340-351, 420-450; k_lambda_transformer_jit.cuh:57-137, identity
 * transform). */
int tm_nonbonded_all_pairs_interpolated_create(int N, const int32_t *lambda_plane_idxs,
                                               const int32_t *lambda_offset_idxs, double beta, double cutoff,
                                               const int32_t *atom_idxs, int n_atom_idxs, tm_potential **out);
/* NonbondedInteractionGroup_f32[_interpolated](kernel_dir, row_atom_idxs, lambda_plane_idxs,
 * lambda_offset_idxs, beta, cutoff): pairs (i, j) with i among the NR row atoms and j among the
 * others (wrap_kernels.cpp:896-954, nonbonded_interaction_group.cu:19-157). interpolated != 0:
 * parameters [2N, 3] as above. The padding / Hilbert-sort / tile-count calls below accept it. */
int tm_nonbonded_interaction_group_create(int N, int NR, const int32_t *row_atom_idxs,
                                          const int32_t *lambda_plane_idxs, const int32_t *lambda_offset_idxs,
                                          double beta, double cutoff, int interpolated, tm_potential **out);
/* NonbondedAllPairs::set_nblist_padding, nonbonded_all_pairs.cu:180-183 */
int tm_nonbonded_all_pairs_set_nblist_padding(tm_potential *p, double padding);
/* NonbondedAllPairs::disable_hilbert_sort, nonbonded_all_pairs.cu:185-187 */
int tm_nonbonded_all_pairs_disable_hilbert_sort(tm_potential *p);
/* Tiles of the current neighborlist (no reference counterpart; diagnostics). */
int tm_nonbonded_all_pairs_num_tiles(tm_potential *p, int *out);

/* NonbondedPairList_f32[_negated](kernel_dir, pair_idxs, scales, plane, offset, beta, cutoff)
 * wrap_kernels.cpp:957-1014, nonbonded_pair_list.cu:10-91. */
int tm_nonbonded_pair_list_create(int N, int M, const int32_t *pair_idxs, const double *scales,
                                  const int32_t *lambda_plane_idxs, const int32_t *lambda_offset_idxs, double beta,
                                  double cutoff, int negated, tm_potential **out);
/* NonbondedPairList_f32[_negated]_interpolated, parameters [2N, 3] as above
 * (nonbonded_pair_list.cu, Interpolated = true). */
int tm_nonbonded_pair_list_interpolated_create(int N, int M, const int32_t *pair_idxs, const double *scales,
                                               const int32_t *lambda_plane_idxs, const int32_t *lambda_offset_idxs,
                                               double beta, double cutoff, int negated, tm_potential **out);

/* HarmonicBond_f32(bond_idxs, lamb_mult, lamb_offset), harmonic_bond.cu:10-56. mult/offset may be NULL. */
int tm_harmonic_bond_create(int B, const int32_t *bond_idxs, const int32_t *lamb_mult, const int32_t *lamb_offset,
                            tm_potential **out);
/* HarmonicAngle_f32, harmonic_angle.cu:10-60 */
int tm_harmonic_angle_create(int A, const int32_t *angle_idxs, const int32_t *lamb_mult, const int32_t *lamb_offset,
                             tm_potential **out);
/* PeriodicTorsion_f32, periodic_torsion.cu:10-62 */
int tm_periodic_torsion_create(int T, const int32_t *torsion_idxs, const int32_t *lamb_mult,
                               const int32_t *lamb_offset, tm_potential **out);
/* SummedPotential(potentials, params_sizes), summed_potential.cu:8-15 */
int tm_summed_potential_create(int n, tm_potential *const *pots, const int32_t *params_sizes, tm_potential **out);
/* FanoutSummedPotential(potentials), fanout_summed_potential.cu:6-9 */
int tm_fanout_summed_potential_create(int n, tm_potential *const *pots, tm_potential **out);
void tm_potential_release(tm_potential *p);

/* Potential::execute_device, potential.hpp:51-62. Device pointers; stream is a hipStream_t (NULL = default). */
int tm_potential_execute_device(tm_potential *p, int N, int P, const double *d_x, const double *d_p,
                                const double *d_box, double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp,
                                uint64_t *d_du_dl, uint64_t *d_u, void *stream);
/* tm_potential_execute_device for a NonbondedAllPairs potential that allocates no device memory.
 * No counterpart in the reference: it is for callers that cannot allocate (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc). The neighborlist and tile
 * arrays live in a fixed 16 MiB static device array per GPU (up to about 21,000 atoms; larger
 * systems fail with TM_ERR_INVALID), so the potential must cover all atoms with plain
 * parameters (no atom subset, row subset or interpolation). Same inputs, outputs and fixed-point
 * results as tm_potential_execute_device. Every call on one GPU shares that array: calls that
 * overlap on one GPU (two streams or threads) race; sequential calls do not affect each other. */
int tm_potential_execute_device_noalloc(tm_potential *p, int N, int P, const double *d_x, const double *d_p,
                                        const double *d_box, double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp,
                                        uint64_t *d_du_dl, uint64_t *d_u, void *stream);
/* Potential::execute_host, potential.cu:110-185 (du_dl / u are [N] partial sums). */
int tm_potential_execute_host(tm_potential *p, int N, int P, const double *x, const double *params,
                              const double *box, double lambda, uint64_t *du_dx, uint64_t *du_dp, uint64_t *du_dl,
                              uint64_t *u);
/* Potential::execute_batch_host, potential.cu:15-108. Outputs [Bc, Bp, Bl, ...]. */
int tm_potential_execute_batch_host(tm_potential *p, int coord_batch, int N, int param_batch, int P,
                                    int lambda_batch, const double *x, const double *params, const double *boxes,
                                    const double *lambdas, uint64_t *du_dx, uint64_t *du_dp, uint64_t *du_dl,
                                    uint64_t *u);
/* Potential::du_dp_fixed_to_float, potential.cu:227-231 / nonbonded_all_pairs.cu:435-450 */
int tm_potential_du_dp_fixed_to_float(tm_potential *p, int N, int P, const uint64_t *du_dp, double *out);
/* d(du/dlambda)/dp, accumulated into a caller-zeroed [P] buffer with the du_dp fixed-point
 * exponents (decode with tm_potential_du_dp_fixed_to_float). No counterpart in the reference's
 * custom ops: the reference obtains this mixed second derivative from its JAX potentials
 * (timemachine/potentials/nonbonded.py:133-264 under jax.grad). It is the kernel behind
 * gradients of lambda-window averages of du/dlambda. */
int tm_potential_execute_dl_dp_device(tm_potential *p, int N, int P, const double *d_x, const double *d_p,
                                      const double *d_box, double lambda, uint64_t *d_dl_dp, void *stream);
int tm_potential_execute_dl_dp_host(tm_potential *p, int N, int P, const double *x, const double *params,
                                    const double *box, double lambda, uint64_t *dl_dp);
/* Hessian-vector product: the directional derivative of (du/dx, du/dp, du/dlambda) along a
 * tangent (vx [N, 3], vp [P], vl) of coordinates, parameters and lambda, accumulated into
 * caller-zeroed hx [N, 3] (fixed point 2^36, like du_dx), hp [P] (du_dp exponents; decode with
 * tm_potential_du_dp_fixed_to_float) and hl [1] (2^36). Every accumulated term must stay below
 * 2^63 / 2^36 ~ 1.3e8 in magnitude: scale the tangent (the product is linear in it) to at most ~1.
 * The host variant does that itself and returns decoded FP64 outputs (hx [N, 3], hp [P],
 * hl [1]). By the symmetry of the Hessian it is also
 * the vector-Jacobian product of the gradient with (vx, vp, vl): the adjoint of du/dx that
 * reverse-mode differentiation of a loss on forces or on a trajectory needs. Null vx / vp: zero
 * tangent; a null output is not requested. No counterpart in the reference's custom ops: the
 * reference differentiates its JAX potentials (timemachine/potentials/nonbonded.py:133-264,
 * bonded.py:150-332) for second derivatives, and describes the trajectory adjoint in
 * attic/docs/paper.tex:27. Interpolated parameters ([2N, 3], used at (1 - lambda) p0 + lambda p1)
 * are differentiated through that interpolation: hp is [2N, 3] and hl includes the parameters'
 * lambda dependence. */
int tm_potential_execute_hvp_device(tm_potential *p, int N, int P, const double *d_x, const double *d_p,
                                    const double *d_box, double lambda, const double *d_vx, const double *d_vp,
                                    double vl, uint64_t *d_hx, uint64_t *d_hp, uint64_t *d_hl, void *stream);
int tm_potential_execute_hvp_host(tm_potential *p, int N, int P, const double *x, const double *params,
                                  const double *box, double lambda, const double *vx, const double *vp, double vl,
                                  double *hx, double *hp, double *hl);

/* BoundPotential(potential, params), bound_potential.cu:6-12 */
int tm_bound_potential_create(tm_potential *p, int ndim, const int32_t *shape, const double *params,
                              tm_bound_potential **out);
void tm_bound_potential_release(tm_bound_potential *bp);
/* BoundPotential.size(), bound_potential.cu:47-57 */
int tm_bound_potential_size(tm_bound_potential *bp, int *out);
/* BoundPotential.execute, wrap_kernels.cpp:667-693 (du_dl / u are [N] partial sums) */
int tm_bound_potential_execute_host(tm_bound_potential *bp, int N, const double *x, const double *box,
                                    double lambda, uint64_t *du_dx, uint64_t *du_dl, uint64_t *u);

/* ---- neighborlist ----------------------------------------------------- */

/* Neighborlist_f32(N), neighborlist.cu:12-38 */
int tm_neighborlist_create(int N, tm_neighborlist **out);
void tm_neighborlist_release(tm_neighborlist *nl);
/* Neighborlist::resize, neighborlist.cu:263-280 */
int tm_neighborlist_resize(tm_neighborlist *nl, int size);
/* Neighborlist::get_nblist_host, neighborlist.cu:83-124: builds the tile list of x
 * (given order, no spatial sort) and returns the tile count. */
int tm_neighborlist_build_host(tm_neighborlist *nl, int N, const double *x, const double *box, double cutoff,
                               int *n_tiles);
/* Copies the tiles of the last build: row block per tile, 64 column atoms per tile
 * (0xFFFFFFFF = padding). With row indices set, row block b holds row_idxs[64 b .. 64 b + 63]
 * (in the order given) and the column atoms are the other atoms. */
int tm_neighborlist_get_tiles(tm_neighborlist *nl, int capacity, int32_t *tile_rows, uint32_t *tile_atoms);
/* Neighborlist::set_row_idxs, neighborlist.cu:215-243: later builds pair only the NR row atoms
 * with the other atoms. Same validation and messages. */
int tm_neighborlist_set_row_idxs(tm_neighborlist *nl, int NR, const uint32_t *row_idxs);
/* Waves per row block of later builds: 0 = chosen from the launch size (8 for
 * one system, 1 for large batches), or 1, 2, 4, 8. The tile set does not
 * depend on it. No reference counterpart (the reference always uses one
 * thread block per row block, k_neighborlist.cuh:168-186); for tests. */
int tm_neighborlist_set_pieces(tm_neighborlist *nl, int pieces);
/* Neighborlist::reset_row_idxs, neighborlist.cu:245-260 */
int tm_neighborlist_reset_row_idxs(tm_neighborlist *nl);
/* Neighborlist::compute_block_bounds_host, neighborlist.cu:56-81 (64-atom blocks; with row
 * indices set, the blocks of the column atoms). */
int tm_neighborlist_compute_block_bounds_host(tm_neighborlist *nl, int N, const double *x, const double *box,
                                              double *ctr, double *ext);

/* ---- minimisation -------------------------------------------------------- */

/* fire_minimize(x0, u_impls, box, lamb_sched), md/minimizer.py:35-73, on the device: x_out [N, 3]
 * after n_sched - 1 FIRE steps of the summed potentials. */
int tm_fire_minimize(int N, const double *x0, const double *box, int n_bps, tm_bound_potential *const *bps,
                     int n_sched, const double *lambda_sched, double *x_out);

/* ---- integrator, barostat, context ------------------------------------ */

/* LangevinIntegrator(dt, ca, cbs, ccs, seed), wrap_kernels.cpp:314-332, integrator.cu:19-29 */
int tm_langevin_integrator_create(int N, double dt, double ca, const double *cbs, const double *ccs, uint64_t seed,
                                  tm_integrator **out);
void tm_integrator_release(tm_integrator *it);

/* MonteCarloBarostat(N, pressure, temperature, group_idxs, interval, bps, seed),
 * wrap_kernels.cpp:1016-1033, barostat.cu:16-116. Groups are given as sizes + concatenated atoms. */
int tm_barostat_create(int N, double pressure, double temperature, int n_groups, const int32_t *group_sizes,
                       const int32_t *group_atoms, int interval, int n_bps, tm_bound_potential *const *bps,
                       uint64_t seed, tm_barostat **out);
int tm_barostat_set_interval(tm_barostat *b, int interval);  /* barostat.cu:375-382 */
int tm_barostat_get_interval(tm_barostat *b, int *interval); /* barostat.cu:384 */
int tm_barostat_set_pressure(tm_barostat *b, double pressure); /* barostat.cu:386-391 */
void tm_barostat_release(tm_barostat *b);

/* Context(x0, v0, box, integrator, bps, barostat), wrap_kernels.cpp:89-121, context.cu:10-37 */
int tm_context_create(int N, const double *x0, const double *v0, const double *box, tm_integrator *intg, int n_bps,
                      tm_bound_potential *const *bps, tm_barostat *barostat, tm_context **out);
void tm_context_release(tm_context *c);
int tm_context_num_atoms(tm_context *c, int *out);
/* Context::step, context.cu:235-238 */
int tm_context_step(tm_context *c, double lambda);
/* Context::multiple_steps, context.cu:55-137. Intervals > 0 (the Python layer maps 0 to n_steps).
 * du_dl_out [n_steps / du_dl_interval], x_out [F, N, 3], box_out [F, 3, 3], F = n_steps / x_interval. */
int tm_context_multiple_steps(tm_context *c, int n_steps, const double *lambda_schedule, int du_dl_interval,
                              int x_interval, double *du_dl_out, double *x_out, double *box_out);
/* Context::multiple_steps_U, context.cu:139-233: steps at `lambda`; every u_interval steps the
 * state after the step is evaluated at each of the n_windows lambdas, u_out [n_steps / u_interval,
 * n_windows]; frames every x_interval steps, x_out [F, N, 3], box_out [F, 3, 3]. Intervals > 0
 * (the Python layer maps 0 to n_steps, wrap_kernels.cpp:197-198). */
int tm_context_multiple_steps_U(tm_context *c, double lambda, int n_steps, int n_windows, const double *windows,
                                int u_interval, int x_interval, double *u_out, double *x_out, double *box_out);
/* Context getters / setters, context.cu:284-302 */
int tm_context_get_x_t(tm_context *c, double *out);
int tm_context_get_v_t(tm_context *c, double *out);
int tm_context_get_box(tm_context *c, double *out);
int tm_context_set_x_t(tm_context *c, const double *in);
int tm_context_get_du_dx_t_minus_1(tm_context *c, uint64_t *out);

/* ---- batched engine (no reference counterpart: the reference runs one
 * Context per lambda window per process, fe/estimator.py:84-169) ---------- */

/* Packs the current state of n contexts (each keeps its own potentials,
 * integrator, barostat and seed) into one engine; lambdas [n] is the lambda
 * of each system. */
int tm_batch_create(int n, tm_context *const *contexts, const double *lambdas, tm_batch **out);
void tm_batch_release(tm_batch *b);
/* Continuous batching (no reference counterpart): a new batch holding systems keep[0..n_keep)
 * of `old` (their x, v, box, step counters and barostat state copied device to device, never
 * through the host; neighborlists rebuilt on the next step) followed by n_new further contexts.
 * lambdas: [n_keep + n_new]. `old` is left unchanged and may then be released. */
int tm_batch_repack(tm_batch *old, int n_keep, const int32_t *keep, int n_new, tm_context *const *new_contexts,
                    const double *lambdas, tm_batch **out);
/* W lambda windows of one system in one engine, without a Context per window (no reference
 * counterpart; what fe/estimator.py:172-281 builds one Context at a time for). Every window
 * starts from x0 [N, 3] and box [3, 3] with its own velocities v0 [W, N, 3], the potentials
 * bps, the integrator's coefficients with seed seeds[w], and, if barostat is not NULL, the
 * barostat's settings with seed barostat_seeds[w]; lambdas [W]. A window is bitwise the
 * Context built from the same arguments. tm_batch_write_back skips these systems. */
int tm_batch_create_windows(int W, int N, const double *x0, const double *v0, const double *box,
                            tm_integrator *intg, const uint64_t *seeds, int n_bps, tm_bound_potential *const *bps,
                            tm_barostat *barostat, const uint64_t *barostat_seeds, const double *lambdas,
                            tm_batch **out);
/* Continuous batching (no reference counterpart): a new batch whose system i is system
 * systems[i] of srcs[i] (state copied device to device; each source system at most once);
 * lambdas [n]. The sources are left unchanged and may then be released. */
int tm_batch_gather(int n, tm_batch *const *srcs, const int32_t *systems, const double *lambdas, tm_batch **out);
/* The barostat move system s made in the last step (no reference counterpart; the values
 * k_decide_move uses, barostat.cu:220-279): out[9] = triggered, accepted, u_before, u_after,
 * u_after - u_before, volume change, Metropolis draw, volume scale after adaptation, length
 * scale. All zero when that step made no move. */
int tm_batch_get_barostat_move(tm_batch *b, int s, double *out);
/* Per-system fault isolation (no reference counterpart; the reference raises for its one system,
 * nonbonded_all_pairs.cu:315-329). Status bits of each system so far, out [n]: 1 box too small for
 * cutoff + padding, 2 box not orthorhombic, 4 neighborlist tile buffer overflow, 8 coordinates or
 * velocities not finite, 16 a fixed-point force sum holds an out-of-range pair term. A faulted
 * system does not change any other system's results. */
int tm_batch_get_status(tm_batch *b, int32_t *out);
/* raise = 0: tm_batch_multiple_steps returns TM_OK with faulted systems (read them with
 * tm_batch_get_status) instead of failing the whole batch; 1 (default): it raises the first
 * system's error. */
int tm_batch_set_raise_faults(tm_batch *b, int raise);
/* Device bytes held by the batch's engine (admission control). */
int tm_batch_device_bytes(tm_batch *b, uint64_t *out);
int tm_batch_num_systems(tm_batch *b, int *out);
int tm_batch_total_atoms(tm_batch *b, int *out);
/* Steps every system n_steps. lambda_sched [n, n_steps] or NULL (keep lambdas).
 * du_dl_out [n, n_steps / du_dl_interval]; x_out concatenated [F, N_s, 3] per system;
 * box_out [n, F, 3, 3]. Intervals of 0 disable the store. */
int tm_batch_multiple_steps(tm_batch *b, int n_steps, const double *lambda_sched, int du_dl_interval,
                            int x_interval, double *du_dl_out, double *x_out, double *box_out, int use_graph);
/* multiple_steps_U for every system of the batch at its own lambda (no reference counterpart for
 * a batch): u_out [n, n_steps / u_interval, n_windows]; x_interval 0 stores no frames. */
int tm_batch_multiple_steps_U(tm_batch *b, int n_steps, int n_windows, const double *windows, int u_interval,
                              int x_interval, double *u_out, double *x_out, double *box_out, int use_graph);
/* u and du/dlambda of every system at its current state. */
int tm_batch_energies(tm_batch *b, double *u_out, double *du_dl_out);
/* fire_minimize (md/minimizer.py:35-73, _vendored/fire.py:60-140) of every system flagged in
 * active ([n] or NULL: all) on the device, in place: the FIRE state starts from the force at
 * lambda_sched[s, 0], then one step per further entry; lambda_sched [n, n_sched] or NULL (the
 * current lambdas, n_sched - 1 steps). Velocities, boxes and step counters are not changed. */
int tm_batch_fire_minimize(tm_batch *b, int n_sched, const double *lambda_sched, const int32_t *active);
/* Largest per-atom norm of the total force of every system at its current state and lambda
 * (the sanity check after minimisation, fe/estimator.py:245-249, on the summed force). */
int tm_batch_max_force(tm_batch *b, double *out);
int tm_batch_get_x(tm_batch *b, int s, double *out);
int tm_batch_get_v(tm_batch *b, int s, double *out);
int tm_batch_get_box(tm_batch *b, int s, double *out);
/* Copies every system's state back into the contexts it came from. */
int tm_batch_write_back(tm_batch *b);
/* Average durations in ms over reps launches, timed with HIP events on the
 * batch's own stream at its current state: [0] neighborlist preparation with a
 * forced rebuild, [1] pair (tile) kernel, [2] pair lists + bonded, [3] integrator. */
int tm_batch_profile(tm_batch *b, int reps, double *ms_out);
/* Tiles in the current neighborlists of all systems. */
int tm_batch_total_tiles(tm_batch *b, int *out);
/* One-wave workgroups of the pair (tile) kernel (0 = default, CUs x 128; at most 2^24). No reference counterpart
 * (the reference launches one block per tile); used by the autotuner. */
int tm_batch_set_tile_grid(tm_batch *b, int grid);
/* Waves per SIMD the pair (tile) kernel is compiled for: 4 (default, at most 128 VGPRs), 3 or 2 (up to 168 / 256
 * VGPRs, fewer waves in flight). No reference counterpart; used by the autotuner. Results do not depend on it. */
int tm_batch_set_tile_waves(tm_batch *b, int waves);
/* Waves per row block in the neighborlist build (1, 2, 4 or 8; 0 = automatic). No reference
 * counterpart; used by the autotuner. Results do not depend on it. */
int tm_batch_set_build_pieces(tm_batch *b, int pieces);
/* Neighborlist padding of every system of the batch (nm); the counterpart of
 * NonbondedAllPairs::set_nblist_padding (nonbonded_all_pairs.cu:180-187) for a batch.
 * Results do not depend on it. */
int tm_batch_set_nblist_padding(tm_batch *b, double padding);

/* Gradient observer (no reference counterpart; the statistics behind gradients of
 * lambda-window averages of A = du/dlambda with respect to forcefield parameters).
 * Every `interval` steps of tm_batch_multiple_steps (0 = off), at the states where du_dl
 * is stored with the same interval, each system's A, g = du/dp and h = d(du/dlambda)/dp
 * are summed on the device.
 * Setting the interval resets the sums. */
int tm_batch_set_grad_observer(tm_batch *b, int interval);
/* One observation of every system at its current state (no step). */
int tm_batch_observe(tm_batch *b);
/* Flat parameters of system s: nonbonded [N, 3], then bond [B, 2], angle [A, 2],
 * torsion [T, 3] terms, in the order the system's potentials were bound. */
int tm_batch_grad_size(tm_batch *b, int s, int *out);
/* n frames, sum A, sum A^2 (scalars) and sum g, sum A g, sum h ([grad_size]). */
int tm_batch_grad_stats(tm_batch *b, int s, int *n_frames, double *sum_a, double *sum_aa, double *sum_g,
                        double *sum_ag, double *sum_h);

/* Trajectory (pathwise) adjoint, no reference counterpart: the reference's design notes
 * describe differentiating through the Langevin trajectory with Hessian-vector products
 * (attic/docs/paper.tex:27); its code computes none. adjoint_forward runs n_steps of every
 * system like tm_batch_multiple_steps (no lambda schedule), stores du_dl every du_dl_interval
 * steps into du_dl_out [S, n_steps / du_dl_interval] and keeps checkpoints of the state every
 * checkpoint_interval steps (a multiple of du_dl_interval). adjoint_backward takes the loss's
 * cotangent of each stored du_dl value, dl_bar [S, n_steps / du_dl_interval], recomputes each
 * segment of frames from its checkpoint (checked bitwise against the forward run) and walks the
 * steps backwards with one Hessian-vector product per step. grad_out: the gradient in every
 * system's flat parameters, systems in order (tm_batch_grad_size each); x0_bar, v0_bar
 * [total atoms, 3] (may be null): cotangents of the initial coordinates and velocities. Not
 * available with a barostat, interpolated parameters or an interaction group. */
int tm_batch_adjoint_forward(tm_batch *b, int n_steps, int du_dl_interval, int checkpoint_interval,
                             double *du_dl_out);
/* horizon > 0: truncated reverse sweep, the state cotangents cleared every `horizon` steps
 * (a stored value's sensitivity followed back at most to the start of its block); 0: full. */
int tm_batch_adjoint_backward(tm_batch *b, const double *dl_bar, double *grad_out, double *x0_bar,
                              double *v0_bar, int horizon);

#ifdef __cplusplus
}
#endif

#endif /* TM_AMD_H */
