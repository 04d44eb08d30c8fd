// Batched MD engine: S independent systems (lambda windows, replicas, or
// unrelated serving jobs of uneven size) packed into one HBM arena and
// stepped together. Every per-step launch covers all systems, sizes are read
// on the device, so a chunk of steps is captured once into a hipGraph and
// replayed.
//
// A reference Context (timemachine/cpp/src/context.cu:10-304) is an Engine
// with S = 1: the same kernels, the same per-system arithmetic and the same
// counter-based noise, which is why one system run alone and the same system
// run inside a batch produce bitwise identical trajectories.
//
// Per-step sequence (context.cu:240-280 restated for S systems):
//   zero du_dx -> nonbonded (rebuild check, sort, gather, neighborlist, tiles)
//   -> pair lists -> bonded -> du_dl store -> BAOAB (integrator.cu:37-86)
//   -> every `interval` steps the MC barostat (barostat.cu:286-373); inside a
//   run, a move whose state nothing reads before the next step is evaluated
//   at the start of that step together with its forces (same values, same
//   order, see Engine::run).
#pragma once

#include <memory>
#include <vector>

#include "bonded.h"
#include "nb_state.h"
#include "potentials.h"

namespace tmamd {

// Per-system status bits raised by the engine, beside the neighborlist's
// NB_STATUS_* bits (nb_state.h). Sticky for the life of the engine.
constexpr int ENGINE_STATUS_NONFINITE = 8;      // coordinates or velocities not finite after a step
constexpr int ENGINE_STATUS_FIXED_INVALID = 16; // a fixed-point force sum holds an out-of-range pair term

struct IntegratorSpec {
    double dt = 0.0;
    double ca = 0.0;
    std::vector<double> cbs, ccs;
    uint64_t seed = 0;
};

struct BarostatSpec {
    int N = 0;
    double pressure = 1.0;    // bar
    double temperature = 300; // K
    int interval = 0;
    uint64_t seed = 0;
    std::vector<std::vector<int>> groups;
};

struct SystemSpec {
    int N = 0;
    std::vector<double> x, v, box;
    IntegratorSpec integ;
    std::vector<BoundPotentialPtr> bps;
    std::shared_ptr<BarostatSpec> baro; // may be null
    // carried state (so a system can move between engines bit-exactly)
    int64_t step = 0;
    double baro_vscale = 0.0;
    int baro_accepted = 0;
    int baro_attempted = 0;
    int baro_step = -1; // barostat calls since construction / set_interval; -1: derive from step
    // x, v, box, step and barostat state come from another engine through
    // Engine::copy_state_from (x, v, box may then be empty here)
    bool state_on_device = false;
};

class Engine {
  public:
    explicit Engine(const std::vector<SystemSpec> &systems);
    ~Engine();

    int num_systems() const { return S_; }
    int num_atoms(int s) const { return h_n_[s]; }
    int total_atoms() const { return A_; }

    // Current lambda of each system (host copy pushed to the device).
    void set_lambdas(const std::vector<double> &lams);

    // Runs n_steps. lambda_sched: [S, n_steps] per-step lambdas or nullptr
    // (keep the current lambdas). du_dl is stored every du_dl_interval steps
    // (0: never) into du_dl_out [S, n_steps / du_dl_interval]; frames every
    // x_interval steps (0: never) into x_out (per system, concatenated
    // [F, N_s, 3] blocks in system order) and box_out [S, F, 9].
    // With u_windows (W lambdas) and u_interval > 0, the state after every
    // u_interval-th step is also evaluated at each window:
    // u_out [S, n_steps / u_interval, W] (Context::multiple_steps_U).
    void run(int n_steps, const double *lambda_sched, int du_dl_interval, int x_interval, double *du_dl_out,
             double *x_out, double *box_out, bool use_graph, const std::vector<double> *u_windows = nullptr,
             int u_interval = 0, double *u_out = nullptr);

    // Device-to-device handover of system src_s of `src` (x, v, box, step
    // counter, barostat state) into system dst_s of this engine; the
    // neighborlists are rebuilt on the next step. Same atom count required.
    void copy_state_from(Engine &src, int src_s, int dst_s);

    // Energies of every system at its current state and lambda (u, du_dl).
    void energies(double *u_out, double *du_dl_out);

    // FIRE minimisation of the coordinates of every active system
    // (md/minimizer.py:35-73: state from the force at schedule entry 0, then one
    // step per further entry). lambda_sched: [S, n_sched] host array, or
    // nullptr for the current lambdas; active: [S] flags or nullptr (all).
    // Velocities, box and step counters are left alone.
    // raise = false: faulted systems are left flagged in the status instead of throwing
    void fire_minimize(int n_sched, const double *lambda_sched, const int *active, bool raise = true);
    // Largest per-atom force norm of every system at its current state and lambda.
    void max_force(double *out, bool raise = true);

    // Gradient observer for lambda-window averages of A = du/dlambda. Every
    // `interval` steps of run() (0: off), at the states where run() stores
    // du_dl with the same interval (start of run step j, (j + 1) % interval
    // == 0), each system's A, g = du/dp and h = d(du/dlambda)/dp are summed
    // on the device: n, sum A, sum A^2, sum g, sum A g, sum h. This resets the
    // sums. Parameters are those of the system's potentials: nonbonded
    // [N, 3], then bond [B, 2], angle [A, 2], torsion [T, 3] terms.
    void set_grad_observer(int interval);
    // One observation at the current state (no step).
    void observe();
    // Number of flat parameters of system s (N*3 + 2B + 2A + 3T).
    int grad_size(int s) const;
    void grad_stats(int s, int *n_frames, double *sum_a, double *sum_aa, double *sum_g, double *sum_ag,
                    double *sum_h);

    // Trajectory (pathwise) adjoint of a loss on the du/dlambda values of a run,
    // by reverse mode through the Langevin steps (attic/docs/paper.tex:27).
    // adjoint_forward runs n_steps like run() (no lambda schedule; du_dl stored
    // every du_dl_interval steps into du_dl_out [S, n_steps / du_dl_interval])
    // and keeps the state at the start of every checkpoint_interval steps (a
    // multiple of du_dl_interval). adjoint_backward takes the loss's cotangent
    // of each stored value, dl_bar [S, F], and walks the steps backwards, one
    // segment at a time: it recomputes the segment's frames from its checkpoint
    // (bitwise the forward run's, checked at the segment end) and takes one
    // Hessian-vector product per step. Outputs: grad_out, per system in system
    // order, the gradient in that system's flat parameters (grad_size(s) each,
    // grad_stats order); x0_bar, v0_bar [A, 3] (may be null) the cotangents of
    // the initial coordinates and velocities. The engine ends in the state after
    // the forward run. Memory: (n_steps / checkpoint_interval + 1) checkpoints
    // of x and v, plus checkpoint_interval frames of x. horizon > 0 truncates
    // the reverse sweep: the state cotangents are cleared at every multiple of
    // `horizon` steps, so a stored value's sensitivity is followed back at most
    // to the start of its block (biased, but bounded where the full adjoint
    // grows with the trajectory's Lyapunov exponent); x0_bar and v0_bar are
    // then those of the first block. Not available with a barostat,
    // interpolated parameters or an interaction group.
    void adjoint_forward(int n_steps, int du_dl_interval, int checkpoint_interval, double *du_dl_out);
    void adjoint_backward(const double *dl_bar, double *grad_out, double *x0_bar, double *v0_bar, int horizon = 0);

    void get_x(int s, double *out);
    void get_v(int s, double *out);
    void get_box(int s, double *out);
    void get_du_dx(int s, uint64_t *out);
    void set_x(int s, const double *in);
    void set_v(int s, const double *in);
    void set_box(int s, const double *in);
    int64_t get_step(int s);
    void get_barostat_state(int s, double *vscale, int *accepted, int *attempted, int *baro_step);
    // The barostat move of system s decided in the last step (all zero when that
    // step made none, k_baro_setup clears it each step): [triggered, accepted,
    // u_before, u_after, u_after - u_before (from the fixed-point sums), volume
    // change, Metropolis draw r2, volume scale after adaptation, length scale].
    void get_barostat_move(int s, double *out);
    // OR of the status bits (NB_STATUS_*, ENGINE_STATUS_*) raised for each
    // system so far: out [S]. A faulted system does not disturb the others
    // (the pair kernel skips systems with a bad box; every other term is per
    // system), so callers may retire it and keep the rest.
    void system_status(int *out);
    // Raises the reference's box messages if a system's box went bad.
    void check_status();
    // Average durations (ms) over `reps` launches, timed with HIP events on the
    // engine's stream at the current state: [0] neighborlist/sort/gather
    // preparation (forced rebuild), [1] pair (tile) kernel, [2] pair lists +
    // bonded, [3] integrator.
    void profile(int reps, double *ms_out);
    // Tiles currently in the neighborlist of every system (sum).
    int total_tiles();
    // Device bytes owned by the engine.
    size_t device_bytes() const;

    hipStream_t stream() const { return stream_; }

    // Neighborlist padding of every system (autotuner); forces a rebuild.
    void set_nblist_padding(double padding);

    int tile_grid = 0; // nonbonded grid override (autotuner), 0 = default
    // Waves per SIMD of the pair kernel build (4, 3 or 2; autotuner).
    void set_tile_waves(int waves);
    // Waves per row block in the neighborlist build (1, 2, 4 or 8; 0 =
    // automatic). Results do not depend on it.
    void set_build_pieces(int pieces);

  private:
    void flat_layout();
    void download_flat(const double *src, int s, double *out);
    struct Impl;
    std::unique_ptr<Impl> impl_;
    int S_ = 0;
    int A_ = 0;
    std::vector<int> h_n_, h_atom0_;
    hipStream_t stream_ = nullptr;
};

} // namespace tmamd
