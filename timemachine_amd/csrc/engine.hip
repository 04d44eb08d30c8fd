// Batched MD engine (see engine.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "engine.h"
#include "hvp.h"
#include "fire.h"
#include "fixed_point.h"
#include "philox.h"
#include "wave.h"

namespace tmamd {

namespace {

// Constants of the reference barostat (timemachine/cpp/src/barostat.cu:11-12).
constexpr double BARO_AVOGADRO = 6.0221367e23;
constexpr double BARO_BOLTZ = 0.008314462618;

__global__ void k_set_lambda(int S, const double *__restrict__ sched, int n_steps, const int *__restrict__ run_step,
                             double *__restrict__ lam) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) {
        return;
    }
    const int r = run_step[s];
    lam[s] = sched[static_cast<size_t>(s) * n_steps + (r < n_steps ? r : n_steps - 1)];
}

// BAOAB rotated by half a step (integrator.cu:37-69, integrator.py:81-88):
//   v_mid = v + cb * F, v' = ca v_mid + cc xi, x += dt/2 (v_mid + v'),  cb = -dt/m.
// Noise: Philox keyed by the system seed, counter (2 * atom + block, step).
// A system whose new coordinates or velocities are not finite, or whose
// fixed-point force holds an out-of-range pair term (a non-finite or huge pair
// contribution converts to 2^63, which leaves |sum| >= 2^62 for any force below
// 2^26 kJ/mol/nm), raises a status bit for that system only.
__global__ void k_langevin(int A, const int *__restrict__ atom_sys, const int *__restrict__ atom0,
                           const int64_t *__restrict__ step, const uint64_t *__restrict__ seed,
                           const double *__restrict__ ca, const double *__restrict__ dt,
                           const double *__restrict__ cbs, const double *__restrict__ ccs,
                           const uint64_t *__restrict__ du_dx, double *__restrict__ x, double *__restrict__ v,
                           int *__restrict__ sys_status) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A) {
        return;
    }
    const int s = atom_sys[a];
    const uint32_t local = static_cast<uint32_t>(a - atom0[s]);
    const uint64_t st = static_cast<uint64_t>(step[s]);
    const uint64_t sd = seed[s];
    const uint32_t k0 = static_cast<uint32_t>(sd), k1 = static_cast<uint32_t>(sd >> 32);
    Philox4 c0{{2 * local + 0, static_cast<uint32_t>(st), static_cast<uint32_t>(st >> 32), RNG_STREAM_LANGEVIN}};
    Philox4 c1{{2 * local + 1, static_cast<uint32_t>(st), static_cast<uint32_t>(st >> 32), RNG_STREAM_LANGEVIN}};
    double z[4];
    normal2(philox4x32_10(c0, k0, k1), z[0], z[1]);
    normal2(philox4x32_10(c1, k0, k1), z[2], z[3]);
    const double cb = cbs[a], cc = ccs[a], cas = ca[s], h = 0.5 * dt[s];
    int bad = 0;
    for (int d = 0; d < 3; d++) {
        const uint64_t fx = du_dx[a * 3 + d];
        const int64_t fi = static_cast<int64_t>(fx);
        if (fi >= (int64_t(1) << 62) || fi <= -(int64_t(1) << 62)) {
            bad |= ENGINE_STATUS_FIXED_INVALID;
        }
        const double f = fixed_to_double(fx, FIXED_SCALE);
        const double v_mid = v[a * 3 + d] + cb * f;
        const double v_new = cas * v_mid + cc * z[d];
        v[a * 3 + d] = v_new;
        const double x_new = x[a * 3 + d] + h * (v_mid + v_new);
        x[a * 3 + d] = x_new;
        if (!isfinite(x_new) || !isfinite(v_new)) {
            bad |= ENGINE_STATUS_NONFINITE;
        }
    }
    if (bad) {
        atomicOr(sys_status + s, bad);
    }
}

// du_dl store, box frame store, per-system step counters.
__global__ void k_step_end(int S, int64_t *__restrict__ step, int *__restrict__ run_step,
                           uint64_t *__restrict__ du_dl_buf, int dl_interval, int K, uint64_t *__restrict__ du_dl_traj,
                           int x_interval, int F, const double *__restrict__ box, double *__restrict__ box_frames) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) {
        return;
    }
    const int r = run_step[s] + 1;
    if (dl_interval > 0) {
        if (r % dl_interval == 0) {
            const int k = r / dl_interval - 1;
            if (k < K) {
                du_dl_traj[static_cast<size_t>(s) * K + k] = du_dl_buf[s];
            }
        }
        du_dl_buf[s] = 0;
    }
    if (x_interval > 0 && r % x_interval == 0) {
        const int f = r / x_interval - 1;
        if (f < F) {
            for (int d = 0; d < 9; d++) {
                box_frames[(static_cast<size_t>(s) * F + f) * 9 + d] = box[s * 9 + d];
            }
        }
    }
    run_step[s] = r;
    step[s] += 1;
}

// Frame store with an explicit per-system atom count table.
__global__ void k_store_frames2(int A, const int *__restrict__ atom_sys, const int *__restrict__ atom0,
                                const int *__restrict__ natoms, const int *__restrict__ run_step, int x_interval,
                                int F, const double *__restrict__ x, double *__restrict__ frames) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A) {
        return;
    }
    const int s = atom_sys[a];
    const int r = run_step[s] + 1;
    if (r % x_interval != 0) {
        return;
    }
    const int f = r / x_interval - 1;
    if (f >= F) {
        return;
    }
    const int a0 = atom0[s];
    const int ns = natoms[s];
    const size_t dst = static_cast<size_t>(F) * a0 * 3 + (static_cast<size_t>(f) * ns + (a - a0)) * 3;
    frames[dst + 0] = x[a * 3 + 0];
    frames[dst + 1] = x[a * 3 + 1];
    frames[dst + 2] = x[a * 3 + 2];
}

// Energies at other lambda windows: every system evaluates window w.
// ---- FIRE minimisation (md/minimizer.py:35-73, _vendored/fire.py:110-140; arithmetic in fire.h)

__global__ void k_fire_init(int A3, const int *__restrict__ atom_sys, const int *__restrict__ active,
                            const uint64_t *__restrict__ du_dx, double *__restrict__ fv, double *__restrict__ ff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A3 || !active[atom_sys[i / 3]]) {
        return;
    }
    fv[i] = 0.0;
    ff[i] = -fixed_to_double(du_dx[i], FIXED_SCALE);
}

__global__ void k_fire_state_init(int S, float *__restrict__ dt, float *__restrict__ alpha, int *__restrict__ npos) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) {
        dt[s] = FIRE_DT_START;
        alpha[s] = FIRE_ALPHA_START;
        npos[s] = 0;
    }
}

__global__ void k_fire_move(int A3, const int *__restrict__ atom_sys, const int *__restrict__ active,
                            const float *__restrict__ dt, const double *__restrict__ fv, const double *__restrict__ ff,
                            double *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A3) {
        return;
    }
    const int s = atom_sys[i / 3];
    if (!active[s]) {
        return;
    }
    x[i] = fire_shift(x[i], fv[i], ff[i], dt[s]);
}

constexpr int FIRE_THREADS = 1024;

// One workgroup per system: new forces, half-step velocities, the three sums
// (fixed tree order, so a system gives the same bits alone or in a batch),
// velocity mixing, then dt / alpha / n_pos.
__global__ __launch_bounds__(FIRE_THREADS) void k_fire_update(
    const int *__restrict__ atom0, const int *__restrict__ natoms, const int *__restrict__ active,
    const uint64_t *__restrict__ du_dx, double *__restrict__ fv, double *__restrict__ ff, float *__restrict__ dt,
    float *__restrict__ alpha, int *__restrict__ npos) {
    const int s = blockIdx.x;
    if (!active[s]) {
        return;
    }
    __shared__ double red[3][FIRE_THREADS];
    __shared__ double stats[3];
    const int t = threadIdx.x;
    const int base = atom0[s] * 3;
    const int n3 = natoms[s] * 3;
    const float dts = dt[s];
    double sf = 0.0, sv = 0.0, sp = 0.0;
    for (int q = t; q < n3; q += FIRE_THREADS) {
        const double f = -fixed_to_double(du_dx[base + q], FIXED_SCALE);
        const double v = fire_velocity(fv[base + q], ff[base + q], f, dts);
        fv[base + q] = v;
        ff[base + q] = f;
        sf += f * f;
        sv += v * v;
        sp += f * v;
    }
    red[0][t] = sf;
    red[1][t] = sv;
    red[2][t] = sp;
    __syncthreads();
    for (int stride = FIRE_THREADS / 2; stride > 0; stride >>= 1) {
        if (t < stride) {
            red[0][t] += red[0][t + stride];
            red[1][t] += red[1][t + stride];
            red[2][t] += red[2][t + stride];
        }
        __syncthreads();
    }
    if (t == 0) {
        stats[0] = std::sqrt(red[0][0] + static_cast<double>(FIRE_F_EPS)); // |F|
        stats[1] = std::sqrt(red[1][0]);                                    // |V|
        stats[2] = red[2][0];                                               // P
    }
    __syncthreads();
    const double f_norm = stats[0], v_norm = stats[1], P = stats[2];
    const float a = alpha[s];
    for (int q = t; q < n3; q += FIRE_THREADS) {
        fv[base + q] = fire_mix(fv[base + q], ff[base + q], v_norm, f_norm, a, P);
    }
    if (t == 0) {
        float d = dts, al = a;
        int np = npos[s];
        fire_adapt(P, d, al, np);
        dt[s] = d;
        alpha[s] = al;
        npos[s] = np;
    }
}

// Lambdas of FIRE step `*done + 1` (the schedule is step-major, [n_sched][S]).
__global__ void k_fire_lambda(int S, int n_sched, const double *__restrict__ sched, const int *__restrict__ done,
                              double *__restrict__ lam_cur) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) {
        const int k = min(*done + 1, n_sched - 1);
        lam_cur[s] = sched[static_cast<size_t>(k) * S + s];
    }
}

__global__ void k_fire_count(int *__restrict__ done) {
    if (threadIdx.x == 0) {
        *done += 1;
    }
}

__global__ void k_max_force(const int *__restrict__ atom0, const int *__restrict__ natoms,
                            const uint64_t *__restrict__ du_dx, double *__restrict__ out) {
    const int s = blockIdx.x;
    __shared__ double red[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < natoms[s]; i += 256) {
        const uint64_t *g = du_dx + (atom0[s] + i) * 3;
        const double fx = fixed_to_double(g[0], FIXED_SCALE), fy = fixed_to_double(g[1], FIXED_SCALE),
                     fz = fixed_to_double(g[2], FIXED_SCALE);
        m = fmax(m, std::sqrt(fx * fx + fy * fy + fz * fz));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (static_cast<int>(threadIdx.x) < stride) {
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + stride]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[s] = red[0];
    }
}

__global__ void k_set_eval_lambda(int S, const double *__restrict__ windows, int w, double *__restrict__ lam_eval) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) {
        lam_eval[s] = windows[w];
    }
}

// u_traj[s, f, w] = u of the state after run step run_step[s] (f-th stored
// frame, (run_step + 1) % u_interval == 0), then clears the reduction slot.
__global__ void k_store_u(int S, const int *__restrict__ run_step, int u_interval, int Fu, int W, int w,
                          uint64_t *__restrict__ u_buf, uint64_t *__restrict__ u_traj) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) {
        return;
    }
    const int f = (run_step[s] + 1) / u_interval - 1;
    if (f >= 0 && f < Fu) {
        u_traj[(static_cast<size_t>(s) * Fu + f) * W + w] = u_buf[s];
    }
    u_buf[s] = 0;
}

// Gradient observer: per flat parameter element e (nonbonded [A, 3] rows, then
// bond / angle / torsion parameters) of system s = elem_sys[e], with
// A_s = du/dlambda, g = du/dp, h = d(du/dlambda)/dp at the observed frame:
//   sum_g += g, sum_ag += A_s g, sum_h += h      (FP64, one thread per element,
// frames in order, so the sums are deterministic).
__global__ void k_obs_accum(int Pg, const int *__restrict__ elem_sys, const double *__restrict__ elem_inv_scale,
                            const uint64_t *__restrict__ dl, const uint64_t *__restrict__ du_dp,
                            const uint64_t *__restrict__ dl_dp, double *__restrict__ sum_g,
                            double *__restrict__ sum_ag, double *__restrict__ sum_h) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Pg) {
        return;
    }
    const double a = fixed_to_double(dl[elem_sys[e]], FIXED_SCALE);
    const double inv = elem_inv_scale[e];
    const double g = static_cast<double>(static_cast<int64_t>(du_dp[e])) * inv;
    const double h = static_cast<double>(static_cast<int64_t>(dl_dp[e])) * inv;
    sum_g[e] += g;
    sum_ag[e] += a * g;
    sum_h[e] += h;
}

__global__ void k_obs_accum_sys(int S, const uint64_t *__restrict__ dl, double *__restrict__ sum_a,
                                double *__restrict__ sum_aa, int *__restrict__ n_frames) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) {
        return;
    }
    const double a = fixed_to_double(dl[s], FIXED_SCALE);
    sum_a[s] += a;
    sum_aa[s] += a * a;
    n_frames[s] += 1;
}

// ------------------------------------------------------ trajectory adjoint
// Reverse mode through the BAOAB step of k_langevin. With the step
//   v_mid = v + cb g(x),  v' = ca v_mid + cc xi,  x' = x + h (v_mid + v'),
// (g = du/dx, h = dt / 2), output cotangents (xb', vb') give
//   vb_mid = ca vb' + h (1 + ca) xb',  vb = vb_mid,  gb = cb vb_mid,
//   xb = xb' + H_xx gb + H_xl lb,  pb += H_px gb + H_pl lb,
// where lb is the cotangent of the du/dlambda stored at x (zero when the
// step stores none). The noise is additive and drops out. The Hessian
// products come from one HVP call with tangent (gb, 0, lb), scaled per system
// by a power of two to at most 1 so the fixed-point accumulators stay in
// range (the product is linear in the tangent; the scale is undone below).

// vb, gb; per-system max |gb| (bit pattern of a non-negative double orders as
// the value) into smax
__global__ void k_adj_pre(int A, const int *__restrict__ atom_sys, const double *__restrict__ ca,
                          const double *__restrict__ dt, const double *__restrict__ cbs,
                          const double *__restrict__ xb, double *__restrict__ vb, double *__restrict__ gb,
                          unsigned long long *__restrict__ smax) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = a < A;
    const int s = act ? atom_sys[a] : -1;
    double m = 0.0;
    if (act) {
        const double c = ca[s], h = 0.5 * dt[s], cb = cbs[a];
        for (int d = 0; d < 3; d++) {
            const double vm = c * vb[a * 3 + d] + h * (1.0 + c) * xb[a * 3 + d];
            vb[a * 3 + d] = vm;
            const double g = cb * vm;
            gb[a * 3 + d] = g;
            m = fmax(m, fabs(g));
        }
    }
    // one atomic per wave when its atoms belong to one system (atoms of a
    // system are contiguous), else one per atom
    uint64_t bits = static_cast<uint64_t>(__double_as_longlong(m));
    const int s0 = __builtin_amdgcn_readfirstlane(s);
    if (__all(!act || s == s0)) {
        for (int o = 32; o >= 1; o >>= 1) {
            const uint64_t other = shfl_xor_u64(bits, o);
            bits = other > bits ? other : bits;
        }
        if (lane_id() == 0 && s0 >= 0 && bits != 0) {
            atomicMax(smax + s0, static_cast<unsigned long long>(bits));
        }
    } else if (act && bits != 0) {
        atomicMax(smax + s, static_cast<unsigned long long>(bits));
    }
}

// lambda cotangent of step j (the du/dlambda stored at the start of step j
// when (j + 1) % interval == 0), the tangent scale 2^-e (scaled tangent at
// most 1 in magnitude), the scaled lambda tangent
__global__ void k_adj_scale(int S, int j, int interval, int F, const double *__restrict__ dl_bar,
                            const unsigned long long *__restrict__ smax, double *__restrict__ scale,
                            double *__restrict__ vls) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) {
        return;
    }
    double lb = 0.0;
    if (interval > 0 && (j + 1) % interval == 0) {
        const int f = (j + 1) / interval - 1;
        if (f < F) {
            lb = dl_bar[static_cast<size_t>(s) * F + f];
        }
    }
    const double m = fmax(__longlong_as_double(static_cast<long long>(smax[s])), fabs(lb));
    int e = 0;
    if (m > 0.0) {
        frexp(m, &e);
    }
    const double sc = ldexp(1.0, -e);
    scale[s] = sc;
    vls[s] = lb * sc;
}

__global__ void k_adj_scale_vec(int A, const int *__restrict__ atom_sys, const double *__restrict__ scale,
                                const double *__restrict__ gb, double *__restrict__ vx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A * 3) {
        return;
    }
    vx[i] = gb[i] * scale[atom_sys[i / 3]];
}

// xb += H gb (decoded and unscaled)
__global__ void k_adj_post_x(int A, const int *__restrict__ atom_sys, const double *__restrict__ scale,
                             const uint64_t *__restrict__ hx, double *__restrict__ xb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A * 3) {
        return;
    }
    xb[i] += fixed_to_double(hx[i], FIXED_SCALE) / scale[atom_sys[i / 3]];
}

// pb += H_p. gb per flat parameter element (one thread per element, steps in
// order: deterministic)
__global__ void k_adj_post_p(int Pg, const int *__restrict__ elem_sys, const double *__restrict__ elem_inv_scale,
                             const double *__restrict__ scale, const uint64_t *__restrict__ hp,
                             double *__restrict__ pb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Pg) {
        return;
    }
    pb[e] += static_cast<double>(static_cast<int64_t>(hp[e])) * elem_inv_scale[e] / scale[elem_sys[e]];
}

// flag = 1 if a differs from b anywhere (bitwise)
__global__ void k_adj_compare(size_t n, const double *__restrict__ a, const double *__restrict__ b,
                              int *__restrict__ flag) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && __double_as_longlong(a[i]) != __double_as_longlong(b[i])) {
        flag[0] = 1;
    }
}

// ---------------------------------------------------------------- barostat
// Restates MonteCarloBarostat::inplace_move (barostat.cu:286-373). The move is
// decided by one thread per system and applied by a separate kernel, so no
// thread reads a box another thread is overwriting.

struct BaroDev {
    int *has;          // [S]
    int *interval;     // [S]
    int *bstep;        // [S] barostat call counter modulo the interval (reset by set_interval)
    uint64_t *seed;    // [S]
    double *pressure;  // [S] internal units (bar * AVOGADRO * 1e-25)
    double *kT;        // [S]
    int *nmol;         // [S]
    double *vscale;    // [S]
    int *accepted;     // [S]
    int *attempted;    // [S]
    int *trig;         // [S]
    int *accept;       // [S]
    double *dvol;      // [S]
    double *lscale;    // [S]
    double *r2;        // [S]
    uint64_t *u_before; // [S]
    uint64_t *u_after;  // [S]
};

// The move's random numbers are keyed on the system's absolute MD step (the
// counter that also keys the Langevin noise, carried across engine rebuilds),
// not on the call counter, which only holds the phase within the interval: a
// rebuilt engine (batch hand-over, set_pressure, set_interval) continues the
// stream instead of repeating its first draws.
__global__ void k_baro_setup(int S, BaroDev b, const int64_t *__restrict__ step, const double *__restrict__ box,
                             double *__restrict__ box_after) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) {
        return;
    }
    b.trig[s] = 0;
    b.accept[s] = 0;
    b.u_before[s] = 0;
    b.u_after[s] = 0;
    for (int d = 0; d < 9; d++) {
        box_after[s * 9 + d] = box[s * 9 + d];
    }
    if (!b.has[s]) {
        return;
    }
    const int call = b.bstep[s] + 1;
    b.bstep[s] = call % b.interval[s];
    if (call % b.interval[s] != 0) {
        return;
    }
    b.trig[s] = 1;
    const uint64_t sd = b.seed[s];
    const uint64_t st = static_cast<uint64_t>(step[s]);
    Philox4 c{{0u, static_cast<uint32_t>(st), static_cast<uint32_t>(st >> 32), RNG_STREAM_BAROSTAT}};
    const Philox4 r = philox4x32_10(c, static_cast<uint32_t>(sd), static_cast<uint32_t>(sd >> 32));
    const double r1 = u01_open0(r.v[0], r.v[1]);
    b.r2[s] = u01_open0(r.v[2], r.v[3]);
    const double volume = box[s * 9 + 0] * box[s * 9 + 4] * box[s * 9 + 8];
    if (b.vscale[s] == 0.0) {
        b.vscale[s] = 0.01 * volume;
    }
    const double dv = b.vscale[s] * 2.0 * (r1 - 0.5);
    const double new_volume = volume + dv;
    b.dvol[s] = dv;
    const double ls = cbrt(new_volume / volume);
    b.lscale[s] = ls;
    box_after[s * 9 + 0] = box[s * 9 + 0] * ls;
    box_after[s * 9 + 4] = box[s * 9 + 4] * ls;
    box_after[s * 9 + 8] = box[s * 9 + 8] * ls;
}

__global__ void k_baro_zero_centroids(int nmol_total, const int *__restrict__ mol_sys, const int *__restrict__ trig,
                                      uint64_t *__restrict__ centroids) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol_total || !trig[mol_sys[m]]) {
        return;
    }
    centroids[m * 3 + 0] = 0;
    centroids[m * 3 + 1] = 0;
    centroids[m * 3 + 2] = 0;
}

// find_group_centroids (barostat.cu:181-197): fixed-point sums of coordinates.
__global__ void k_baro_centroids(int G, const int *__restrict__ g_atom, const int *__restrict__ g_mol,
                                 const int *__restrict__ mol_sys, const int *__restrict__ trig,
                                 const double *__restrict__ x, uint64_t *__restrict__ centroids) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) {
        return;
    }
    const int m = g_mol[i];
    if (!trig[mol_sys[m]]) {
        return;
    }
    const int a = g_atom[i];
    for (int d = 0; d < 3; d++) {
        atomicAdd(reinterpret_cast<unsigned long long *>(centroids + m * 3 + d),
                  static_cast<unsigned long long>(f64_to_fixed(x[a * 3 + d], FIXED_SCALE)));
    }
}

// rescale_positions (barostat.cu:139-179): move each molecule's centroid
// towards/away from the box centre by the length scale.
__global__ void k_baro_rescale(int A, const int *__restrict__ atom_sys, const int *__restrict__ atom_mol,
                               const int *__restrict__ mol_count, const int *__restrict__ trig,
                               const double *__restrict__ lscale, const double *__restrict__ box,
                               const uint64_t *__restrict__ centroids, const double *__restrict__ x,
                               double *__restrict__ x_after) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A) {
        return;
    }
    const int s = atom_sys[a];
    const int m = atom_mol[a];
    if (!trig[s] || m < 0) {
        x_after[a * 3 + 0] = x[a * 3 + 0];
        x_after[a * 3 + 1] = x[a * 3 + 1];
        x_after[a * 3 + 2] = x[a * 3 + 2];
        return;
    }
    const double n = static_cast<double>(mol_count[m]);
    const double ls = lscale[s];
    for (int d = 0; d < 3; d++) {
        const double center = box[s * 9 + d * 4] * 0.5;
        const double centroid = fixed_to_double(centroids[m * 3 + d], FIXED_SCALE) / n;
        const double disp = ((centroid - center) * ls) + center - centroid;
        x_after[a * 3 + d] = x[a * 3 + d] + disp;
    }
}

// k_decide_move (barostat.cu:220-279) for one system per thread.
__global__ void k_baro_decide(int S, BaroDev b, double *__restrict__ box, const double *__restrict__ box_after) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || !b.trig[s]) {
        return;
    }
    const double volume = box[s * 9 + 0] * box[s * 9 + 4] * box[s * 9 + 8];
    const double new_volume = volume + b.dvol[s];
    const double de = fixed_to_double(b.u_after[s] - b.u_before[s], FIXED_SCALE);
    const double kT = b.kT[s];
    const double w = de + b.pressure[s] * b.dvol[s] - b.nmol[s] * kT * log(new_volume / volume);
    const bool rejected = w > 0 && b.r2[s] > exp(-w / kT);
    if (!rejected) {
        b.accepted[s] += 1;
    }
    b.attempted[s] += 1;
    if (b.attempted[s] >= 10) {
        if (b.accepted[s] < 0.25 * b.attempted[s]) {
            b.vscale[s] /= 1.1;
            b.attempted[s] = 0;
            b.accepted[s] = 0;
        } else if (b.accepted[s] > 0.75 * b.attempted[s]) {
            b.vscale[s] = fmin(b.vscale[s] * 1.1, volume * 0.3);
            b.attempted[s] = 0;
            b.accepted[s] = 0;
        }
    }
    b.accept[s] = rejected ? 0 : 1;
    if (!rejected) {
        for (int d = 0; d < 9; d++) {
            box[s * 9 + d] = box_after[s * 9 + d];
        }
    }
}

__global__ void k_baro_apply(int A, const int *__restrict__ atom_sys, const int *__restrict__ accept,
                             const double *__restrict__ x_after, double *__restrict__ x) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A || !accept[atom_sys[a]]) {
        return;
    }
    x[a * 3 + 0] = x_after[a * 3 + 0];
    x[a * 3 + 1] = x_after[a * 3 + 1];
    x[a * 3 + 2] = x_after[a * 3 + 2];
}

// A barostat move evaluated with forces (see Engine::run): systems that
// accepted the move take the forces and du/dlambda of the moved state.
__global__ void k_baro_take_trial(int A, int S, const int *__restrict__ atom_sys, const int *__restrict__ accept,
                                  const uint64_t *__restrict__ dx_trial, uint64_t *__restrict__ dx,
                                  const uint64_t *__restrict__ dl_trial, uint64_t *__restrict__ dl) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < S && dl && accept[a]) {
        dl[a] = dl_trial[a];
    }
    if (a >= A || !accept[atom_sys[a]]) {
        return;
    }
    dx[a * 3 + 0] = dx_trial[a * 3 + 0];
    dx[a * 3 + 1] = dx_trial[a * 3 + 1];
    dx[a * 3 + 2] = dx_trial[a * 3 + 2];
}

// Flattened potential terms of one system.
struct Flat {
    std::shared_ptr<NonbondedAllPairs> nb;
    std::vector<double> nb_params;
    std::shared_ptr<NonbondedAllPairs> ig; // a NonbondedInteractionGroup (rows then columns)
    std::vector<double> ig_params;
    std::vector<std::pair<std::shared_ptr<NonbondedPairList>, std::vector<double>>> pls;
    std::vector<std::pair<std::shared_ptr<BondedPotential>, std::vector<double>>> bonded;
};

void flatten(const PotentialPtr &p, const double *params, int P, Flat &out) {
    switch (p->kind()) {
    case PotKind::NonbondedAllPairs: {
        if (out.nb) {
            throw Error("Engine: at most one NonbondedAllPairs per system is supported");
        }
        out.nb = std::static_pointer_cast<NonbondedAllPairs>(p);
        out.nb_params.assign(params, params + P);
        break;
    }
    case PotKind::NonbondedInteractionGroup: {
        if (out.ig) {
            throw Error("Engine: at most one NonbondedInteractionGroup per system is supported");
        }
        out.ig = std::static_pointer_cast<NonbondedAllPairs>(p);
        out.ig_params.assign(params, params + P);
        break;
    }
    case PotKind::NonbondedPairList:
        out.pls.emplace_back(std::static_pointer_cast<NonbondedPairList>(p), std::vector<double>(params, params + P));
        break;
    case PotKind::HarmonicBond:
    case PotKind::HarmonicAngle:
    case PotKind::PeriodicTorsion:
        out.bonded.emplace_back(std::static_pointer_cast<BondedPotential>(p), std::vector<double>(params, params + P));
        break;
    case PotKind::Fanout: {
        auto f = std::static_pointer_cast<FanoutSummedPotential>(p);
        for (auto &c : f->pots_) {
            flatten(c, params, P, out);
        }
        break;
    }
    case PotKind::Summed: {
        auto f = std::static_pointer_cast<SummedPotential>(p);
        int off = 0;
        for (size_t i = 0; i < f->pots_.size(); i++) {
            flatten(f->pots_[i], params + off, f->sizes_[i], out);
            off += f->sizes_[i];
        }
        break;
    }
    }
}

template <typename T> void upload_vec(DeviceArray<T> &d, const std::vector<T> &h) {
    d.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty()) {
        d.upload(h.data(), h.size());
    }
}

} // namespace

struct Engine::Impl {
    int S = 0, A = 0;
    std::vector<int> atom0, natoms;

    DeviceArray<double> x, v, x_after, box, box_after, cbs, ccs, ca, dt, lam;
    DeviceArray<uint64_t> du_dx, seed, u_buf, du_dl_buf;
    DeviceArray<int64_t> step;
    DeviceArray<int> atom_sys, d_atom0, d_natoms, red_idx, run_step;
    DeviceArray<int> sys_status; // [S] ENGINE_STATUS_* bits (sticky)

    // nonbonded
    bool has_nb = false;
    NbState nb;
    DeviceArray<double> nb_params, nb_params1; // nb_params1: final parameters of interpolated systems
    DeviceArray<int> nb_plane, nb_offset;
    // an interaction group per system (row-subset layout), same arrays
    DeviceArray<uint64_t> dx_trial, dl_trial; // forces and du/dl of a deferred barostat move's state
    // Graph executables of run() kept across calls (runs without a lambda
    // schedule, frames or window energies): keyed by the step pattern of a
    // graph chunk and every setting its launches bake in. They write du/dl
    // into run_dl, one row of run_dl_cap entries per system.
    std::map<std::vector<int>, hipGraphExec_t> graphs;
    DeviceArray<uint64_t> run_dl;
    int run_dl_cap = 0;
    DeviceArray<double> run_box_frames; // never written by those runs (no frames), its address is baked in
    void clear_graphs() {
        if (graphs.empty()) {
            return;
        }
        DeviceLock lock;
        for (auto &kv : graphs) {
            (void)hipGraphExecDestroy(kv.second);
        }
        graphs.clear();
    }
    bool has_ig = false;
    NbState ig;
    DeviceArray<double> ig_params, ig_params1;
    DeviceArray<int> ig_plane, ig_offset;

    // pair lists (index 0: plain, 1: negated)
    int pl_n[2] = {0, 0};
    DeviceArray<int> pl_pairs[2], pl_sys[2];
    DeviceArray<double> pl_scales[2];
    DeviceArray<double> pl_params, pl_params1;
    bool interp = false; // interpolated nonbonded parameters (gradient observer unavailable)
    DeviceArray<int> pl_plane, pl_offset;
    DeviceArray<double> pl_cutoffs;
    DeviceArray<float> pl_betas;
    bool has_pl = false;

    // bonded
    int bt_n[3] = {0, 0, 0};
    DeviceArray<int> bt_idx[3], bt_mult[3], bt_offset[3], bt_sys[3];
    // force slots and their incidence per atom: the force path gathers per atom
    // instead of one global atomic per term, atom and component
    DeviceArray<int> bt_inc_off;
    DeviceArray<uint32_t> bt_inc_items;
    DeviceArray<uint64_t> bt_slots;
    std::vector<int> bt_slot0;      // first slot of bonds, angles, torsions, pair list, negated pair list
    std::vector<int> h_pl_pairs[2]; // host copy of the pair lists (incidence only)
    bool bt_gather = false;
    DeviceArray<double> bt_params[3];

    // barostat
    bool has_baro = false;
    std::vector<int> h_has_baro, h_interval, h_bstep;
    DeviceArray<int> b_has, b_interval, b_bstep, b_nmol, b_accepted, b_attempted, b_trig, b_accept;
    DeviceArray<uint64_t> b_seed, b_u_before, b_u_after;
    DeviceArray<double> b_pressure, b_kT, b_vscale, b_dvol, b_lscale, b_r2;
    int G = 0, nmol_total = 0;
    DeviceArray<int> g_atom, g_mol, mol_sys, mol_count, atom_mol;
    DeviceArray<uint64_t> centroids;

    // gradient observer
    bool pl_shares_nb = true;            // pair lists use the all-pairs parameters
    int obs_interval = 0;
    int Pg = 0;                          // flat parameter elements
    int bt_off[3] = {0, 0, 0};           // offset of each bonded kind in the flat space
    std::vector<int> bt_first[3], bt_count[3]; // per-system term ranges
    DeviceArray<uint64_t> obs_dudp, obs_dldp, obs_dl;
    DeviceArray<int> obs_sys, obs_n;
    DeviceArray<double> obs_inv_scale, obs_sum_g, obs_sum_ag, obs_sum_h, obs_sum_a, obs_sum_aa;

    // trajectory adjoint: checkpoints of the last adjoint_forward (state at the
    // start of every segment of ck_interval steps)
    int adj_steps = 0, adj_dl_interval = 0, adj_ck_interval = 0;
    std::vector<DeviceArray<double>> ck_x, ck_v;
    std::vector<DeviceArray<int64_t>> ck_step;

    BaroDev baro_dev() {
        return BaroDev{b_has.ptr,      b_interval.ptr, b_bstep.ptr,    b_seed.ptr,     b_pressure.ptr,  b_kT.ptr,
                       b_nmol.ptr,     b_vscale.ptr,   b_accepted.ptr, b_attempted.ptr, b_trig.ptr,     b_accept.ptr,
                       b_dvol.ptr,     b_lscale.ptr,   b_r2.ptr,       b_u_before.ptr, b_u_after.ptr};
    }

    void eval(const double *xp, const double *boxp, uint64_t *ddx, uint64_t *ddl, uint64_t *uu, hipStream_t st,
              int tile_grid, const double *lamp = nullptr) {
        if (!lamp) {
            lamp = lam.ptr;
        }
        if (has_nb) {
            nb.prepare(xp, nb_params.ptr, nb_params1.n ? nb_params1.ptr : nullptr, boxp, nb_plane.ptr, nb_offset.ptr,
                       lamp, 0.0, st);
            nb.compute(boxp, ddx, nullptr, ddl, uu, st, tile_grid);
        }
        if (has_ig) {
            ig.prepare(xp, ig_params.ptr, ig_params1.n ? ig_params1.ptr : nullptr, boxp, ig_plane.ptr, ig_offset.ptr,
                       lamp, 0.0, st);
            ig.compute(boxp, ddx, nullptr, ddl, uu, st, tile_grid);
        }
        const bool gather = bt_gather && ddx;
        for (int neg = 0; neg < 2; neg++) {
            if (pl_n[neg] == 0) {
                continue;
            }
            PairListArgs a = pl_args(neg, xp, boxp);
            a.lambdas = lamp;
            if (gather) {
                a.slots = bt_slots.ptr + static_cast<size_t>(bt_slot0[3 + neg]) * 3;
            }
            launch_pair_list(a, ddx, nullptr, ddl, uu, st);
        }
        for (int k = 0; k < 3; k++) {
            if (bt_n[k] == 0) {
                continue;
            }
            BondedTerms t{bt_n[k], bt_idx[k].ptr, bt_params[k].ptr, bt_mult[k].ptr, bt_offset[k].ptr, bt_sys[k].ptr};
            launch_bonded(static_cast<BondedKind>(k), t, xp, lamp, 0.0, red_idx.ptr, gather ? nullptr : ddx, nullptr,
                          ddl, uu, st, false, gather ? bt_slots.ptr + static_cast<size_t>(bt_slot0[k]) * 3 : nullptr);
        }
        if (gather) {
            launch_bonded_gather(BondedIncidence{A, bt_inc_off.ptr, bt_inc_items.ptr}, bt_slots.ptr, ddx, st);
        }
    }

    PairListArgs pl_args(int neg, const double *xp, const double *boxp) const {
        PairListArgs a{};
        a.M = pl_n[neg];
        a.pairs = pl_pairs[neg].ptr;
        a.scales = pl_scales[neg].ptr;
        a.pair_sys = pl_sys[neg].ptr;
        a.red_idx = red_idx.ptr;
        a.x = xp;
        a.params = pl_params.ptr;
        a.boxes = boxp;
        a.plane = pl_plane.ptr;
        a.offset = pl_offset.ptr;
        a.lambdas = lam.ptr;
        a.params1 = pl_params1.n ? pl_params1.ptr : nullptr;
        a.cutoffs = pl_cutoffs.ptr;
        a.betas = pl_betas.ptr;
        a.negated = neg;
        return a;
    }

    // One observed frame at the current state of every system: du/dlambda,
    // du/dp and d(du/dlambda)/dp into the flat buffers, then the running sums.
    void observe(hipStream_t st, int tile_grid) {
        const int TPB = 256;
        launch_fill_u64(obs_dudp.ptr, Pg, 0, st);
        launch_fill_u64(obs_dldp.ptr, Pg, 0, st);
        launch_fill_u64(obs_dl.ptr, S, 0, st);
        if (has_nb) {
            nb.prepare(x.ptr, nb_params.ptr, nb_params1.ptr, box.ptr, nb_plane.ptr, nb_offset.ptr, lam.ptr, 0.0, st);
            nb.compute(box.ptr, nullptr, obs_dudp.ptr, obs_dl.ptr, nullptr, st, tile_grid);
            nb.compute_dl_dp(box.ptr, obs_dldp.ptr, st, tile_grid);
        }
        for (int neg = 0; neg < 2; neg++) {
            if (pl_n[neg] == 0) {
                continue;
            }
            const PairListArgs a = pl_args(neg, x.ptr, box.ptr);
            launch_pair_list(a, nullptr, obs_dudp.ptr, obs_dl.ptr, nullptr, st);
            launch_pair_list_dl_dp(a, obs_dldp.ptr, st);
        }
        for (int k = 0; k < 3; k++) {
            if (bt_n[k] == 0) {
                continue;
            }
            BondedTerms t{bt_n[k], bt_idx[k].ptr, bt_params[k].ptr, bt_mult[k].ptr, bt_offset[k].ptr, bt_sys[k].ptr};
            launch_bonded(static_cast<BondedKind>(k), t, x.ptr, lam.ptr, 0.0, red_idx.ptr, nullptr,
                          obs_dudp.ptr + bt_off[k], obs_dl.ptr, nullptr, st);
            launch_bonded(static_cast<BondedKind>(k), t, x.ptr, lam.ptr, 0.0, red_idx.ptr, nullptr,
                          obs_dldp.ptr + bt_off[k], nullptr, nullptr, st, true);
        }
        k_obs_accum<<<ceil_div(Pg, TPB), TPB, 0, st>>>(Pg, obs_sys.ptr, obs_inv_scale.ptr, obs_dl.ptr, obs_dudp.ptr,
                                                      obs_dldp.ptr, obs_sum_g.ptr, obs_sum_ag.ptr, obs_sum_h.ptr);
        TM_LAUNCH_CHECK();
        k_obs_accum_sys<<<ceil_div(S, 64), 64, 0, st>>>(S, obs_dl.ptr, obs_sum_a.ptr, obs_sum_aa.ptr, obs_n.ptr);
        TM_LAUNCH_CHECK();
    }
};

Engine::Engine(const std::vector<SystemSpec> &systems) : impl_(new Impl) {
    DeviceLock lock; // allocations, null-stream uploads, stream creation
    Impl &I = *impl_;
    S_ = static_cast<int>(systems.size());
    if (S_ == 0) {
        throw Error("Engine: no systems");
    }
    TM_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    I.S = S_;
    // layout
    int a0 = 0;
    for (auto &sp : systems) {
        if (!sp.state_on_device && (static_cast<int>(sp.x.size()) != sp.N * 3 ||
                                    static_cast<int>(sp.v.size()) != sp.N * 3 || sp.box.size() != 9)) {
            throw Error("Engine: bad system state shapes");
        }
        if (static_cast<int>(sp.integ.cbs.size()) != sp.N || static_cast<int>(sp.integ.ccs.size()) != sp.N) {
            throw Error("Engine: integrator coefficient size does not match N");
        }
        I.atom0.push_back(a0);
        I.natoms.push_back(sp.N);
        a0 += sp.N;
    }
    A_ = I.A = a0;
    h_n_ = I.natoms;
    h_atom0_ = I.atom0;

    std::vector<double> hx(A_ * 3), hv(A_ * 3), hbox(S_ * 9), hcbs(A_), hccs(A_), hca(S_), hdt(S_), hlam(S_, 0.0);
    std::vector<uint64_t> hseed(S_);
    std::vector<int64_t> hstep(S_);
    std::vector<int> hatom_sys(A_), hred(S_);
    for (int s = 0; s < S_; s++) {
        const SystemSpec &sp = systems[s];
        if (!sp.state_on_device) {
            std::copy(sp.x.begin(), sp.x.end(), hx.begin() + I.atom0[s] * 3);
            std::copy(sp.v.begin(), sp.v.end(), hv.begin() + I.atom0[s] * 3);
            std::copy(sp.box.begin(), sp.box.end(), hbox.begin() + s * 9);
        }
        std::copy(sp.integ.cbs.begin(), sp.integ.cbs.end(), hcbs.begin() + I.atom0[s]);
        std::copy(sp.integ.ccs.begin(), sp.integ.ccs.end(), hccs.begin() + I.atom0[s]);
        hca[s] = sp.integ.ca;
        hdt[s] = sp.integ.dt;
        hseed[s] = sp.integ.seed;
        hstep[s] = sp.step;
        hred[s] = s;
        for (int i = 0; i < sp.N; i++) {
            hatom_sys[I.atom0[s] + i] = s;
        }
    }
    // state of systems handed over from another engine stays on the device
    // (copy_state_from); only the others are uploaded
    I.x.alloc(std::max(A_ * 3, 1));
    I.v.alloc(std::max(A_ * 3, 1));
    I.box.alloc(S_ * 9);
    for (int s = 0; s < S_; s++) {
        if (systems[s].state_on_device) {
            continue;
        }
        const size_t a3 = static_cast<size_t>(I.atom0[s]) * 3, n3 = static_cast<size_t>(I.natoms[s]) * 3;
        TM_HIP_CHECK(hipMemcpy(I.x.ptr + a3, hx.data() + a3, n3 * sizeof(double), hipMemcpyHostToDevice));
        TM_HIP_CHECK(hipMemcpy(I.v.ptr + a3, hv.data() + a3, n3 * sizeof(double), hipMemcpyHostToDevice));
        TM_HIP_CHECK(hipMemcpy(I.box.ptr + s * 9, hbox.data() + s * 9, 9 * sizeof(double), hipMemcpyHostToDevice));
    }
    I.x_after.alloc(std::max(A_ * 3, 1));
    I.box_after.alloc(S_ * 9);
    upload_vec(I.cbs, hcbs);
    upload_vec(I.ccs, hccs);
    upload_vec(I.ca, hca);
    upload_vec(I.dt, hdt);
    upload_vec(I.lam, hlam);
    upload_vec(I.seed, hseed);
    upload_vec(I.step, hstep);
    upload_vec(I.atom_sys, hatom_sys);
    upload_vec(I.d_atom0, I.atom0);
    upload_vec(I.d_natoms, I.natoms);
    upload_vec(I.red_idx, hred);
    I.sys_status.alloc(S_);
    I.sys_status.zero();
    I.run_step.alloc(S_);
    I.run_step.zero();
    I.du_dx.alloc(std::max(A_ * 3, 1));
    I.du_dx.zero();
    I.u_buf.alloc(S_);
    I.du_dl_buf.alloc(S_);
    I.u_buf.zero();
    I.du_dl_buf.zero();

    // potentials
    std::vector<double> h_nb_params;
    std::vector<Flat> flats(S_);
    for (int s = 0; s < S_; s++) {
        for (auto &bp : systems[s].bps) {
            flatten(bp->potential, bp->h_params.data(), bp->size(), flats[s]);
        }
    }
    // One device nonbonded state over every system's term of one kind (the
    // all-pairs potential, or the interaction group): parameters, plane and
    // offset indices packed in atom order, one descriptor per system.
    // Returns whether the parameters are interpolated.
    auto build_nb = [&](auto term, const char *what, NbState &st, DeviceArray<double> &d_params,
                        DeviceArray<double> &d_params1, DeviceArray<int> &d_plane, DeviceArray<int> &d_offset,
                        std::vector<double> &h_params) {
        std::vector<NbSys> ys(S_);
        std::vector<int> subset;
        std::vector<double> params(A_ * 3, 0.0), params1;
        std::vector<int> plane(A_, 0), offset(A_, 0);
        const bool interp0 = term(0).first->interp_;
        if (interp0) {
            params1.assign(A_ * 3, 0.0);
        }
        for (int s = 0; s < S_; s++) {
            const auto &nb = term(s).first;
            const std::vector<double> &fp = *term(s).second;
            if (nb->N_ != I.natoms[s]) {
                throw Error(std::string("Engine: ") + what + " N does not match the system");
            }
            if (nb->interp_ != interp0) {
                throw Error("Engine: either every system or none has interpolated nonbonded parameters");
            }
            if (fp.size() != static_cast<size_t>((interp0 ? 6 : 3) * I.natoms[s])) {
                throw Error("Engine: nonbonded parameter size mismatch");
            }
            NbSys y{};
            y.atom0 = I.atom0[s];
            y.n_atoms = I.natoms[s];
            y.k = nb->K_;
            y.red_idx = s;
            y.subset0 = nb->subset_.empty() ? -1 : static_cast<int>(subset.size());
            subset.insert(subset.end(), nb->subset_.begin(), nb->subset_.end());
            y.nrow = nb->nrow_;
            y.sorted = nb->sorted_ ? 1 : 0;
            y.interp = interp0 ? 1 : 0;
            y.beta = static_cast<float>(nb->beta_);
            y.cutoff_d = nb->cutoff_;
            y.padding_d = nb->padding_;
            ys[s] = y;
            const size_t n3 = static_cast<size_t>(I.natoms[s]) * 3;
            std::copy(fp.begin(), fp.begin() + n3, params.begin() + I.atom0[s] * 3);
            if (interp0) {
                std::copy(fp.begin() + n3, fp.begin() + 2 * n3, params1.begin() + I.atom0[s] * 3);
            }
            std::copy(nb->plane_.begin(), nb->plane_.end(), plane.begin() + I.atom0[s]);
            std::copy(nb->offset_.begin(), nb->offset_.end(), offset.begin() + I.atom0[s]);
        }
        st.nb_margin = NB_LIST_MARGIN;
        st.init(ys, subset);
        h_params = params;
        upload_vec(d_params, params);
        if (interp0) {
            upload_vec(d_params1, params1);
        }
        upload_vec(d_plane, plane);
        upload_vec(d_offset, offset);
        return interp0;
    };
    int n_nb = 0, n_ig = 0;
    for (auto &f : flats) {
        n_nb += f.nb ? 1 : 0;
        n_ig += f.ig ? 1 : 0;
    }
    if (n_nb != 0 && n_nb != S_) {
        throw Error("Engine: either every system or none has a NonbondedAllPairs");
    }
    if (n_ig != 0 && n_ig != S_) {
        throw Error("Engine: either every system or none has a NonbondedInteractionGroup");
    }
    I.has_nb = n_nb == S_;
    if (I.has_nb) {
        I.interp = build_nb([&](int s) { return std::make_pair(flats[s].nb, &flats[s].nb_params); },
                            "NonbondedAllPairs", I.nb, I.nb_params, I.nb_params1, I.nb_plane, I.nb_offset,
                            h_nb_params);
    }
    I.has_ig = n_ig == S_;
    if (I.has_ig) {
        std::vector<double> h_ig_params;
        const bool ig_interp =
            build_nb([&](int s) { return std::make_pair(flats[s].ig, &flats[s].ig_params); },
                     "NonbondedInteractionGroup", I.ig, I.ig_params, I.ig_params1, I.ig_plane, I.ig_offset,
                     h_ig_params);
        I.interp = I.interp || ig_interp;
    }
    // pair lists
    {
        std::vector<int> pairs[2], psys[2];
        std::vector<double> scales[2];
        std::vector<double> params(A_ * 3, 0.0), params1, cutoffs(S_, 1.0);
        std::vector<float> betas(S_, 0.0f);
        std::vector<int> plane(A_, 0), offset(A_, 0);
        int pl_interp = -1; // unknown, 0 or 1: every pair list of the batch must agree
        for (int s = 0; s < S_; s++) {
            const std::vector<double> *first = nullptr;
            const NonbondedPairList *first_pl = nullptr;
            for (auto &pr : flats[s].pls) {
                const auto &pl = pr.first;
                if (pl->N_ != I.natoms[s]) {
                    throw Error("Engine: NonbondedPairList N does not match the system");
                }
                if (pl_interp < 0) {
                    pl_interp = pl->interp_ ? 1 : 0;
                    if (pl_interp) {
                        params1.assign(A_ * 3, 0.0);
                    }
                } else if (pl_interp != (pl->interp_ ? 1 : 0)) {
                    throw Error("Engine: either every pair list or none has interpolated parameters");
                }
                const size_t n3 = static_cast<size_t>(I.natoms[s]) * 3;
                if (pr.second.size() != (pl_interp ? 2 : 1) * n3) {
                    throw Error("Engine: pair-list parameter size mismatch");
                }
                if (!first) {
                    first = &pr.second;
                    first_pl = pl.get();
                    std::copy(pr.second.begin(), pr.second.begin() + n3, params.begin() + I.atom0[s] * 3);
                    if (pl_interp) {
                        std::copy(pr.second.begin() + n3, pr.second.begin() + 2 * n3, params1.begin() + I.atom0[s] * 3);
                    }
                    std::copy(pl->plane_.begin(), pl->plane_.end(), plane.begin() + I.atom0[s]);
                    std::copy(pl->offset_.begin(), pl->offset_.end(), offset.begin() + I.atom0[s]);
                    cutoffs[s] = pl->cutoff_;
                    betas[s] = static_cast<float>(pl->beta_);
                } else if (*first != pr.second || pl->cutoff_ != cutoffs[s] ||
                           static_cast<float>(pl->beta_) != betas[s] || pl->plane_ != first_pl->plane_ ||
                           pl->offset_ != first_pl->offset_) {
                    // the merged lists share one per-atom parameter, plane and offset array
                    throw Error("Engine: pair lists of one system must share parameters, lambda plane and offset "
                                "indices, beta and cutoff");
                }
                const int neg = pl->negated_ ? 1 : 0;
                for (int m = 0; m < pl->M_; m++) {
                    pairs[neg].push_back(pl->pairs_[m * 2 + 0] + I.atom0[s]);
                    pairs[neg].push_back(pl->pairs_[m * 2 + 1] + I.atom0[s]);
                    scales[neg].push_back(pl->scales_[m * 2 + 0]);
                    scales[neg].push_back(pl->scales_[m * 2 + 1]);
                    psys[neg].push_back(s);
                }
            }
        }
        for (int neg = 0; neg < 2; neg++) {
            I.pl_n[neg] = static_cast<int>(psys[neg].size());
            I.h_pl_pairs[neg] = pairs[neg];
            upload_vec(I.pl_pairs[neg], pairs[neg]);
            upload_vec(I.pl_scales[neg], scales[neg]);
            upload_vec(I.pl_sys[neg], psys[neg]);
        }
        I.has_pl = I.pl_n[0] + I.pl_n[1] > 0;
        if (I.has_pl) {
            // every system with pair lists must share the all-pairs parameters for the
            // observer's du/dp of the pair lists to land on the nonbonded rows
            for (int s = 0; s < S_ && I.pl_shares_nb; s++) {
                if (flats[s].pls.empty()) {
                    continue;
                }
                if (!I.has_nb || !std::equal(params.begin() + I.atom0[s] * 3,
                                             params.begin() + (I.atom0[s] + I.natoms[s]) * 3,
                                             h_nb_params.begin() + I.atom0[s] * 3)) {
                    I.pl_shares_nb = false;
                }
            }
        }
        upload_vec(I.pl_params, params);
        if (pl_interp == 1) {
            upload_vec(I.pl_params1, params1);
            I.interp = true;
        }
        upload_vec(I.pl_plane, plane);
        upload_vec(I.pl_offset, offset);
        upload_vec(I.pl_cutoffs, cutoffs);
        upload_vec(I.pl_betas, betas);
    }
    // bonded
    {
        std::vector<int> idx[3], mult[3], off[3], sys[3];
        std::vector<double> prm[3];
        for (int s = 0; s < S_; s++) {
            for (auto &pr : flats[s].bonded) {
                const auto &b = pr.first;
                const int k = static_cast<int>(b->bk_);
                const int ar = bonded_arity(b->bk_);
                for (int t = 0; t < b->n_; t++) {
                    for (int q = 0; q < ar; q++) {
                        const int ai = b->idxs_[t * ar + q];
                        if (ai < 0 || ai >= I.natoms[s]) {
                            throw Error("Engine: bonded index out of range");
                        }
                        idx[k].push_back(ai + I.atom0[s]);
                    }
                    mult[k].push_back(b->mult_[t]);
                    off[k].push_back(b->offset_[t]);
                    sys[k].push_back(s);
                }
                if (static_cast<int>(pr.second.size()) != b->n_ * bonded_nparams(b->bk_)) {
                    throw Error("Engine: bonded parameter size mismatch");
                }
                prm[k].insert(prm[k].end(), pr.second.begin(), pr.second.end());
            }
        }
        for (int k = 0; k < 3; k++) {
            I.bt_first[k].assign(S_, 0);
            I.bt_count[k].assign(S_, 0);
            for (size_t t = 0; t < sys[k].size(); t++) {
                I.bt_count[k][sys[k][t]] += 1;
            }
            for (int s = 1; s < S_; s++) {
                I.bt_first[k][s] = I.bt_first[k][s - 1] + I.bt_count[k][s - 1];
            }
            I.bt_n[k] = static_cast<int>(sys[k].size());
            upload_vec(I.bt_idx[k], idx[k]);
            upload_vec(I.bt_mult[k], mult[k]);
            upload_vec(I.bt_offset[k], off[k]);
            upload_vec(I.bt_sys[k], sys[k]);
            upload_vec(I.bt_params[k], prm[k]);
        }
        if (I.bt_n[0] + I.bt_n[1] + I.bt_n[2] + I.pl_n[0] + I.pl_n[1] > 0) {
            // force slots of bonded terms and pair lists, gathered per atom
            std::vector<int> inc_off;
            std::vector<uint32_t> inc_items;
            build_bonded_incidence(A_, {&idx[0], &idx[1], &idx[2], &I.h_pl_pairs[0], &I.h_pl_pairs[1]}, inc_off,
                                   inc_items, I.bt_slot0);
            upload_vec(I.bt_inc_off, inc_off);
            upload_vec(I.bt_inc_items, inc_items);
            I.bt_slots.alloc((idx[0].size() + idx[1].size() + idx[2].size() + I.h_pl_pairs[0].size() +
                              I.h_pl_pairs[1].size()) *
                             3);
            I.bt_gather = true;
        }
        I.h_pl_pairs[0].clear();
        I.h_pl_pairs[1].clear();
    }
    // barostat
    {
        std::vector<int> has(S_, 0), interval(S_, 1), bstep(S_, 0), nmol(S_, 0), acc(S_, 0), att(S_, 0);
        std::vector<uint64_t> bseed(S_, 0);
        std::vector<double> pressure(S_, 0.0), kT(S_, 0.0), vscale(S_, 0.0);
        std::vector<int> gatom, gmol, msys, mcount, amol(A_, -1);
        for (int s = 0; s < S_; s++) {
            const auto &b = systems[s].baro;
            if (!b) {
                continue;
            }
            if (b->interval <= 0) {
                throw Error("Barostat interval must be greater than 0");
            }
            if (b->N != I.natoms[s]) {
                throw Error("Engine: barostat N does not match the system");
            }
            has[s] = 1;
            interval[s] = b->interval;
            bstep[s] = systems[s].baro_step >= 0 ? systems[s].baro_step % b->interval
                                                 : static_cast<int>(systems[s].step % b->interval);
            bseed[s] = b->seed;
            pressure[s] = b->pressure * BARO_AVOGADRO * 1e-25;
            kT[s] = BARO_BOLTZ * b->temperature;
            nmol[s] = static_cast<int>(b->groups.size());
            vscale[s] = systems[s].baro_vscale;
            acc[s] = systems[s].baro_accepted;
            att[s] = systems[s].baro_attempted;
            for (auto &grp : b->groups) {
                const int m = static_cast<int>(msys.size());
                msys.push_back(s);
                mcount.push_back(static_cast<int>(grp.size()));
                for (int a : grp) {
                    if (a < 0 || a >= I.natoms[s]) {
                        throw Error("Grouped indices must be between 0 and N");
                    }
                    gatom.push_back(a + I.atom0[s]);
                    gmol.push_back(m);
                    if (amol[a + I.atom0[s]] != -1) {
                        throw Error("All grouped indices must be unique");
                    }
                    amol[a + I.atom0[s]] = m;
                }
            }
        }
        I.has_baro = std::any_of(has.begin(), has.end(), [](int h) { return h != 0; });
        I.h_has_baro = has;
        I.h_interval = interval;
        I.h_bstep = bstep;
        I.G = static_cast<int>(gatom.size());
        I.nmol_total = static_cast<int>(msys.size());
        upload_vec(I.b_has, has);
        upload_vec(I.b_interval, interval);
        upload_vec(I.b_bstep, bstep);
        upload_vec(I.b_nmol, nmol);
        upload_vec(I.b_accepted, acc);
        upload_vec(I.b_attempted, att);
        upload_vec(I.b_seed, bseed);
        upload_vec(I.b_pressure, pressure);
        upload_vec(I.b_kT, kT);
        upload_vec(I.b_vscale, vscale);
        I.b_trig.alloc(S_);
        I.b_accept.alloc(S_);
        I.b_trig.zero();
        I.b_accept.zero();
        I.b_u_before.alloc(S_);
        I.b_u_after.alloc(S_);
        I.b_dvol.alloc(S_);
        I.b_lscale.alloc(S_);
        I.b_r2.alloc(S_);
        upload_vec(I.g_atom, gatom);
        upload_vec(I.g_mol, gmol);
        upload_vec(I.mol_sys, msys);
        upload_vec(I.mol_count, mcount);
        upload_vec(I.atom_mol, amol);
        I.centroids.alloc(std::max(I.nmol_total * 3, 1));
    }
    device_synchronize();
}

Engine::~Engine() {
    DeviceLock lock;
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
    if (impl_) {
        impl_->clear_graphs();
    }
    impl_.reset();
}

void Engine::set_lambdas(const std::vector<double> &lams) {
    if (static_cast<int>(lams.size()) != S_) {
        throw Error("Engine::set_lambdas: expected one lambda per system");
    }
    impl_->lam.upload(lams.data(), S_, stream_);
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::run(int n_steps, const double *lambda_sched, int du_dl_interval, int x_interval, double *du_dl_out,
                 double *x_out, double *box_out, bool use_graph, const std::vector<double> *u_windows,
                 int u_interval, double *u_out) {
    Impl &I = *impl_;
    if (n_steps <= 0) {
        return;
    }
    const int W = u_windows ? static_cast<int>(u_windows->size()) : 0;
    const int Fu = (W > 0 && u_interval > 0) ? n_steps / u_interval : 0;
    DeviceArray<double> d_windows(std::max(W, 1)), lam_eval(S_);
    DeviceArray<uint64_t> u_traj(std::max<size_t>(static_cast<size_t>(S_) * Fu * W, 1));
    if (Fu > 0) {
        d_windows.upload(u_windows->data(), W, stream_);
        launch_fill_u64(I.u_buf.ptr, S_, 0, stream_);
    }
    const int K = du_dl_interval > 0 ? n_steps / du_dl_interval : 0;
    const int F = x_interval > 0 ? n_steps / x_interval : 0;
    DeviceArray<double> sched;
    if (lambda_sched) {
        sched.alloc(static_cast<size_t>(S_) * n_steps);
        sched.upload(lambda_sched, static_cast<size_t>(S_) * n_steps, stream_);
    }
    // Runs without a lambda schedule, frames or window energies keep their graph
    // executables on the engine (the serving loop and the benchmark call run()
    // once per chunk of steps); du/dl then goes to the engine's run_dl, with a
    // row stride (Kst) that does not change from run to run.
    const bool keep_graphs = use_graph && !lambda_sched && Fu == 0 && x_interval <= 0;
    DeviceArray<uint64_t> dl_local;
    DeviceArray<double> frames(keep_graphs ? 1 : std::max<size_t>(static_cast<size_t>(F) * A_ * 3, 1));
    DeviceArray<double> box_local;
    uint64_t *dl_ptr = nullptr;
    double *box_frames_ptr = nullptr;
    int Kst = K;
    if (keep_graphs) {
        if (I.run_dl_cap < std::max(K, 1)) {
            const int cap = std::max({K, 2 * I.run_dl_cap, 64});
            I.clear_graphs(); // their launches point at the old buffer
            I.run_dl.alloc(static_cast<size_t>(S_) * cap);
            I.run_dl_cap = cap;
        }
        if (I.run_box_frames.n == 0) {
            I.run_box_frames.alloc(9);
        }
        dl_ptr = I.run_dl.ptr;
        Kst = I.run_dl_cap;
        box_frames_ptr = I.run_box_frames.ptr;
    } else {
        dl_local.alloc(std::max<size_t>(static_cast<size_t>(S_) * K, 1));
        box_local.alloc(std::max<size_t>(static_cast<size_t>(S_) * F * 9, 1));
        dl_ptr = dl_local.ptr;
        box_frames_ptr = box_local.ptr;
    }
    launch_fill_i32(I.run_step.ptr, S_, 0, stream_);
    launch_fill_u64(I.du_dl_buf.ptr, S_, 0, stream_);

    // Deferred barostat moves. A move due after step j needs U(x) and U(x')
    // at the state after the Langevin update, and step j + 1 then computes the
    // forces at x or x'. When nothing reads the state in between (no frame or
    // window energies stored at step j, not the run's last step) and no
    // system's lambda changes between the steps, the move is evaluated at the
    // start of step j + 1 instead, with forces: U(x) comes with the forces at
    // x, U(x') with the forces at x', and each system keeps the forces of the
    // state it ends in. The same values in the same order as evaluating the
    // move first, two force evaluations instead of one plus two energy-only
    // ones.
    const bool can_defer = I.has_baro;
    if (can_defer && I.dx_trial.n == 0) {
        I.dx_trial.alloc(static_cast<size_t>(A_) * 3);
        I.dl_trial.alloc(S_);
    }
    DeviceArray<uint64_t> &dx_trial = I.dx_trial;
    DeviceArray<uint64_t> &dl_trial = I.dl_trial;
    uint64_t *const dl_buf = du_dl_interval > 0 ? I.du_dl_buf.ptr : nullptr;

    const int TPB = 256;
    const int obs_k = I.obs_interval;
    auto one_step = [&](bool baro_now, bool obs_now, bool u_now, bool defer_now, bool resolve_now) {
        if (lambda_sched) {
            k_set_lambda<<<ceil_div(S_, 64), 64, 0, stream_>>>(S_, sched.ptr, n_steps, I.run_step.ptr, I.lam.ptr);
            TM_LAUNCH_CHECK();
        }
        launch_fill_u64(I.du_dx.ptr, static_cast<size_t>(A_) * 3, 0, stream_);
        if (resolve_now) {
            // the move set up at the end of the previous step
            const BaroDev b = I.baro_dev();
            launch_fill_u64(dx_trial.ptr, static_cast<size_t>(A_) * 3, 0, stream_);
            launch_fill_u64(dl_trial.ptr, S_, 0, stream_);
            I.eval(I.x.ptr, I.box.ptr, I.du_dx.ptr, dl_buf, I.b_u_before.ptr, stream_, tile_grid);
            if (I.nmol_total > 0) {
                k_baro_zero_centroids<<<ceil_div(I.nmol_total, TPB), TPB, 0, stream_>>>(
                    I.nmol_total, I.mol_sys.ptr, I.b_trig.ptr, I.centroids.ptr);
                TM_LAUNCH_CHECK();
                k_baro_centroids<<<ceil_div(I.G, TPB), TPB, 0, stream_>>>(I.G, I.g_atom.ptr, I.g_mol.ptr,
                                                                       I.mol_sys.ptr, I.b_trig.ptr, I.x.ptr,
                                                                       I.centroids.ptr);
                TM_LAUNCH_CHECK();
            }
            k_baro_rescale<<<ceil_div(A_, TPB), TPB, 0, stream_>>>(
                A_, I.atom_sys.ptr, I.atom_mol.ptr, I.mol_count.ptr, I.b_trig.ptr, I.b_lscale.ptr, I.box.ptr,
                I.centroids.ptr, I.x.ptr, I.x_after.ptr);
            TM_LAUNCH_CHECK();
            I.eval(I.x_after.ptr, I.box_after.ptr, dx_trial.ptr, dl_buf ? dl_trial.ptr : nullptr, I.b_u_after.ptr,
                   stream_, tile_grid);
            k_baro_decide<<<ceil_div(S_, 64), 64, 0, stream_>>>(S_, b, I.box.ptr, I.box_after.ptr);
            TM_LAUNCH_CHECK();
            k_baro_apply<<<ceil_div(A_, TPB), TPB, 0, stream_>>>(A_, I.atom_sys.ptr, I.b_accept.ptr, I.x_after.ptr,
                                                               I.x.ptr);
            TM_LAUNCH_CHECK();
            k_baro_take_trial<<<ceil_div(std::max(A_, S_), TPB), TPB, 0, stream_>>>(
                A_, S_, I.atom_sys.ptr, I.b_accept.ptr, dx_trial.ptr, I.du_dx.ptr, dl_trial.ptr, dl_buf);
            TM_LAUNCH_CHECK();
        } else {
            I.eval(I.x.ptr, I.box.ptr, I.du_dx.ptr, dl_buf, nullptr, stream_, tile_grid);
        }
        if (obs_now) {
            I.observe(stream_, tile_grid);
        }
        k_langevin<<<ceil_div(A_, TPB), TPB, 0, stream_>>>(A_, I.atom_sys.ptr, I.d_atom0.ptr, I.step.ptr, I.seed.ptr,
                                                         I.ca.ptr, I.dt.ptr, I.cbs.ptr, I.ccs.ptr, I.du_dx.ptr,
                                                         I.x.ptr, I.v.ptr, I.sys_status.ptr);
        TM_LAUNCH_CHECK();
        if (I.has_baro) {
            const BaroDev b = I.baro_dev();
            k_baro_setup<<<ceil_div(S_, 64), 64, 0, stream_>>>(S_, b, I.step.ptr, I.box.ptr, I.box_after.ptr);
            TM_LAUNCH_CHECK();
            if (baro_now && !defer_now) {
                I.eval(I.x.ptr, I.box.ptr, nullptr, nullptr, I.b_u_before.ptr, stream_, tile_grid);
                if (I.nmol_total > 0) {
                    k_baro_zero_centroids<<<ceil_div(I.nmol_total, TPB), TPB, 0, stream_>>>(
                        I.nmol_total, I.mol_sys.ptr, I.b_trig.ptr, I.centroids.ptr);
                    TM_LAUNCH_CHECK();
                    k_baro_centroids<<<ceil_div(I.G, TPB), TPB, 0, stream_>>>(I.G, I.g_atom.ptr, I.g_mol.ptr,
                                                                           I.mol_sys.ptr, I.b_trig.ptr, I.x.ptr,
                                                                           I.centroids.ptr);
                    TM_LAUNCH_CHECK();
                }
                k_baro_rescale<<<ceil_div(A_, TPB), TPB, 0, stream_>>>(
                    A_, I.atom_sys.ptr, I.atom_mol.ptr, I.mol_count.ptr, I.b_trig.ptr, I.b_lscale.ptr, I.box.ptr,
                    I.centroids.ptr, I.x.ptr, I.x_after.ptr);
                TM_LAUNCH_CHECK();
                I.eval(I.x_after.ptr, I.box_after.ptr, nullptr, nullptr, I.b_u_after.ptr, stream_, tile_grid);
                k_baro_decide<<<ceil_div(S_, 64), 64, 0, stream_>>>(S_, b, I.box.ptr, I.box_after.ptr);
                TM_LAUNCH_CHECK();
                k_baro_apply<<<ceil_div(A_, TPB), TPB, 0, stream_>>>(A_, I.atom_sys.ptr, I.b_accept.ptr,
                                                                   I.x_after.ptr, I.x.ptr);
                TM_LAUNCH_CHECK();
            }
        }
        if (u_now) {
            // energies of the state after this step at every window (context.cu:139-233)
            for (int w = 0; w < W; w++) {
                k_set_eval_lambda<<<ceil_div(S_, 64), 64, 0, stream_>>>(S_, d_windows.ptr, w, lam_eval.ptr);
                TM_LAUNCH_CHECK();
                I.eval(I.x.ptr, I.box.ptr, nullptr, nullptr, I.u_buf.ptr, stream_, tile_grid, lam_eval.ptr);
                k_store_u<<<ceil_div(S_, 64), 64, 0, stream_>>>(S_, I.run_step.ptr, u_interval, Fu, W, w, I.u_buf.ptr,
                                                              u_traj.ptr);
                TM_LAUNCH_CHECK();
            }
        }
        if (x_interval > 0) {
            k_store_frames2<<<ceil_div(A_, TPB), TPB, 0, stream_>>>(A_, I.atom_sys.ptr, I.d_atom0.ptr,
                                                                  I.d_natoms.ptr, I.run_step.ptr, x_interval, F,
                                                                  I.x.ptr, frames.ptr);
            TM_LAUNCH_CHECK();
        }
        k_step_end<<<ceil_div(S_, 64), 64, 0, stream_>>>(S_, I.step.ptr, I.run_step.ptr, I.du_dl_buf.ptr,
                                                        du_dl_interval, Kst, dl_ptr, x_interval, F, I.box.ptr,
                                                        box_frames_ptr);
        TM_LAUNCH_CHECK();
    };

    // which steps of the run need the barostat energies (host mirror of the
    // per-system barostat call counters)
    auto baro_at = [&](int j) {
        if (!I.has_baro) {
            return false;
        }
        for (int s = 0; s < S_; s++) {
            if (I.h_has_baro[s] && (I.h_bstep[s] + j + 1) % I.h_interval[s] == 0) {
                return true;
            }
        }
        return false;
    };

    // observed frames: the states at which du_dl is stored with interval obs_k
    // (the state at the start of run step j, (j + 1) % obs_k == 0)
    auto obs_at = [&](int j) { return obs_k > 0 && (j + 1) % obs_k == 0; };
    // energies at the windows: the state after run step j, (j + 1) % u_interval == 0
    auto u_at = [&](int j) { return Fu > 0 && (j + 1) % u_interval == 0; };
    // a barostat move after step j evaluated at the start of step j + 1 (above)
    auto same_lambda = [&](int j) {
        if (!lambda_sched) {
            return true;
        }
        for (int s = 0; s < S_; s++) {
            if (lambda_sched[static_cast<size_t>(s) * n_steps + j] !=
                lambda_sched[static_cast<size_t>(s) * n_steps + j + 1]) {
                return false;
            }
        }
        return true;
    };
    auto defer_at = [&](int j) {
        return can_defer && baro_at(j) && j + 1 < n_steps && !u_at(j) &&
               !(x_interval > 0 && (j + 1) % x_interval == 0) && same_lambda(j);
    };
    auto resolve_at = [&](int j) { return j > 0 && defer_at(j - 1); };

    int j = 0;
    if (use_graph) {
        int chunk = 10;
        if (I.has_baro) {
            chunk = I.h_interval[0];
            for (int s = 0; s < S_; s++) {
                if (I.h_has_baro[s]) {
                    chunk = std::max(chunk, I.h_interval[s]);
                }
            }
        }
        // per-run executables unless this run keeps them on the engine
        std::map<std::vector<int>, hipGraphExec_t> run_cache;
        std::map<std::vector<int>, hipGraphExec_t> &cache = keep_graphs ? I.graphs : run_cache;
        while (j + chunk <= n_steps) {
            // the step pattern and every value the captured launches bake in
            std::vector<int> pattern(chunk);
            for (int q = 0; q < chunk; q++) {
                pattern[q] = (baro_at(j + q) ? 1 : 0) | (obs_at(j + q) ? 2 : 0) | (u_at(j + q) ? 4 : 0) |
                             (defer_at(j + q) ? 8 : 0) | (resolve_at(j + q) ? 16 : 0);
            }
            pattern.insert(pattern.end(), {du_dl_interval, Kst, tile_grid, I.nb.tile_waves, obs_k});
            auto it = cache.find(pattern);
            if (it == cache.end()) {
                DeviceLock lock; // no other thread may allocate or synchronise the device meanwhile
                hipGraph_t g;
                TM_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                for (int q = 0; q < chunk; q++) {
                    one_step((pattern[q] & 1) != 0, (pattern[q] & 2) != 0, (pattern[q] & 4) != 0,
                             (pattern[q] & 8) != 0, (pattern[q] & 16) != 0);
                }
                TM_HIP_CHECK(hipStreamEndCapture(stream_, &g));
                hipGraphExec_t ge;
                TM_HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
                TM_HIP_CHECK(hipGraphDestroy(g));
                it = cache.emplace(pattern, ge).first;
            }
            TM_HIP_CHECK(hipGraphLaunch(it->second, stream_));
            j += chunk;
        }
        TM_HIP_CHECK(hipStreamSynchronize(stream_));
        DeviceLock lock;
        for (auto &kv : run_cache) {
            TM_HIP_CHECK(hipGraphExecDestroy(kv.second));
        }
    }
    for (; j < n_steps; j++) {
        one_step(baro_at(j), obs_at(j), u_at(j), defer_at(j), resolve_at(j));
    }
    for (int s = 0; s < S_; s++) {
        if (I.h_has_baro[s]) {
            I.h_bstep[s] = (I.h_bstep[s] + n_steps) % I.h_interval[s];
        }
    }
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    if (du_dl_out && K > 0) {
        std::vector<uint64_t> h(static_cast<size_t>(S_) * K);
        TM_HIP_CHECK(hipMemcpy2DAsync(h.data(), sizeof(uint64_t) * K, dl_ptr, sizeof(uint64_t) * Kst,
                                      sizeof(uint64_t) * K, S_, hipMemcpyDeviceToHost, stream_));
        TM_HIP_CHECK(hipStreamSynchronize(stream_));
        for (size_t i = 0; i < h.size(); i++) {
            du_dl_out[i] = fixed_to_double(h[i], FIXED_SCALE);
        }
    }
    if (x_out && F > 0) {
        frames.download(x_out, static_cast<size_t>(F) * A_ * 3, stream_);
    }
    if (u_out && Fu > 0) {
        std::vector<uint64_t> h(static_cast<size_t>(S_) * Fu * W);
        u_traj.download(h.data(), h.size(), stream_);
        TM_HIP_CHECK(hipStreamSynchronize(stream_));
        for (size_t i = 0; i < h.size(); i++) {
            u_out[i] = fixed_to_double(h[i], FIXED_SCALE);
        }
    }
    if (box_out && F > 0) {
        box_local.download(box_out, static_cast<size_t>(S_) * F * 9, stream_);
    }
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::set_tile_waves(int waves) {
    if (waves < 2 || waves > 4) {
        throw Error("tile kernel waves per SIMD must be 2, 3 or 4");
    }
    Impl &I = *impl_;
    I.nb.tile_waves = waves;
    I.ig.tile_waves = waves;
}

void Engine::set_build_pieces(int pieces) {
    if (pieces != 0 && pieces != 1 && pieces != 2 && pieces != 4 && pieces != 8) {
        throw Error("neighborlist build pieces must be 0 (automatic), 1, 2, 4 or 8");
    }
    Impl &I = *impl_;
    I.nb.pieces_override = pieces;
    I.ig.pieces_override = pieces;
}

void Engine::set_nblist_padding(double padding) {
    Impl &I = *impl_;
    if (I.has_nb) {
        I.nb.set_padding(padding, stream_);
    }
    if (I.has_ig) {
        I.ig.set_padding(padding, stream_);
    }
}

void Engine::copy_state_from(Engine &src, int src_s, int dst_s) {
    Impl &I = *impl_;
    Impl &J = *src.impl_;
    if (src_s < 0 || src_s >= src.S_ || dst_s < 0 || dst_s >= S_) {
        throw Error("Engine::copy_state_from: system index out of range");
    }
    if (src.h_n_[src_s] != h_n_[dst_s]) {
        throw Error("Engine::copy_state_from: atom counts differ");
    }
    if ((J.h_has_baro[src_s] != 0) != (I.h_has_baro[dst_s] != 0)) {
        throw Error("Engine::copy_state_from: barostat presence differs");
    }
    TM_HIP_CHECK(hipStreamSynchronize(src.stream_));
    const size_t n3 = static_cast<size_t>(h_n_[dst_s]) * 3;
    auto d2d = [&](void *dst, const void *from, size_t bytes) {
        TM_HIP_CHECK(hipMemcpyAsync(dst, from, bytes, hipMemcpyDeviceToDevice, stream_));
    };
    d2d(I.x.ptr + static_cast<size_t>(h_atom0_[dst_s]) * 3, J.x.ptr + static_cast<size_t>(src.h_atom0_[src_s]) * 3,
        n3 * sizeof(double));
    d2d(I.v.ptr + static_cast<size_t>(h_atom0_[dst_s]) * 3, J.v.ptr + static_cast<size_t>(src.h_atom0_[src_s]) * 3,
        n3 * sizeof(double));
    d2d(I.box.ptr + dst_s * 9, J.box.ptr + src_s * 9, 9 * sizeof(double));
    d2d(I.step.ptr + dst_s, J.step.ptr + src_s, sizeof(int64_t));
    if (I.h_has_baro[dst_s]) {
        d2d(I.b_vscale.ptr + dst_s, J.b_vscale.ptr + src_s, sizeof(double));
        d2d(I.b_accepted.ptr + dst_s, J.b_accepted.ptr + src_s, sizeof(int));
        d2d(I.b_attempted.ptr + dst_s, J.b_attempted.ptr + src_s, sizeof(int));
        d2d(I.b_bstep.ptr + dst_s, J.b_bstep.ptr + src_s, sizeof(int));
        I.h_bstep[dst_s] = J.h_bstep[src_s];
    }
    if (I.has_nb) {
        I.nb.invalidate(stream_); // the new arena layout needs new neighborlists
    }
    if (I.has_ig) {
        I.ig.invalidate(stream_);
    }
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}

// Flat parameter space of the gradient observer and the trajectory adjoint:
// nonbonded [A, 3] rows, then bond [B, 2], angle [A, 2], torsion [T, 3] terms;
// the system and the fixed-point exponent of every element.
void Engine::flat_layout() {
    Impl &I = *impl_;
    if (I.Pg > 0) {
        return;
    }
    static const int np[3] = {2, 2, 3};
    I.bt_off[0] = A_ * 3;
    I.bt_off[1] = I.bt_off[0] + I.bt_n[0] * np[0];
    I.bt_off[2] = I.bt_off[1] + I.bt_n[1] * np[1];
    const int Pg = I.bt_off[2] + I.bt_n[2] * np[2];
    std::vector<int> esys(Pg);
    std::vector<double> inv(Pg, 1.0 / FIXED_SCALE);
    for (int a = 0; a < A_; a++) {
        const int s = static_cast<int>(std::upper_bound(I.atom0.begin(), I.atom0.end(), a) - I.atom0.begin()) - 1;
        for (int c = 0; c < 3; c++) {
            esys[a * 3 + c] = s;
        }
        inv[a * 3 + 1] = 1.0 / FIXED_SCALE_DSIG;
        inv[a * 3 + 2] = 1.0 / FIXED_SCALE_DEPS;
    }
    for (int k = 0; k < 3; k++) {
        for (int s = 0; s < S_; s++) {
            for (int t = I.bt_first[k][s]; t < I.bt_first[k][s] + I.bt_count[k][s]; t++) {
                for (int c = 0; c < np[k]; c++) {
                    esys[I.bt_off[k] + t * np[k] + c] = s;
                }
            }
        }
    }
    upload_vec(I.obs_sys, esys);
    upload_vec(I.obs_inv_scale, inv);
    I.Pg = Pg;
}

// System s's elements of a flat [Pg] device array, in grad_stats order.
void Engine::download_flat(const double *src, int s, double *out) {
    Impl &I = *impl_;
    static const int np[3] = {2, 2, 3};
    std::vector<std::pair<size_t, size_t>> ranges; // (device offset, count) in output order
    ranges.emplace_back(static_cast<size_t>(h_atom0_[s]) * 3, static_cast<size_t>(h_n_[s]) * 3);
    for (int k = 0; k < 3; k++) {
        ranges.emplace_back(static_cast<size_t>(I.bt_off[k]) + static_cast<size_t>(I.bt_first[k][s]) * np[k],
                            static_cast<size_t>(I.bt_count[k][s]) * np[k]);
    }
    size_t o = 0;
    for (auto &r : ranges) {
        if (r.second > 0) {
            TM_HIP_CHECK(hipMemcpyAsync(out + o, src + r.first, r.second * sizeof(double), hipMemcpyDeviceToHost,
                                        stream_));
        }
        o += r.second;
    }
}

void Engine::set_grad_observer(int interval) {
    Impl &I = *impl_;
    if (interval < 0) {
        throw Error("Engine::set_grad_observer: interval must be >= 0");
    }
    if (interval > 0 && !I.pl_shares_nb) {
        throw Error("Engine::set_grad_observer: pair lists must share the nonbonded parameters");
    }
    if (interval > 0 && I.interp) {
        throw Error("Engine::set_grad_observer: not available for interpolated nonbonded parameters");
    }
    if (interval > 0 && I.has_ig) {
        throw Error("Engine::set_grad_observer: not available with a NonbondedInteractionGroup");
    }
    flat_layout();
    if (I.obs_sum_g.n == 0) {
        I.clear_graphs(); // kept graphs point at the observer's old buffers
        I.obs_dudp.alloc(I.Pg);
        I.obs_dldp.alloc(I.Pg);
        I.obs_dl.alloc(S_);
        I.obs_sum_g.alloc(I.Pg);
        I.obs_sum_ag.alloc(I.Pg);
        I.obs_sum_h.alloc(I.Pg);
        I.obs_sum_a.alloc(S_);
        I.obs_sum_aa.alloc(S_);
        I.obs_n.alloc(S_);
    }
    I.obs_interval = interval;
    I.obs_sum_g.zero();
    I.obs_sum_ag.zero();
    I.obs_sum_h.zero();
    I.obs_sum_a.zero();
    I.obs_sum_aa.zero();
    I.obs_n.zero();
    device_synchronize();
}

void Engine::observe() {
    Impl &I = *impl_;
    if (I.obs_sum_g.n == 0) {
        throw Error("Engine::observe: call set_grad_observer first");
    }
    I.observe(stream_, tile_grid);
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}

int Engine::grad_size(int s) const {
    const Impl &I = *impl_;
    static const int np[3] = {2, 2, 3};
    int P = h_n_[s] * 3;
    for (int k = 0; k < 3; k++) {
        P += I.bt_count[k][s] * np[k];
    }
    return P;
}

void Engine::grad_stats(int s, int *n_frames, double *sum_a, double *sum_aa, double *sum_g, double *sum_ag,
                        double *sum_h) {
    Impl &I = *impl_;
    if (I.obs_sum_g.n == 0) {
        throw Error("Engine::grad_stats: call set_grad_observer first");
    }
    download_flat(I.obs_sum_g.ptr, s, sum_g);
    download_flat(I.obs_sum_ag.ptr, s, sum_ag);
    download_flat(I.obs_sum_h.ptr, s, sum_h);
    TM_HIP_CHECK(hipMemcpyAsync(n_frames, I.obs_n.ptr + s, sizeof(int), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(sum_a, I.obs_sum_a.ptr + s, sizeof(double), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(sum_aa, I.obs_sum_aa.ptr + s, sizeof(double), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::adjoint_forward(int n_steps, int du_dl_interval, int checkpoint_interval, double *du_dl_out) {
    Impl &I = *impl_;
    if (I.has_baro) {
        throw Error("Engine::adjoint: not available with a barostat (its moves are not differentiable)");
    }
    if (I.interp || I.has_ig || !I.pl_shares_nb) {
        throw Error("Engine::adjoint: needs plain nonbonded parameters shared by the pair lists and no "
                    "interaction group");
    }
    if (n_steps <= 0 || du_dl_interval <= 0 || checkpoint_interval <= 0 ||
        checkpoint_interval % du_dl_interval != 0) {
        throw Error("Engine::adjoint_forward: need n_steps > 0 and a checkpoint interval that is a positive "
                    "multiple of the du_dl interval");
    }
    flat_layout();
    const int K = checkpoint_interval;
    const int nseg = ceil_div(n_steps, K);
    const int F = n_steps / du_dl_interval;
    const size_t n3 = static_cast<size_t>(A_) * 3;
    I.ck_x.resize(nseg + 1);
    I.ck_v.resize(nseg + 1);
    I.ck_step.resize(nseg + 1);
    auto save = [&](int c) {
        if (I.ck_x[c].n != std::max<size_t>(n3, 1)) {
            I.ck_x[c].alloc(std::max<size_t>(n3, 1));
            I.ck_v[c].alloc(std::max<size_t>(n3, 1));
            I.ck_step[c].alloc(S_);
        }
        TM_HIP_CHECK(hipMemcpyAsync(I.ck_x[c].ptr, I.x.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        TM_HIP_CHECK(hipMemcpyAsync(I.ck_v[c].ptr, I.v.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        TM_HIP_CHECK(hipMemcpyAsync(I.ck_step[c].ptr, I.step.ptr, S_ * sizeof(int64_t), hipMemcpyDeviceToDevice,
                                    stream_));
    };
    I.adj_steps = 0; // invalid until the run completes
    std::vector<double> dl;
    for (int c = 0; c < nseg; c++) {
        save(c);
        const int Kc = std::min(K, n_steps - c * K);
        const int Fc = Kc / du_dl_interval;
        dl.assign(static_cast<size_t>(S_) * std::max(Fc, 1), 0.0);
        run(Kc, nullptr, du_dl_interval, 0, Fc > 0 ? dl.data() : nullptr, nullptr, nullptr, true);
        const int f0 = c * (K / du_dl_interval);
        for (int s = 0; du_dl_out && s < S_; s++) {
            for (int f = 0; f < Fc; f++) {
                du_dl_out[static_cast<size_t>(s) * F + f0 + f] = dl[static_cast<size_t>(s) * Fc + f];
            }
        }
    }
    save(nseg);
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    I.adj_steps = n_steps;
    I.adj_dl_interval = du_dl_interval;
    I.adj_ck_interval = K;
}

void Engine::adjoint_backward(const double *dl_bar, double *grad_out, double *x0_bar, double *v0_bar, int horizon) {
    Impl &I = *impl_;
    if (I.adj_steps == 0) {
        throw Error("Engine::adjoint_backward: call adjoint_forward first");
    }
    if (horizon < 0) {
        throw Error("Engine::adjoint_backward: horizon must be >= 0");
    }
    const int n_steps = I.adj_steps, K = I.adj_ck_interval, interval = I.adj_dl_interval;
    const int nseg = ceil_div(n_steps, K);
    const int F = n_steps / interval;
    const size_t n3 = static_cast<size_t>(A_) * 3;
    const int TPB = 256;
    const int gA = ceil_div(std::max(A_, 1), TPB), gA3 = ceil_div(std::max<size_t>(n3, 1), TPB);
    const int gP = ceil_div(std::max(I.Pg, 1), TPB), gS = ceil_div(S_, 64);
    DeviceArray<double> d_dl_bar(std::max<size_t>(static_cast<size_t>(S_) * F, 1));
    if (F > 0) {
        d_dl_bar.upload(dl_bar, static_cast<size_t>(S_) * F, stream_);
    }
    DeviceArray<double> xb(std::max<size_t>(n3, 1)), vb(std::max<size_t>(n3, 1)), gb(std::max<size_t>(n3, 1)),
        vx(std::max<size_t>(n3, 1)), pb(std::max(I.Pg, 1)), scale(S_), vls(S_);
    DeviceArray<uint64_t> hx(std::max<size_t>(n3, 1)), hp(std::max(I.Pg, 1)), smax(S_);
    DeviceArray<double> seg(std::max<size_t>(static_cast<size_t>(std::min(K, n_steps)) * n3, 1));
    DeviceArray<int> flag(1);
    xb.zero(stream_);
    vb.zero(stream_);
    pb.zero(stream_);
    flag.zero(stream_);
    auto restore = [&](int c) {
        TM_HIP_CHECK(hipMemcpyAsync(I.x.ptr, I.ck_x[c].ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        TM_HIP_CHECK(hipMemcpyAsync(I.v.ptr, I.ck_v[c].ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        TM_HIP_CHECK(hipMemcpyAsync(I.step.ptr, I.ck_step[c].ptr, S_ * sizeof(int64_t), hipMemcpyDeviceToDevice,
                                    stream_));
    };
    // one step of run() without barostat, frames or observer (the same launches)
    auto plain_step = [&]() {
        launch_fill_u64(I.du_dx.ptr, n3, 0, stream_);
        I.eval(I.x.ptr, I.box.ptr, I.du_dx.ptr, I.du_dl_buf.ptr, nullptr, stream_, tile_grid);
        k_langevin<<<gA, TPB, 0, stream_>>>(A_, I.atom_sys.ptr, I.d_atom0.ptr, I.step.ptr, I.seed.ptr, I.ca.ptr,
                                            I.dt.ptr, I.cbs.ptr, I.ccs.ptr, I.du_dx.ptr, I.x.ptr, I.v.ptr,
                                            I.sys_status.ptr);
        TM_LAUNCH_CHECK();
        k_step_end<<<gS, 64, 0, stream_>>>(S_, I.step.ptr, I.run_step.ptr, I.du_dl_buf.ptr, 1, 0, nullptr, 0, 0,
                                          I.box.ptr, nullptr);
        TM_LAUNCH_CHECK();
    };
    HvpTangent tan;
    tan.vx = vx.ptr;
    tan.vls = vls.ptr;
    for (int c = nseg - 1; c >= 0; c--) {
        const int Kc = std::min(K, n_steps - c * K);
        restore(c);
        launch_fill_i32(I.run_step.ptr, S_, 0, stream_);
        launch_fill_u64(I.du_dl_buf.ptr, S_, 0, stream_);
        for (int t = 0; t < Kc; t++) {
            TM_HIP_CHECK(hipMemcpyAsync(seg.ptr + static_cast<size_t>(t) * n3, I.x.ptr, n3 * sizeof(double),
                                        hipMemcpyDeviceToDevice, stream_));
            plain_step();
        }
        // the recomputed segment ends where the forward run's did
        k_adj_compare<<<gA3, TPB, 0, stream_>>>(n3, I.x.ptr, I.ck_x[c + 1].ptr, flag.ptr);
        TM_LAUNCH_CHECK();
        for (int t = Kc - 1; t >= 0; t--) {
            const int j = c * K + t;
            const double *xt = seg.ptr + static_cast<size_t>(t) * n3;
            launch_fill_u64(smax.ptr, S_, 0, stream_);
            k_adj_pre<<<gA, TPB, 0, stream_>>>(A_, I.atom_sys.ptr, I.ca.ptr, I.dt.ptr, I.cbs.ptr, xb.ptr, vb.ptr,
                                               gb.ptr, reinterpret_cast<unsigned long long *>(smax.ptr));
            TM_LAUNCH_CHECK();
            k_adj_scale<<<gS, 64, 0, stream_>>>(S_, j, interval, F, d_dl_bar.ptr,
                                               reinterpret_cast<const unsigned long long *>(smax.ptr), scale.ptr,
                                               vls.ptr);
            TM_LAUNCH_CHECK();
            k_adj_scale_vec<<<gA3, TPB, 0, stream_>>>(A_, I.atom_sys.ptr, scale.ptr, gb.ptr, vx.ptr);
            TM_LAUNCH_CHECK();
            launch_fill_u64(hx.ptr, n3, 0, stream_);
            launch_fill_u64(hp.ptr, I.Pg, 0, stream_);
            if (I.has_nb) {
                I.nb.prepare(xt, I.nb_params.ptr, nullptr, I.box.ptr, I.nb_plane.ptr, I.nb_offset.ptr, I.lam.ptr,
                             0.0, stream_);
                launch_tiles_hvp(I.nb, I.nb_params.ptr, I.nb_offset.ptr, I.box.ptr, tan, HvpOut{hx.ptr, hp.ptr, nullptr},
                                 stream_);
            }
            for (int neg = 0; neg < 2; neg++) {
                if (I.pl_n[neg] > 0) {
                    launch_pair_list_hvp(I.pl_args(neg, xt, I.box.ptr), tan, HvpOut{hx.ptr, hp.ptr, nullptr}, stream_);
                }
            }
            for (int k = 0; k < 3; k++) {
                if (I.bt_n[k] == 0) {
                    continue;
                }
                BondedTerms bt{I.bt_n[k], I.bt_idx[k].ptr, I.bt_params[k].ptr, I.bt_mult[k].ptr, I.bt_offset[k].ptr,
                               I.bt_sys[k].ptr};
                launch_bonded_hvp(static_cast<BondedKind>(k), bt, xt, I.lam.ptr, 0.0, I.red_idx.ptr, tan,
                                  HvpOut{hx.ptr, hp.ptr + I.bt_off[k], nullptr}, stream_);
            }
            k_adj_post_x<<<gA3, TPB, 0, stream_>>>(A_, I.atom_sys.ptr, scale.ptr, hx.ptr, xb.ptr);
            TM_LAUNCH_CHECK();
            k_adj_post_p<<<gP, TPB, 0, stream_>>>(I.Pg, I.obs_sys.ptr, I.obs_inv_scale.ptr, scale.ptr, hp.ptr, pb.ptr);
            TM_LAUNCH_CHECK();
            if (horizon > 0 && j > 0 && j % horizon == 0) {
                // truncated sweep: sensitivities stop at the start of the block
                launch_fill_u64(reinterpret_cast<uint64_t *>(xb.ptr), n3, 0, stream_);
                launch_fill_u64(reinterpret_cast<uint64_t *>(vb.ptr), n3, 0, stream_);
            }
        }
    }
    restore(nseg); // the state after the forward run
    int mismatch = 0;
    flag.download(&mismatch, 1, stream_);
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    if (mismatch) {
        throw Error("Engine::adjoint_backward: a recomputed segment differs from the forward run");
    }
    if (grad_out) {
        size_t o = 0;
        for (int s = 0; s < S_; s++) {
            download_flat(pb.ptr, s, grad_out + o);
            o += grad_size(s);
        }
    }
    if (x0_bar) {
        xb.download(x0_bar, n3, stream_);
    }
    if (v0_bar) {
        vb.download(v0_bar, n3, stream_);
    }
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::fire_minimize(int n_sched, const double *lambda_sched, const int *active, bool raise) {
    Impl &I = *impl_;
    if (n_sched < 1) {
        throw Error("fire_minimize: the lambda schedule needs at least one entry");
    }
    const int A3 = A_ * 3;
    // schedule, step-major on the device: entry k of system s at k * S + s
    std::vector<double> sched(static_cast<size_t>(n_sched) * S_);
    std::vector<double> cur(S_);
    if (!lambda_sched) {
        I.lam.download(cur.data(), S_, stream_);
        TM_HIP_CHECK(hipStreamSynchronize(stream_));
    }
    for (int k = 0; k < n_sched; k++) {
        for (int s = 0; s < S_; s++) {
            sched[static_cast<size_t>(k) * S_ + s] = lambda_sched ? lambda_sched[static_cast<size_t>(s) * n_sched + k]
                                                                  : cur[s];
        }
    }
    std::vector<int> act(S_, 1);
    if (active) {
        for (int s = 0; s < S_; s++) {
            act[s] = active[s] ? 1 : 0;
        }
    }
    DeviceArray<double> d_sched, fv, ff;
    DeviceArray<float> dt, alpha;
    DeviceArray<int> npos, d_act;
    upload_vec(d_sched, sched);
    upload_vec(d_act, act);
    fv.alloc(std::max(A3, 1));
    ff.alloc(std::max(A3, 1));
    dt.alloc(S_);
    alpha.alloc(S_);
    npos.alloc(S_);
    const int TPB = 256;
    const int gA = ceil_div(std::max(A3, 1), TPB);
    k_fire_state_init<<<ceil_div(S_, 64), 64, 0, stream_>>>(S_, dt.ptr, alpha.ptr, npos.ptr);
    TM_LAUNCH_CHECK();
    launch_fill_u64(I.du_dx.ptr, A3, 0, stream_);
    I.eval(I.x.ptr, I.box.ptr, I.du_dx.ptr, nullptr, nullptr, stream_, tile_grid, d_sched.ptr);
    k_fire_init<<<gA, TPB, 0, stream_>>>(A3, I.atom_sys.ptr, d_act.ptr, I.du_dx.ptr, fv.ptr, ff.ptr);
    TM_LAUNCH_CHECK();
    if (n_sched > 1) {
        // one step as a graph, replayed: the step's lambdas come from a device
        // step counter, so every replay is the same launch sequence
        DeviceArray<int> done(1);
        DeviceArray<double> lam_cur(S_);
        done.zero(stream_);
        hipGraph_t g;
        hipGraphExec_t ge;
        {
            DeviceLock lock;
            TM_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            k_fire_move<<<gA, TPB, 0, stream_>>>(A3, I.atom_sys.ptr, d_act.ptr, dt.ptr, fv.ptr, ff.ptr, I.x.ptr);
            TM_LAUNCH_CHECK();
            launch_fill_u64(I.du_dx.ptr, A3, 0, stream_);
            k_fire_lambda<<<ceil_div(S_, 64), 64, 0, stream_>>>(S_, n_sched, d_sched.ptr, done.ptr, lam_cur.ptr);
            TM_LAUNCH_CHECK();
            I.eval(I.x.ptr, I.box.ptr, I.du_dx.ptr, nullptr, nullptr, stream_, tile_grid, lam_cur.ptr);
            k_fire_update<<<S_, FIRE_THREADS, 0, stream_>>>(I.d_atom0.ptr, I.d_natoms.ptr, d_act.ptr, I.du_dx.ptr,
                                                            fv.ptr, ff.ptr, dt.ptr, alpha.ptr, npos.ptr);
            TM_LAUNCH_CHECK();
            k_fire_count<<<1, 64, 0, stream_>>>(done.ptr);
            TM_LAUNCH_CHECK();
            TM_HIP_CHECK(hipStreamEndCapture(stream_, &g));
            TM_HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            TM_HIP_CHECK(hipGraphDestroy(g));
        }
        for (int k = 1; k < n_sched; k++) {
            TM_HIP_CHECK(hipGraphLaunch(ge, stream_));
        }
        TM_HIP_CHECK(hipStreamSynchronize(stream_));
        DeviceLock lock;
        TM_HIP_CHECK(hipGraphExecDestroy(ge));
    }
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    if (raise) {
        check_status();
    }
}

void Engine::max_force(double *out, bool raise) {
    Impl &I = *impl_;
    launch_fill_u64(I.du_dx.ptr, A_ * 3, 0, stream_);
    I.eval(I.x.ptr, I.box.ptr, I.du_dx.ptr, nullptr, nullptr, stream_, tile_grid);
    DeviceArray<double> d(S_);
    k_max_force<<<S_, 256, 0, stream_>>>(I.d_atom0.ptr, I.d_natoms.ptr, I.du_dx.ptr, d.ptr);
    TM_LAUNCH_CHECK();
    d.download(out, S_, stream_);
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    if (raise) {
        check_status();
    }
}

void Engine::energies(double *u_out, double *du_dl_out) {
    Impl &I = *impl_;
    launch_fill_u64(I.u_buf.ptr, S_, 0, stream_);
    launch_fill_u64(I.du_dl_buf.ptr, S_, 0, stream_);
    I.eval(I.x.ptr, I.box.ptr, nullptr, I.du_dl_buf.ptr, I.u_buf.ptr, stream_, tile_grid);
    std::vector<uint64_t> hu(S_), hl(S_);
    I.u_buf.download(hu.data(), S_, stream_);
    I.du_dl_buf.download(hl.data(), S_, stream_);
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    launch_fill_u64(I.du_dl_buf.ptr, S_, 0, stream_);
    for (int s = 0; s < S_; s++) {
        if (u_out) {
            u_out[s] = fixed_to_double(hu[s], FIXED_SCALE);
        }
        if (du_dl_out) {
            du_dl_out[s] = fixed_to_double(hl[s], FIXED_SCALE);
        }
    }
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::get_x(int s, double *out) {
    TM_HIP_CHECK(hipMemcpyAsync(out, impl_->x.ptr + h_atom0_[s] * 3, sizeof(double) * h_n_[s] * 3,
                                hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}
void Engine::get_v(int s, double *out) {
    TM_HIP_CHECK(hipMemcpyAsync(out, impl_->v.ptr + h_atom0_[s] * 3, sizeof(double) * h_n_[s] * 3,
                                hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}
void Engine::get_box(int s, double *out) {
    TM_HIP_CHECK(hipMemcpyAsync(out, impl_->box.ptr + s * 9, sizeof(double) * 9, hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}
void Engine::get_du_dx(int s, uint64_t *out) {
    TM_HIP_CHECK(hipMemcpyAsync(out, impl_->du_dx.ptr + h_atom0_[s] * 3, sizeof(uint64_t) * h_n_[s] * 3,
                                hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}
void Engine::set_x(int s, const double *in) {
    TM_HIP_CHECK(hipMemcpyAsync(impl_->x.ptr + h_atom0_[s] * 3, in, sizeof(double) * h_n_[s] * 3,
                                hipMemcpyHostToDevice, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}
void Engine::set_v(int s, const double *in) {
    TM_HIP_CHECK(hipMemcpyAsync(impl_->v.ptr + h_atom0_[s] * 3, in, sizeof(double) * h_n_[s] * 3,
                                hipMemcpyHostToDevice, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}
void Engine::set_box(int s, const double *in) {
    TM_HIP_CHECK(
        hipMemcpyAsync(impl_->box.ptr + s * 9, in, sizeof(double) * 9, hipMemcpyHostToDevice, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}
int64_t Engine::get_step(int s) {
    int64_t v = 0;
    TM_HIP_CHECK(hipMemcpyAsync(&v, impl_->step.ptr + s, sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    return v;
}
void Engine::get_barostat_state(int s, double *vscale, int *accepted, int *attempted, int *baro_step) {
    Impl &I = *impl_;
    *baro_step = I.h_bstep[s];
    TM_HIP_CHECK(hipMemcpyAsync(vscale, I.b_vscale.ptr + s, sizeof(double), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(accepted, I.b_accepted.ptr + s, sizeof(int), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(attempted, I.b_attempted.ptr + s, sizeof(int), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Engine::get_barostat_move(int s, double *out) {
    Impl &I = *impl_;
    for (int k = 0; k < 9; k++) {
        out[k] = 0.0;
    }
    if (s < 0 || s >= S_) {
        throw Error("system index out of range");
    }
    if (!I.has_baro || !I.h_has_baro[s]) {
        return;
    }
    int trig = 0, acc = 0;
    uint64_t ub = 0, ua = 0;
    double dv = 0.0, r2 = 0.0, vs = 0.0, ls = 0.0;
    TM_HIP_CHECK(hipMemcpyAsync(&trig, I.b_trig.ptr + s, sizeof(int), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(&acc, I.b_accept.ptr + s, sizeof(int), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(&ub, I.b_u_before.ptr + s, sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(&ua, I.b_u_after.ptr + s, sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(&dv, I.b_dvol.ptr + s, sizeof(double), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(&r2, I.b_r2.ptr + s, sizeof(double), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(&vs, I.b_vscale.ptr + s, sizeof(double), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(&ls, I.b_lscale.ptr + s, sizeof(double), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    out[0] = trig;
    out[1] = acc;
    out[2] = fixed_to_double(ub, FIXED_SCALE);
    out[3] = fixed_to_double(ua, FIXED_SCALE);
    out[4] = fixed_to_double(ua - ub, FIXED_SCALE); // the energy change exactly as the decision saw it
    out[5] = dv;
    out[6] = r2;
    out[7] = vs;
    out[8] = ls;
}

void Engine::system_status(int *out) {
    Impl &I = *impl_;
    std::vector<int> st(S_, 0), st_ig(S_, 0), st_sys(S_, 0);
    if (I.has_nb) {
        I.nb.status.download(st.data(), S_, stream_);
    }
    if (I.has_ig) {
        I.ig.status.download(st_ig.data(), S_, stream_);
    }
    I.sys_status.download(st_sys.data(), S_, stream_);
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    for (int s = 0; s < S_; s++) {
        out[s] = st[s] | st_ig[s] | st_sys[s];
    }
}

void Engine::check_status() {
    std::vector<int> st(S_, 0);
    system_status(st.data());
    for (int s = 0; s < S_; s++) {
        if (st[s] & NB_STATUS_BOX_TOO_SMALL) {
            throw Error("Cutoff with padding is more than half of the box width, neighborlist is no longer reliable");
        }
        if (st[s] & NB_STATUS_BOX_NOT_ORTHO) {
            throw Error("Provided non-ortholinear box, unable to compute nonbonded energy");
        }
        if (st[s] & NB_STATUS_TILE_OVERFLOW) {
            throw Error("Engine: neighborlist tile buffer overflow");
        }
        if (st[s] & ENGINE_STATUS_FIXED_INVALID) {
            throw Error("Engine: system " + std::to_string(s) +
                        ": a force left the fixed-point range (non-finite or huge pair terms)");
        }
        if (st[s] & ENGINE_STATUS_NONFINITE) {
            throw Error("Engine: system " + std::to_string(s) + ": coordinates or velocities are not finite");
        }
    }
}

void Engine::profile(int reps, double *ms_out) {
    Impl &I = *impl_;
    hipEvent_t e0, e1;
    TM_HIP_CHECK(hipEventCreate(&e0));
    TM_HIP_CHECK(hipEventCreate(&e1));
    auto timed = [&](auto &&fn) {
        fn(); // warm
        TM_HIP_CHECK(hipEventRecord(e0, stream_));
        for (int r = 0; r < reps; r++) {
            fn();
        }
        TM_HIP_CHECK(hipEventRecord(e1, stream_));
        TM_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        TM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        return static_cast<double>(ms) / reps;
    };
    ms_out[0] = ms_out[1] = ms_out[2] = ms_out[3] = 0.0;
    if (I.has_nb) {
        ms_out[0] = timed([&] {
            I.nb.invalidate(stream_);
            I.nb.prepare(I.x.ptr, I.nb_params.ptr, I.nb_params1.ptr, I.box.ptr, I.nb_plane.ptr, I.nb_offset.ptr,
                         I.lam.ptr, 0.0, stream_);
        });
        ms_out[1] = timed([&] { I.nb.compute(I.box.ptr, I.du_dx.ptr, nullptr, nullptr, nullptr, stream_, tile_grid); });
    }
    ms_out[2] = timed([&] {
        const bool had = I.has_nb;
        I.has_nb = false;
        I.eval(I.x.ptr, I.box.ptr, I.du_dx.ptr, nullptr, nullptr, stream_, tile_grid);
        I.has_nb = had;
    });
    DeviceArray<double> xs(static_cast<size_t>(A_) * 3), vs(static_cast<size_t>(A_) * 3);
    DeviceArray<int> scratch_status(S_); // the timed copies must not flag the systems
    TM_HIP_CHECK(hipMemcpyAsync(xs.ptr, I.x.ptr, xs.bytes(), hipMemcpyDeviceToDevice, stream_));
    TM_HIP_CHECK(hipMemcpyAsync(vs.ptr, I.v.ptr, vs.bytes(), hipMemcpyDeviceToDevice, stream_));
    ms_out[3] = timed([&] {
        k_langevin<<<ceil_div(A_, 256), 256, 0, stream_>>>(A_, I.atom_sys.ptr, I.d_atom0.ptr, I.step.ptr, I.seed.ptr,
                                                         I.ca.ptr, I.dt.ptr, I.cbs.ptr, I.ccs.ptr, I.du_dx.ptr,
                                                         xs.ptr, vs.ptr, scratch_status.ptr);
    });
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    launch_fill_u64(I.du_dx.ptr, static_cast<size_t>(A_) * 3, 0, stream_);
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    TM_HIP_CHECK(hipEventDestroy(e0));
    TM_HIP_CHECK(hipEventDestroy(e1));
}

int Engine::total_tiles() {
    Impl &I = *impl_;
    if (!I.has_nb) {
        return 0;
    }
    int n = 0;
    TM_HIP_CHECK(hipMemcpyAsync(&n, I.nb.tile_prefix.ptr + S_, sizeof(int), hipMemcpyDeviceToHost, stream_));
    TM_HIP_CHECK(hipStreamSynchronize(stream_));
    return n;
}

size_t Engine::device_bytes() const {
    const Impl &I = *impl_;
    size_t b = I.x.bytes() * 3 + I.v.bytes() + I.du_dx.bytes() + I.cbs.bytes() * 2 + I.atom_sys.bytes();
    if (I.has_nb) {
        b += I.nb.tile_atoms.bytes() + I.nb.tile_row.bytes() + I.nb.pos.bytes() * 2 + I.nb.perm.bytes() * 2 +
             I.nb.nblist_x.bytes() + I.nb_params.bytes();
    }
    if (I.has_ig) {
        b += I.ig.tile_atoms.bytes() + I.ig.tile_row.bytes() + I.ig.pos.bytes() * 2 + I.ig.perm.bytes() * 2 +
             I.ig.nblist_x.bytes() + I.ig_params.bytes();
    }
    return b;
}

} // namespace tmamd
