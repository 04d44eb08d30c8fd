// Host-side potential classes (see potentials.h).
#include <cmath>
#include <cstring>
#include <numeric>
#include <set>

#include "fixed_point.h"
#include "hvp.h"
#include "potentials.h"

namespace tmamd {

static HvpTangent hvp_tangent(const double *vx, const double *vp, double vl) {
    HvpTangent t;
    t.vx = vx;
    t.vp = vp;
    t.vl = vl;
    return t;
}

void Potential::du_dp_fixed_to_float(int N, int P, const uint64_t *du_dp, double *out) const {
    for (int i = 0; i < P; i++) {
        out[i] = fixed_to_double(du_dp[i], FIXED_SCALE);
    }
}

void Potential::execute_host(int N, int P, const double *h_x, const double *h_p, const double *h_box, double lambda,
                             uint64_t *h_du_dx, uint64_t *h_du_dp, uint64_t *h_du_dl, uint64_t *h_u) {
    validate_host(N, P, h_box);
    DeviceArray<double> x(static_cast<size_t>(N) * 3), box(9), p(std::max(P, 1));
    x.upload(h_x, static_cast<size_t>(N) * 3);
    box.upload(h_box, 9);
    if (P > 0) {
        p.upload(h_p, P);
    }
    DeviceArray<uint64_t> du_dx, du_dp, du_dl, u;
    if (h_du_dx) {
        du_dx.alloc(static_cast<size_t>(N) * 3);
        du_dx.zero();
    }
    if (h_du_dp) {
        du_dp.alloc(std::max(P, 1));
        du_dp.zero();
    }
    if (h_du_dl) {
        du_dl.alloc(std::max(N, 1));
        du_dl.zero();
    }
    if (h_u) {
        u.alloc(std::max(N, 1));
        u.zero();
    }
    execute_device(N, P, x.ptr, P > 0 ? p.ptr : nullptr, box.ptr, lambda, du_dx.ptr, du_dp.ptr, du_dl.ptr, u.ptr, 0);
    check_device_status(0);
    if (h_du_dx) {
        du_dx.download(h_du_dx, static_cast<size_t>(N) * 3);
    }
    if (h_du_dp) {
        du_dp.download(h_du_dp, P);
    }
    if (h_du_dl) {
        du_dl.download(h_du_dl, N);
    }
    if (h_u) {
        u.download(h_u, N);
    }
    TM_HIP_CHECK(hipStreamSynchronize(0));
}

void Potential::execute_dl_dp_host(int N, int P, const double *h_x, const double *h_p, const double *h_box,
                                   double lambda, uint64_t *h_dl_dp) {
    validate_host(N, P, h_box);
    DeviceArray<double> x(static_cast<size_t>(N) * 3), box(9), p(std::max(P, 1));
    x.upload(h_x, static_cast<size_t>(N) * 3);
    box.upload(h_box, 9);
    if (P > 0) {
        p.upload(h_p, P);
    }
    DeviceArray<uint64_t> dl_dp(std::max(P, 1));
    dl_dp.zero();
    execute_dl_dp_device(N, P, x.ptr, P > 0 ? p.ptr : nullptr, box.ptr, lambda, dl_dp.ptr, 0);
    check_device_status(0);
    dl_dp.download(h_dl_dp, P);
    TM_HIP_CHECK(hipStreamSynchronize(0));
}

void Potential::execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                   double lambda, const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx,
                                   uint64_t *d_hp, uint64_t *d_hl, hipStream_t stream) {
    throw Error("execute_hvp: not available for this potential");
}

void Potential::hvp_at_interpolated(int N, const double *d_x, const double *p_lam, const double *d_box, double lambda,
                                    const double *d_vx, const double *vp_lam, double vl, uint64_t *d_hx, uint64_t *g,
                                    uint64_t *dg, uint64_t *d_hl, hipStream_t stream) {
    throw Error("execute_hvp: not an interpolated nonbonded potential");
}

void interpolated_hvp(const std::vector<Potential *> &pots, InterpHvpScratch &tmp, int N, const double *d_x,
                      const double *d_p, const double *d_box, double lambda, const double *d_vx, const double *d_vp,
                      double vl, uint64_t *d_hx, uint64_t *d_hp, uint64_t *d_hl, hipStream_t stream) {
    if (N == 0) {
        return;
    }
    const size_t n3 = static_cast<size_t>(N) * 3;
    tmp.reserve(N);
    double *p_lam = tmp.par.ptr, *vp_lam = tmp.par.ptr + n3;
    const bool chain = d_hp || d_hl; // du/dlambda needs the du/dp part too
    uint64_t *g = chain ? tmp.grad.ptr : nullptr, *dg = chain ? tmp.grad.ptr + n3 : nullptr;
    launch_interp_hvp_params(N, lambda, d_p, d_vp, vl, p_lam, vp_lam, stream);
    if (chain) {
        launch_fill_u64(tmp.grad.ptr, 2 * n3, 0, stream);
    }
    for (Potential *p : pots) {
        p->hvp_at_interpolated(N, d_x, p_lam, d_box, lambda, d_vx, vp_lam, vl, d_hx, g, dg, d_hl, stream);
    }
    if (chain) {
        launch_interp_hvp_chain(N, lambda, d_p, d_vp, vl, g, dg, d_hp, d_hl, stream);
    }
}

void Potential::execute_hvp_host(int N, int P, const double *h_x, const double *h_p, const double *h_box,
                                 double lambda, const double *h_vx, const double *h_vp, double vl, double *h_hx,
                                 double *h_hp, double *h_hl) {
    validate_host(N, P, h_box);
    const size_t n3 = static_cast<size_t>(N) * 3;
    // tangent scale 2^-e with the scaled tangent at most 1 in magnitude (exact)
    double vmax = std::fabs(vl);
    for (size_t i = 0; h_vx && i < n3; i++) {
        vmax = std::max(vmax, std::fabs(h_vx[i]));
    }
    for (int i = 0; h_vp && i < P; i++) {
        vmax = std::max(vmax, std::fabs(h_vp[i]));
    }
    if (!std::isfinite(vmax)) {
        throw Error("execute_hvp: tangent is not finite");
    }
    int e = 0;
    if (vmax > 0.0) {
        std::frexp(vmax, &e); // vmax = m 2^e, 0.5 <= m < 1
    }
    const double scale = std::ldexp(1.0, -e);
    DeviceArray<double> x(std::max<size_t>(n3, 1)), box(9), p(std::max(P, 1)), vx, vp;
    x.upload(h_x, n3);
    box.upload(h_box, 9);
    if (P > 0) {
        p.upload(h_p, P);
    }
    std::vector<double> tmp;
    if (h_vx) {
        tmp.resize(n3);
        for (size_t i = 0; i < n3; i++) {
            tmp[i] = h_vx[i] * scale;
        }
        vx.alloc(std::max<size_t>(n3, 1));
        vx.upload(tmp.data(), n3);
        TM_HIP_CHECK(hipStreamSynchronize(0)); // tmp is reused below
    }
    if (h_vp && P > 0) {
        tmp.resize(P);
        for (int i = 0; i < P; i++) {
            tmp[i] = h_vp[i] * scale;
        }
        vp.alloc(P);
        vp.upload(tmp.data(), P);
        TM_HIP_CHECK(hipStreamSynchronize(0));
    }
    DeviceArray<uint64_t> hx, hp, hl;
    if (h_hx) {
        hx.alloc(std::max<size_t>(n3, 1));
        hx.zero();
    }
    if (h_hp) {
        hp.alloc(std::max(P, 1));
        hp.zero();
    }
    if (h_hl) {
        hl.alloc(1);
        hl.zero();
    }
    execute_hvp_device(N, P, x.ptr, P > 0 ? p.ptr : nullptr, box.ptr, lambda, vx.ptr, vp.ptr, vl * scale, hx.ptr,
                       hp.ptr, hl.ptr, 0);
    check_device_status(0);
    std::vector<uint64_t> fixed;
    if (h_hx) {
        fixed.resize(std::max<size_t>(n3, 1));
        hx.download(fixed.data(), n3);
        TM_HIP_CHECK(hipStreamSynchronize(0));
        for (size_t i = 0; i < n3; i++) {
            h_hx[i] = fixed_to_double(fixed[i], FIXED_SCALE) / scale;
        }
    }
    if (h_hp) {
        fixed.resize(std::max(P, 1));
        hp.download(fixed.data(), P);
        TM_HIP_CHECK(hipStreamSynchronize(0));
        du_dp_fixed_to_float(N, P, fixed.data(), h_hp);
        for (int i = 0; i < P; i++) {
            h_hp[i] /= scale;
        }
    }
    if (h_hl) {
        uint64_t f = 0;
        hl.download(&f, 1);
        TM_HIP_CHECK(hipStreamSynchronize(0));
        h_hl[0] = fixed_to_double(f, FIXED_SCALE) / scale;
    }
    TM_HIP_CHECK(hipStreamSynchronize(0));
}

void Potential::execute_batch_host(int coord_batch_size, int N, int param_batch_size, int P, int lambda_batch_size,
                                   const double *h_x, const double *h_p, const double *h_box, const double *h_lambdas,
                                   uint64_t *h_du_dx, uint64_t *h_du_dp, uint64_t *h_du_dl, uint64_t *h_u) {
    for (int i = 0; i < coord_batch_size; i++) {
        validate_host(N, P, h_box + i * 9);
    }
    const size_t total = static_cast<size_t>(coord_batch_size) * param_batch_size * lambda_batch_size;
    DeviceArray<double> x(static_cast<size_t>(coord_batch_size) * N * 3), box(static_cast<size_t>(coord_batch_size) * 9);
    DeviceArray<double> p(std::max<size_t>(static_cast<size_t>(param_batch_size) * P, 1));
    x.upload(h_x, static_cast<size_t>(coord_batch_size) * N * 3);
    box.upload(h_box, static_cast<size_t>(coord_batch_size) * 9);
    if (P > 0) {
        p.upload(h_p, static_cast<size_t>(param_batch_size) * P);
    }
    DeviceArray<uint64_t> du_dx, du_dp, du_dl, u;
    if (h_du_dx) {
        du_dx.alloc(total * N * 3);
        du_dx.zero();
    }
    if (h_du_dp) {
        du_dp.alloc(std::max<size_t>(total * P, 1));
        du_dp.zero();
    }
    if (h_du_dl) {
        du_dl.alloc(total * N);
        du_dl.zero();
    }
    if (h_u) {
        u.alloc(total * N);
        u.zero();
    }
    for (int i = 0; i < coord_batch_size; i++) {
        for (int j = 0; j < param_batch_size; j++) {
            for (int k = 0; k < lambda_batch_size; k++) {
                const size_t o = (static_cast<size_t>(i) * param_batch_size + j) * lambda_batch_size + k;
                execute_device(N, P, x.ptr + static_cast<size_t>(i) * N * 3,
                               P > 0 ? p.ptr + static_cast<size_t>(j) * P : nullptr, box.ptr + i * 9, h_lambdas[k],
                               h_du_dx ? du_dx.ptr + o * N * 3 : nullptr, h_du_dp ? du_dp.ptr + o * P : nullptr,
                               h_du_dl ? du_dl.ptr + o * N : nullptr, h_u ? u.ptr + o * N : nullptr, 0);
            }
        }
    }
    check_device_status(0);
    if (h_du_dx) {
        du_dx.download(h_du_dx, total * N * 3);
    }
    if (h_du_dp) {
        du_dp.download(h_du_dp, total * P);
    }
    if (h_du_dl) {
        du_dl.download(h_du_dl, total * N);
    }
    if (h_u) {
        u.download(h_u, total * N);
    }
    TM_HIP_CHECK(hipStreamSynchronize(0));
}

void validate_nonbonded_box(const double *h_box, double cutoff, double padding) {
    const double db_cutoff = (cutoff + padding) * 2;
    for (int i = 0; i < 9; i++) {
        if (i == 0 || i == 4 || i == 8) {
            if (h_box[i] < db_cutoff) {
                throw Error("Cutoff with padding is more than half of the box width, neighborlist is no longer reliable");
            }
        } else if (h_box[i] != 0.0) {
            throw Error("Provided non-ortholinear box, unable to compute nonbonded energy");
        }
    }
}

// ---------------------------------------------------------------- all pairs

NonbondedAllPairs::NonbondedAllPairs(const std::vector<int> &plane, const std::vector<int> &offset, double beta,
                                     double cutoff, const std::vector<int> *atom_idxs, bool interpolated)
    : N_(static_cast<int>(offset.size())), K_(atom_idxs ? static_cast<int>(atom_idxs->size()) : N_), beta_(beta),
      cutoff_(cutoff), interp_(interpolated), plane_(plane), offset_(offset) {
    if (plane.size() != offset.size()) {
        throw Error("lambda offset idxs and plane idxs need to be equivalent");
    }
    if (atom_idxs) {
        std::set<int> uniq(atom_idxs->begin(), atom_idxs->end());
        if (uniq.size() != atom_idxs->size()) {
            throw Error("atom indices must be unique");
        }
        for (int a : *atom_idxs) {
            if (a < 0 || a >= N_) {
                throw Error("atom indices must be in [0, N)");
            }
        }
        subset_ = *atom_idxs;
    }
    d_plane_.alloc(std::max(N_, 1));
    d_offset_.alloc(std::max(N_, 1));
    d_plane_.upload(plane.data(), N_);
    d_offset_.upload(offset.data(), N_);
    device_synchronize();
}

void NonbondedAllPairs::set_nblist_padding(double padding) {
    padding_ = padding;
    dirty_ = true;
}

void NonbondedAllPairs::disable_hilbert_sort() {
    sorted_ = false;
    dirty_ = true;
}

NbSys NonbondedAllPairs::descriptor() const {
    NbSys y{};
    y.atom0 = 0;
    y.n_atoms = N_;
    y.k = K_;
    y.red_idx = 0;
    y.subset0 = subset_.empty() ? -1 : 0;
    y.nrow = nrow_;
    y.sorted = sorted_ ? 1 : 0;
    y.interp = interp_ ? 1 : 0;
    y.beta = static_cast<float>(beta_);
    y.cutoff_d = cutoff_;
    y.padding_d = padding_;
    return y;
}

void NonbondedAllPairs::ensure_state() {
    if (!dirty_) {
        return;
    }
    st_.nb_margin = NB_LIST_MARGIN;
    st_.init({descriptor()}, subset_);
    dirty_ = false;
}

void NonbondedAllPairs::execute_device_in(void *arena, size_t arena_bytes, int N, int P, const double *d_x,
                                          const double *d_p, const double *d_box, double lambda, uint64_t *d_du_dx,
                                          uint64_t *d_du_dp, uint64_t *d_du_dl, uint64_t *d_u, hipStream_t stream) {
    validate_sizes(N, P);
    if (interp_ || !subset_.empty() || nrow_ > 0) {
        throw Error(name_ + "::execute_device_noalloc(): takes all atoms with plain parameters "
                            "(no atom subset, row subset or interpolation)");
    }
    if (K_ == 0) {
        return;
    }
    const NbSys y = descriptor();
    const size_t need = NbState::view_bytes(y);
    if (need > arena_bytes) {
        throw Error(name_ + "::execute_device_noalloc(): N=" + std::to_string(N) + " needs " + std::to_string(need) +
                    " bytes of scratch, more than the " + std::to_string(arena_bytes) + "-byte maximum");
    }
    // a state whose arrays all lie in the arena: nothing is allocated, and the
    // first prepare() rebuilds (the arena holds the previous caller's data)
    NbState st;
    st.nb_margin = NB_LIST_MARGIN;
    st.init_view(y, arena, arena_bytes, stream);
    st.prepare(d_x, d_p, nullptr, d_box, d_plane_.ptr, d_offset_.ptr, nullptr, lambda, stream);
    st.compute(d_box, d_du_dx, d_du_dp, d_du_dl, d_u, stream);
}

void NonbondedAllPairs::validate_host(int N, int P, const double *h_box) const {
    validate_sizes(N, P);
    validate_nonbonded_box(h_box, cutoff_, padding_);
}

void NonbondedAllPairs::validate_sizes(int N, int P) const {
    if (N != N_) {
        throw Error(name_ + "::execute_device(): expected N == N_, got N=" + std::to_string(N) +
                    ", N_=" + std::to_string(N_));
    }
    const int M = interp_ ? 2 : 1;
    if (P != M * N_ * 3) {
        throw Error(name_ + "::execute_device(): expected P == M*N_*3, got P=" + std::to_string(P) +
                    ", M*N_*3=" + std::to_string(M * N_ * 3));
    }
}

void NonbondedAllPairs::execute_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                       double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl,
                                       uint64_t *d_u, hipStream_t stream) {
    validate_sizes(N, P);
    ensure_state();
    if (K_ == 0) {
        return;
    }
    st_.prepare(d_x, d_p, interp_ ? d_p + N_ * 3 : nullptr, d_box, d_plane_.ptr, d_offset_.ptr, nullptr, lambda,
                stream);
    if (interp_ && d_du_dp) {
        // du/dp at p(lambda), then split onto the initial and final parameters
        if (dp_tmp_.n != static_cast<size_t>(N_) * 3) {
            dp_tmp_.alloc(static_cast<size_t>(N_) * 3);
        }
        launch_fill_u64(dp_tmp_.ptr, static_cast<size_t>(N_) * 3, 0, stream);
        st_.compute(d_box, d_du_dx, dp_tmp_.ptr, d_du_dl, d_u, stream);
        launch_split_interpolated_du_dp(N_, lambda, dp_tmp_.ptr, d_du_dp, stream);
    } else {
        st_.compute(d_box, d_du_dx, d_du_dp, d_du_dl, d_u, stream);
    }
}

void NonbondedAllPairs::execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                             double lambda, uint64_t *d_dl_dp, hipStream_t stream) {
    validate_sizes(N, P);
    ensure_state();
    if (K_ == 0) {
        return;
    }
    st_.prepare(d_x, d_p, interp_ ? d_p + N_ * 3 : nullptr, d_box, d_plane_.ptr, d_offset_.ptr, nullptr, lambda,
                stream);
    st_.compute_dl_dp(d_box, d_dl_dp, stream);
}

void NonbondedAllPairs::execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                           double lambda, const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx,
                                           uint64_t *d_hp, uint64_t *d_hl, hipStream_t stream) {
    validate_sizes(N, P);
    if (interp_) {
        interpolated_hvp({this}, hvp_tmp_, N, d_x, d_p, d_box, lambda, d_vx, d_vp, vl, d_hx, d_hp, d_hl, stream);
        return;
    }
    ensure_state();
    if (K_ == 0) {
        return;
    }
    st_.prepare(d_x, d_p, nullptr, d_box, d_plane_.ptr, d_offset_.ptr, nullptr, lambda, stream);
    launch_tiles_hvp(st_, d_p, d_offset_.ptr, d_box, hvp_tangent(d_vx, d_vp, vl), HvpOut{d_hx, d_hp, d_hl}, stream);
}

void NonbondedAllPairs::hvp_at_interpolated(int N, const double *d_x, const double *p_lam, const double *d_box,
                                            double lambda, const double *d_vx, const double *vp_lam, double vl,
                                            uint64_t *d_hx, uint64_t *g, uint64_t *dg, uint64_t *d_hl,
                                            hipStream_t stream) {
    validate_sizes(N, N * 6);
    ensure_state();
    if (K_ == 0) {
        return;
    }
    // p(lambda) gathered as plain parameters (the interpolation's du/dlambda
    // part is the caller's chain rule)
    st_.prepare(d_x, p_lam, nullptr, d_box, d_plane_.ptr, d_offset_.ptr, nullptr, lambda, stream);
    launch_tiles_hvp(st_, p_lam, d_offset_.ptr, d_box, hvp_tangent(d_vx, vp_lam, vl), HvpOut{d_hx, dg, d_hl, g},
                     stream);
}

void NonbondedAllPairs::check_device_status(hipStream_t stream) {
    if (dirty_ || K_ == 0) {
        return;
    }
    int status = 0;
    TM_HIP_CHECK(hipMemcpyAsync(&status, st_.status.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
    TM_HIP_CHECK(hipStreamSynchronize(stream));
    if (status) {
        TM_HIP_CHECK(hipMemsetAsync(st_.status.ptr, 0, sizeof(int), stream));
        TM_HIP_CHECK(hipStreamSynchronize(stream));
        if (status & NB_STATUS_BOX_TOO_SMALL) {
            throw Error("Cutoff with padding is more than half of the box width, neighborlist is no longer reliable");
        }
        if (status & NB_STATUS_BOX_NOT_ORTHO) {
            throw Error("Provided non-ortholinear box, unable to compute nonbonded energy");
        }
        if (status & NB_STATUS_TILE_OVERFLOW) {
            throw Error(name_ + ": neighborlist tile buffer overflow");
        }
    }
}

int NonbondedAllPairs::num_tiles(hipStream_t stream) {
    ensure_state();
    int n = 0;
    TM_HIP_CHECK(hipMemcpyAsync(&n, st_.tile_prefix.ptr + 1, sizeof(int), hipMemcpyDeviceToHost, stream));
    TM_HIP_CHECK(hipStreamSynchronize(stream));
    return n;
}

static void nonbonded_du_dp_to_float(int P, const uint64_t *du_dp, double *out) {
    for (int i = 0; i + 2 < P; i += 3) {
        out[i + 0] = fixed_to_double(du_dp[i + 0], FIXED_SCALE);
        out[i + 1] = fixed_to_double(du_dp[i + 1], FIXED_SCALE_DSIG);
        out[i + 2] = fixed_to_double(du_dp[i + 2], FIXED_SCALE_DEPS);
    }
}

void NonbondedAllPairs::du_dp_fixed_to_float(int N, int P, const uint64_t *du_dp, double *out) const {
    nonbonded_du_dp_to_float(P, du_dp, out);
}

// ---------------------------------------------------------------- interaction group

namespace {
// rows ascending (the reference keeps them in a std::set), then every other atom ascending
std::vector<int> rows_then_columns(const std::vector<int> &rows, int N) {
    if (rows.empty()) {
        throw Error("row_atom_idxs must be nonempty");
    }
    std::set<int> uniq(rows.begin(), rows.end());
    if (uniq.size() != rows.size()) {
        throw Error("atom indices must be unique");
    }
    if (*uniq.begin() < 0 || *uniq.rbegin() >= N) {
        throw Error("indices values must be less than N");
    }
    if (static_cast<int>(rows.size()) >= N) {
        throw Error("number of idxs must be less than N");
    }
    std::vector<int> out(uniq.begin(), uniq.end());
    for (int a = 0; a < N; a++) {
        if (!uniq.count(a)) {
            out.push_back(a);
        }
    }
    return out;
}
} // namespace

NonbondedInteractionGroup::NonbondedInteractionGroup(const std::vector<int> &rows, const std::vector<int> &plane,
                                                     const std::vector<int> &offset, double beta, double cutoff,
                                                     bool interpolated)
    : NonbondedAllPairs(plane, offset, beta, cutoff, nullptr, interpolated) {
    name_ = "NonbondedInteractionGroup";
    subset_ = rows_then_columns(rows, N_);
    K_ = N_;
    nrow_ = static_cast<int>(rows.size());
}

// ---------------------------------------------------------------- pair list

NonbondedPairList::NonbondedPairList(const std::vector<int> &pair_idxs, const std::vector<double> &scales,
                                     const std::vector<int> &plane, const std::vector<int> &offset, double beta,
                                     double cutoff, bool negated, bool interpolated)
    : N_(static_cast<int>(offset.size())), M_(static_cast<int>(pair_idxs.size() / 2)), beta_(beta), cutoff_(cutoff),
      negated_(negated), interp_(interpolated), pairs_(pair_idxs), plane_(plane), offset_(offset), scales_(scales) {
    if (pair_idxs.size() % 2 != 0) {
        throw Error("pair_idxs.size() must be even, but got " + std::to_string(pair_idxs.size()));
    }
    for (int i = 0; i < M_; i++) {
        const int src = pair_idxs[i * 2 + 0];
        const int dst = pair_idxs[i * 2 + 1];
        if (src == dst) {
            throw Error("illegal pair with src == dst: " + std::to_string(src) + ", " + std::to_string(dst));
        }
    }
    if (static_cast<int>(scales.size() / 2) != M_) {
        throw Error("expected same number of pairs and scale tuples, but got " + std::to_string(M_) +
                    " != " + std::to_string(scales.size() / 2));
    }
    if (plane.size() != offset.size()) {
        throw Error("lambda offset idxs and plane idxs need to be equivalent");
    }
    d_pairs_.alloc(std::max(M_ * 2, 1));
    d_scales_.alloc(std::max(M_ * 2, 1));
    d_plane_.alloc(std::max(N_, 1));
    d_offset_.alloc(std::max(N_, 1));
    d_pairs_.upload(pair_idxs.data(), M_ * 2);
    d_scales_.upload(scales.data(), M_ * 2);
    d_plane_.upload(plane.data(), N_);
    d_offset_.upload(offset.data(), N_);
    device_synchronize();
}

void NonbondedPairList::execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                             double lambda, uint64_t *d_dl_dp, hipStream_t stream) {
    launch_pair_list_dl_dp(args(N, P, d_x, d_p, d_box, lambda), d_dl_dp, stream);
}

void NonbondedPairList::execute_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                       double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl,
                                       uint64_t *d_u, hipStream_t stream) {
    const PairListArgs a = args(N, P, d_x, d_p, d_box, lambda);
    if (interp_ && d_du_dp) {
        if (dp_tmp_.n != static_cast<size_t>(N_) * 3) {
            dp_tmp_.alloc(static_cast<size_t>(N_) * 3);
        }
        launch_fill_u64(dp_tmp_.ptr, static_cast<size_t>(N_) * 3, 0, stream);
        launch_pair_list(a, d_du_dx, dp_tmp_.ptr, d_du_dl, d_u, stream);
        launch_split_interpolated_du_dp(N_, lambda, dp_tmp_.ptr, d_du_dp, stream);
    } else {
        launch_pair_list(a, d_du_dx, d_du_dp, d_du_dl, d_u, stream);
    }
}

void NonbondedPairList::execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                           double lambda, const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx,
                                           uint64_t *d_hp, uint64_t *d_hl, hipStream_t stream) {
    if (interp_) {
        args(N, P, d_x, d_p, d_box, lambda); // size checks
        interpolated_hvp({this}, hvp_tmp_, N, d_x, d_p, d_box, lambda, d_vx, d_vp, vl, d_hx, d_hp, d_hl, stream);
        return;
    }
    launch_pair_list_hvp(args(N, P, d_x, d_p, d_box, lambda), hvp_tangent(d_vx, d_vp, vl), HvpOut{d_hx, d_hp, d_hl},
                         stream);
}

void NonbondedPairList::hvp_at_interpolated(int N, const double *d_x, const double *p_lam, const double *d_box,
                                            double lambda, const double *d_vx, const double *vp_lam, double vl,
                                            uint64_t *d_hx, uint64_t *g, uint64_t *dg, uint64_t *d_hl,
                                            hipStream_t stream) {
    PairListArgs a = args(N, N * 6, d_x, p_lam, d_box, lambda);
    a.params1 = nullptr; // p(lambda) as plain parameters
    launch_pair_list_hvp(a, hvp_tangent(d_vx, vp_lam, vl), HvpOut{d_hx, dg, d_hl, g}, stream);
}

PairListArgs NonbondedPairList::args(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                     double lambda) const {
    if (N != N_) {
        throw Error("NonbondedPairList::execute_device(): expected N == N_, got N=" + std::to_string(N) +
                    ", N_=" + std::to_string(N_));
    }
    const int Mp = interp_ ? 2 : 1;
    if (P != Mp * N_ * 3) {
        throw Error("NonbondedPairList::execute_device(): expected P == M*N_*3, got P=" + std::to_string(P) +
                    ", M*N_*3=" + std::to_string(Mp * N_ * 3));
    }
    PairListArgs a{};
    a.M = M_;
    a.pairs = d_pairs_.ptr;
    a.scales = d_scales_.ptr;
    a.pair_sys = nullptr;
    a.red_idx = nullptr;
    a.x = d_x;
    a.params = d_p;
    a.boxes = d_box;
    a.plane = d_plane_.ptr;
    a.offset = d_offset_.ptr;
    a.lambdas = nullptr;
    a.lambda0 = lambda;
    a.params1 = interp_ ? d_p + N_ * 3 : nullptr;
    a.cutoffs = nullptr;
    a.cutoff0 = cutoff_;
    a.betas = nullptr;
    a.beta0 = static_cast<float>(beta_);
    a.negated = negated_ ? 1 : 0;
    return a;
}

void NonbondedPairList::du_dp_fixed_to_float(int N, int P, const uint64_t *du_dp, double *out) const {
    nonbonded_du_dp_to_float(P, du_dp, out);
}

// ---------------------------------------------------------------- bonded

static const char *bonded_name(BondedKind k) {
    return k == BondedKind::Bond ? "HarmonicBond" : (k == BondedKind::Angle ? "HarmonicAngle" : "PeriodicTorsion");
}

BondedPotential::BondedPotential(BondedKind bk, const std::vector<int> &idxs, const std::vector<int> *mult,
                                 const std::vector<int> *offset)
    : bk_(bk), idxs_(idxs) {
    const int ar = bonded_arity(bk);
    if (idxs.size() % ar != 0) {
        if (bk == BondedKind::Bond) {
            throw Error("bond_idxs.size() must be exactly 2*k!");
        } else if (bk == BondedKind::Angle) {
            throw Error("angle_idxs.size() must be exactly 3*A");
        } else {
            throw Error("torsion_idxs.size() must be exactly 4*k");
        }
    }
    n_ = static_cast<int>(idxs.size()) / ar;
    if (mult && !mult->empty() && static_cast<int>(mult->size()) != n_) {
        throw Error("bad lambda_mult size()");
    }
    if (offset && !offset->empty() && static_cast<int>(offset->size()) != n_) {
        throw Error("bad lambda_offset size()");
    }
    for (int t = 0; t < n_; t++) {
        const int *q = idxs.data() + t * ar;
        bool dup = false;
        for (int a = 0; a < ar; a++) {
            for (int b = a + 1; b < ar; b++) {
                dup = dup || q[a] == q[b];
            }
        }
        if (dup) {
            if (bk == BondedKind::Bond) {
                throw Error("src == dst");
            } else if (bk == BondedKind::Angle) {
                throw Error("angle triplets must be unique");
            } else {
                throw Error("torsion quads must be unique");
            }
        }
    }
    mult_ = (mult && !mult->empty()) ? *mult : std::vector<int>(n_, 0);
    offset_ = (offset && !offset->empty()) ? *offset : std::vector<int>(n_, 1);
    d_idxs_.alloc(std::max<size_t>(idxs.size(), 1));
    d_mult_.alloc(std::max(n_, 1));
    d_offset_.alloc(std::max(n_, 1));
    d_idxs_.upload(idxs.data(), idxs.size());
    d_mult_.upload(mult_.data(), n_);
    d_offset_.upload(offset_.data(), n_);
    device_synchronize();
}

PotKind BondedPotential::kind() const {
    return bk_ == BondedKind::Bond ? PotKind::HarmonicBond
                                   : (bk_ == BondedKind::Angle ? PotKind::HarmonicAngle : PotKind::PeriodicTorsion);
}

void BondedPotential::execute_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                     double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl,
                                     uint64_t *d_u, hipStream_t stream) {
    if (n_ == 0) {
        return;
    }
    const int np = bonded_nparams(bk_);
    if (P != np * n_) {
        const char *nm = bonded_name(bk_);
        const char *sym = bk_ == BondedKind::Bond ? "B_" : (bk_ == BondedKind::Angle ? "A_" : "T_");
        throw Error(std::string(nm) + "::execute_device(): expected P == " + std::to_string(np) + "*" + sym +
                    ", got P=" + std::to_string(P) + ", " + std::to_string(np) + "*" + sym + "=" +
                    std::to_string(np * n_));
    }
    BondedTerms t{n_, d_idxs_.ptr, d_p, d_mult_.ptr, d_offset_.ptr, nullptr};
    launch_bonded(bk_, t, d_x, nullptr, lambda, nullptr, d_du_dx, d_du_dp, d_du_dl, d_u, stream);
}

void BondedPotential::execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                           double lambda, uint64_t *d_dl_dp, hipStream_t stream) {
    if (n_ == 0) {
        return;
    }
    // same size validation as execute_device
    execute_device(N, P, d_x, d_p, d_box, lambda, nullptr, nullptr, nullptr, nullptr, stream);
    BondedTerms t{n_, d_idxs_.ptr, d_p, d_mult_.ptr, d_offset_.ptr, nullptr};
    launch_bonded(bk_, t, d_x, nullptr, lambda, nullptr, nullptr, d_dl_dp, nullptr, nullptr, stream, true);
}

void BondedPotential::execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                         double lambda, const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx,
                                         uint64_t *d_hp, uint64_t *d_hl, hipStream_t stream) {
    if (n_ == 0) {
        return;
    }
    execute_device(N, P, d_x, d_p, d_box, lambda, nullptr, nullptr, nullptr, nullptr, stream); // size validation
    BondedTerms t{n_, d_idxs_.ptr, d_p, d_mult_.ptr, d_offset_.ptr, nullptr};
    launch_bonded_hvp(bk_, t, d_x, nullptr, lambda, nullptr, hvp_tangent(d_vx, d_vp, vl), HvpOut{d_hx, d_hp, d_hl},
                      stream);
}

// ---------------------------------------------------------------- sums

SummedPotential::SummedPotential(const std::vector<PotentialPtr> &pots, const std::vector<int> &sizes)
    : pots_(pots), sizes_(sizes), P_(std::accumulate(sizes.begin(), sizes.end(), 0)) {
    if (pots.size() != sizes.size()) {
        throw Error("number of potentials != number of parameter sizes");
    }
}

void SummedPotential::execute_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                     double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl,
                                     uint64_t *d_u, hipStream_t stream) {
    if (P != P_) {
        throw Error("SummedPotential::execute_device(): expected " + std::to_string(P_) + " parameters, got " +
                    std::to_string(P));
    }
    int off = 0;
    for (size_t i = 0; i < pots_.size(); i++) {
        pots_[i]->execute_device(N, sizes_[i], d_x, d_p + off, d_box, lambda, d_du_dx,
                                 d_du_dp ? d_du_dp + off : nullptr, d_du_dl, d_u, stream);
        off += sizes_[i];
    }
}

void SummedPotential::execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                           double lambda, uint64_t *d_dl_dp, hipStream_t stream) {
    if (P != P_) {
        throw Error("SummedPotential::execute_device(): expected " + std::to_string(P_) + " parameters, got " +
                    std::to_string(P));
    }
    int off = 0;
    for (size_t i = 0; i < pots_.size(); i++) {
        pots_[i]->execute_dl_dp_device(N, sizes_[i], d_x, d_p + off, d_box, lambda, d_dl_dp + off, stream);
        off += sizes_[i];
    }
}

void SummedPotential::execute_hvp_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                         double lambda, const double *d_vx, const double *d_vp, double vl, uint64_t *d_hx,
                                         uint64_t *d_hp, uint64_t *d_hl, hipStream_t stream) {
    if (P != P_) {
        throw Error("SummedPotential::execute_device(): expected " + std::to_string(P_) + " parameters, got " +
                    std::to_string(P));
    }
    int off = 0;
    for (size_t i = 0; i < pots_.size(); i++) {
        pots_[i]->execute_hvp_device(N, sizes_[i], d_x, d_p + off, d_box, lambda, d_vx, d_vp ? d_vp + off : nullptr,
                                     vl, d_hx, d_hp ? d_hp + off : nullptr, d_hl, stream);
        off += sizes_[i];
    }
}

void SummedPotential::validate_host(int N, int P, const double *h_box) const {
    for (size_t i = 0; i < pots_.size(); i++) {
        pots_[i]->validate_host(N, sizes_[i], h_box);
    }
}

void SummedPotential::check_device_status(hipStream_t stream) {
    for (auto &p : pots_) {
        p->check_device_status(stream);
    }
}

void SummedPotential::du_dp_fixed_to_float(int N, int P, const uint64_t *du_dp, double *out) const {
    int off = 0;
    for (size_t i = 0; i < pots_.size(); i++) {
        pots_[i]->du_dp_fixed_to_float(N, sizes_[i], du_dp + off, out + off);
        off += sizes_[i];
    }
}

FanoutSummedPotential::FanoutSummedPotential(const std::vector<PotentialPtr> &pots) : pots_(pots) {}

void FanoutSummedPotential::execute_device(int N, int P, const double *d_x, const double *d_p, const double *d_box,
                                           double lambda, uint64_t *d_du_dx, uint64_t *d_du_dp, uint64_t *d_du_dl,
                                           uint64_t *d_u, hipStream_t stream) {
    for (auto &p : pots_) {
        p->execute_device(N, P, d_x, d_p, d_box, lambda, d_du_dx, d_du_dp, d_du_dl, d_u, stream);
    }
}

void FanoutSummedPotential::execute_dl_dp_device(int N, int P, const double *d_x, const double *d_p,
                                                 const double *d_box, double lambda, uint64_t *d_dl_dp,
                                                 hipStream_t stream) {
    for (auto &p : pots_) {
        p->execute_dl_dp_device(N, P, d_x, d_p, d_box, lambda, d_dl_dp, stream);
    }
}

void FanoutSummedPotential::execute_hvp_device(int N, int P, const double *d_x, const double *d_p,
                                               const double *d_box, double lambda, const double *d_vx,
                                               const double *d_vp, double vl, uint64_t *d_hx, uint64_t *d_hp,
                                               uint64_t *d_hl, hipStream_t stream) {
    std::vector<Potential *> interp;
    for (auto &p : pots_) {
        if (p->interpolated_nonbonded()) {
            interp.push_back(p.get());
        }
    }
    if (!interp.empty() && interp.size() == pots_.size()) {
        // one chain rule over the summed products (the all-pairs and exclusion
        // terms of a Nonbonded potential cancel before it)
        if (P != N * 6) {
            throw Error("FanoutSummedPotential::execute_hvp(): expected P == 2*N*3 for interpolated parameters");
        }
        interpolated_hvp(interp, hvp_tmp_, N, d_x, d_p, d_box, lambda, d_vx, d_vp, vl, d_hx, d_hp, d_hl, stream);
        return;
    }
    for (auto &p : pots_) {
        p->execute_hvp_device(N, P, d_x, d_p, d_box, lambda, d_vx, d_vp, vl, d_hx, d_hp, d_hl, stream);
    }
}

void FanoutSummedPotential::validate_host(int N, int P, const double *h_box) const {
    for (auto &p : pots_) {
        p->validate_host(N, P, h_box);
    }
}

void FanoutSummedPotential::check_device_status(hipStream_t stream) {
    for (auto &p : pots_) {
        p->check_device_status(stream);
    }
}

void FanoutSummedPotential::du_dp_fixed_to_float(int N, int P, const uint64_t *du_dp, double *out) const {
    if (pots_.empty()) {
        Potential::du_dp_fixed_to_float(N, P, du_dp, out);
    } else {
        pots_[0]->du_dp_fixed_to_float(N, P, du_dp, out);
    }
}

// ---------------------------------------------------------------- bound

BoundPotential::BoundPotential(PotentialPtr p, const std::vector<int> &shp, const double *h_p)
    : potential(std::move(p)), shape(shp) {
    const int n = size();
    h_params.assign(h_p, h_p + n);
    d_params.alloc(std::max(n, 1));
    d_params.upload(h_params.data(), n);
    device_synchronize();
}

int BoundPotential::size() const {
    if (shape.empty()) {
        return 0;
    }
    int total = 1;
    for (int s : shape) {
        total *= s;
    }
    return total;
}

void BoundPotential::execute_device(int N, const double *d_x, const double *d_box, double lambda, uint64_t *d_du_dx,
                                    uint64_t *d_du_dp, uint64_t *d_du_dl, uint64_t *d_u, hipStream_t stream) {
    potential->execute_device(N, size(), d_x, size() > 0 ? d_params.ptr : nullptr, d_box, lambda, d_du_dx, d_du_dp,
                              d_du_dl, d_u, stream);
}

void BoundPotential::execute_host(int N, const double *h_x, const double *h_box, double lambda, uint64_t *h_du_dx,
                                  uint64_t *h_du_dl, uint64_t *h_u) {
    potential->validate_host(N, size(), h_box);
    DeviceArray<double> x(static_cast<size_t>(N) * 3), box(9);
    x.upload(h_x, static_cast<size_t>(N) * 3);
    box.upload(h_box, 9);
    DeviceArray<uint64_t> du_dx(static_cast<size_t>(N) * 3), du_dl(std::max(N, 1)), u(std::max(N, 1));
    du_dx.zero();
    du_dl.zero();
    u.zero();
    execute_device(N, x.ptr, box.ptr, lambda, du_dx.ptr, nullptr, du_dl.ptr, u.ptr, 0);
    potential->check_device_status(0);
    du_dx.download(h_du_dx, static_cast<size_t>(N) * 3);
    du_dl.download(h_du_dl, N);
    u.download(h_u, N);
    TM_HIP_CHECK(hipStreamSynchronize(0));
}

} // namespace tmamd
