// Directional derivatives of the gradients (Hessian-vector products) of the
// bonded and nonbonded potentials. With tangents vx on the coordinates, vp on
// the parameters and vl on lambda, the tangents of du/dx, du/dp and du/dlambda
// are the rows of H (vx, vp, vl), H the Hessian of U in (x, p, lambda)
// (dual.h). By the symmetry of H they are also the vector-Jacobian product of
// the gradient with (vx, vp, vl): the adjoint of du/dx that reverse-mode
// differentiation through forces and MD trajectories needs, and the
// double-backward of the torch.autograd.Function over the potentials.
//
// The reference has no second-order custom ops; its design notes describe
// differentiating through Langevin dynamics with Hessians and mixed second
// derivatives of the energy (attic/docs/paper.tex:27). The energies here are
// the ones of the force kernels: the bonded terms of bonded_terms.h evaluated
// in dual arithmetic, and the nonbonded pair energy of nonbonded.py:133-264
// (erfc-damped Coulomb plus LJ 12-6 on the 4D distance, parameters at their
// FP32 values as the force kernels see them) with exact derivatives in FP64.
// Tangents accumulate in 64-bit fixed point with the exponents of du_dx
// (2^36), du_dp (bonded 2^36; nonbonded 2^36 / 2^37 / 2^38 for q / sigma /
// epsilon) and du_dl (2^36), so they decode like the gradients and do not
// depend on the order of the terms.
#include <algorithm>

#include "bonded.h"
#include "bonded_terms.h"
#include "dual.h"
#include "fixed_point.h"
#include "hvp.h"
#include "nb_pair.h"
#include "nb_state.h"
#include "wave.h"

namespace tmamd {

namespace {

__device__ __forceinline__ void atomic_add_fixed(uint64_t *p, uint64_t f) {
    if (f != 0) {
        atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(f));
    }
}

__device__ __forceinline__ uint64_t fixed_signed(double v, double scale, bool negate) {
    const uint64_t f = f64_to_fixed(v, scale);
    return negate ? 0ull - f : f;
}

// ------------------------------------------------------------------ bonded

template <int KIND>
__global__ __launch_bounds__(128) void k_bonded_hvp(const BondedTerms t, const double *__restrict__ x,
                                                    const double *__restrict__ lambdas, double lambda0,
                                                    const int *__restrict__ red_idx, const HvpTangent v,
                                                    const HvpOut out) {
    constexpr int AR = KIND == 0 ? 2 : (KIND == 1 ? 3 : 4);
    constexpr int NP = KIND == 2 ? 3 : 2;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t fl = 0;
    int red = -1;
    if (q < t.n) {
        const int s = t.term_sys ? t.term_sys[q] : 0;
        const double lambda = lambdas ? lambdas[s] : lambda0;
        const double vl = v.vls ? v.vls[s] : v.vl;
        red = red_idx ? red_idx[s] : 0;
        Dual pos[AR][3];
        int at[AR];
        for (int a = 0; a < AR; a++) {
            at[a] = t.idx[q * AR + a];
            for (int d = 0; d < 3; d++) {
                pos[a][d] = Dual(x[at[a] * 3 + d], v.vx ? v.vx[at[a] * 3 + d] : 0.0);
            }
        }
        Dual prm[NP];
        for (int c = 0; c < NP; c++) {
            prm[c] = Dual(t.params[q * NP + c], v.vp ? v.vp[q * NP + c] : 0.0);
        }
        const int mult = t.mult ? t.mult[q] : 0;
        const int off = t.offset ? t.offset[q] : 1;
        // prefactor offset + mult lambda (the force kernels' value), tangent mult vl
        const Dual pf(lam_prefactor<double>(mult, off, lambda), mult * vl);
        Dual g[AR][3];
        Dual dp[NP];
        Dual dl;
        if constexpr (KIND == 0) {
            const BondOut<Dual> o = bond_term<Dual, Dual>(pos[0], pos[1], prm[0], prm[1], mult, pf);
            for (int a = 0; a < AR; a++) {
                for (int d = 0; d < 3; d++) {
                    g[a][d] = o.g[a][d];
                }
            }
            dp[0] = o.dkb;
            dp[1] = o.db0;
            dl = o.dl;
        } else if constexpr (KIND == 1) {
            const AngleOut<Dual> o = angle_term<Dual, Dual>(pos[0], pos[1], pos[2], prm[0], prm[1], mult, pf);
            for (int a = 0; a < AR; a++) {
                for (int d = 0; d < 3; d++) {
                    g[a][d] = o.g[a][d];
                }
            }
            dp[0] = o.dka;
            dp[1] = o.da0;
            dl = o.dl;
        } else {
            const TorsionOut<Dual> o =
                torsion_term<Dual, Dual>(pos[0], pos[1], pos[2], pos[3], prm[0], prm[1], prm[2], mult, pf);
            for (int a = 0; a < AR; a++) {
                for (int d = 0; d < 3; d++) {
                    g[a][d] = o.g[a][d];
                }
            }
            dp[0] = o.dk;
            dp[1] = o.dphase;
            dp[2] = o.dperiod;
            dl = o.dl;
        }
        if (out.hx) {
            for (int a = 0; a < AR; a++) {
                for (int d = 0; d < 3; d++) {
                    atomic_add_fixed(out.hx + at[a] * 3 + d, f64_to_fixed(g[a][d].d, FIXED_SCALE));
                }
            }
        }
        if (out.hp) {
            for (int c = 0; c < NP; c++) {
                atomic_add_fixed(out.hp + q * NP + c, f64_to_fixed(dp[c].d, FIXED_SCALE));
            }
        }
        fl = f64_to_fixed(dl.d, FIXED_SCALE);
    }
    if (out.hl) {
        wave_add_by_key(out.hl, red, fl);
    }
}

// --------------------------------------------------------------- nonbonded

// Gradient of one pair's energy in dual arithmetic. dv = 3D minimum-image
// displacement x_i - x_j, dw = w_i - w_j (w = (plane + offset lambda) cutoff),
// dwdl = d(dw)/dlambda. du/dx_i = pre dv, du/dx_j = -pre dv; the charge,
// sigma (one value for both atoms: sig_ij = sig_i + sig_j) and epsilon
// derivatives of each atom; du/dlambda = pre dw dwdl.
struct PairGradDual {
    Dual pre;
    Dual dq_i, dq_j, dsig, deps_i, deps_j;
    Dual dl;
    bool lj;
};

__device__ __forceinline__ PairGradDual pair_grad_dual(const Dual *dv, const Dual &dw, double dwdl, const Dual &qi,
                                                       const Dual &qj, const Dual &si, const Dual &sj, const Dual &ei,
                                                       const Dual &ej, double beta, double sq, double slj) {
    PairGradDual o;
    const Dual d2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2] + dw * dw;
    const Dual r = sqrt(d2);
    const Dual inv_r = Dual(1.0) / r;
    const Dual bx = Dual(beta) * r;
    // exp(-x^2) once for both: erfc'(x) = -2 / sqrt(pi) exp(-x^2), (exp(-x^2))' = -2 x exp(-x^2)
    const double e = std::exp(-(bx.v * bx.v));
    const Dual ec(std::erfc(bx.v), -TWO_OVER_SQRT_PI * e * bx.d);
    const Dual ex(e, -2.0 * bx.v * e * bx.d);
    // u_es = sq q_i q_j erfc(beta r) / r;  (du/dr) / r = -sq q_i q_j (erfc / r + 2 beta / sqrt(pi) e^{-b^2 r^2}) / r^2
    const Dual qij = Dual(sq) * qi * qj;
    o.pre = -(qij * (ec * inv_r + Dual(TWO_OVER_SQRT_PI * beta) * ex) * inv_r * inv_r);
    const Dual es_unit = Dual(sq) * ec * inv_r;
    o.dq_i = qj * es_unit;
    o.dq_j = qi * es_unit;
    o.lj = ei.v != 0.0 && ej.v != 0.0;
    if (o.lj) {
        const Dual sig_ij = si + sj;
        const Dual s1 = sig_ij * inv_r;
        const Dual s2 = s1 * s1;
        const Dual s6 = s2 * s2 * s2;
        const Dual s12 = s6 * s6;
        const Dual se = Dual(slj) * ei * ej;
        // u_lj = 4 se (s12 - s6); (du/dr) / r = 4 se (6 s6 - 12 s12) / r^2
        o.pre = o.pre + Dual(4.0) * se * (Dual(6.0) * s6 - Dual(12.0) * s12) * inv_r * inv_r;
        o.dsig = Dual(4.0) * se * (Dual(12.0) * s12 - Dual(6.0) * s6) / sig_ij;
        const Dual lj_shape = Dual(4.0 * slj) * (s12 - s6);
        o.deps_i = lj_shape * ej;
        o.deps_j = lj_shape * ei;
    }
    o.dl = o.pre * dw * Dual(dwdl);
    return o;
}

// parameters as the force kernels see them (their FP32 values), tangents in FP64
__device__ __forceinline__ Dual dual_param(const double *p, const double *vp, int atom, int c) {
    return Dual(static_cast<double>(static_cast<float>(p[atom * 3 + c])), vp ? vp[atom * 3 + c] : 0.0);
}

// Fixed-point tangents of one pair's gradient: du_dx (3) and du_dp (3) of
// atoms i and j, and du_dl. Zero for pairs outside the cutoff.
// With G, also the values of du_dp (gi, gj).
template <bool G> struct PairTangents {
    uint64_t xi[3], xj[3], pi[3], pj[3], l;
    uint64_t gi[G ? 3 : 1], gj[G ? 3 : 1];
};

template <bool G> __device__ __forceinline__ PairTangents<G> zero_tangents() {
    PairTangents<G> t;
    for (int d = 0; d < 3; d++) {
        t.xi[d] = t.xj[d] = t.pi[d] = t.pj[d] = 0;
        if constexpr (G) {
            t.gi[d] = t.gj[d] = 0;
        }
    }
    t.l = 0;
    return t;
}

template <bool G>
__device__ __forceinline__ PairTangents<G> pair_tangents(const PairGradDual &o, const Dual *dv, bool neg) {
    PairTangents<G> t = zero_tangents<G>();
    if constexpr (G) {
        t.gi[0] = fixed_signed(o.dq_i.v, FIXED_SCALE, neg);
        t.gj[0] = fixed_signed(o.dq_j.v, FIXED_SCALE, neg);
        if (o.lj) {
            t.gi[1] = t.gj[1] = fixed_signed(o.dsig.v, FIXED_SCALE_DSIG, neg);
            t.gi[2] = fixed_signed(o.deps_i.v, FIXED_SCALE_DEPS, neg);
            t.gj[2] = fixed_signed(o.deps_j.v, FIXED_SCALE_DEPS, neg);
        }
    }
    for (int d = 0; d < 3; d++) {
        const Dual f = o.pre * dv[d];
        t.xi[d] = fixed_signed(f.d, FIXED_SCALE, neg);
        t.xj[d] = 0ull - t.xi[d];
    }
    t.pi[0] = fixed_signed(o.dq_i.d, FIXED_SCALE, neg);
    t.pj[0] = fixed_signed(o.dq_j.d, FIXED_SCALE, neg);
    if (o.lj) {
        t.pi[1] = t.pj[1] = fixed_signed(o.dsig.d, FIXED_SCALE_DSIG, neg);
        t.pi[2] = fixed_signed(o.deps_i.d, FIXED_SCALE_DEPS, neg);
        t.pj[2] = fixed_signed(o.deps_j.d, FIXED_SCALE_DEPS, neg);
    }
    t.l = fixed_signed(o.dl.d, FIXED_SCALE, neg);
    return t;
}

// One thread per pair of the list, as k_pair_list (one or many systems).
template <bool G>
__global__ __launch_bounds__(256) void k_pair_list_hvp(const PairListArgs a, const HvpTangent v, const HvpOut out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    PairTangents<G> t = zero_tangents<G>();
    int gi = 0, gj = 0, red = -1;
    bool in = false;
    if (p < a.M) {
        const int s = a.pair_sys ? a.pair_sys[p] : 0;
        gi = a.pairs[p * 2 + 0];
        gj = a.pairs[p * 2 + 1];
        red = a.red_idx ? a.red_idx[s] : 0;
        const double lam = a.lambdas ? a.lambdas[s] : a.lambda0;
        const double cutoff = a.cutoffs ? a.cutoffs[s] : a.cutoff0;
        const double beta = static_cast<double>(a.betas ? a.betas[s] : a.beta0);
        const double vl = v.vls ? v.vls[s] : v.vl;
        const int pli = a.plane ? a.plane[gi] : 0, plj = a.plane ? a.plane[gj] : 0;
        const int ofi = a.offset ? a.offset[gi] : 0, ofj = a.offset ? a.offset[gj] : 0;
        const double dwdl = (ofi - ofj) * cutoff;
        const Dual dw((pli + ofi * lam) * cutoff - (plj + ofj * lam) * cutoff, dwdl * vl);
        const double *box = a.boxes + s * 9;
        const double L[3] = {box[0], box[4], box[8]};
        Dual dv[3];
        for (int d = 0; d < 3; d++) {
            // the image shift is a constant: the tangent is vx_i - vx_j
            const double tan = v.vx ? v.vx[gi * 3 + d] - v.vx[gj * 3 + d] : 0.0;
            dv[d] = Dual(min_image_d(a.x[gi * 3 + d] - a.x[gj * 3 + d], L[d], 1.0 / L[d]), tan);
        }
        const double d2 = __builtin_fma(dw.v, dw.v, pair_d2(dv[0].v, dv[1].v, dv[2].v));
        in = d2 < cutoff * cutoff;
        if (in) {
            const PairGradDual o = pair_grad_dual(
                dv, dw, dwdl, dual_param(a.params, v.vp, gi, 0), dual_param(a.params, v.vp, gj, 0),
                dual_param(a.params, v.vp, gi, 1), dual_param(a.params, v.vp, gj, 1),
                dual_param(a.params, v.vp, gi, 2), dual_param(a.params, v.vp, gj, 2), beta, a.scales[p * 2 + 0],
                a.scales[p * 2 + 1]);
            t = pair_tangents<G>(o, dv, a.negated != 0);
        }
    }
    if (in) {
        for (int c = 0; c < 3; c++) {
            if (out.hx) {
                atomic_add_fixed(out.hx + gi * 3 + c, t.xi[c]);
                atomic_add_fixed(out.hx + gj * 3 + c, t.xj[c]);
            }
            if (out.hp) {
                atomic_add_fixed(out.hp + gi * 3 + c, t.pi[c]);
                atomic_add_fixed(out.hp + gj * 3 + c, t.pj[c]);
            }
            if constexpr (G) {
                atomic_add_fixed(out.gp + gi * 3 + c, t.gi[c]);
                atomic_add_fixed(out.gp + gj * 3 + c, t.gj[c]);
            }
        }
    }
    if (out.hl) {
        wave_add_by_key(out.hl, red, t.l);
    }
}

// All pairs of the tile lists of the last neighborlist build (one or many
// systems), one wave per tile, in the force kernel's two passes: the screen
// pairs row lane l with column slot (l + r) mod 64 in round r and compacts the
// pairs inside the cutoff (FP64 test, as the force kernel) into an LDS queue
// (ballot / mbcnt); every 64 queued pairs are evaluated one per lane, their
// fixed-point tangents added to per-atom LDS accumulators of the tile's rows
// and columns, which go to global memory once per tile and atom.
template <int NACC> struct HvpTileLds {
    // [0] rows, [1] columns: position (x, y, z, w), parameters at their FP32
    // values, tangents, lambda offset, sorted slot and global atom index
    double px[2][TILE], py[2][TILE], pz[2][TILE], pw[2][TILE];
    double q[2][TILE], sg[2][TILE], ep[2][TILE];
    double tq[2][TILE], ts[2][TILE], te[2][TILE];
    double tx[2][TILE], ty[2][TILE], tz[2][TILE];
    int off[2][TILE];
    uint32_t slot[2][TILE];
    int gidx[2][TILE];
    uint64_t acc[2][NACC][TILE]; // du_dx (3) and du_dp (3) tangents, [du_dp (3)], of each row / column atom
    uint16_t pq[2 * TILE];    // queued pair: row lane | column lane << 6
};

__device__ __forceinline__ void lds_add(uint64_t *p, uint64_t v) {
    if (v != 0) {
        atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(v));
    }
}

template <bool G>
__global__ __launch_bounds__(64) void k_tiles_hvp(
    const NbSys *__restrict__ sys, int S, const int *__restrict__ tile_prefix, const int *__restrict__ tile_order,
    const int *__restrict__ tile_row, const uint32_t *__restrict__ tile_atoms, const uint32_t *__restrict__ perm, const double4 *__restrict__ pos,
    const double *__restrict__ params, const int *__restrict__ offset, const double *__restrict__ boxes,
    const float4 *__restrict__ bb_ctr, const float4 *__restrict__ bb_ext, const int *__restrict__ status,
    const HvpTangent v, const HvpOut out) {
    constexpr int NACC = G ? 9 : 6;
    __shared__ HvpTileLds<NACC> L;
    const int total = tile_prefix[S];
    const int lane = threadIdx.x;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = S - 1; // the system of tile t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tile_prefix[mid] <= t) {
                lo = mid;
            } else {
                hi = mid - 1;
            }
        }
        const int s = lo;
        if (status[s] & (NB_STATUS_BOX_TOO_SMALL | NB_STATUS_BOX_NOT_ORTHO)) {
            continue;
        }
        const NbSys Sd = sys[s];
        const double bx = boxes[s * 9 + 0], by = boxes[s * 9 + 4], bz = boxes[s * 9 + 8];
        const double ibx = 1.0 / bx, iby = 1.0 / by, ibz = 1.0 / bz;
        const double beta = static_cast<double>(Sd.beta);
        const double cut2 = Sd.cutoff_d * Sd.cutoff_d;
        const double vl = v.vls ? v.vls[s] : v.vl;
        const int kt = Sd.tile0 + tile_order[Sd.tile0 + (t - tile_prefix[s])];
        const int r = tile_row[kt];

        // stage the tile's rows and columns
        wave_lds_sync(); // the previous tile's LDS reads are done
        for (int side = 0; side < 2; side++) {
            const uint32_t sl = side == 0 ? static_cast<uint32_t>(r * TILE + lane)
                                          : tile_atoms[static_cast<size_t>(kt) * TILE + lane];
            const bool valid = side == 0 ? sl < static_cast<uint32_t>(Sd.row_end)
                                         : sl < static_cast<uint32_t>(Sd.slot_end);
            double4 p4 = make_double4(0.0, 0.0, 0.0, 0.0);
            int g = 0;
            if (valid) {
                p4 = pos[Sd.slot0 + sl];
                g = Sd.atom0 + static_cast<int>(perm[Sd.slot0 + sl]);
            }
            L.px[side][lane] = p4.x;
            L.py[side][lane] = p4.y;
            L.pz[side][lane] = p4.z;
            L.pw[side][lane] = p4.w;
            L.q[side][lane] = valid ? static_cast<double>(static_cast<float>(params[g * 3 + 0])) : 0.0;
            L.sg[side][lane] = valid ? static_cast<double>(static_cast<float>(params[g * 3 + 1])) : 0.0;
            L.ep[side][lane] = valid ? static_cast<double>(static_cast<float>(params[g * 3 + 2])) : 0.0;
            L.tq[side][lane] = (valid && v.vp) ? v.vp[g * 3 + 0] : 0.0;
            L.ts[side][lane] = (valid && v.vp) ? v.vp[g * 3 + 1] : 0.0;
            L.te[side][lane] = (valid && v.vp) ? v.vp[g * 3 + 2] : 0.0;
            L.tx[side][lane] = (valid && v.vx) ? v.vx[g * 3 + 0] : 0.0;
            L.ty[side][lane] = (valid && v.vx) ? v.vx[g * 3 + 1] : 0.0;
            L.tz[side][lane] = (valid && v.vx) ? v.vx[g * 3 + 2] : 0.0;
            L.off[side][lane] = (valid && offset) ? offset[g] : 0;
            L.slot[side][lane] = valid ? sl : NB_INVALID;
            L.gidx[side][lane] = g;
            for (int k = 0; k < NACC; k++) {
                L.acc[side][k][lane] = 0;
            }
        }
        wave_lds_sync();

        uint64_t acc_l = 0;
        // pass 2: the first n queued pairs, one per lane
        auto eval = [&](int n) {
            if (lane < n) {
                const int e = L.pq[lane];
                const int il = e & 63, jl = e >> 6;
                Dual dv[3];
                dv[0] = Dual(min_image_d(L.px[0][il] - L.px[1][jl], bx, ibx), L.tx[0][il] - L.tx[1][jl]);
                dv[1] = Dual(min_image_d(L.py[0][il] - L.py[1][jl], by, iby), L.ty[0][il] - L.ty[1][jl]);
                dv[2] = Dual(min_image_d(L.pz[0][il] - L.pz[1][jl], bz, ibz), L.tz[0][il] - L.tz[1][jl]);
                const double dwdl = (L.off[0][il] - L.off[1][jl]) * Sd.cutoff_d;
                const Dual dw(L.pw[0][il] - L.pw[1][jl], dwdl * vl);
                // the exact FP64 test of the force kernels (the screen is a superset)
                if (!(__builtin_fma(dw.v, dw.v, pair_d2(dv[0].v, dv[1].v, dv[2].v)) < cut2)) {
                    return;
                }
                const PairGradDual o = pair_grad_dual(
                    dv, dw, dwdl, Dual(L.q[0][il], L.tq[0][il]), Dual(L.q[1][jl], L.tq[1][jl]),
                    Dual(L.sg[0][il], L.ts[0][il]), Dual(L.sg[1][jl], L.ts[1][jl]), Dual(L.ep[0][il], L.te[0][il]),
                    Dual(L.ep[1][jl], L.te[1][jl]), beta, 1.0, 1.0);
                const PairTangents<G> tg = pair_tangents<G>(o, dv, false);
                for (int k = 0; k < 3; k++) {
                    lds_add(&L.acc[0][k][il], tg.xi[k]);
                    lds_add(&L.acc[1][k][jl], tg.xj[k]);
                    lds_add(&L.acc[0][3 + k][il], tg.pi[k]);
                    lds_add(&L.acc[1][3 + k][jl], tg.pj[k]);
                    if constexpr (G) {
                        lds_add(&L.acc[0][6 + k][il], tg.gi[k]);
                        lds_add(&L.acc[1][6 + k][jl], tg.gj[k]);
                    }
                }
                acc_l += tg.l;
            }
        };
        int queued = 0;
        const bool vi = L.slot[0][lane] != NB_INVALID;
        const uint32_t islot = L.slot[0][lane];
        // Screen in FP32 with the column values rotating through the wave, on
        // positions relative to the row block's centre from the last build,
        // imaged once more per pair when the block is too wide for one image
        // (as the force kernel's pass 1, same margins); eval repeats the exact
        // FP64 test.
        const float4 bc = bb_ctr[Sd.block0 + r];
        const float4 be = bb_ext[Sd.block0 + r];
        const double slack = 0.5 * Sd.padding_d + 1e-4;
        const double reach = Sd.cutoff_d + 1e-3;
        const bool local = static_cast<double>(be.x) + slack + reach < 0.5 * bx &&
                           static_cast<double>(be.y) + slack + reach < 0.5 * by &&
                           static_cast<double>(be.z) + slack + reach < 0.5 * bz;
        const double cx = bc.x, cy = bc.y, cz = bc.z, cw = L.pw[0][0];
        const float nan = __builtin_nanf("");
        const bool vj0 = L.slot[1][lane] != NB_INVALID;
        float qx = nan, qy = nan, qz = nan, qw = 0.0f;
        if (vj0) {
            qx = static_cast<float>(min_image_d(L.px[1][lane] - cx, bx, ibx));
            qy = static_cast<float>(min_image_d(L.py[1][lane] - cy, by, iby));
            qz = static_cast<float>(min_image_d(L.pz[1][lane] - cz, bz, ibz));
            qw = static_cast<float>(L.pw[1][lane] - cw);
        }
        uint32_t qc = L.slot[1][lane];
        const float fx = static_cast<float>(min_image_d(L.px[0][lane] - cx, bx, ibx));
        const float fy = static_cast<float>(min_image_d(L.py[0][lane] - cy, by, iby));
        const float fz = static_cast<float>(min_image_d(L.pz[0][lane] - cz, bz, ibz));
        const float fw = static_cast<float>(L.pw[0][lane] - cw);
        const float lx = uniform_f(static_cast<float>(bx)), ly = uniform_f(static_cast<float>(by)),
                    lz = uniform_f(static_cast<float>(bz));
        const float cut2v = vi ? static_cast<float>(cut2 * (local ? 1.0 + 1e-5 : 1.0 + 1e-4)) : -1.0f;
        const bool wrap = !local;
        for (int round = 0; round < TILE; round++) {
            const int c = (lane + round) & (TILE - 1); // the column slot this lane holds
            float dx = qx - fx;
            float dy = qy - fy;
            float dz = qz - fz;
            if (wrap) {
                dx = __builtin_fminf(__builtin_fabsf(dx), lx - __builtin_fabsf(dx));
                dy = __builtin_fminf(__builtin_fabsf(dy), ly - __builtin_fabsf(dy));
                dz = __builtin_fminf(__builtin_fabsf(dz), lz - __builtin_fabsf(dz));
            }
            const float dw = qw - fw;
            const float d2 = __builtin_fmaf(dw, dw, __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)));
            const bool in = d2 < cut2v && islot < qc; // padding columns: NaN d2 and NB_INVALID slot
            const uint64_t m = ballot(in);
            if (in) {
                const int at = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), queued));
                L.pq[at] = static_cast<uint16_t>(lane | (c << 6));
            }
            queued += __popcll(m);
            if (queued >= TILE) {
                wave_lds_sync();
                eval(TILE);
                wave_lds_sync(); // pass 2 has read the first 64 entries
                queued -= TILE;
                const uint16_t rest = L.pq[TILE + lane];
                if (lane < queued) {
                    L.pq[lane] = rest;
                }
                wave_lds_sync();
            }
            qx = rot1(qx);
            qy = rot1(qy);
            qz = rot1(qz);
            qw = rot1(qw);
            qc = rot1(qc);
        }
        if (queued > 0) {
            wave_lds_sync();
            eval(queued);
        }
        wave_lds_sync();
        // the tile's per-atom tangents to global memory
        for (int side = 0; side < 2; side++) {
            if (L.slot[side][lane] != NB_INVALID) {
                const int g = L.gidx[side][lane];
                for (int k = 0; k < 3; k++) {
                    if (out.hx) {
                        atomic_add_fixed(out.hx + g * 3 + k, L.acc[side][k][lane]);
                    }
                    if (out.hp) {
                        atomic_add_fixed(out.hp + g * 3 + k, L.acc[side][3 + k][lane]);
                    }
                    if constexpr (G) {
                        atomic_add_fixed(out.gp + g * 3 + k, L.acc[side][6 + k][lane]);
                    }
                }
            }
        }
        if (out.hl) {
            const uint64_t tot = wave_sum_u64(acc_l);
            if (lane == 0) {
                atomic_add_fixed(out.hl + Sd.red_idx, tot);
            }
        }
    }
}

// p(lambda) and its tangent, one thread per parameter word.
__global__ __launch_bounds__(256) void k_interp_hvp_params(int n3, double lam, const double *__restrict__ p,
                                                           const double *__restrict__ vp, double vl,
                                                           double *__restrict__ p_lam, double *__restrict__ vp_lam) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) {
        return;
    }
    const double a = p[i], b = p[n3 + i];
    p_lam[i] = (1.0 - lam) * a + lam * b;
    double t = (b - a) * vl;
    if (vp) {
        t += (1.0 - lam) * vp[i] + lam * vp[n3 + i];
    }
    vp_lam[i] = t;
}

// Chain rule through p(lambda), one thread per parameter word; the du/dlambda
// part is summed per wave in fixed point.
__global__ __launch_bounds__(256) void k_interp_hvp_chain(int n3, double lam, const double *__restrict__ p,
                                                          const double *__restrict__ vp, double vl,
                                                          const uint64_t *__restrict__ g, const uint64_t *__restrict__ dg,
                                                          uint64_t *__restrict__ hp, uint64_t *__restrict__ hl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t l = 0;
    if (i < n3) {
        const int c = i % 3;
        const double scale = c == 0 ? FIXED_SCALE : (c == 1 ? FIXED_SCALE_DSIG : FIXED_SCALE_DEPS);
        const double gv = fixed_to_double(g[i], scale);
        const double dv = fixed_to_double(dg[i], scale);
        if (hp) {
            hp[i] += f64_to_fixed((1.0 - lam) * dv - vl * gv, scale);
            hp[n3 + i] += f64_to_fixed(lam * dv + vl * gv, scale);
        }
        const double dvp = vp ? vp[n3 + i] - vp[i] : 0.0;
        l = f64_to_fixed(dv * (p[n3 + i] - p[i]) + gv * dvp, FIXED_SCALE);
    }
    if (hl) {
        const uint64_t tot = wave_sum_u64(l);
        if ((threadIdx.x & 63) == 0) {
            atomic_add_fixed(hl, tot);
        }
    }
}

} // namespace

void launch_interp_hvp_params(int N, double lambda, const double *params, const double *vp, double vl, double *p_lam,
                              double *vp_lam, hipStream_t stream) {
    const int n3 = N * 3;
    if (n3 == 0) {
        return;
    }
    k_interp_hvp_params<<<ceil_div(n3, 256), 256, 0, stream>>>(n3, lambda, params, vp, vl, p_lam, vp_lam);
    TM_LAUNCH_CHECK();
}

void launch_interp_hvp_chain(int N, double lambda, const double *params, const double *vp, double vl, const uint64_t *g,
                             const uint64_t *dg, uint64_t *hp, uint64_t *hl, hipStream_t stream) {
    const int n3 = N * 3;
    if (n3 == 0 || (!hp && !hl)) {
        return;
    }
    k_interp_hvp_chain<<<ceil_div(n3, 256), 256, 0, stream>>>(n3, lambda, params, vp, vl, g, dg, hp, hl);
    TM_LAUNCH_CHECK();
}

void launch_bonded_hvp(BondedKind kind, const BondedTerms &t, const double *x, const double *lambdas, double lambda0,
                       const int *red_idx, const HvpTangent &v, const HvpOut &o, hipStream_t stream) {
    if (t.n == 0 || (!o.hx && !o.hp && !o.hl)) {
        return;
    }
    const int grid = ceil_div(t.n, 128);
    switch (kind) {
    case BondedKind::Bond:
        k_bonded_hvp<0><<<grid, 128, 0, stream>>>(t, x, lambdas, lambda0, red_idx, v, o);
        break;
    case BondedKind::Angle:
        k_bonded_hvp<1><<<grid, 128, 0, stream>>>(t, x, lambdas, lambda0, red_idx, v, o);
        break;
    case BondedKind::Torsion:
        k_bonded_hvp<2><<<grid, 128, 0, stream>>>(t, x, lambdas, lambda0, red_idx, v, o);
        break;
    }
    TM_LAUNCH_CHECK();
}

void launch_pair_list_hvp(const PairListArgs &a, const HvpTangent &v, const HvpOut &o, hipStream_t stream) {
    if (a.M == 0 || (!o.hx && !o.hp && !o.hl && !o.gp)) {
        return;
    }
    if (a.params1 || a.slots) {
        throw Error("Hessian-vector products: plain parameters only");
    }
    if (o.gp) {
        k_pair_list_hvp<true><<<ceil_div(a.M, 256), 256, 0, stream>>>(a, v, o);
    } else {
        k_pair_list_hvp<false><<<ceil_div(a.M, 256), 256, 0, stream>>>(a, v, o);
    }
    TM_LAUNCH_CHECK();
}

void launch_tiles_hvp(const NbState &st, const double *params, const int *offset, const double *boxes,
                      const HvpTangent &v, const HvpOut &o, hipStream_t stream) {
    if (st.S == 0 || (!o.hx && !o.hp && !o.hl && !o.gp)) {
        return;
    }
    // one wave per tile, the grid striding over the tile count read on the device
    const int grid = std::max(1, std::min(st.total_tiles, 16384));
    if (o.gp) {
        k_tiles_hvp<true><<<grid, TILE, 0, stream>>>(st.sys.ptr, st.S, st.tile_prefix.ptr, st.tile_order.ptr,
                                                      st.tile_row.ptr,
                                                      st.tile_atoms.ptr, st.perm.ptr, st.pos.ptr, params, offset,
                                                      boxes, st.bb_ctr.ptr, st.bb_ext.ptr, st.status.ptr, v, o);
    } else {
        k_tiles_hvp<false><<<grid, TILE, 0, stream>>>(st.sys.ptr, st.S, st.tile_prefix.ptr, st.tile_order.ptr,
                                                       st.tile_row.ptr,
                                                       st.tile_atoms.ptr, st.perm.ptr, st.pos.ptr, params, offset,
                                                       boxes, st.bb_ctr.ptr, st.bb_ext.ptr, st.status.ptr, v, o);
    }
    TM_LAUNCH_CHECK();
}

} // namespace tmamd
