// HarmonicBond / HarmonicAngle / PeriodicTorsion kernels, one thread per term,
// over the terms of one or many systems (term_sys maps a term to its system
// for lambda and the reduction index). Atom indices are global. Arithmetic is
// FP64 (bonded_terms.h), accumulation 64-bit fixed point at 2^36. u and
// du/dlambda are summed over the wave per system before one atomic each
// (a per-term atomic on one address per system serialises the whole launch).
#include "bonded.h"
#include "bonded_terms.h"
#include "fixed_point.h"
#include "wave.h"

namespace tmamd {

namespace {

__device__ __forceinline__ void add_fixed(uint64_t *p, double v) {
    const uint64_t f = f64_to_fixed(v, FIXED_SCALE);
    if (f != 0) {
        atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(f));
    }
}

__device__ __forceinline__ void term_context(const BondedTerms &t, int term, const double *lambdas, double lambda0,
                                             const int *red_idx, double &lam, int &red) {
    const int s = t.term_sys ? t.term_sys[term] : 0;
    lam = lambdas ? lambdas[s] : lambda0;
    red = red_idx ? red_idx[s] : 0;
}

__global__ void k_harmonic_bond(const BondedTerms t, const double *__restrict__ x, const double *__restrict__ lambdas,
                                double lambda0, const int *__restrict__ red_idx, uint64_t *__restrict__ du_dx,
                                uint64_t *__restrict__ du_dp, uint64_t *__restrict__ du_dl, uint64_t *__restrict__ u, int dl_dp,
                                uint64_t *__restrict__ slots) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = b < t.n;
    double lam = 0.0;
    int red = -1;
    uint64_t fu = 0, fdl = 0;
    if (active) {
        term_context(t, b, lambdas, lambda0, red_idx, lam, red);
        const int i = t.idx[b * 2 + 0];
        const int j = t.idx[b * 2 + 1];
        const int mult = t.mult ? t.mult[b] : 0;
        const int off = t.offset ? t.offset[b] : 1;
        // dl_dp: d(du/dlambda)/dp = mult * d(shape)/dp, i.e. du/dp at prefactor mult
        const double pf = dl_dp ? static_cast<double>(mult) : lam_prefactor<double>(mult, off, lam);
        const BondOut<double> o =
            bond_term<double>(x + i * 3, x + j * 3, t.params[b * 2 + 0], t.params[b * 2 + 1], mult, pf);
        if (slots) {
            for (int d = 0; d < 3; d++) {
                slots[(b * 2 + 0) * 3 + d] = f64_to_fixed(o.g[0][d], FIXED_SCALE);
                slots[(b * 2 + 1) * 3 + d] = f64_to_fixed(o.g[1][d], FIXED_SCALE);
            }
        } else if (du_dx) {
            for (int d = 0; d < 3; d++) {
                add_fixed(du_dx + i * 3 + d, o.g[0][d]);
                add_fixed(du_dx + j * 3 + d, o.g[1][d]);
            }
        }
        if (du_dp) {
            add_fixed(du_dp + b * 2 + 0, o.dkb);
            add_fixed(du_dp + b * 2 + 1, o.db0);
        }
        fu = f64_to_fixed(o.u, FIXED_SCALE);
        fdl = f64_to_fixed(o.dl, FIXED_SCALE);
    }
    if (u) {
        wave_add_by_key(u, red, fu);
    }
    if (du_dl) {
        wave_add_by_key(du_dl, red, fdl);
    }
}

__global__ void k_harmonic_angle(const BondedTerms t, const double *__restrict__ x, const double *__restrict__ lambdas,
                                 double lambda0, const int *__restrict__ red_idx, uint64_t *__restrict__ du_dx,
                                 uint64_t *__restrict__ du_dp, uint64_t *__restrict__ du_dl,
                                 uint64_t *__restrict__ u, int dl_dp, uint64_t *__restrict__ slots) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = a < t.n;
    double lam = 0.0;
    int red = -1;
    uint64_t fu = 0, fdl = 0;
    if (active) {
        term_context(t, a, lambdas, lambda0, red_idx, lam, red);
        const int i = t.idx[a * 3 + 0];
        const int j = t.idx[a * 3 + 1];
        const int k = t.idx[a * 3 + 2];
        const int mult = t.mult ? t.mult[a] : 0;
        const int off = t.offset ? t.offset[a] : 1;
        // dl_dp: d(du/dlambda)/dp = mult * d(shape)/dp, i.e. du/dp at prefactor mult
        const double pf = dl_dp ? static_cast<double>(mult) : lam_prefactor<double>(mult, off, lam);
        const AngleOut<double> o =
            angle_term<double>(x + i * 3, x + j * 3, x + k * 3, t.params[a * 2 + 0], t.params[a * 2 + 1], mult, pf);
        if (slots) {
            for (int d = 0; d < 3; d++) {
                slots[(a * 3 + 0) * 3 + d] = f64_to_fixed(o.g[0][d], FIXED_SCALE);
                slots[(a * 3 + 1) * 3 + d] = f64_to_fixed(o.g[1][d], FIXED_SCALE);
                slots[(a * 3 + 2) * 3 + d] = f64_to_fixed(o.g[2][d], FIXED_SCALE);
            }
        } else if (du_dx) {
            for (int d = 0; d < 3; d++) {
                add_fixed(du_dx + i * 3 + d, o.g[0][d]);
                add_fixed(du_dx + j * 3 + d, o.g[1][d]);
                add_fixed(du_dx + k * 3 + d, o.g[2][d]);
            }
        }
        if (du_dp) {
            add_fixed(du_dp + a * 2 + 0, o.dka);
            add_fixed(du_dp + a * 2 + 1, o.da0);
        }
        fu = f64_to_fixed(o.u, FIXED_SCALE);
        fdl = f64_to_fixed(o.dl, FIXED_SCALE);
    }
    if (u) {
        wave_add_by_key(u, red, fu);
    }
    if (du_dl) {
        wave_add_by_key(du_dl, red, fdl);
    }
}

__global__ void k_periodic_torsion(const BondedTerms t, const double *__restrict__ x,
                                   const double *__restrict__ lambdas, double lambda0, const int *__restrict__ red_idx,
                                   uint64_t *__restrict__ du_dx, uint64_t *__restrict__ du_dp,
                                   uint64_t *__restrict__ du_dl, uint64_t *__restrict__ u, int dl_dp,
                                   uint64_t *__restrict__ slots) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = q < t.n;
    double lam = 0.0;
    int red = -1;
    uint64_t fu = 0, fdl = 0;
    if (active) {
        term_context(t, q, lambdas, lambda0, red_idx, lam, red);
        const int i = t.idx[q * 4 + 0];
        const int j = t.idx[q * 4 + 1];
        const int k = t.idx[q * 4 + 2];
        const int l = t.idx[q * 4 + 3];
        const int mult = t.mult ? t.mult[q] : 0;
        const int off = t.offset ? t.offset[q] : 1;
        // dl_dp: d(du/dlambda)/dp = mult * d(shape)/dp, i.e. du/dp at prefactor mult
        const double pf = dl_dp ? static_cast<double>(mult) : lam_prefactor<double>(mult, off, lam);
        const TorsionOut<double> o = torsion_term<double>(x + i * 3, x + j * 3, x + k * 3, x + l * 3,
                                                          t.params[q * 3 + 0], t.params[q * 3 + 1],
                                                          t.params[q * 3 + 2], mult, pf);
        if (slots) {
            for (int a = 0; a < 4; a++) {
                for (int d = 0; d < 3; d++) {
                    slots[(q * 4 + a) * 3 + d] = f64_to_fixed(o.g[a][d], FIXED_SCALE);
                }
            }
        } else if (du_dx) {
            const int at[4] = {i, j, k, l};
            for (int a = 0; a < 4; a++) {
                for (int d = 0; d < 3; d++) {
                    add_fixed(du_dx + at[a] * 3 + d, o.g[a][d]);
                }
            }
        }
        if (du_dp) {
            add_fixed(du_dp + q * 3 + 0, o.dk);
            add_fixed(du_dp + q * 3 + 1, o.dphase);
            add_fixed(du_dp + q * 3 + 2, o.dperiod);
        }
        fu = f64_to_fixed(o.u, FIXED_SCALE);
        fdl = f64_to_fixed(o.dl, FIXED_SCALE);
    }
    if (u) {
        wave_add_by_key(u, red, fu);
    }
    if (du_dl) {
        wave_add_by_key(du_dl, red, fdl);
    }
}

// Sum of the fixed-point force slots incident to each atom, added to du_dx
// with a plain read-modify-write (one thread per atom owns its three words).
__global__ void k_bonded_gather(const BondedIncidence inc, const uint64_t *__restrict__ slots,
                                uint64_t *__restrict__ du_dx) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= inc.A) {
        return;
    }
    const int e0 = inc.off[a], e1 = inc.off[a + 1];
    if (e0 == e1) {
        return;
    }
    uint64_t acc[3] = {0, 0, 0};
    for (int e = e0; e < e1; e++) {
        const uint64_t *g = slots + static_cast<size_t>(inc.items[e]) * 3;
        acc[0] += g[0];
        acc[1] += g[1];
        acc[2] += g[2];
    }
    for (int d = 0; d < 3; d++) {
        if (acc[d] != 0) {
            du_dx[a * 3 + d] += acc[d];
        }
    }
}

} // namespace

void build_bonded_incidence(int A, const std::vector<const std::vector<int> *> &idx, std::vector<int> &off,
                            std::vector<uint32_t> &items, std::vector<int> &slot0) {
    off.assign(static_cast<size_t>(A) + 1, 0);
    slot0.assign(idx.size(), 0);
    size_t n_slots = 0;
    for (size_t k = 0; k < idx.size(); k++) {
        slot0[k] = static_cast<int>(n_slots);
        n_slots += idx[k]->size();
        for (int a : *idx[k]) {
            off[a + 1] += 1;
        }
    }
    if (n_slots >= (1u << 31)) {
        throw Error("bonded incidence: too many terms");
    }
    for (int a = 0; a < A; a++) {
        off[a + 1] += off[a];
    }
    items.assign(off[A], 0);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (size_t k = 0; k < idx.size(); k++) {
        const std::vector<int> &v = *idx[k];
        for (size_t e = 0; e < v.size(); e++) {
            items[fill[v[e]]++] = static_cast<uint32_t>(slot0[k] + e); // term e / arity, position e % arity
        }
    }
}

void launch_bonded_gather(const BondedIncidence &inc, const uint64_t *slots, uint64_t *du_dx, hipStream_t stream) {
    if (inc.A == 0) {
        return;
    }
    k_bonded_gather<<<ceil_div(inc.A, 256), 256, 0, stream>>>(inc, slots, du_dx);
    TM_LAUNCH_CHECK();
}

void launch_bonded(BondedKind kind, const BondedTerms &t, const double *x, const double *lambdas, double lambda0,
                   const int *red_idx, uint64_t *du_dx, uint64_t *du_dp, uint64_t *du_dl, uint64_t *u,
                   hipStream_t stream, bool dl_dp, uint64_t *slots) {
    if (t.n == 0 || (!du_dx && !du_dp && !du_dl && !u && !slots)) {
        return;
    }
    if (dl_dp && (du_dx || du_dl || u || !du_dp)) {
        throw Error("launch_bonded: the dl_dp mode writes du_dp only");
    }
    const int grid = ceil_div(t.n, 128);
    switch (kind) {
    case BondedKind::Bond:
        k_harmonic_bond<<<grid, 128, 0, stream>>>(t, x, lambdas, lambda0, red_idx, du_dx, du_dp, du_dl, u, dl_dp ? 1 : 0,
                                                  slots);
        break;
    case BondedKind::Angle:
        k_harmonic_angle<<<grid, 128, 0, stream>>>(t, x, lambdas, lambda0, red_idx, du_dx, du_dp, du_dl, u, dl_dp ? 1 : 0,
                                                   slots);
        break;
    case BondedKind::Torsion:
        k_periodic_torsion<<<grid, 128, 0, stream>>>(t, x, lambdas, lambda0, red_idx, du_dx, du_dp, du_dl, u, dl_dp ? 1 : 0,
                                                     slots);
        break;
    }
    TM_LAUNCH_CHECK();
}

} // namespace tmamd
