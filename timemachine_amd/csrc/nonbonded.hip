// Nonbonded pair kernels: the 64x64 tile kernel over the neighborlist and the
// pair-list kernel used for exclusions.
//
// Tile kernel (restates the semantics of kernels/k_nonbonded.cuh:311-671 for
// wave64): one wave per tile, the row atom of the tile in each lane and the
// 64 column atoms rotated one lane per round with DPP wave_rol:1, so after
// 64 rounds every (row, column) pair has met exactly once and every column
// accumulator is back in its own lane. Force, du/dp, du/dl and energy terms
// are rounded to 64-bit fixed point per pair and summed in registers;
// results leave through 64-bit global atomics (row and column atoms) and one
// wave-reduced atomic per tile for u and du/dl.
//
// Pair-list kernel (kernels/k_nonbonded_pair_list.cuh:17-216): one thread
// per pair; with negated=1 it subtracts exactly what the tile kernel added,
// which is how exclusions cancel.
#include "nb_pair.h"
#include "nb_state.h"
#include "wave.h"

namespace tmamd {

namespace {

constexpr int NB_BLOCK = 256;
constexpr int NB_BAD_BOX = NB_STATUS_BOX_TOO_SMALL | NB_STATUS_BOX_NOT_ORTHO;

__device__ __forceinline__ void atomic_add_u64(uint64_t *p, uint64_t v) {
    atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(v));
}

template <bool ALCH, bool U, bool DX, bool DL, bool DP>
__device__ __forceinline__ void tile_body(
    const NbSys &S,
    const int s,
    const int r,
    const uint32_t jl0,
    const double4 pi,
    const float4 ai,
    double4 pj,
    float4 aj,
    const uint32_t *__restrict__ perm,
    const double *__restrict__ boxes,
    uint64_t *__restrict__ du_dx,
    uint64_t *__restrict__ du_dp,
    uint64_t *__restrict__ du_dl,
    uint64_t *__restrict__ u) {
    const int lane = threadIdx.x & 63;
    const uint32_t i = static_cast<uint32_t>(r * TILE + lane);
    const uint32_t k = static_cast<uint32_t>(S.k);
    const bool vi = i < k;

    const double bx = boxes[s * 9 + 0];
    const double by = boxes[s * 9 + 4];
    const double bz = boxes[s * 9 + 8];
    const double ibx = 1.0 / bx;
    const double iby = 1.0 / by;
    const double ibz = 1.0 / bz;
    const double cut2 = S.cutoff_d * S.cutoff_d;
    const double beta = static_cast<double>(S.beta);

    uint64_t gix = 0, giy = 0, giz = 0;
    uint64_t gjx = 0, gjy = 0, gjz = 0;
    uint64_t gqi = 0, gsi = 0, gei = 0;
    uint64_t gqj = 0, gsj = 0, gej = 0;
    uint64_t uacc = 0, dlacc = 0;
    uint32_t jc = jl0;

#pragma unroll 2
    for (int round = 0; round < TILE; round++) {
        const double dx = min_image_d(pi.x - pj.x, bx, ibx);
        const double dy = min_image_d(pi.y - pj.y, by, iby);
        const double dz = min_image_d(pi.z - pj.z, bz, ibz);
        double d2 = dx * dx + dy * dy + dz * dz;
        double dw = 0.0;
        if (ALCH) {
            dw = pi.w - pj.w;
            d2 = d2 + dw * dw;
        }
        const bool valid = vi && (jc < k) && (i < jc);
        if (valid && d2 < cut2) {
            const PairTerms T = nb_pair_terms<U, DP>(d2, ai.x, aj.x, ai.y, aj.y, ai.z, aj.z, beta, 1.0, 1.0);
            if (DX) {
                const uint64_t fx = f64_to_fixed(T.prefactor * dx, FIXED_SCALE);
                const uint64_t fy = f64_to_fixed(T.prefactor * dy, FIXED_SCALE);
                const uint64_t fz = f64_to_fixed(T.prefactor * dz, FIXED_SCALE);
                gix += fx;
                giy += fy;
                giz += fz;
                gjx -= fx;
                gjy -= fy;
                gjz -= fz;
            }
            if (DP) {
                gqi += f64_to_fixed((1.0 * aj.x) * T.erfc_inv_d, FIXED_SCALE);
                gqj += f64_to_fixed((1.0 * ai.x) * T.erfc_inv_d, FIXED_SCALE);
                if (T.lj) {
                    const uint64_t gs = f64_to_fixed(T.sig_grad, FIXED_SCALE_DSIG);
                    gsi += gs;
                    gsj += gs;
                    gei += f64_to_fixed(T.eps_grad * aj.z, FIXED_SCALE_DEPS);
                    gej += f64_to_fixed(T.eps_grad * ai.z, FIXED_SCALE_DEPS);
                }
            }
            if (DL && ALCH) {
                dlacc += f64_to_fixed(dw * T.prefactor * (static_cast<double>(ai.w) - static_cast<double>(aj.w)),
                                      FIXED_SCALE);
            }
            if (U) {
                uacc += f64_to_fixed(T.u, FIXED_SCALE);
            }
        }
        pj.x = rot1(pj.x);
        pj.y = rot1(pj.y);
        pj.z = rot1(pj.z);
        aj.x = rot1(aj.x);
        aj.y = rot1(aj.y);
        aj.z = rot1(aj.z);
        jc = rot1(jc);
        if (ALCH) {
            pj.w = rot1(pj.w);
            if (DL) {
                aj.w = rot1(aj.w);
            }
        }
        if (DX) {
            gjx = rot1(gjx);
            gjy = rot1(gjy);
            gjz = rot1(gjz);
        }
        if (DP) {
            gqj = rot1(gqj);
            gsj = rot1(gsj);
            gej = rot1(gej);
        }
    }

    const bool vj = jl0 < k;
    if (DX || DP) {
        if (vi) {
            const int gi = S.atom0 + static_cast<int>(perm[S.slot0 + i]);
            if (DX) {
                atomic_add_u64(du_dx + gi * 3 + 0, gix);
                atomic_add_u64(du_dx + gi * 3 + 1, giy);
                atomic_add_u64(du_dx + gi * 3 + 2, giz);
            }
            if (DP) {
                atomic_add_u64(du_dp + gi * 3 + 0, gqi);
                atomic_add_u64(du_dp + gi * 3 + 1, gsi);
                atomic_add_u64(du_dp + gi * 3 + 2, gei);
            }
        }
        if (vj) {
            const int gj = S.atom0 + static_cast<int>(perm[S.slot0 + jl0]);
            if (DX) {
                atomic_add_u64(du_dx + gj * 3 + 0, gjx);
                atomic_add_u64(du_dx + gj * 3 + 1, gjy);
                atomic_add_u64(du_dx + gj * 3 + 2, gjz);
            }
            if (DP) {
                atomic_add_u64(du_dp + gj * 3 + 0, gqj);
                atomic_add_u64(du_dp + gj * 3 + 1, gsj);
                atomic_add_u64(du_dp + gj * 3 + 2, gej);
            }
        }
    }
    if (U) {
        const uint64_t tot = wave_sum_u64(uacc);
        if (lane == 0 && tot != 0) {
            atomic_add_u64(u + S.red_idx, tot);
        }
    }
    if (DL && ALCH) {
        const uint64_t tot = wave_sum_u64(dlacc);
        if (lane == 0 && tot != 0) {
            atomic_add_u64(du_dl + S.red_idx, tot);
        }
    }
}

// Grid-stride over all tiles of all systems (the grid is sized to the
// machine, the tile count is read on the device).
template <bool U, bool DX, bool DL, bool DP>
__global__ __launch_bounds__(NB_BLOCK) void k_nonbonded_tiles(
    const NbSys *__restrict__ sys,
    const int S,
    const int *__restrict__ tile_prefix,
    const int *__restrict__ tile_row,
    const uint32_t *__restrict__ tile_atoms,
    const uint32_t *__restrict__ perm,
    const double4 *__restrict__ pos,
    const float4 *__restrict__ par,
    const double *__restrict__ boxes,
    const int *__restrict__ status,
    uint64_t *__restrict__ du_dx,
    uint64_t *__restrict__ du_dp,
    uint64_t *__restrict__ du_dl,
    uint64_t *__restrict__ u) {
    const int lane = threadIdx.x & 63;
    const int waves_per_block = NB_BLOCK / 64;
    const int first = uniform(blockIdx.x * waves_per_block + (threadIdx.x >> 6));
    const int stride = gridDim.x * waves_per_block;
    const int total = tile_prefix[S];
    for (int t = first; t < total; t += stride) {
        int lo = 0, hi = S - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tile_prefix[mid] <= t) {
                lo = mid;
            } else {
                hi = mid - 1;
            }
        }
        const int s = lo;
        if (status[s] & NB_BAD_BOX) {
            continue;
        }
        const NbSys Sd = sys[s];
        const int kt = Sd.tile0 + (t - tile_prefix[s]);
        const int r = tile_row[kt];
        const uint32_t jl = tile_atoms[static_cast<size_t>(kt) * TILE + lane];
        const int i = r * TILE + lane;
        const double4 pi = pos[Sd.slot0 + i];
        const float4 ai = par[Sd.slot0 + i];
        const bool vj = jl < static_cast<uint32_t>(Sd.k);
        const double4 pj = vj ? pos[Sd.slot0 + jl] : make_double4(0.0, 0.0, 0.0, 0.0);
        const float4 aj = vj ? par[Sd.slot0 + jl] : make_float4(0.f, 0.f, 0.f, 0.f);
        const bool alch = ballot(pi.w != 0.0 || pj.w != 0.0) != 0;
        if (alch) {
            tile_body<true, U, DX, DL, DP>(Sd, s, r, jl, pi, ai, pj, aj, perm, boxes, du_dx, du_dp, du_dl, u);
        } else {
            tile_body<false, U, DX, DL, DP>(Sd, s, r, jl, pi, ai, pj, aj, perm, boxes, du_dx, du_dp, du_dl, u);
        }
    }
}

using TileKernel = void (*)(const NbSys *, int, const int *, const int *, const uint32_t *, const uint32_t *,
                            const double4 *, const float4 *, const double *, const int *, uint64_t *, uint64_t *,
                            uint64_t *, uint64_t *);

// index bits: 8 = u, 4 = du_dx, 2 = du_dl, 1 = du_dp
const TileKernel tile_kernels[16] = {
    k_nonbonded_tiles<0, 0, 0, 0>, k_nonbonded_tiles<0, 0, 0, 1>, k_nonbonded_tiles<0, 0, 1, 0>,
    k_nonbonded_tiles<0, 0, 1, 1>, k_nonbonded_tiles<0, 1, 0, 0>, k_nonbonded_tiles<0, 1, 0, 1>,
    k_nonbonded_tiles<0, 1, 1, 0>, k_nonbonded_tiles<0, 1, 1, 1>, k_nonbonded_tiles<1, 0, 0, 0>,
    k_nonbonded_tiles<1, 0, 0, 1>, k_nonbonded_tiles<1, 0, 1, 0>, k_nonbonded_tiles<1, 0, 1, 1>,
    k_nonbonded_tiles<1, 1, 0, 0>, k_nonbonded_tiles<1, 1, 0, 1>, k_nonbonded_tiles<1, 1, 1, 0>,
    k_nonbonded_tiles<1, 1, 1, 1>,
};

template <bool U, bool DX, bool DL, bool DP>
__global__ __launch_bounds__(256) void k_pair_list(
    const PairListArgs a, uint64_t *__restrict__ du_dx, uint64_t *__restrict__ du_dp, uint64_t *__restrict__ du_dl,
    uint64_t *__restrict__ u) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.M) {
        return;
    }
    const int s = a.pair_sys ? a.pair_sys[p] : 0;
    const int gi = a.pairs[p * 2 + 0];
    const int gj = a.pairs[p * 2 + 1];
    const double *box = a.boxes + s * 9;
    const double lam = a.lambdas ? a.lambdas[s] : a.lambda0;
    const double cutoff_d = a.cutoffs ? a.cutoffs[s] : a.cutoff0;
    const float beta = a.betas ? a.betas[s] : a.beta0;
    const int red = a.red_idx ? a.red_idx[s] : 0;

    const double bx = box[0];
    const double by = box[4];
    const double bz = box[8];
    const double ibx = 1.0 / bx;
    const double iby = 1.0 / by;
    const double ibz = 1.0 / bz;

    const int pli = a.plane ? a.plane[gi] : 0, plj = a.plane ? a.plane[gj] : 0;
    const int ofi = a.offset ? a.offset[gi] : 0, ofj = a.offset ? a.offset[gj] : 0;
    const double wi = (pli + ofi * lam) * cutoff_d;
    const double wj = (plj + ofj * lam) * cutoff_d;
    const double dwdl_i = static_cast<double>(static_cast<float>(ofi * cutoff_d));
    const double dwdl_j = static_cast<double>(static_cast<float>(ofj * cutoff_d));

    const double dx = min_image_d(a.x[gi * 3 + 0] - a.x[gj * 3 + 0], bx, ibx);
    const double dy = min_image_d(a.x[gi * 3 + 1] - a.x[gj * 3 + 1], by, iby);
    const double dz = min_image_d(a.x[gi * 3 + 2] - a.x[gj * 3 + 2], bz, ibz);
    const double dw = wi - wj;
    double d2 = dx * dx + dy * dy + dz * dz;
    d2 = d2 + dw * dw;
    if (!(d2 < cutoff_d * cutoff_d)) {
        return;
    }
    const double qi = static_cast<float>(a.params[gi * 3 + 0]), qj = static_cast<float>(a.params[gj * 3 + 0]);
    const double si = static_cast<float>(a.params[gi * 3 + 1]), sj = static_cast<float>(a.params[gj * 3 + 1]);
    const double ei = static_cast<float>(a.params[gi * 3 + 2]), ej = static_cast<float>(a.params[gj * 3 + 2]);
    const double scale_q = a.scales[p * 2 + 0];
    const double scale_lj = a.scales[p * 2 + 1];
    const PairTerms T = nb_pair_terms<U, DP>(d2, qi, qj, si, sj, ei, ej, static_cast<double>(beta), scale_q, scale_lj);
    const bool neg = a.negated != 0;
    auto acc = [neg](uint64_t *dst, uint64_t v) { atomic_add_u64(dst, neg ? (0ull - v) : v); };
    if (DX) {
        const uint64_t fx = f64_to_fixed(T.prefactor * dx, FIXED_SCALE);
        const uint64_t fy = f64_to_fixed(T.prefactor * dy, FIXED_SCALE);
        const uint64_t fz = f64_to_fixed(T.prefactor * dz, FIXED_SCALE);
        acc(du_dx + gi * 3 + 0, fx);
        acc(du_dx + gi * 3 + 1, fy);
        acc(du_dx + gi * 3 + 2, fz);
        acc(du_dx + gj * 3 + 0, 0ull - fx);
        acc(du_dx + gj * 3 + 1, 0ull - fy);
        acc(du_dx + gj * 3 + 2, 0ull - fz);
    }
    if (DP) {
        acc(du_dp + gi * 3 + 0, f64_to_fixed((scale_q * qj) * T.erfc_inv_d, FIXED_SCALE));
        acc(du_dp + gj * 3 + 0, f64_to_fixed((scale_q * qi) * T.erfc_inv_d, FIXED_SCALE));
        if (T.lj) {
            const uint64_t gs = f64_to_fixed(T.sig_grad, FIXED_SCALE_DSIG);
            acc(du_dp + gi * 3 + 1, gs);
            acc(du_dp + gj * 3 + 1, gs);
            acc(du_dp + gi * 3 + 2, f64_to_fixed(T.eps_grad * ej, FIXED_SCALE_DEPS));
            acc(du_dp + gj * 3 + 2, f64_to_fixed(T.eps_grad * ei, FIXED_SCALE_DEPS));
        }
    }
    if (DL) {
        const bool vanilla = (wi == 0.0) && (wj == 0.0);
        if (!vanilla) {
            acc(du_dl + red, f64_to_fixed(dw * T.prefactor * (dwdl_i - dwdl_j), FIXED_SCALE));
        }
    }
    if (U) {
        acc(u + red, f64_to_fixed(T.u, FIXED_SCALE));
    }
}

using PairKernel = void (*)(const PairListArgs, uint64_t *, uint64_t *, uint64_t *, uint64_t *);
const PairKernel pair_kernels[16] = {
    k_pair_list<0, 0, 0, 0>, k_pair_list<0, 0, 0, 1>, k_pair_list<0, 0, 1, 0>, k_pair_list<0, 0, 1, 1>,
    k_pair_list<0, 1, 0, 0>, k_pair_list<0, 1, 0, 1>, k_pair_list<0, 1, 1, 0>, k_pair_list<0, 1, 1, 1>,
    k_pair_list<1, 0, 0, 0>, k_pair_list<1, 0, 0, 1>, k_pair_list<1, 0, 1, 0>, k_pair_list<1, 0, 1, 1>,
    k_pair_list<1, 1, 0, 0>, k_pair_list<1, 1, 0, 1>, k_pair_list<1, 1, 1, 0>, k_pair_list<1, 1, 1, 1>,
};

inline int kernel_index(const void *du_dx, const void *du_dp, const void *du_dl, const void *u) {
    return (u ? 8 : 0) | (du_dx ? 4 : 0) | (du_dl ? 2 : 0) | (du_dp ? 1 : 0);
}

int default_tile_grid() {
    static int grid = 0;
    if (grid == 0) {
        int dev = 0;
        TM_HIP_CHECK(hipGetDevice(&dev));
        int cus = 0;
        TM_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        grid = cus * 8; // 8 workgroups (32 waves) per CU: enough to hide latency, no tail blow-up
    }
    return grid;
}

} // namespace

void NbState::compute(
    const double *boxes,
    uint64_t *du_dx,
    uint64_t *du_dp,
    uint64_t *du_dl,
    uint64_t *u,
    hipStream_t stream,
    int grid_override) const {
    const int idx = kernel_index(du_dx, du_dp, du_dl, u);
    if (idx == 0) {
        return;
    }
    const int grid = grid_override > 0 ? grid_override : default_tile_grid();
    tile_kernels[idx]<<<grid, NB_BLOCK, 0, stream>>>(sys.ptr, S, tile_prefix.ptr, tile_row.ptr, tile_atoms.ptr,
                                                      perm.ptr, pos.ptr, par.ptr, boxes, status.ptr, du_dx, du_dp,
                                                      du_dl, u);
    TM_LAUNCH_CHECK();
}

void launch_pair_list(
    const PairListArgs &a, uint64_t *du_dx, uint64_t *du_dp, uint64_t *du_dl, uint64_t *u, hipStream_t stream) {
    const int idx = kernel_index(du_dx, du_dp, du_dl, u);
    if (idx == 0 || a.M == 0) {
        return;
    }
    pair_kernels[idx]<<<ceil_div(a.M, 256), 256, 0, stream>>>(a, du_dx, du_dp, du_dl, u);
    TM_LAUNCH_CHECK();
}

} // namespace tmamd
