// Nonbonded pair kernels: the 64x64 tile kernel over the neighborlist and the
// pair-list kernel used for exclusions.
//
// Tile kernel (restates the semantics of kernels/k_nonbonded.cuh:311-671 for
// wave64): one wave per 64 x 64 tile. The reference evaluates every
// (row, column) slot of a tile; here a first pass over the 64 rounds
// (lane l meets column slot (l + r) mod 64 in round r) screens the pairs
// (FP32 in a frame local to the row block, or the exact FP64 test) and
// compacts those inside the cutoff into an LDS queue, and a second pass
// evaluates the queued pairs 64 at a time with every lane busy, repeating the
// exact FP64 cutoff test. Force, du/dp, du/dl and energy terms are rounded to
// 64-bit fixed point per pair and summed exactly (LDS u64 atomics per atom,
// registers for u and du/dl). Column atoms leave through 64-bit global
// atomics at the end of each tile, row atoms at the end of a run of tiles of
// the same row block, u and du/dl with one wave-reduced atomic per system
// and wave. Integer sums do not depend on the order, so the outputs are
// bitwise those of evaluating the pairs one by one.
//
// Pair-list kernel (kernels/k_nonbonded_pair_list.cuh:17-216): one thread
// per pair; with negated=1 it subtracts exactly what the tile kernel added,
// which is how exclusions cancel.
#include <type_traits>

#include "nb_pair.h"
#include "nb_state.h"
#include "wave.h"

namespace tmamd {

namespace {

// One wave per workgroup: waves share nothing but the output atomics, and a
// one-wave workgroup is scheduled and retired on its own (measured on MI355X at
// an equal wave count: 3% faster per batched step than two-wave workgroups,
// 5% faster than four). Per-wave LDS is 10 KB (forces) to 15 KB (parameter
// derivatives of interpolated parameters).
constexpr int NB_BLOCK = 64;
constexpr int NB_BAD_BOX = NB_STATUS_BOX_TOO_SMALL | NB_STATUS_BOX_NOT_ORTHO;

__device__ __forceinline__ void atomic_add_u64(uint64_t *p, uint64_t v) {
    atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(v));
}

// Queued pairs per wave: fewer than 64 wait between rounds and a round adds
// at most 64, so 128 entries always suffice.
constexpr int PAIR_QUEUE = 2 * TILE;

// Per-wave LDS of the tile kernel. NACC accumulators per atom: x, y, z
// forces, then q, sig, eps parameter derivatives when requested.
// dq/dl, dsig/dl, deps/dl of row and column atoms (interpolated parameters only)
template <bool INTERP> struct TileDp {
    float v[2][3][TILE];
};
template <> struct TileDp<false> {};

// Structure of arrays: consecutive lanes read consecutive 8- or 4-byte words,
// which keeps LDS bank conflicts away (a 32-byte record per lane did not).
template <int NACC, bool INTERP> struct TileLds {
    double pos[2][4][TILE];      // [row, column][x, y, z, w][lane / slot]
    float par[2][4][TILE];       // [row, column][q, sig, eps, dw/dlambda][lane / slot]
    uint32_t colidx[TILE];       // column sorted index (NB_INVALID for padding)
    uint32_t pij[PAIR_QUEUE];    // queued pair: row lane | screen round << 8 (column slot = lane + round mod 64)
    uint64_t acc[2][NACC][TILE]; // [row, column][component][lane / slot]
    TileDp<INTERP> dp;

    __device__ __forceinline__ void put(int side, int l, const double4 &p, const float4 &a) {
        pos[side][0][l] = p.x;
        pos[side][1][l] = p.y;
        pos[side][2][l] = p.z;
        pos[side][3][l] = p.w;
        par[side][0][l] = a.x;
        par[side][1][l] = a.y;
        par[side][2][l] = a.z;
        par[side][3][l] = a.w;
    }
    template <bool W> __device__ __forceinline__ double4 get_pos(int side, int l) const {
        return make_double4(pos[side][0][l], pos[side][1][l], pos[side][2][l], W ? pos[side][3][l] : 0.0);
    }
    __device__ __forceinline__ float4 get_par(int side, int l) const {
        return make_float4(par[side][0][l], par[side][1][l], par[side][2][l], par[side][3][l]);
    }
};

__device__ __forceinline__ void lds_add_u64(uint64_t *p, uint64_t v) {
    atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(v));
}
// p -= v modulo 2^64 (one LDS atomic, no separate negation)
__device__ __forceinline__ void lds_sub_u64(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Minimum-image displacement and squared 4D distance of a pair. Both passes
// and the pair-list kernel evaluate exactly these expressions, so the pass-1
// cutoff test and the pass-2 terms see identical bits.
template <bool ALCH>
__device__ __forceinline__ double pair_delta(const double4 &pi, const double4 &pj, const double bx, const double by,
                                             const double bz, const double ibx, const double iby, const double ibz,
                                             double &dx, double &dy, double &dz, double &dw) {
    dx = min_image_d(pi.x - pj.x, bx, ibx);
    dy = min_image_d(pi.y - pj.y, by, iby);
    dz = min_image_d(pi.z - pj.z, bz, ibz);
    double d2 = pair_d2(dx, dy, dz);
    dw = 0.0;
    if (ALCH) {
        dw = pi.w - pj.w;
        d2 = __builtin_fma(dw, dw, d2);
    }
    return d2;
}

struct BoxDev {
    double bx, by, bz, ibx, iby, ibz;
    double cut2; // squared cutoff: the exact test of pass 2
};

// Pass 2: evaluates the first n <= 64 queued pairs, one
// pair per lane, and adds the fixed-point terms to the LDS accumulators of
// the pair's row and column atom (u and du/dl to per-lane registers).
// DLDP: instead of du/dp, accumulate d(du/dlambda)/dp into the du_dp slots
// (same fixed-point exponents); only alchemical tiles contribute.
template <bool ALCH, bool U, bool DX, bool DL, bool DP, bool DLDP, int NACC, bool INTERP>
__device__ __forceinline__ void eval_pairs(TileLds<NACC, INTERP> &L, const int n, const BoxDev &B,
                                           const double beta, uint64_t &uacc, uint64_t &dlacc) {
    constexpr int PQ = DX ? 3 : 0; // first parameter-derivative component
    // interpolated parameters add du/dp . dp/dl to du/dl, so the LJ parameter gradients are needed
    constexpr bool NEED_DP = DP || (INTERP && ALCH && DL);
    const int lane = threadIdx.x & 63;
    if (lane < n) {
        const uint32_t ij = L.pij[lane];
        const int il = static_cast<int>(ij & 0xFFu);
        const int sl = static_cast<int>((ij + (ij >> 8)) & (TILE - 1));
        const double4 pi = L.template get_pos<ALCH>(0, il);
        const double4 pj = L.template get_pos<ALCH>(1, sl);
        const float4 ai = L.get_par(0, il);
        const float4 aj = L.get_par(1, sl);
        double dx, dy, dz, dw;
        const double d2 = pair_delta<ALCH>(pi, pj, B.bx, B.by, B.bz, B.ibx, B.iby, B.ibz, dx, dy, dz, dw);
        if (!(d2 < B.cut2)) {
            return; // queued by the FP32 screen of pass 1, outside the exact cutoff
        }
        const PairTerms T = nb_pair_terms<U, NEED_DP>(d2, ai.x, aj.x, ai.y, aj.y, ai.z, aj.z, beta, 1.0, 1.0);
        if (DX) {
            // fixed(prefactor * d) per component; scaling by 2^36 is exact, so
            // (prefactor 2^36) d rounds exactly as (prefactor d) 2^36
            const double ps = T.prefactor * FIXED_SCALE;
            uint64_t fx, fy, fz;
            round_to_fixed3(ps * dx, ps * dy, ps * dz, fx, fy, fz);
            lds_add_u64(&L.acc[0][0][il], fx);
            lds_add_u64(&L.acc[0][1][il], fy);
            lds_add_u64(&L.acc[0][2][il], fz);
            lds_sub_u64(&L.acc[1][0][sl], fx);
            lds_sub_u64(&L.acc[1][1][sl], fy);
            lds_sub_u64(&L.acc[1][2][sl], fz);
        }
        if (DP) {
            lds_add_u64(&L.acc[0][PQ + 0][il], f64_to_fixed((1.0 * aj.x) * T.erfc_inv_d, FIXED_SCALE));
            lds_add_u64(&L.acc[1][PQ + 0][sl], f64_to_fixed((1.0 * ai.x) * T.erfc_inv_d, FIXED_SCALE));
            if (T.lj) {
                const uint64_t gs = f64_to_fixed(T.sig_grad, FIXED_SCALE_DSIG);
                lds_add_u64(&L.acc[0][PQ + 1][il], gs);
                lds_add_u64(&L.acc[1][PQ + 1][sl], gs);
                lds_add_u64(&L.acc[0][PQ + 2][il], f64_to_fixed(T.eps_grad * aj.z, FIXED_SCALE_DEPS));
                lds_add_u64(&L.acc[1][PQ + 2][sl], f64_to_fixed(T.eps_grad * ai.z, FIXED_SCALE_DEPS));
            }
        }
        if (ALCH && (DL || DLDP)) {
            if constexpr (DL) {
                if constexpr (INTERP) {
                    dlacc += f64_to_fixed(
                        pair_du_dl(T, dw, ai.w, aj.w, true, 1.0, ai.x, aj.x, ai.z, aj.z, L.dp.v[0][0][il],
                                   L.dp.v[1][0][sl], L.dp.v[0][1][il], L.dp.v[1][1][sl], L.dp.v[0][2][il],
                                   L.dp.v[1][2][sl]),
                        FIXED_SCALE);
                } else {
                    dlacc += f64_to_fixed(
                        dw * T.prefactor * (static_cast<double>(ai.w) - static_cast<double>(aj.w)), FIXED_SCALE);
                }
            }
            if (DLDP) {
                const PairDlDp D = nb_pair_dl_dp(d2, ai.y, aj.y, ai.z, aj.z, beta, 1.0);
                const double dd = dw * (static_cast<double>(ai.w) - static_cast<double>(aj.w));
                lds_add_u64(&L.acc[0][PQ + 0][il],
                            f64_to_fixed((static_cast<double>(aj.x) * D.es_unit) * dd, FIXED_SCALE));
                lds_add_u64(&L.acc[1][PQ + 0][sl],
                            f64_to_fixed((static_cast<double>(ai.x) * D.es_unit) * dd, FIXED_SCALE));
                if (D.lj) {
                    const uint64_t gs = f64_to_fixed(D.lj_dsig * dd, FIXED_SCALE_DSIG);
                    lds_add_u64(&L.acc[0][PQ + 1][il], gs);
                    lds_add_u64(&L.acc[1][PQ + 1][sl], gs);
                    lds_add_u64(&L.acc[0][PQ + 2][il],
                                f64_to_fixed((D.lj_unit * static_cast<double>(aj.z)) * dd, FIXED_SCALE_DEPS));
                    lds_add_u64(&L.acc[1][PQ + 2][sl],
                                f64_to_fixed((D.lj_unit * static_cast<double>(ai.z)) * dd, FIXED_SCALE_DEPS));
                }
            }
        }
        if (U) {
            uacc += f64_to_fixed(T.u, FIXED_SCALE);
        }
    }
}

// One 64 x 64 tile on one wave.
// Pass 1: 64 rounds; in round r lane l pairs its row atom with column slot
// (l + r) mod 64 (read from LDS one round ahead), computes the exact FP64 4D
// distance and queues the indices of the pairs inside the cutoff (ballot +
// mbcnt compaction into an LDS queue). Whenever 64 pairs are queued, pass 2
// recomputes their displacements and evaluates them with every lane busy.
// Fixed-point sums are exact, so the order of evaluation does not change
// any output bit.
template <bool ALCH, bool U, bool DX, bool DL, bool DP, bool DLDP, int NACC, bool INTERP>
__device__ __forceinline__ void tile_body(
    TileLds<NACC, INTERP> &L,
    const NbSys &S,
    const int s,
    const int r,
    const uint32_t jl0,
    const double4 pi,
    const float4 ai,
    const double4 pj0,
    const float4 aj0,
    const float4 dpi,
    const float4 dpj,
    const uint32_t *__restrict__ perm,
    const double *__restrict__ boxes,
    const float4 *__restrict__ bb_ctr,
    const float4 *__restrict__ bb_ext,
    const bool row_fresh,
    const bool row_flush,
    uint64_t &uacc,
    uint64_t &dlacc,
    uint64_t *__restrict__ du_dx,
    uint64_t *__restrict__ du_dp) {
    constexpr int PQ = DX ? 3 : 0;
    const int lane = threadIdx.x & 63;
    const uint32_t i = static_cast<uint32_t>(r * TILE + lane);
    const uint32_t k = static_cast<uint32_t>(S.slot_end); // column slots (NB_INVALID pads) are below this
    const bool vi = i < static_cast<uint32_t>(S.row_end);

    BoxDev B;
    B.bx = boxes[s * 9 + 0];
    B.by = boxes[s * 9 + 4];
    B.bz = boxes[s * 9 + 8];
    B.ibx = 1.0 / B.bx;
    B.iby = 1.0 / B.by;
    B.ibz = 1.0 / B.bz;
    const double cut2 = S.cutoff_d * S.cutoff_d;
    B.cut2 = cut2;
    const double beta = static_cast<double>(S.beta);

    wave_lds_sync(); // the previous tile's LDS reads are complete
    L.put(0, lane, pi, ai);
    L.put(1, lane, pj0, aj0);
    L.colidx[lane] = jl0;
    if constexpr (INTERP) {
        L.dp.v[0][0][lane] = dpi.x;
        L.dp.v[0][1][lane] = dpi.y;
        L.dp.v[0][2][lane] = dpi.z;
        L.dp.v[1][0][lane] = dpj.x;
        L.dp.v[1][1][lane] = dpj.y;
        L.dp.v[1][2][lane] = dpj.z;
    }
    // row accumulators carry over the tiles of one row run (see k_nonbonded_tiles)
#pragma unroll
    for (int c = 0; c < NACC; c++) {
        if (row_fresh) {
            L.acc[0][c][lane] = 0;
        }
        L.acc[1][c][lane] = 0;
    }
    wave_lds_sync();

    // Entries [0, queued) of L.pij wait for pass 2 (wave-uniform count, below
    // 64 between rounds).
    int queued = 0;
    // queue this lane's pair of the round when `in`; at 64 or more, evaluate
    // the first 64 and move the rest (fewer than 64) to the front
    auto queue = [&](const int round, const bool in) {
        const uint64_t m = ballot(in);
        if (in) {
            const int at = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), queued));
            L.pij[at] = static_cast<uint32_t>(lane) | (static_cast<uint32_t>(round) << 8);
        }
        queued += __popcll(m);
        if (queued >= TILE) {
            wave_lds_sync();
            eval_pairs<ALCH, U, DX, DL, DP, DLDP, NACC, INTERP>(L, TILE, B, beta, uacc, dlacc);
            wave_lds_sync(); // pass 2 has read the first 64 entries
            queued -= TILE;
            const uint32_t rest = L.pij[TILE + (lane & (TILE - 1))];
            if (lane < queued) {
                L.pij[lane] = rest;
            }
            wave_lds_sync();
        }
    };

    // Pass 1, round r pairs lane l with column slot (l + r) mod 64.
    //
    // Screen in FP32 on positions relative to the centre c of the row block,
    // each atom imaged once (per axis a = min_image(x - c)). If every row atom
    // is within e of c and e + cutoff < L / 2 on every axis ("local"), a pair
    // within the cutoff has a_j - a_i equal to its minimum-image displacement:
    // any other image differs by a multiple of L, so |a_j - a_i| >= L - cutoff
    // > L / 2 + e, which a_j - a_i cannot reach. Otherwise (small boxes, wide
    // blocks) the screen images the difference once more in FP32: with both
    // a in [-L/2, L/2], t = a_j - a_i has |t| <= L, so its nearest image is
    // the nearer of t and t -+ L, at min(|t|, L - |t|) per axis. The 4D
    // coordinate w is taken relative to row lane 0's. For a pair
    // within the cutoff the FP32 rounding of a_i, a_j, the image shift and d2
    // stays near 1e-6 relative to the squared cutoff; the screen keeps a
    // margin of 1e-5 (local) or 1e-4 (imaged, whose coordinates reach L / 2)
    // and pass 2 repeats the exact FP64 test. The column values travel around
    // the wave in registers (one lane rotation per round), so the screen reads
    // no LDS.
    // c and e come from the block's bounding box of the last neighborlist
    // build (k_block_bounds, FP32): since then the box is unchanged and no atom
    // has moved more than padding / 2 (else the list would have been rebuilt),
    // so every row atom is within the box's half-extent + padding / 2 of its
    // centre, plus 1e-4 nm for the FP32 rounding of the build.
    const double cw = L.pos[0][3][0]; // row lane 0 holds an atom
    const float4 bc = bb_ctr[S.block0 + r];
    const float4 be = bb_ext[S.block0 + r];
    const double slack = 0.5 * S.padding_d + 1e-4;
    const double reach = S.cutoff_d + 1e-3; // 1e-3 nm covers the FP32 extents
    const bool local = static_cast<double>(be.x) + slack + reach < 0.5 * B.bx &&
                       static_cast<double>(be.y) + slack + reach < 0.5 * B.by &&
                       static_cast<double>(be.z) + slack + reach < 0.5 * B.bz;
    {
        const double cx = bc.x, cy = bc.y, cz = bc.z;
        const double rx = min_image_d(pi.x - cx, B.bx, B.ibx);
        const double ry = min_image_d(pi.y - cy, B.by, B.iby);
        const double rz = min_image_d(pi.z - cz, B.bz, B.ibz);
        const bool vj0 = jl0 < k;
        // this lane's column; padding columns are NaN, whose d2 (NaN) fails
        // every comparison (IEEE semantics: no fast-math flags in this build)
        const float nan = __builtin_nanf("");
        float qx = nan, qy = nan, qz = nan, qw = 0.0f;
        if (vj0) {
            qx = static_cast<float>(min_image_d(pj0.x - cx, B.bx, B.ibx));
            qy = static_cast<float>(min_image_d(pj0.y - cy, B.by, B.iby));
            qz = static_cast<float>(min_image_d(pj0.z - cz, B.bz, B.ibz));
            qw = ALCH ? static_cast<float>(pj0.w - cw) : 0.0f;
        }
        uint32_t qc = jl0;
        const float fx = static_cast<float>(rx), fy = static_cast<float>(ry), fz = static_cast<float>(rz);
        const float fw = ALCH ? static_cast<float>(pi.w - cw) : 0.0f;
        // box edges of the imaged screen in scalar registers (wave-uniform)
        const float lx = uniform_f(static_cast<float>(B.bx)), ly = uniform_f(static_cast<float>(B.by)),
                    lz = uniform_f(static_cast<float>(B.bz));
        // columns of the row's own block need i < j (upper triangle); later blocks always pass
        const bool tri = uniform(ballot(vj0 && jl0 < static_cast<uint32_t>((r + 1) * TILE)) != 0 ? 1 : 0);
        const float cut2f = static_cast<float>(cut2 * (local ? 1.0 + 1e-5 : 1.0 + 1e-4));
        const float cut2v = vi ? cut2f : -1.0f; // padding rows accept nothing
        const bool wrap = !local;
        auto screen = [&](auto tri_tag) {
            constexpr bool TRI = decltype(tri_tag)::value;
            for (int round = 0; round < TILE; round++) {
                // lane l holds column slot (l + round) mod 64
                float dx = qx - fx;
                float dy = qy - fy;
                float dz = qz - fz;
                if (wrap) {
                    // |dx| <= L: the nearer of the two images
                    dx = __builtin_fminf(__builtin_fabsf(dx), lx - __builtin_fabsf(dx));
                    dy = __builtin_fminf(__builtin_fabsf(dy), ly - __builtin_fabsf(dy));
                    dz = __builtin_fminf(__builtin_fabsf(dz), lz - __builtin_fabsf(dz));
                }
                float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                if (ALCH) {
                    const float dw = qw - fw;
                    d2 = __builtin_fmaf(dw, dw, d2);
                }
                bool in = d2 < cut2v;
                if constexpr (TRI) {
                    in = in && i < qc;
                }
                queue(round, in);
                qx = rot1(qx);
                qy = rot1(qy);
                qz = rot1(qz);
                if (ALCH) {
                    qw = rot1(qw);
                }
                if constexpr (TRI) {
                    qc = rot1(qc);
                }
            }
        };
        if (tri) {
            screen(std::true_type{});
        } else {
            screen(std::false_type{});
        }
    }
    if (queued > 0) {
        wave_lds_sync();
        eval_pairs<ALCH, U, DX, DL, DP, DLDP, NACC, INTERP>(L, queued, B, beta, uacc, dlacc);
    }
    wave_lds_sync();

    const bool vj = jl0 < k;
    if (DX || DP || DLDP) {
        if (vi && row_flush) {
            const int gi = S.atom0 + static_cast<int>(perm[S.slot0 + i]);
            if (DX) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const uint64_t v = L.acc[0][c][lane];
                    if (v != 0) {
                        atomic_add_u64(du_dx + gi * 3 + c, v);
                    }
                }
            }
            if (DP || DLDP) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const uint64_t v = L.acc[0][PQ + c][lane];
                    if (v != 0) {
                        atomic_add_u64(du_dp + gi * 3 + c, v);
                    }
                }
            }
        }
        if (vj) {
            const int gj = S.atom0 + static_cast<int>(perm[S.slot0 + jl0]);
            if (DX) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const uint64_t v = L.acc[1][c][lane];
                    if (v != 0) {
                        atomic_add_u64(du_dx + gj * 3 + c, v);
                    }
                }
            }
            if (DP || DLDP) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const uint64_t v = L.acc[1][PQ + c][lane];
                    if (v != 0) {
                        atomic_add_u64(du_dp + gj * 3 + c, v);
                    }
                }
            }
        }
    }
}

// Every wave takes a contiguous range of the tile list of all systems (the
// grid is sized to the machine, the tile count is read on the device).
// WAVES_PER_SIMD = 4 (the default build): at most 128 VGPRs, so four waves
// share a SIMD (the 10 KB of LDS per wave allows four as well); 3 and 2 let
// the compiler use up to 168 and 256 VGPRs with fewer waves to hide latency.
// The engine's instances exist in all three forms and the autotuner picks
// one (NbState::tile_waves).
template <bool U, bool DX, bool DL, bool DP, bool DLDP = false, bool INTERP = false, int WAVES_PER_SIMD = 4>
__global__ __launch_bounds__(NB_BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES_PER_SIMD))) void k_nonbonded_tiles(
    const NbSys *__restrict__ sys,
    const int S,
    const int *__restrict__ tile_prefix,
    const int *__restrict__ tile_order,
    const int *__restrict__ tile_row,
    const uint32_t *__restrict__ tile_atoms,
    const uint32_t *__restrict__ perm,
    const double4 *__restrict__ pos,
    const float4 *__restrict__ par,
    const float4 *__restrict__ dpar,
    const double *__restrict__ boxes,
    const float4 *__restrict__ bb_ctr,
    const float4 *__restrict__ bb_ext,
    const int *__restrict__ status,
    uint64_t *__restrict__ du_dx,
    uint64_t *__restrict__ du_dp,
    uint64_t *__restrict__ du_dl,
    uint64_t *__restrict__ u) {
    constexpr int NACC = (DX ? 3 : 0) + ((DP || DLDP) ? 3 : 0) > 0 ? (DX ? 3 : 0) + ((DP || DLDP) ? 3 : 0) : 1;
    constexpr int WAVES = NB_BLOCK / 64;
    __shared__ TileLds<NACC, INTERP> lds[WAVES];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    TileLds<NACC, INTERP> &L = lds[wave];
    // Each wave takes a contiguous range of the tile list. Tiles of a system
    // are listed grouped by row block (tile_order), so consecutive tiles of a
    // wave mostly share their row atoms: the row accumulators stay in LDS over
    // such a run and go out with one set of global atomics at its end instead
    // of one per tile (device-scope 64-bit atomics are the costly part of a
    // tile's output). The d(du/dlambda)/dp kernel, which skips tiles, writes
    // per tile.
    const int w = uniform(blockIdx.x * WAVES + wave);
    const int nwaves = gridDim.x * WAVES;
    const int total = tile_prefix[S];
    const int chunk = (total + nwaves - 1) / nwaves;
    int t = w * chunk;
    const int t_end = min(total, t + chunk);
    if (t >= t_end) {
        return;
    }
    int s = 0;
    {
        int lo = 0, hi = S - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tile_prefix[mid] <= t) {
                lo = mid;
            } else {
                hi = mid - 1;
            }
        }
        s = lo;
    }
    bool open = false; // the row accumulators hold a run that is not written yet
    // u and du/dl of the system's tiles so far (per lane); written with one
    // wave-reduced atomic per system and wave, not per tile (every tile of a
    // system adds to the same word)
    uint64_t uacc = 0, dlacc = 0;
    for (; t < t_end; t++) {
        while (tile_prefix[s + 1] <= t) {
            s++; // tile_prefix[S] = total > t
        }
        if (status[s] & NB_BAD_BOX) {
            continue;
        }
        const NbSys Sd = sys[s];
        const int local = t - tile_prefix[s];
        const int kt = Sd.tile0 + tile_order[Sd.tile0 + local];
        const int r = tile_row[kt];
        bool last = true; // the next tile of this wave has another row (or system)
        if (!DLDP && t + 1 < t_end && t + 1 < tile_prefix[s + 1]) {
            last = tile_row[Sd.tile0 + tile_order[Sd.tile0 + local + 1]] != r;
        }
        const bool fresh = !open;
        open = !last;
        const bool sys_last = t + 1 >= t_end || t + 1 >= tile_prefix[s + 1];
        const uint32_t jl = tile_atoms[static_cast<size_t>(kt) * TILE + lane];
        const int i = r * TILE + lane;
        const double4 pi = pos[Sd.slot0 + i];
        const float4 ai = par[Sd.slot0 + i];
        const bool vj = jl < static_cast<uint32_t>(Sd.slot_end);
        const double4 pj = vj ? pos[Sd.slot0 + jl] : make_double4(0.0, 0.0, 0.0, 0.0);
        const float4 aj = vj ? par[Sd.slot0 + jl] : make_float4(0.f, 0.f, 0.f, 0.f);
        float4 dpi = make_float4(0.f, 0.f, 0.f, 0.f), dpj = dpi;
        bool moving = pi.w != 0.0 || pj.w != 0.0;
        if (INTERP && Sd.interp) {
            dpi = dpar[Sd.slot0 + i];
            if (vj) {
                dpj = dpar[Sd.slot0 + jl];
            }
            moving = moving || dpi.x != 0.f || dpi.y != 0.f || dpi.z != 0.f || dpj.x != 0.f || dpj.y != 0.f ||
                     dpj.z != 0.f;
        }
        // "vanilla" tiles (no lambda dependence, k_nonbonded.cuh:626-630) skip the 4D terms
        const bool alch = ballot(moving) != 0;
        if (alch) {
            tile_body<true, U, DX, DL, DP, DLDP, NACC, INTERP>(L, Sd, s, r, jl, pi, ai, pj, aj, dpi, dpj, perm, boxes,
                                                               bb_ctr, bb_ext, fresh, last, uacc, dlacc, du_dx, du_dp);
        } else if (!DLDP) {
            tile_body<false, U, DX, DL, DP, false, NACC, INTERP>(L, Sd, s, r, jl, pi, ai, pj, aj, dpi, dpj, perm,
                                                                 boxes, bb_ctr, bb_ext, fresh, last, uacc, dlacc, du_dx, du_dp);
        }
        if (sys_last && (U || DL)) {
            if (U) {
                const uint64_t tot = wave_sum_u64(uacc);
                if (lane == 0 && tot != 0) {
                    atomic_add_u64(u + Sd.red_idx, tot);
                }
            }
            if (DL) {
                const uint64_t tot = wave_sum_u64(dlacc);
                if (lane == 0 && tot != 0) {
                    atomic_add_u64(du_dl + Sd.red_idx, tot);
                }
            }
            uacc = 0;
            dlacc = 0;
        }
    }
}

using TileKernel = void (*)(const NbSys *, int, const int *, const int *, const int *, const uint32_t *, const uint32_t *,
                            const double4 *, const float4 *, const float4 *, const double *, const float4 *,
                            const float4 *, const int *, uint64_t *, uint64_t *, uint64_t *, uint64_t *);

// index bits: 8 = u, 4 = du_dx, 2 = du_dl, 1 = du_dp
const TileKernel tile_kernels[16] = {
    k_nonbonded_tiles<0, 0, 0, 0>, k_nonbonded_tiles<0, 0, 0, 1>, k_nonbonded_tiles<0, 0, 1, 0>,
    k_nonbonded_tiles<0, 0, 1, 1>, k_nonbonded_tiles<0, 1, 0, 0>, k_nonbonded_tiles<0, 1, 0, 1>,
    k_nonbonded_tiles<0, 1, 1, 0>, k_nonbonded_tiles<0, 1, 1, 1>, k_nonbonded_tiles<1, 0, 0, 0>,
    k_nonbonded_tiles<1, 0, 0, 1>, k_nonbonded_tiles<1, 0, 1, 0>, k_nonbonded_tiles<1, 0, 1, 1>,
    k_nonbonded_tiles<1, 1, 0, 0>, k_nonbonded_tiles<1, 1, 0, 1>, k_nonbonded_tiles<1, 1, 1, 0>,
    k_nonbonded_tiles<1, 1, 1, 1>,
};

// interpolated parameters: only the outputs that see dp/dl (du_dl) differ, but
// the whole table is instantiated for uniform dispatch
const TileKernel tile_kernels_interp[16] = {
    k_nonbonded_tiles<0, 0, 0, 0, 0, 1>, k_nonbonded_tiles<0, 0, 0, 1, 0, 1>, k_nonbonded_tiles<0, 0, 1, 0, 0, 1>,
    k_nonbonded_tiles<0, 0, 1, 1, 0, 1>, k_nonbonded_tiles<0, 1, 0, 0, 0, 1>, k_nonbonded_tiles<0, 1, 0, 1, 0, 1>,
    k_nonbonded_tiles<0, 1, 1, 0, 0, 1>, k_nonbonded_tiles<0, 1, 1, 1, 0, 1>, k_nonbonded_tiles<1, 0, 0, 0, 0, 1>,
    k_nonbonded_tiles<1, 0, 0, 1, 0, 1>, k_nonbonded_tiles<1, 0, 1, 0, 0, 1>, k_nonbonded_tiles<1, 0, 1, 1, 0, 1>,
    k_nonbonded_tiles<1, 1, 0, 0, 0, 1>, k_nonbonded_tiles<1, 1, 0, 1, 0, 1>, k_nonbonded_tiles<1, 1, 1, 0, 0, 1>,
    k_nonbonded_tiles<1, 1, 1, 1, 0, 1>,
};

// The pair kernel for output set idx (bits as above) at the given waves per
// SIMD: the instances the MD engine launches (du_dx, du_dx + du_dl, u + du_dx,
// u + du_dx + du_dl, plain parameters) exist at 4, 3 and 2; everything else
// only at 4.
TileKernel tile_kernel(const int idx, const bool interp, const int waves) {
    if (!interp && waves == 3) {
        switch (idx) {
        case 4: return k_nonbonded_tiles<0, 1, 0, 0, false, false, 3>;
        case 6: return k_nonbonded_tiles<0, 1, 1, 0, false, false, 3>;
        case 12: return k_nonbonded_tiles<1, 1, 0, 0, false, false, 3>;
        case 14: return k_nonbonded_tiles<1, 1, 1, 0, false, false, 3>;
        default: break;
        }
    }
    if (!interp && waves == 2) {
        switch (idx) {
        case 4: return k_nonbonded_tiles<0, 1, 0, 0, false, false, 2>;
        case 6: return k_nonbonded_tiles<0, 1, 1, 0, false, false, 2>;
        case 12: return k_nonbonded_tiles<1, 1, 0, 0, false, false, 2>;
        case 14: return k_nonbonded_tiles<1, 1, 1, 0, false, false, 2>;
        default: break;
        }
    }
    return interp ? tile_kernels_interp[idx] : tile_kernels[idx];
}

template <bool U, bool DX, bool DL, bool DP, bool DLDP = false>
__global__ __launch_bounds__(256) void k_pair_list(
    const PairListArgs a, uint64_t *__restrict__ du_dx, uint64_t *__restrict__ du_dp, uint64_t *__restrict__ du_dl,
    uint64_t *__restrict__ u) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool neg = a.negated != 0;
    int red = -1;
    uint64_t fu = 0, fdl = 0; // this lane's u and du/dl terms (reduced over the wave below)
    if (p < a.M) {
        const int s = a.pair_sys ? a.pair_sys[p] : 0;
        const int gi = a.pairs[p * 2 + 0];
        const int gj = a.pairs[p * 2 + 1];
        const double *box = a.boxes + s * 9;
        const double lam = a.lambdas ? a.lambdas[s] : a.lambda0;
        const double cutoff_d = a.cutoffs ? a.cutoffs[s] : a.cutoff0;
        const float beta = a.betas ? a.betas[s] : a.beta0;
        red = a.red_idx ? a.red_idx[s] : 0;

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
        const double d2 = __builtin_fma(dw, dw, pair_d2(dx, dy, dz));
        uint64_t fs[3] = {0, 0, 0}; // force on atom 0 for the slots
        if (d2 < cutoff_d * cutoff_d) {
            // parameters as the tile kernel sees them: FP32 values of p, or of
            // p = (1 - lambda) p0 + lambda p1 when interpolated (k_gather)
            double pv[2][3], dp[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
            const int ga[2] = {gi, gj};
            for (int t = 0; t < 2; t++) {
                for (int c = 0; c < 3; c++) {
                    const double p0 = a.params[ga[t] * 3 + c];
                    if (a.params1) {
                        const double p1 = a.params1[ga[t] * 3 + c];
                        pv[t][c] = static_cast<float>((1.0 - lam) * p0 + lam * p1);
                        dp[t][c] = static_cast<float>(p1 - p0);
                    } else {
                        pv[t][c] = static_cast<float>(p0);
                    }
                }
            }
            const double qi = pv[0][0], qj = pv[1][0];
            const double si = pv[0][1], sj = pv[1][1];
            const double ei = pv[0][2], ej = pv[1][2];
            const bool interp = a.params1 != nullptr;
            const double scale_q = a.scales[p * 2 + 0];
            const double scale_lj = a.scales[p * 2 + 1];
            const PairTerms T = (DL && interp) ? nb_pair_terms<U, true>(d2, qi, qj, si, sj, ei, ej,
                                                                      static_cast<double>(beta), scale_q, scale_lj)
                                               : nb_pair_terms<U, DP>(d2, qi, qj, si, sj, ei, ej,
                                                                      static_cast<double>(beta), scale_q, scale_lj);
            auto acc = [neg](uint64_t *dst, uint64_t v) {
                if (v != 0) {
                    atomic_add_u64(dst, neg ? (0ull - v) : v);
                }
            };
            if (DX) {
                const double ps = T.prefactor * FIXED_SCALE; // as the tile kernel rounds
                const uint64_t fx = round_to_fixed(ps * dx);
                const uint64_t fy = round_to_fixed(ps * dy);
                const uint64_t fz = round_to_fixed(ps * dz);
                if (a.slots) {
                    fs[0] = neg ? 0ull - fx : fx;
                    fs[1] = neg ? 0ull - fy : fy;
                    fs[2] = neg ? 0ull - fz : fz;
                } else {
                    acc(du_dx + gi * 3 + 0, fx);
                    acc(du_dx + gi * 3 + 1, fy);
                    acc(du_dx + gi * 3 + 2, fz);
                    acc(du_dx + gj * 3 + 0, 0ull - fx);
                    acc(du_dx + gj * 3 + 1, 0ull - fy);
                    acc(du_dx + gj * 3 + 2, 0ull - fz);
                }
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
            const bool vanilla = (wi == 0.0) && (wj == 0.0) && !interp;
            if (DL && !vanilla) {
                if (interp) {
                    fdl = f64_to_fixed(pair_du_dl(T, dw, dwdl_i, dwdl_j, true, scale_q, qi, qj, ei, ej, dp[0][0],
                                                  dp[1][0], dp[0][1], dp[1][1], dp[0][2], dp[1][2]),
                                       FIXED_SCALE);
                } else {
                    fdl = f64_to_fixed(dw * T.prefactor * (dwdl_i - dwdl_j), FIXED_SCALE);
                }
            }
            if (DLDP && !vanilla && !interp) {
                const PairDlDp D = nb_pair_dl_dp(d2, si, sj, ei, ej, static_cast<double>(beta), scale_lj);
                const double dd = dw * (dwdl_i - dwdl_j);
                acc(du_dp + gi * 3 + 0, f64_to_fixed(((scale_q * qj) * D.es_unit) * dd, FIXED_SCALE));
                acc(du_dp + gj * 3 + 0, f64_to_fixed(((scale_q * qi) * D.es_unit) * dd, FIXED_SCALE));
                if (D.lj) {
                    const uint64_t gs = f64_to_fixed(D.lj_dsig * dd, FIXED_SCALE_DSIG);
                    acc(du_dp + gi * 3 + 1, gs);
                    acc(du_dp + gj * 3 + 1, gs);
                    acc(du_dp + gi * 3 + 2, f64_to_fixed((D.lj_unit * ej) * dd, FIXED_SCALE_DEPS));
                    acc(du_dp + gj * 3 + 2, f64_to_fixed((D.lj_unit * ei) * dd, FIXED_SCALE_DEPS));
                }
            }
            if (U) {
                fu = f64_to_fixed(T.u, FIXED_SCALE);
            }
            if (neg) {
                fu = 0ull - fu;
                fdl = 0ull - fdl;
            }
        }
        if (DX && a.slots) {
            for (int c = 0; c < 3; c++) {
                a.slots[static_cast<size_t>(p) * 6 + c] = fs[c];
                a.slots[static_cast<size_t>(p) * 6 + 3 + c] = 0ull - fs[c];
            }
        }
    }
    if (U) {
        wave_add_by_key(u, red, fu);
    }
    if (DL) {
        wave_add_by_key(du_dl, red, fdl);
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
    // initialised once, thread-safe (engines may be built from several host threads)
    static const int grid = [] {
        int dev = 0;
        TM_HIP_CHECK(hipGetDevice(&dev));
        int cus = 0;
        TM_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        // 128 one-wave workgroups per CU: up to 16 are resident per CU (128 VGPRs,
        // 10 KB LDS per wave for the force kernels) and the rest queue behind
        // them, which balances the uneven per-tile pair counts better than a
        // fully persistent grid (measured on MI355X: -15%)
        return cus * 128;
    }();
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
    const TileKernel kern = tile_kernel(idx, any_interp, tile_waves);
    kern<<<grid, NB_BLOCK, 0, stream>>>(sys.ptr, S, tile_prefix.ptr, tile_order.ptr, tile_row.ptr, tile_atoms.ptr, perm.ptr, pos.ptr,
                                        par.ptr, any_interp ? dpar.ptr : nullptr, boxes, bb_ctr.ptr, bb_ext.ptr,
                                        status.ptr, du_dx, du_dp, du_dl, u);
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

void NbState::compute_dl_dp(const double *boxes, uint64_t *dl_dp, hipStream_t stream, int grid_override) const {
    const int grid = grid_override > 0 ? grid_override : default_tile_grid();
    if (any_interp) {
        throw Error("d(du/dlambda)/dp is not available for interpolated parameters");
    }
    k_nonbonded_tiles<false, false, false, false, true><<<grid, NB_BLOCK, 0, stream>>>(
        sys.ptr, S, tile_prefix.ptr, tile_order.ptr, tile_row.ptr, tile_atoms.ptr, perm.ptr, pos.ptr, par.ptr, nullptr, boxes,
        bb_ctr.ptr, bb_ext.ptr, status.ptr, nullptr, dl_dp, nullptr, nullptr);
    TM_LAUNCH_CHECK();
}

void launch_pair_list_dl_dp(const PairListArgs &a, uint64_t *dl_dp, hipStream_t stream) {
    if (a.M == 0) {
        return;
    }
    if (a.params1) {
        throw Error("d(du/dlambda)/dp is not available for interpolated parameters");
    }
    k_pair_list<false, false, false, false, true><<<ceil_div(a.M, 256), 256, 0, stream>>>(a, nullptr, dl_dp, nullptr,
                                                                                          nullptr);
    TM_LAUNCH_CHECK();
}

} // namespace tmamd
