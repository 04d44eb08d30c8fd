// Rebuild check, spatial sort, sorted FP32 gather, block bounds and the
// 64-atom tile neighborlist for S systems at once.
//
// Behaviour follows the reference (timemachine/cpp/src/nonbonded_all_pairs.cu:222-433,
// timemachine/cpp/src/neighborlist.cu:126-185, kernels/k_neighborlist.cuh:9-427):
//  * rebuild when the box changed or any atom moved more than padding/2 since
//    the last build;
//  * atoms are permuted along a Hilbert curve before building (any order gives
//    the same results, only the tile occupancy changes);
//  * row block r of the upper-triangular list holds every column atom j >= 64 r
//    (sorted order) within nb_cutoff of at least one atom of the block.
// The design is wave64-native: one wave per 64-atom row block (or per piece
// of its column range when few systems rebuild), column atoms tested
// lane-parallel, interacting atoms compacted with ballot/mbcnt into LDS and
// emitted as 64-atom tiles; each rebuilt system's tiles are then listed by
// row block for the tile kernel.
#include <algorithm>

#include "nb_pair.h"
#include "nb_state.h"
#include "wave.h"

namespace tmamd {

namespace {

constexpr int SORT_THREADS = 1024;
constexpr int HILBERT_BITS = 4;
constexpr int HILBERT_GRID = 1 << HILBERT_BITS;
constexpr int HILBERT_CELLS = HILBERT_GRID * HILBERT_GRID * HILBERT_GRID;

// 3D Hilbert index of a cell, Skilling's transpose algorithm
// (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004).
__device__ uint32_t hilbert_index(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t X[3] = {x, y, z};
    const uint32_t M = 1u << (HILBERT_BITS - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; i++) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        if (X[2] & Q) {
            t ^= Q - 1;
        }
    }
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    uint32_t key = 0;
    for (int b = HILBERT_BITS - 1; b >= 0; b--) {
        for (int i = 0; i < 3; i++) {
            key = (key << 1) | ((X[i] >> b) & 1u);
        }
    }
    return key;
}

// FP32 coordinates of the neighborlist predicate (the oracle rounds the same way).
__device__ __forceinline__ float4 to_f32(const double4 &p) {
    return make_float4(static_cast<float>(p.x), static_cast<float>(p.y), static_cast<float>(p.z),
                       static_cast<float>(p.w));
}

__device__ __forceinline__ uint32_t cell_of(double v, double box) {
    double f = v / box;
    f -= floor(f);
    int c = static_cast<int>(f * HILBERT_GRID);
    c = c < 0 ? 0 : (c >= HILBERT_GRID ? HILBERT_GRID - 1 : c);
    return static_cast<uint32_t>(c);
}

__global__ void k_fill_u64(uint64_t *p, size_t n, uint64_t v) {
    size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        p[i] = v;
    }
}

__global__ void k_fill_i32(int *p, size_t n, int v) {
    size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        p[i] = v;
    }
}

// grid (ceil(max_atoms / 256), S). Sets rebuild[s] when the box changed or an
// atom of the nonbonded set moved more than padding / 2 since the last build;
// validates the box as the reference does on rebuild.
__global__ void k_check_rebuild(
    const NbSys *__restrict__ sys,
    const int *__restrict__ subset,
    const double *__restrict__ x,
    const double *__restrict__ nblist_x,
    const double *__restrict__ boxes,
    const double *__restrict__ nblist_box,
    int *__restrict__ rebuild,
    int *__restrict__ status) {
    const int s = blockIdx.y;
    const NbSys S = sys[s];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const double *box = boxes + s * 9;
    if (a < 9) {
        if (box[a] != nblist_box[s * 9 + a]) {
            rebuild[s] = 1;
        }
        if (a == 0) {
            const double db_cutoff = 2.0 * (S.cutoff_d + S.padding_d);
            int st = 0;
            for (int d = 0; d < 9; d++) {
                if (d == 0 || d == 4 || d == 8) {
                    if (box[d] < db_cutoff) {
                        st |= NB_STATUS_BOX_TOO_SMALL;
                    }
                } else if (box[d] != 0.0) {
                    st |= NB_STATUS_BOX_NOT_ORTHO;
                }
            }
            if (st) {
                atomicOr(status + s, st);
            }
        }
    }
    if (a >= S.k) {
        return;
    }
    const int atom = S.subset0 >= 0 ? subset[S.subset0 + a] : a;
    const int g = S.atom0 + atom;
    const double dx = x[g * 3 + 0] - nblist_x[g * 3 + 0];
    const double dy = x[g * 3 + 1] - nblist_x[g * 3 + 1];
    const double dz = x[g * 3 + 2] - nblist_x[g * 3 + 2];
    const double half = 0.5 * S.padding_d;
    if (dx * dx + dy * dy + dz * dz > half * half) {
        rebuild[s] = 1; // every writer stores the same value
    }
}

// Exclusive scan of HILBERT_CELLS counts held in LDS, SORT_THREADS threads.
__device__ void block_exclusive_scan_cells(int *cells, int *wave_tot) {
    constexpr int PER = HILBERT_CELLS / SORT_THREADS;
    const int t = threadIdx.x;
    int local[PER];
    int sum = 0;
    for (int i = 0; i < PER; i++) {
        local[i] = cells[t * PER + i];
        sum += local[i];
    }
    // inclusive scan of per-thread sums within the wave
    int inc = sum;
    const int lane = t & 63;
    for (int off = 1; off < 64; off <<= 1) {
        int v = __shfl_up(inc, off, 64);
        if (lane >= off) {
            inc += v;
        }
    }
    const int w = t >> 6;
    if (lane == 63) {
        wave_tot[w] = inc;
    }
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int i = 0; i < SORT_THREADS / 64; i++) {
            int v = wave_tot[i];
            wave_tot[i] = acc;
            acc += v;
        }
    }
    __syncthreads();
    int run = wave_tot[w] + inc - sum;
    for (int i = 0; i < PER; i++) {
        cells[t * PER + i] = run;
        run += local[i];
    }
    __syncthreads();
}

// Ascending bitonic sorting network over SORT_NET values in registers (every
// index is a compile-time constant once the loops unroll).
constexpr int SORT_NET = 16;
__device__ __forceinline__ void bitonic_sort(uint32_t (&v)[SORT_NET]) {
#pragma unroll
    for (int k = 2; k <= SORT_NET; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < SORT_NET; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t lo = min(v[i], v[l]);
                    const uint32_t hi = max(v[i], v[l]);
                    const bool up = (i & k) == 0;
                    v[i] = up ? lo : hi;
                    v[l] = up ? hi : lo;
                }
            }
        }
    }
}

// One workgroup per system; only systems flagged for rebuild do work.
// Counting sort of the nonbonded atoms by Hilbert cell, then atom index order
// inside each cell so the permutation is deterministic: a sorting network in
// registers for cells of up to 16 atoms (a 34k-atom system has ~8 per cell;
// the insertion sort in global memory it replaces took ~170 us there), an
// insertion sort for larger cells.
// Also records the coordinates/box the neighborlist is built from and
// resets the tile counter.
__global__ __launch_bounds__(SORT_THREADS) void k_sort(
    const NbSys *__restrict__ sys,
    const int *__restrict__ subset,
    const double *__restrict__ x,
    const double *__restrict__ boxes,
    const int *__restrict__ rebuild,
    uint32_t *__restrict__ keys,
    uint32_t *__restrict__ perm,
    double *__restrict__ nblist_x,
    double *__restrict__ nblist_box,
    int *__restrict__ tile_count) {
    const int s = blockIdx.x;
    if (!rebuild[s]) {
        return;
    }
    const NbSys S = sys[s];
    const int t = threadIdx.x;
    __shared__ int count[HILBERT_CELLS];
    __shared__ int start[HILBERT_CELLS];
    __shared__ int wave_tot[SORT_THREADS / 64];

    const double bx = boxes[s * 9 + 0];
    const double by = boxes[s * 9 + 4];
    const double bz = boxes[s * 9 + 8];

    // one segment, or rows then columns with a row subset; each is sorted on its own
    const int nseg = S.nrow > 0 ? 2 : 1;
    for (int sg = 0; sg < nseg; sg++) {
        const int a_lo = sg == 1 ? S.nrow : 0;
        const int a_hi = (nseg == 2 && sg == 0) ? S.nrow : S.k;
        uint32_t *seg_perm = perm + S.slot0 + (sg == 1 ? S.row_blocks * TILE : 0);
        if (!S.sorted) {
            for (int a = a_lo + t; a < a_hi; a += SORT_THREADS) {
                seg_perm[a - a_lo] = static_cast<uint32_t>(S.subset0 >= 0 ? subset[S.subset0 + a] : a);
            }
            continue;
        }
        for (int c = t; c < HILBERT_CELLS; c += SORT_THREADS) {
            count[c] = 0;
        }
        __syncthreads();
        for (int a = a_lo + t; a < a_hi; a += SORT_THREADS) {
            const int atom = S.subset0 >= 0 ? subset[S.subset0 + a] : a;
            const int g = S.atom0 + atom;
            const uint32_t key =
                hilbert_index(cell_of(x[g * 3 + 0], bx), cell_of(x[g * 3 + 1], by), cell_of(x[g * 3 + 2], bz));
            keys[S.slot0 + a] = key;
            atomicAdd(&count[key], 1);
        }
        __syncthreads();
        for (int c = t; c < HILBERT_CELLS; c += SORT_THREADS) {
            start[c] = count[c];
        }
        __syncthreads();
        block_exclusive_scan_cells(start, wave_tot);
        // count[] becomes the running fill position
        for (int c = t; c < HILBERT_CELLS; c += SORT_THREADS) {
            count[c] = start[c];
        }
        __syncthreads();
        for (int a = a_lo + t; a < a_hi; a += SORT_THREADS) {
            const int atom = S.subset0 >= 0 ? subset[S.subset0 + a] : a;
            const uint32_t key = keys[S.slot0 + a];
            const int p = atomicAdd(&count[key], 1);
            seg_perm[p] = static_cast<uint32_t>(atom);
        }
        __syncthreads();
        for (int c = t; c < HILBERT_CELLS; c += SORT_THREADS) {
            const int b = start[c];
            const int e = count[c];
            if (e - b <= 1) {
                continue;
            }
            if (e - b <= SORT_NET) {
                // in registers: a bitonic network over SORT_NET entries, the
                // unused ones past the cell's end (NB_INVALID sorts last)
                uint32_t v[SORT_NET];
#pragma unroll
                for (int i = 0; i < SORT_NET; i++) {
                    v[i] = b + i < e ? seg_perm[b + i] : NB_INVALID;
                }
                bitonic_sort(v);
#pragma unroll
                for (int i = 0; i < SORT_NET; i++) {
                    if (b + i < e) {
                        seg_perm[b + i] = v[i];
                    }
                }
                continue;
            }
            for (int i = b + 1; i < e; i++) {
                const uint32_t v = seg_perm[i];
                int j = i - 1;
                while (j >= b && seg_perm[j] > v) {
                    seg_perm[j + 1] = seg_perm[j];
                    j--;
                }
                seg_perm[j + 1] = v;
            }
        }
        __syncthreads(); // count / start are reused by the next segment
    }
    for (int a = t; a < S.nblocks * TILE; a += SORT_THREADS) {
        if (a >= nb_block_end(S, a / TILE)) {
            perm[S.slot0 + a] = NB_INVALID;
        }
    }
    for (int i = t; i < S.n_atoms * 3; i += SORT_THREADS) {
        nblist_x[S.atom0 * 3 + i] = x[S.atom0 * 3 + i];
    }
    if (t < 9) {
        nblist_box[s * 9 + t] = boxes[s * 9 + t];
    }
    if (t < NB_TILE_LANES) {
        tile_count[s * NB_TILE_LANES + t] = 0;
    }
}

// grid (ceil(max_slots / 256), S): sorted FP32 coordinates and parameters.
// w = (plane + offset * lambda) * cutoff and dw/dlambda = offset * cutoff, in
// f64 before rounding to FP32 (kernels/k_lambda_transformer_jit.cuh:25-55 with
// the identity transform).
__global__ void k_gather(
    const NbSys *__restrict__ sys,
    const double *__restrict__ x,
    const double *__restrict__ params,
    const double *__restrict__ params1,
    const int *__restrict__ plane,
    const int *__restrict__ offset,
    const double *__restrict__ lambdas,
    const double lambda0,
    const uint32_t *__restrict__ perm,
    double4 *__restrict__ pos,
    float4 *__restrict__ par,
    float4 *__restrict__ dpar) {
    const int s = blockIdx.y;
    const NbSys S = sys[s];
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= S.nblocks * TILE) {
        return;
    }
    const uint32_t atom = perm[S.slot0 + slot];
    if (atom == NB_INVALID) {
        pos[S.slot0 + slot] = make_double4(0.0, 0.0, 0.0, 0.0);
        par[S.slot0 + slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (dpar) {
            dpar[S.slot0 + slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        return;
    }
    const int g = S.atom0 + static_cast<int>(atom);
    const double lam = lambdas ? lambdas[s] : lambda0;
    const int pl = plane ? plane[g] : 0;
    const int of = offset ? offset[g] : 0;
    const double w = (pl + of * lam) * S.cutoff_d;
    const double dw_dl = of * S.cutoff_d;
    pos[S.slot0 + slot] = make_double4(x[g * 3 + 0], x[g * 3 + 1], x[g * 3 + 2], w);
    if (params) {
        if (S.interp && params1) {
            // parameter interpolation with the identity transform
            // (k_lambda_transformer_jit.cuh:57-98): p = (1 - lambda) p0 + lambda p1, dp/dl = p1 - p0
            double p[3], d[3];
            for (int c = 0; c < 3; c++) {
                const double a = params[g * 3 + c], b = params1[g * 3 + c];
                p[c] = (1.0 - lam) * a + lam * b;
                d[c] = b - a;
            }
            par[S.slot0 + slot] = make_float4(static_cast<float>(p[0]), static_cast<float>(p[1]),
                                              static_cast<float>(p[2]), static_cast<float>(dw_dl));
            dpar[S.slot0 + slot] =
                make_float4(static_cast<float>(d[0]), static_cast<float>(d[1]), static_cast<float>(d[2]), 0.f);
        } else {
            par[S.slot0 + slot] =
                make_float4(static_cast<float>(params[g * 3 + 0]), static_cast<float>(params[g * 3 + 1]),
                            static_cast<float>(params[g * 3 + 2]), static_cast<float>(dw_dl));
            if (dpar) {
                dpar[S.slot0 + slot] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
    }
}

// du/dp split of interpolated parameters (see nb_state.h), grid (ceil(N / 256), 3).
__global__ void k_split_interpolated_du_dp(int N, double lambda, const uint64_t *__restrict__ src,
                                           uint64_t *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= N) {
        return;
    }
    const double scale = c == 0 ? FIXED_SCALE : (c == 1 ? FIXED_SCALE_DSIG : FIXED_SCALE_DEPS);
    const double g = fixed_to_double(src[i * 3 + c], scale);
    dst[i * 3 + c] += f64_to_fixed((1.0 - lambda) * g, scale);
    dst[(N + i) * 3 + c] += f64_to_fixed(lambda * g, scale);
}

// grid (ceil(max_blocks / 4), S), 256 threads: one wave per 64-atom block.
// Bounding box of the block with every atom imaged next to the block's first
// atom (a valid periodic box for any block narrower than half the box).
__global__ __launch_bounds__(256) void k_block_bounds(
    const NbSys *__restrict__ sys,
    const double *__restrict__ boxes,
    const int *__restrict__ rebuild,
    const double4 *__restrict__ pos,
    float4 *__restrict__ bb_ctr,
    float4 *__restrict__ bb_ext) {
    const int s = blockIdx.y;
    if (!rebuild[s]) {
        return;
    }
    const NbSys S = sys[s];
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= S.nblocks) {
        return;
    }
    const int lane = threadIdx.x & 63;
    const float bx = static_cast<float>(boxes[s * 9 + 0]);
    const float by = static_cast<float>(boxes[s * 9 + 4]);
    const float bz = static_cast<float>(boxes[s * 9 + 8]);
    const float ibx = 1.0f / bx;
    const float iby = 1.0f / by;
    const float ibz = 1.0f / bz;
    const float4 ref = to_f32(pos[S.slot0 + b * TILE]);
    const int slot = b * TILE + lane;
    float px = ref.x, py = ref.y, pz = ref.z;
    if (slot < nb_block_end(S, b)) {
        const float4 p = to_f32(pos[S.slot0 + slot]);
        px = ref.x + min_image_f(p.x - ref.x, bx, ibx);
        py = ref.y + min_image_f(p.y - ref.y, by, iby);
        pz = ref.z + min_image_f(p.z - ref.z, bz, ibz);
    }
    const float lox = wave_min(px), hix = wave_max(px);
    const float loy = wave_min(py), hiy = wave_max(py);
    const float loz = wave_min(pz), hiz = wave_max(pz);
    if (lane == 0) {
        bb_ctr[S.block0 + b] = make_float4(0.5f * (lox + hix), 0.5f * (loy + hiy), 0.5f * (loz + hiz), 0.f);
        bb_ext[S.block0 + b] = make_float4(0.5f * (hix - lox), 0.5f * (hiy - loy), 0.5f * (hiz - loz), 0.f);
    }
}

__device__ __forceinline__ float box_gap(float d, float ext, float b, float ib) {
    d = min_image_f(d, b, ib);
    return fmaxf(0.0f, fabsf(d) - ext);
}

// grid (S, ceil(max_blocks / 4) x pieces), 256 threads: one wave per row
// block (or piece of one). The system index runs fastest so that the first
// row blocks of every system, which scan the most column blocks (upper
// triangle), are dispatched first and the launch does not end on a tail of
// long waves.
// Column blocks c >= r are screened by bounding boxes (with a small margin
// so the screen never rejects what the exact test would accept); candidate
// column atoms are then tested exactly against the row atoms with the FP32
// min-image predicate d^2 < nb_cutoff^2 (oracle/nblist_ref.c restates it).
__global__ __launch_bounds__(256) void k_find_tiles(
    const NbSys *__restrict__ sys,
    const double *__restrict__ boxes,
    const int *__restrict__ rebuild,
    const double4 *__restrict__ pos,
    const float4 *__restrict__ bb_ctr,
    const float4 *__restrict__ bb_ext,
    int *__restrict__ tile_count,
    int *__restrict__ tile_row,
    uint32_t *__restrict__ tile_atoms,
    int *__restrict__ status,
    const int pieces,
    uint32_t *__restrict__ piece_buf,
    int *__restrict__ piece_cnt) {
    const int s = blockIdx.x;
    if (!rebuild[s]) {
        return;
    }
    const NbSys S = sys[s];
    // the 4 waves of a workgroup take the same piece of 4 consecutive row
    // blocks: a row block's first piece holds most of its candidates (those
    // near it in the Hilbert order), and a workgroup's slot is free only when
    // all its waves are done, so its waves should carry similar work
    const int wave = threadIdx.x >> 6;
    const int group = blockIdx.y / pieces;
    const int piece = blockIdx.y - group * pieces;
    const int r = group * 4 + wave;
    if (r >= S.row_blocks) {
        return;
    }
    const int tile_lane = r & (NB_TILE_LANES - 1); // this row block's tile counter
    const int lane = threadIdx.x & 63;
    constexpr int SUB = 8;
    constexpr int SUBS = TILE / SUB;
    __shared__ float4 row_lds[4][TILE];
    __shared__ uint32_t buf_lds[4][2 * TILE];
    __shared__ float4 sub_ctr[4][SUBS];
    __shared__ float4 sub_ext[4][SUBS];
    float4 *rows = row_lds[wave];
    uint32_t *buf = buf_lds[wave];
    float4 *sctr = sub_ctr[wave];
    float4 *sext = sub_ext[wave];

    const float bx = static_cast<float>(boxes[s * 9 + 0]);
    const float by = static_cast<float>(boxes[s * 9 + 4]);
    const float bz = static_cast<float>(boxes[s * 9 + 8]);
    const float ibx = 1.0f / bx;
    const float iby = 1.0f / by;
    const float ibz = 1.0f / bz;
    const float cut = S.nb_cutoff;
    const float cut2 = cut * cut;
    const float scr = cut + 1e-3f; // screening radius, strictly larger
    const float scr2 = scr * scr;

    const int i = r * TILE + lane;
    const bool row_valid = i < S.row_end;
    const float4 pi = row_valid ? to_f32(pos[S.slot0 + i]) : make_float4(0.f, 0.f, 0.f, 0.f);
    rows[lane] = pi;
    wave_lds_sync();
    const float4 rc = bb_ctr[S.block0 + r];
    const float4 re = bb_ext[S.block0 + r];

    // bounding boxes of the 8 row sub-blocks of 8 atoms (lane q < 8 builds
    // sub-block q; extents are relative to its first atom, min-image)
    if (lane < SUBS) {
        const int a0 = r * TILE + lane * SUB;
        float4 c = make_float4(0.f, 0.f, 0.f, 0.f), e = make_float4(-1.f, -1.f, -1.f, 0.f); // empty
        if (a0 < S.row_end) {
            const float4 ref = rows[lane * SUB];
            float lox = 0.f, loy = 0.f, loz = 0.f, hix = 0.f, hiy = 0.f, hiz = 0.f;
            for (int a = 1; a < SUB && a0 + a < S.row_end; a++) {
                const float4 q = rows[lane * SUB + a];
                const float dx = min_image_f(q.x - ref.x, bx, ibx);
                const float dy = min_image_f(q.y - ref.y, by, iby);
                const float dz = min_image_f(q.z - ref.z, bz, ibz);
                lox = fminf(lox, dx);
                hix = fmaxf(hix, dx);
                loy = fminf(loy, dy);
                hiy = fmaxf(hiy, dy);
                loz = fminf(loz, dz);
                hiz = fmaxf(hiz, dz);
            }
            c = make_float4(ref.x + 0.5f * (lox + hix), ref.y + 0.5f * (loy + hiy), ref.z + 0.5f * (loz + hiz), 0.f);
            e = make_float4(0.5f * (hix - lox), 0.5f * (hiy - loy), 0.5f * (hiz - loz), 0.f);
        }
        sctr[lane] = c;
        sext[lane] = e;
    }
    wave_lds_sync();

    int nbuf = 0;
    // the upper triangle c >= r, or every column block after the rows; with
    // several pieces per row block, this wave's contiguous share of them
    const int c_first = S.nrow > 0 ? S.row_blocks : r;
    const int n_cols = S.nblocks - c_first;
    const int c_begin = c_first + static_cast<int>(static_cast<long long>(n_cols) * piece / pieces);
    const int c_end = c_first + static_cast<int>(static_cast<long long>(n_cols) * (piece + 1) / pieces);
    for (int cbase = c_begin; cbase < c_end; cbase += TILE) {
        const int c = cbase + lane;
        bool cand = false;
        if (c < c_end) {
            const float4 cc = bb_ctr[S.block0 + c];
            const float4 ce = bb_ext[S.block0 + c];
            const float gx = box_gap(rc.x - cc.x, re.x + ce.x, bx, ibx);
            const float gy = box_gap(rc.y - cc.y, re.y + ce.y, by, iby);
            const float gz = box_gap(rc.z - cc.z, re.z + ce.z, bz, ibz);
            cand = gx * gx + gy * gy + gz * gz < scr2;
        }
        uint64_t cmask = ballot(cand);
        // the next candidate's bounding box and column atoms are loaded while the
        // current one is tested (a row block visits tens of candidates in order)
        auto load = [&](int cbx, float4 &cc, float4 &ce, float4 &pj) {
            cc = bb_ctr[S.block0 + cbx];
            ce = bb_ext[S.block0 + cbx];
            const int jx = cbx * TILE + lane;
            pj = jx < S.slot_end ? to_f32(pos[S.slot0 + jx]) : make_float4(0.f, 0.f, 0.f, 0.f);
        };
        int cb_next = -1;
        float4 cc_next, ce_next, pj_next;
        if (cmask) {
            cb_next = cbase + __builtin_ctzll(cmask);
            cmask &= cmask - 1;
            load(cb_next, cc_next, ce_next, pj_next);
        }
        while (cb_next >= 0) {
            const int cb = cb_next;
            const float4 cc = cc_next;
            const float4 ce = ce_next;
            const float4 pj_cur = pj_next;
            if (cmask) {
                cb_next = cbase + __builtin_ctzll(cmask);
                cmask &= cmask - 1;
                load(cb_next, cc_next, ce_next, pj_next);
            } else {
                cb_next = -1;
            }
            // row atoms near the column block
            bool row_near = false;
            if (row_valid) {
                const float gx = box_gap(pi.x - cc.x, ce.x, bx, ibx);
                const float gy = box_gap(pi.y - cc.y, ce.y, by, iby);
                const float gz = box_gap(pi.z - cc.z, ce.z, bz, ibz);
                row_near = gx * gx + gy * gy + gz * gz < scr2;
            }
            const uint64_t rmask = ballot(row_near);
            const int j = cb * TILE + lane;
            const bool col_valid = j < S.slot_end;
            bool pending = false;
            const float4 pj = pj_cur;
            if (col_valid && rmask) {
                const float gx = box_gap(pj.x - rc.x, re.x, bx, ibx);
                const float gy = box_gap(pj.y - rc.y, re.y, by, iby);
                const float gz = box_gap(pj.z - rc.z, re.z, bz, ibz);
                pending = gx * gx + gy * gy + gz * gz < scr2;
            }
            // Exact test, per lane: a column atom visits only the row sub-blocks
            // within the screening radius, nearest first, and stops at its first
            // row atom inside the cutoff. Every row atom that can be within the
            // cutoff lies in such a sub-block, so the accepted set is the one of
            // testing all row atoms.
            bool found = false;
            if (pending) {
                uint32_t near = 0;
                int first = 0;
                float best = 3.0e38f;
                for (int q = 0; q < SUBS; q++) {
                    const float4 se = sext[q];
                    if (se.x < 0.f) {
                        continue; // sub-block without atoms
                    }
                    const float4 sc = sctr[q];
                    const float gx = box_gap(pj.x - sc.x, se.x, bx, ibx);
                    const float gy = box_gap(pj.y - sc.y, se.y, by, iby);
                    const float gz = box_gap(pj.z - sc.z, se.z, bz, ibz);
                    const float g2 = gx * gx + gy * gy + gz * gz;
                    if (g2 < scr2) {
                        near |= 1u << q;
                        if (g2 < best) {
                            best = g2;
                            first = q;
                        }
                    }
                }
                int q = first;
                while (near != 0 && !found) {
                    near &= ~(1u << q);
                    const int a_end = min(SUB, S.row_end - (r * TILE + q * SUB));
                    for (int a = 0; a < a_end && !found; a++) {
                        const float4 rq = rows[q * SUB + a];
                        const float dx = min_image_f(rq.x - pj.x, bx, ibx);
                        const float dy = min_image_f(rq.y - pj.y, by, iby);
                        const float dz = min_image_f(rq.z - pj.z, bz, ibz);
                        found = dx * dx + dy * dy + dz * dz < cut2;
                    }
                    if (near != 0) {
                        q = __builtin_ctz(near);
                    }
                }
            }
            const uint64_t fmask = ballot(found);
            if (found) {
                buf[nbuf + lanes_below(fmask)] = static_cast<uint32_t>(j);
            }
            nbuf += __popcll(fmask);
            wave_lds_sync();
            if (nbuf >= TILE) {
                int tile = 0;
                if (lane == 0) {
                    tile = tile_lane + NB_TILE_LANES * atomicAdd(tile_count + s * NB_TILE_LANES + tile_lane, 1);
                }
                tile = __shfl(tile, 0, 64);
                if (tile < S.tile_cap) {
                    tile_atoms[static_cast<size_t>(S.tile0 + tile) * TILE + lane] = buf[lane];
                    if (lane == 0) {
                        tile_row[S.tile0 + tile] = r;
                    }
                } else if (lane == 0) {
                    atomicOr(status + s, NB_STATUS_TILE_OVERFLOW);
                }
                const uint32_t rest = buf[TILE + lane];
                wave_lds_sync();
                buf[lane] = rest;
                nbuf -= TILE;
                wave_lds_sync();
            }
        }
    }
    if (pieces > 1) {
        // this piece's last (< 64) column atoms: k_merge_pieces packs the
        // leftovers of a row block's pieces into full tiles and one partial
        const size_t slot = static_cast<size_t>(S.block0 + r) * NB_MAX_PIECES + piece;
        if (lane < nbuf) {
            piece_buf[slot * TILE + lane] = buf[lane];
        }
        if (lane == 0) {
            piece_cnt[slot] = nbuf;
        }
    } else if (nbuf > 0) {
        const uint32_t v = lane < nbuf ? buf[lane] : NB_INVALID;
        int tile = 0;
        if (lane == 0) {
            tile = tile_lane + NB_TILE_LANES * atomicAdd(tile_count + s * NB_TILE_LANES + tile_lane, 1);
        }
        tile = __shfl(tile, 0, 64);
        if (tile < S.tile_cap) {
            tile_atoms[static_cast<size_t>(S.tile0 + tile) * TILE + lane] = v;
            if (lane == 0) {
                tile_row[S.tile0 + tile] = r;
            }
        } else if (lane == 0) {
            atomicOr(status + s, NB_STATUS_TILE_OVERFLOW);
        }
    }
}

// The leftovers of a row block's pieces (each < 64 column atoms) as full
// tiles and at most one partial tile, so splitting a row block's scan over
// several waves adds no tiles. grid (S, ceil(max_blocks / 4)), one wave per
// row block of a rebuilt system.
__global__ __launch_bounds__(256) void k_merge_pieces(
    const NbSys *__restrict__ sys,
    const int *__restrict__ rebuild,
    const int pieces,
    const uint32_t *__restrict__ piece_buf,
    const int *__restrict__ piece_cnt,
    int *__restrict__ tile_count,
    int *__restrict__ tile_row,
    uint32_t *__restrict__ tile_atoms,
    int *__restrict__ status) {
    const int s = blockIdx.x;
    if (!rebuild[s]) {
        return;
    }
    const NbSys S = sys[s];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.y * 4 + wave;
    if (r >= S.row_blocks) {
        return;
    }
    __shared__ uint32_t merged[4][NB_MAX_PIECES * TILE];
    uint32_t *b = merged[wave];
    int n = 0;
    for (int p = 0; p < pieces; p++) {
        const size_t slot = static_cast<size_t>(S.block0 + r) * NB_MAX_PIECES + p;
        const int c = piece_cnt[slot];
        if (lane < c) {
            b[n + lane] = piece_buf[slot * TILE + lane];
        }
        n += c;
    }
    wave_lds_sync();
    for (int base = 0; base < n; base += TILE) {
        const uint32_t v = base + lane < n ? b[base + lane] : NB_INVALID;
        int tile = 0;
        if (lane == 0) {
            const int j = r & (NB_TILE_LANES - 1);
            tile = j + NB_TILE_LANES * atomicAdd(tile_count + s * NB_TILE_LANES + j, 1);
        }
        tile = __shfl(tile, 0, 64);
        if (tile < S.tile_cap) {
            tile_atoms[static_cast<size_t>(S.tile0 + tile) * TILE + lane] = v;
            if (lane == 0) {
                tile_row[S.tile0 + tile] = r;
            }
        } else if (lane == 0) {
            atomicOr(status + s, NB_STATUS_TILE_OVERFLOW);
        }
    }
}

// Tiles of each rebuilt system listed by row block: tile_order[tile0 + p] is
// the local index of the p-th tile in row order (tiles of one row in any
// order; the fixed-point sums do not depend on it). The tile kernel walks the
// list in this order so that one wave meets the tiles of a row block one after
// another. The valid local tiles are those each counter handed out
// (j + NB_TILE_LANES c, c below the counter's count); systems with more row
// blocks than the LDS histogram holds list them counter by counter.
constexpr int ORDER_ROWS = 4096;

__global__ __launch_bounds__(256) void k_tile_order(
    const NbSys *__restrict__ sys, const int *__restrict__ rebuild, const int *__restrict__ tile_count,
    const int *__restrict__ tile_row, int *__restrict__ tile_order) {
    const int s = blockIdx.x;
    if (!rebuild[s]) {
        return;
    }
    const NbSys S = sys[s];
    const int rb = S.row_blocks;
    const int t = threadIdx.x;
    int *order = tile_order + S.tile0;
    const int *rows = tile_row + S.tile0;
    __shared__ int lane_n[NB_TILE_LANES], lane_start[NB_TILE_LANES + 1];
    if (t == 0) {
        int acc = 0;
        for (int j = 0; j < NB_TILE_LANES; j++) {
            lane_n[j] = nb_lane_tiles(tile_count[s * NB_TILE_LANES + j], j, S.tile_cap);
            lane_start[j] = acc;
            acc += lane_n[j];
        }
        lane_start[NB_TILE_LANES] = acc;
    }
    __syncthreads();
    const int n = lane_start[NB_TILE_LANES];
    // the q-th valid tile (counter by counter) as a local tile index, q < n
    auto local_tile = [&](int q) {
        int j = 0;
        while (q >= lane_start[j + 1]) {
            j++;
        }
        return j + NB_TILE_LANES * (q - lane_start[j]);
    };
    if (rb > ORDER_ROWS) {
        for (int q = t; q < n; q += 256) {
            order[q] = local_tile(q);
        }
        return;
    }
    __shared__ int cnt[ORDER_ROWS];
    __shared__ int part[256];
    for (int r = t; r < rb; r += 256) {
        cnt[r] = 0;
    }
    __syncthreads();
    for (int q = t; q < n; q += 256) {
        atomicAdd(&cnt[rows[local_tile(q)]], 1);
    }
    __syncthreads();
    // exclusive scan of cnt[0, rb): thread t owns a contiguous range
    const int per = (rb + 255) / 256;
    const int lo = min(t * per, rb);
    const int hi = min(lo + per, rb);
    int sum = 0;
    for (int r = lo; r < hi; r++) {
        sum += cnt[r];
    }
    part[t] = sum;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int i = 0; i < 256; i++) {
            const int v = part[i];
            part[i] = acc;
            acc += v;
        }
    }
    __syncthreads();
    int run = part[t];
    for (int r = lo; r < hi; r++) {
        const int v = cnt[r];
        cnt[r] = run;
        run += v;
    }
    __syncthreads();
    for (int q = t; q < n; q += 256) {
        const int k = local_tile(q);
        order[atomicAdd(&cnt[rows[k]], 1)] = k;
    }
}

// One workgroup: valid tiles of each system (counts limited to the capacity),
// exclusive prefix over systems
// (tile_prefix[S] = total), clear the rebuild flags.
__global__ __launch_bounds__(256) void k_tile_prefix(
    const NbSys *__restrict__ sys, int S, int *__restrict__ tile_count, int *__restrict__ tile_prefix,
    int *__restrict__ rebuild) {
    __shared__ int part[256];
    const int t = threadIdx.x;
    const int per = (S + 255) / 256;
    int sum = 0;
    auto tiles_of = [&](int s) {
        const int cap = sys[s].tile_cap;
        int c = 0;
        for (int j = 0; j < NB_TILE_LANES; j++) {
            c += nb_lane_tiles(tile_count[s * NB_TILE_LANES + j], j, cap);
        }
        return c;
    };
    for (int k = 0; k < per; k++) {
        const int s = t * per + k;
        if (s < S) {
            sum += tiles_of(s);
        }
    }
    part[t] = sum;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int i = 0; i < 256; i++) {
            int v = part[i];
            part[i] = acc;
            acc += v;
        }
        tile_prefix[S] = acc;
    }
    __syncthreads();
    int run = part[t];
    for (int k = 0; k < per; k++) {
        const int s = t * per + k;
        if (s < S) {
            tile_prefix[s] = run;
            run += tiles_of(s);
            rebuild[s] = 0;
        }
    }
}

} // namespace

void launch_fill_u64(uint64_t *p, size_t n, uint64_t v, hipStream_t stream) {
    if (n == 0) {
        return;
    }
    k_fill_u64<<<ceil_div(static_cast<long long>(n), 256), 256, 0, stream>>>(p, n, v);
    TM_LAUNCH_CHECK();
}

void launch_fill_i32(int *p, size_t n, int v, hipStream_t stream) {
    if (n == 0) {
        return;
    }
    k_fill_i32<<<ceil_div(static_cast<long long>(n), 256), 256, 0, stream>>>(p, n, v);
    TM_LAUNCH_CHECK();
}

// Block, slot and tile layout of one system placed at the given offsets.
static void lay_out(NbSys &y, int slot0, int block0, int tile0, double nb_margin) {
    if (y.nrow > 0) {
        if (y.nrow >= y.k || y.subset0 < 0) {
            throw Error("NbState: a row subset needs column atoms and a subset list");
        }
        y.row_blocks = ceil_div(y.nrow, TILE);
        y.nblocks = y.row_blocks + ceil_div(y.k - y.nrow, TILE);
        y.row_end = y.nrow;
        y.slot_end = y.row_blocks * TILE + (y.k - y.nrow);
    } else {
        y.nrow = 0;
        y.nblocks = ceil_div(y.k, TILE);
        y.row_blocks = y.nblocks;
        y.row_end = y.k;
        y.slot_end = y.k;
    }
    y.slot0 = slot0;
    y.block0 = block0;
    y.tile0 = tile0;
    y.tile_cap = nb_tile_capacity(y.k, y.nrow);
    y.nb_cutoff = static_cast<float>(y.cutoff_d + y.padding_d + nb_margin);
    y.cutoff = static_cast<float>(y.cutoff_d);
}

// Writes the descriptor of the one system of a view state and clears its
// counters (the memory holds whatever the last user left); the first
// prepare() rebuilds.
__global__ void k_init_view(const NbSys y, NbSys *sys, int *rebuild, int *status, int *tile_count,
                            int *tile_prefix) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sys[0] = y;
        rebuild[0] = 1;
        status[0] = 0;
        for (int j = 0; j < NB_TILE_LANES; j++) {
            tile_count[j] = 0;
        }
        tile_prefix[0] = 0;
        tile_prefix[1] = 0;
    }
}

namespace {
// Carves consecutive 256-byte aligned arrays out of an arena; with a null
// arena it only counts the bytes.
struct Carver {
    unsigned char *base;
    size_t used = 0;
    template <typename T> T *take(size_t count) {
        used = (used + 255) & ~static_cast<size_t>(255);
        T *p = base ? reinterpret_cast<T *>(base + used) : nullptr;
        used += count * sizeof(T);
        return p;
    }
};
} // namespace

size_t NbState::view_bytes(const NbSys &sys_in) {
    NbSys y = sys_in;
    lay_out(y, 0, 0, 0, 0.0);
    Carver c{nullptr};
    const size_t slots = static_cast<size_t>(y.nblocks) * TILE;
    c.take<NbSys>(1);
    c.take<int>(1);  // rebuild
    c.take<int>(1);  // status
    c.take<int>(NB_TILE_LANES); // tile_count
    c.take<int>(2);  // tile_prefix
    c.take<int>(y.tile_cap);
    c.take<int>(y.tile_cap);
    c.take<uint32_t>(static_cast<size_t>(y.tile_cap) * TILE);
    c.take<uint32_t>(slots); // perm
    c.take<uint32_t>(slots); // keys
    c.take<double4>(slots);
    c.take<float4>(slots);
    c.take<float4>(y.nblocks);
    c.take<float4>(y.nblocks);
    c.take<double>(static_cast<size_t>(y.n_atoms) * 3);
    c.take<double>(9);
    return c.used;
}

void NbState::init_view(const NbSys &sys_in, void *arena, size_t arena_bytes, hipStream_t stream) {
    if (sys_in.subset0 >= 0 || sys_in.nrow > 0 || sys_in.interp != 0 || sys_in.atom0 != 0) {
        throw Error("NbState::init_view: one system of all atoms, no subset, no interpolation");
    }
    const size_t need = view_bytes(sys_in);
    if (need > arena_bytes) {
        throw Error("NbState::init_view: " + std::to_string(need) + " bytes needed, " + std::to_string(arena_bytes) +
                    " available");
    }
    NbSys y = sys_in;
    lay_out(y, 0, 0, 0, nb_margin);
    S = 1;
    total_slots = max_slots = y.nblocks * TILE;
    total_blocks = max_blocks = y.nblocks;
    total_tiles = y.tile_cap;
    max_atoms = y.n_atoms;
    any_interp = false;
    Carver c{static_cast<unsigned char *>(arena)};
    const size_t slots = static_cast<size_t>(total_slots);
    sys.view(c.take<NbSys>(1), 1);
    rebuild.view(c.take<int>(1), 1);
    status.view(c.take<int>(1), 1);
    tile_count.view(c.take<int>(NB_TILE_LANES), NB_TILE_LANES);
    tile_prefix.view(c.take<int>(2), 2);
    tile_row.view(c.take<int>(total_tiles), total_tiles);
    tile_order.view(c.take<int>(total_tiles), total_tiles);
    tile_atoms.view(c.take<uint32_t>(static_cast<size_t>(total_tiles) * TILE), static_cast<size_t>(total_tiles) * TILE);
    perm.view(c.take<uint32_t>(slots), slots);
    keys.view(c.take<uint32_t>(slots), slots);
    pos.view(c.take<double4>(slots), slots);
    par.view(c.take<float4>(slots), slots);
    bb_ctr.view(c.take<float4>(total_blocks), total_blocks);
    bb_ext.view(c.take<float4>(total_blocks), total_blocks);
    nblist_x.view(c.take<double>(static_cast<size_t>(y.n_atoms) * 3), static_cast<size_t>(y.n_atoms) * 3);
    nblist_box.view(c.take<double>(9), 9);
    subset.view(nullptr, 0);
    dpar.view(nullptr, 0);
    k_init_view<<<1, 64, 0, stream>>>(y, sys.ptr, rebuild.ptr, status.ptr, tile_count.ptr, tile_prefix.ptr);
    TM_LAUNCH_CHECK();
}

void NbState::init(const std::vector<NbSys> &sys_in, const std::vector<int> &subset_atoms) {
    S = static_cast<int>(sys_in.size());
    h_sys = sys_in;
    total_slots = total_blocks = total_tiles = 0;
    max_blocks = max_slots = max_atoms = 0;
    int total_atoms_end = 0;
    for (int s = 0; s < S; s++) {
        NbSys &y = h_sys[s];
        lay_out(y, total_slots, total_blocks, total_tiles, nb_margin);
        total_slots += y.nblocks * TILE;
        total_blocks += y.nblocks;
        total_tiles += y.tile_cap;
        max_blocks = std::max(max_blocks, y.nblocks);
        max_slots = std::max(max_slots, y.nblocks * TILE);
        max_atoms = std::max(max_atoms, y.n_atoms);
        total_atoms_end = std::max(total_atoms_end, y.atom0 + y.n_atoms);
    }
    sys.alloc(S);
    sys.upload(h_sys.data(), S);
    subset.alloc(std::max<size_t>(subset_atoms.size(), 1));
    if (!subset_atoms.empty()) {
        subset.upload(subset_atoms.data(), subset_atoms.size());
    }
    rebuild.alloc(S);
    status.alloc(S);
    tile_count.alloc(static_cast<size_t>(S) * NB_TILE_LANES);
    tile_prefix.alloc(S + 1);
    tile_row.alloc(total_tiles);
    tile_order.alloc(total_tiles);
    tile_atoms.alloc(static_cast<size_t>(total_tiles) * TILE);
    perm.alloc(total_slots);
    keys.alloc(total_slots);
    pos.alloc(total_slots);
    par.alloc(total_slots);
    any_interp = false;
    for (const auto &y : h_sys) {
        any_interp = any_interp || y.interp != 0;
    }
    if (any_interp) {
        dpar.alloc(total_slots);
    } else {
        dpar.release();
    }
    bb_ctr.alloc(total_blocks);
    bb_ext.alloc(total_blocks);
    piece_buf.alloc(static_cast<size_t>(total_blocks) * NB_MAX_PIECES * TILE);
    piece_cnt.alloc(static_cast<size_t>(total_blocks) * NB_MAX_PIECES);
    nblist_x.alloc(static_cast<size_t>(total_atoms_end) * 3);
    nblist_box.alloc(static_cast<size_t>(S) * 9);
    status.zero();
    tile_count.zero();
    tile_prefix.zero();
    nblist_x.zero();
    nblist_box.zero();
    launch_fill_i32(rebuild.ptr, S, 1, 0);
    device_synchronize();
}

void NbState::invalidate(hipStream_t stream) { launch_fill_i32(rebuild.ptr, S, 1, stream); }

// Waves per row block in the neighborlist build. One wave scans a row
// block's column blocks one candidate after another, so a build takes as
// long as its longest row block (the first ones, upper triangle) however few
// systems rebuild, while most of the GPU idles. Each row block's column range
// is split into up to 8 contiguous pieces, one wave each; the pieces'
// leftovers are merged (k_merge_pieces), so the tile list is as short as with
// one wave. Only batches far beyond the GPU's wave slots (where every wave
// would queue anyway) use fewer pieces. Views (no staging arrays) use one.
int NbState::find_pieces() const {
    if (piece_cnt.n == 0) {
        return 1;
    }
    if (pieces_override > 0) {
        return pieces_override;
    }
    const long long waves = static_cast<long long>(S) * max_blocks;
    int q = NB_MAX_PIECES;
    while (q > 1 && waves * q > 262144) {
        q /= 2;
    }
    return q;
}

void NbState::set_padding(double padding, hipStream_t stream) {
    if (!(padding >= 0.0)) {
        throw Error("neighborlist padding must be >= 0");
    }
    for (auto &y : h_sys) {
        y.padding_d = padding;
        y.nb_cutoff = static_cast<float>(y.cutoff_d + y.padding_d + nb_margin);
    }
    TM_HIP_CHECK(hipStreamSynchronize(stream));
    sys.upload(h_sys.data(), S, stream);
    invalidate(stream);
    TM_HIP_CHECK(hipStreamSynchronize(stream));
}

void NbState::find_tiles(const double *boxes, hipStream_t stream) {
    const int pieces = find_pieces();
    k_find_tiles<<<dim3(S, ceil_div(max_blocks, 4) * pieces), 256, 0, stream>>>(
        sys.ptr, boxes, rebuild.ptr, pos.ptr, bb_ctr.ptr, bb_ext.ptr, tile_count.ptr, tile_row.ptr, tile_atoms.ptr,
        status.ptr, pieces, piece_buf.ptr, piece_cnt.ptr);
    TM_LAUNCH_CHECK();
    if (pieces > 1) {
        k_merge_pieces<<<dim3(S, ceil_div(max_blocks, 4)), 256, 0, stream>>>(
            sys.ptr, rebuild.ptr, pieces, piece_buf.ptr, piece_cnt.ptr, tile_count.ptr, tile_row.ptr, tile_atoms.ptr,
            status.ptr);
        TM_LAUNCH_CHECK();
    }
}

void NbState::build_only(const double *x, const double *boxes, hipStream_t stream) {
    const int gx_atoms = ceil_div(std::max(max_atoms, 9), 256);
    k_check_rebuild<<<dim3(gx_atoms, S), 256, 0, stream>>>(sys.ptr, subset.ptr, x, nblist_x.ptr, boxes,
                                                          nblist_box.ptr, rebuild.ptr, status.ptr);
    TM_LAUNCH_CHECK();
    k_sort<<<S, SORT_THREADS, 0, stream>>>(sys.ptr, subset.ptr, x, boxes, rebuild.ptr, keys.ptr, perm.ptr, nblist_x.ptr,
                                          nblist_box.ptr, tile_count.ptr);
    TM_LAUNCH_CHECK();
    k_gather<<<dim3(ceil_div(max_slots, 256), S), 256, 0, stream>>>(sys.ptr, x, nullptr, nullptr, nullptr, nullptr,
                                                                    nullptr, 0.0, perm.ptr, pos.ptr, par.ptr, nullptr);
    TM_LAUNCH_CHECK();
    k_block_bounds<<<dim3(ceil_div(max_blocks, 4), S), 256, 0, stream>>>(sys.ptr, boxes, rebuild.ptr, pos.ptr,
                                                                         bb_ctr.ptr, bb_ext.ptr);
    TM_LAUNCH_CHECK();
    find_tiles(boxes, stream);
    k_tile_order<<<S, 256, 0, stream>>>(sys.ptr, rebuild.ptr, tile_count.ptr, tile_row.ptr, tile_order.ptr);
    TM_LAUNCH_CHECK();
    k_tile_prefix<<<1, 256, 0, stream>>>(sys.ptr, S, tile_count.ptr, tile_prefix.ptr, rebuild.ptr);
    TM_LAUNCH_CHECK();
}

void launch_split_interpolated_du_dp(int N, double lambda, const uint64_t *src, uint64_t *dst, hipStream_t stream) {
    if (N <= 0) {
        return;
    }
    k_split_interpolated_du_dp<<<dim3(ceil_div(N, 256), 3), 256, 0, stream>>>(N, lambda, src, dst);
    TM_LAUNCH_CHECK();
}

void NbState::prepare(
    const double *x,
    const double *params,
    const double *params1,
    const double *boxes,
    const int *plane,
    const int *offset,
    const double *lambdas,
    double lambda0,
    hipStream_t stream) {
    const int gx_atoms = ceil_div(std::max(max_atoms, 9), 256);
    k_check_rebuild<<<dim3(gx_atoms, S), 256, 0, stream>>>(sys.ptr, subset.ptr, x, nblist_x.ptr, boxes,
                                                          nblist_box.ptr, rebuild.ptr, status.ptr);
    TM_LAUNCH_CHECK();
    k_sort<<<S, SORT_THREADS, 0, stream>>>(sys.ptr, subset.ptr, x, boxes, rebuild.ptr, keys.ptr, perm.ptr, nblist_x.ptr,
                                          nblist_box.ptr, tile_count.ptr);
    TM_LAUNCH_CHECK();
    k_gather<<<dim3(ceil_div(max_slots, 256), S), 256, 0, stream>>>(
        sys.ptr, x, params, params1, plane, offset, lambdas, lambda0, perm.ptr, pos.ptr, par.ptr,
        any_interp ? dpar.ptr : nullptr);
    TM_LAUNCH_CHECK();
    k_block_bounds<<<dim3(ceil_div(max_blocks, 4), S), 256, 0, stream>>>(sys.ptr, boxes, rebuild.ptr, pos.ptr,
                                                                         bb_ctr.ptr, bb_ext.ptr);
    TM_LAUNCH_CHECK();
    find_tiles(boxes, stream);
    k_tile_order<<<S, 256, 0, stream>>>(sys.ptr, rebuild.ptr, tile_count.ptr, tile_row.ptr, tile_order.ptr);
    TM_LAUNCH_CHECK();
    k_tile_prefix<<<1, 256, 0, stream>>>(sys.ptr, S, tile_count.ptr, tile_prefix.ptr, rebuild.ptr);
    TM_LAUNCH_CHECK();
}

} // namespace tmamd
